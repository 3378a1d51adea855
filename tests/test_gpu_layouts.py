"""GPU tests of the derived layouts (assembly plan from connec, element-major
He and gpvol) for arrays reached by raw device pointer, through the C ABI.

The reference re-reads every array on every call (elem_convec.f90:228).
Device memory outside the data region has no version, so its derived layouts
must be rebuilt on each call: a caller that rewrites He or connec in place
with its own kernel or a device-to-device copy, and tells the library
nothing, still gets the residuals of the new arrays.  A device array entered
with msod_data_register is cached, and msod_data_invalidate after a rewrite
makes the next call rebuild.  Bar: relative L2 <= 1e-12 per residual against
the oracle on the arrays the call saw.
"""
import ctypes
import sys

import numpy as np
import pytest

import minisod2d_amd as m
from oracle_lib import Oracle, rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class DeviceArrays:
    """hipMalloc'd copies of host arrays, freed on close (memory the library
    does not know about)."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.ptrs = []

    def put(self, a):
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(a.nbytes)) == 0
        self.ptrs.append(p)
        self.write(p.value, a)
        return p.value

    def write(self, dst, a):
        assert self.hip.hipMemcpy(ctypes.c_void_p(dst), ctypes.c_void_p(a.ctypes.data),
                                  ctypes.c_size_t(a.nbytes), 1) == 0  # host to device

    def copy(self, dst, src, nbytes):
        assert self.hip.hipMemcpy(ctypes.c_void_p(dst), ctypes.c_void_p(src),
                                  ctypes.c_size_t(nbytes), 3) == 0  # device to device
        assert self.hip.hipDeviceSynchronize() == 0

    def get(self, src, like):
        out = np.empty_like(like)
        assert self.hip.hipMemcpy(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(src),
                                  ctypes.c_size_t(out.nbytes), 2) == 0  # device to host
        return out

    def close(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def residuals(prob, dev, d):
    """Both residuals through the device addresses in d (the tables stay host
    arrays); returns the six outputs."""
    t = prob.tables
    R, D = prob.outputs(), prob.outputs()
    o = {k: dev.put(a) for k, a in zip(("Rm", "Rmo", "Re", "Dm", "Dmo", "De"), R + D)}
    m.full_convec_ijk(prob.nelem, prob.npoin, d["connec"], t.Ngp, t.dNgp, d["He"], d["gpvol"],
                      t.dlxigp_ip, t.xgp, t.invAtoIJK, t.AtoI, t.AtoJ, t.AtoK, d["u"], d["q"],
                      d["rho"], d["pr"], d["E"], o["Rm"], o["Rmo"], o["Re"])
    m.full_diffusion_ijk(prob.nelem, prob.npoin, d["connec"], t.Ngp, t.dNgp, d["He"], d["gpvol"],
                         t.dlxigp_ip, t.xgp, t.invAtoIJK, t.AtoI, t.AtoJ, t.AtoK, prob.Cp, prob.Pr,
                         d["rho"], d["u"], d["Tem"], d["mu_fluid"], d["mu_e"], d["mu_sgs"],
                         o["Dm"], o["Dmo"], o["De"])
    return [dev.get(o[k], a) for k, a in zip(("Rm", "Rmo", "Re", "Dm", "Dmo", "De"), R + D)]


def check(prob, got, what):
    o = Oracle()
    want = o.convec(prob) + o.diffu(prob)
    errs = [rel_l2(g, w) for g, w in zip(got, want)]
    say(f"    {what}: rel L2 " + " ".join(f"{e:.1e}" for e in errs))
    assert max(errs) <= TOL, errs


def jittered(n, p, ny, nz, seed):
    return m.box_problem(n, p, ny=ny, nz=nz, jitter=0.05,
                         seeds=(m.SEED_RHO, m.SEED_MUE, m.SEED_MUSGS, seed))


def upload(prob, dev):
    names = ("connec", "He", "gpvol", "u", "q", "rho", "pr", "E", "Tem", "mu_fluid", "mu_e", "mu_sgs")
    return {k: dev.put(getattr(prob, k)) for k in names}


def test_raw_device_arrays_rewritten_in_place(gpu):
    """He rewritten by a device-to-device copy (another jitter of the same
    mesh), then connec (elements renumbered), with no call to the library in
    between: each call sees the arrays as they are."""
    say("[gpu] raw device pointers: He and connec rewritten in place, p=3 5x4x3")
    a = jittered(5, 3, 4, 3, 11)
    b = jittered(5, 3, 4, 3, 12)
    assert not np.array_equal(a.He, b.He)
    dev = DeviceArrays()
    try:
        d = upload(a, dev)
        check(a, residuals(a, dev, d), "first arrays")
        hb = dev.put(b.He)
        dev.copy(d["He"], hb, a.He.nbytes)                 # the library is not told
        a2 = jittered(5, 3, 4, 3, 11)
        a2.He = b.He
        check(a2, residuals(a2, dev, d), "He rewritten")
        perm = np.random.default_rng(5).permutation(a.nelem)
        a2.connec = np.asfortranarray(a.connec[perm])
        a2.He = np.asfortranarray(b.He[:, :, :, perm])
        a2.gpvol = np.asfortranarray(a.gpvol[:, :, perm])
        a2.mu_e = np.asfortranarray(a.mu_e[perm])
        a2.mu_sgs = np.asfortranarray(a.mu_sgs[perm])
        for k in ("connec", "He", "gpvol", "mu_e", "mu_sgs"):
            dev.write(d[k], getattr(a2, k))
        check(a2, residuals(a2, dev, d), "elements renumbered")
    finally:
        dev.close()


def test_registered_device_arrays_cache_until_invalidated(gpu):
    """msod_data_register caches the layouts of a device array; after an
    in-place rewrite, msod_data_invalidate makes the next call rebuild them."""
    say("[gpu] registered device arrays: He rewritten, then invalidated")
    L = m.hip_lib()
    a = jittered(4, 4, 3, 2, 21)
    b = jittered(4, 4, 3, 2, 22)
    dev = DeviceArrays()
    try:
        d = upload(a, dev)
        for k in ("connec", "He", "gpvol"):
            assert L.msod_data_register(d[k], getattr(a, k).nbytes) == 0, L.msod_last_error()
        assert L.msod_data_present(d["He"]) == 1
        check(a, residuals(a, dev, d), "registered arrays")
        hb = dev.put(b.He)
        dev.copy(d["He"], hb, a.He.nbytes)
        assert L.msod_data_invalidate(d["He"]) == 0
        a.He = b.He
        check(a, residuals(a, dev, d), "He rewritten and invalidated")
        # registered arrays have no host copy to update
        assert L.msod_data_update_host(d["He"], 8) == m.MSOD_ERR_ARG
        for k in ("connec", "He", "gpvol"):
            assert L.msod_data_delete(d[k]) == 0  # unregisters, does not free
        assert L.msod_data_present(d["He"]) == 0
        check(a, residuals(a, dev, d), "after unregistering")
    finally:
        dev.close()


def test_shuffled_mesh_past_32_colours_runs_atomic(gpu):
    """Elements in random order: more than 32 batch colours.  The atomic
    assembly (default) runs; the coloured one is refused with a message."""
    say("[gpu] shuffled 16^3 p=3 mesh, atomic assembly")
    prob = m.box_problem(16, 3, jitter=0.05)
    perm = np.random.default_rng(3).permutation(prob.nelem)
    for k in ("connec", "mu_e", "mu_sgs"):
        setattr(prob, k, np.asfortranarray(getattr(prob, k)[perm]))
    prob.He = np.asfortranarray(prob.He[:, :, :, perm])
    prob.gpvol = np.asfortranarray(prob.gpvol[:, :, perm])
    R, D = prob.outputs(), prob.outputs()
    ins = prob.device_inputs()
    m.copyin(*ins)
    m.create(*R, *D)
    try:
        m.full_convec_ijk(*prob.convec_args(*R))
        m.full_diffusion_ijk(*prob.diffu_args(*D))
        m.update_host(*R, *D)
        m.set_assembly(m.ASSEMBLY_COLORED)
        try:
            with pytest.raises(m.MsodError, match="more than 32 colours"):
                m.full_convec_ijk(*prob.convec_args(*R))
        finally:
            m.set_assembly(m.ASSEMBLY_ATOMIC)
    finally:
        m.delete(*ins, *R, *D)
    check(prob, list(R + D), "shuffled, atomic")

"""GPU tests of the drop-in for callers built at the reference's default real
kind rp=4 (msod_full_convec_ijk_r4 / msod_full_diffusion_ijk_r4, and the
generic Fortran procedures full_convec_ijk / full_diffusion_ijk with real(4)
arrays), against the fixtures of the reference compiled at rp=4
(tests/golden/r4_*.npz).

Bar, per residual array, stated from the rp=4 reference's own distance d_ref
to the FP64 oracle on the same float32 inputs (2e-7..7e-7, pinned on the CPU
by tests/test_r4_golden.py):
  * the drop-in is no farther from the FP64 oracle than the reference
    (its kernels compute in FP64 and round each node's value once);
  * the drop-in and the reference differ by at most 2 d_ref.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import minisod2d_amd as m
from r4_common import OUTS, R4_CASES, distances, load_case, oracle64

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def outputs32(npoin):
    z = lambda shape: np.zeros(shape, np.float32, order="F")
    return [z(npoin), z((npoin, 3)), z(npoin), z(npoin), z((npoin, 3)), z(npoin)]


def check(name, got, ref, want):
    for k, (dg, dr, dgr) in zip(OUTS, distances(got, ref, want)):
        say(f"    {k}: drop-in vs FP64 oracle {dg:.2e}, rp=4 reference vs oracle {dr:.2e}, "
            f"drop-in vs reference {dgr:.2e}")
        assert dg <= dr, (name, k, dg, dr)
        assert dgr <= 2 * dr, (name, k, dgr, dr)


@pytest.mark.parametrize("name", list(R4_CASES))
def test_r4_entry_points_against_rp4_reference(gpu, name):
    say(f"[gpu] rp=4 drop-in, C ABI: {name}")
    c, ref, prob = load_case(name)
    out = outputs32(c["npoin"])
    ins = [c[k] for k in ("connec", "He", "gpvol", "u", "q", "rho", "pr", "E", "Tem", "mu_fluid", "mu_e",
                          "mu_sgs")]
    m.copyin(*ins)
    m.create(*out)
    try:
        m.full_convec_ijk_r4(c["nelem"], c["npoin"], c["connec"], c["Ngp"], c["dNgp"], c["He"], c["gpvol"],
                             c["dlxigp_ip"], c["xgp"], c["invAtoIJK"], c["AtoI"], c["AtoJ"], c["AtoK"],
                             c["u"], c["q"], c["rho"], c["pr"], c["E"], *out[:3])
        m.full_diffusion_ijk_r4(c["nelem"], c["npoin"], c["connec"], c["Ngp"], c["dNgp"], c["He"],
                                c["gpvol"], c["dlxigp_ip"], c["xgp"], c["invAtoIJK"], c["AtoI"], c["AtoJ"],
                                c["AtoK"], c["Cp"], c["Pr"], c["rho"], c["u"], c["Tem"], c["mu_fluid"],
                                c["mu_e"], c["mu_sgs"], *out[3:])
        m.update_host(*out)
    finally:
        m.delete(*ins, *out)
    assert all(o.dtype == np.float32 for o in out)
    check(name, out, ref, oracle64(c, prob))


def write_caller_input(path, c):
    with open(path, "wb") as fh:
        np.array([c["porder"], c["nelem"], c["npoin"]], np.int32).tofile(fh)
        np.array([c["Cp"], c["Pr"]], np.float32).tofile(fh)
        for k in ("connec", "invAtoIJK", "AtoI", "AtoJ", "AtoK"):
            np.asfortranarray(c[k], np.int32).ravel(order="F").tofile(fh)
        for k in ("dlxigp_ip", "xgp", "Ngp", "dNgp", "He", "gpvol", "u", "q", "rho", "pr", "E", "Tem",
                  "mu_fluid", "mu_e", "mu_sgs"):
            np.asfortranarray(c[k], np.float32).ravel(order="F").tofile(fh)


@pytest.mark.parametrize("name", ["r4_box8_p3", "r4_box2_p7_jit_jump"])
def test_fortran_caller_built_with_real4(gpu, tmp_path, name):
    """A Fortran program with real(4) arrays calls full_convec_ijk and
    full_diffusion_ijk through the generic drop-in modules, as a caller of the
    reference built at rp=4 does (minisod2d_amd/fortran/msod_r4_caller.f90)."""
    say(f"[gpu] rp=4 drop-in, Fortran caller: {name}")
    c, ref, prob = load_case(name)
    exe = os.path.join(m.LIB_DIR, "msod_r4_caller")
    assert os.path.exists(exe), "msod_r4_caller not built (make -C minisod2d_amd)"
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    write_caller_input(fin, c)
    r = subprocess.run([exe, fin, fout], capture_output=True, text=True, timeout=120)
    say("    " + r.stdout.strip())
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.fromfile(fout, np.float32)
    n = c["npoin"]
    sizes = (n, 3 * n, n, n, 3 * n, n)
    assert raw.size == sum(sizes)
    got, off = [], 0
    for s, like in zip(sizes, ref):
        got.append(raw[off:off + s].reshape(like.shape, order="F"))
        off += s
    check(name, got, ref, oracle64(c, prob))

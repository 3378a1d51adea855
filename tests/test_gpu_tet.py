"""GPU parity of the linear-tetrahedron kernels (msod_fem_convec,
msod_fem_diffu, msod_tet_jacobian) through the C ABI:

  * the reference's own outputs (golden fixtures of fem_convec, fem_diffu,
    tet_jacobian), relative L2 <= 1e-12;
  * the oracle on jittered meshes with ragged sizes (element counts around the
    64-element workgroup, npoin not a multiple of 4: unreferenced nodes stay 0);
  * the full benchmark size (the 128^3 porder-4 box's 135 M nodal values under
    33.75 M tetrahedra): sampled elements against the oracle, and a uniform
    state whose residuals vanish;
  * error paths: connec outside 1..npoin, bad sizes.
"""
import json
import os
import sys

import numpy as np
import pytest

import minisod2d_amd as m
from oracle_lib import Oracle, rel_l2

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TET_CASES = sorted(f[:-4] for f in os.listdir(GOLD) if f.startswith("tet_case_"))
KEYS = ("Rmass", "Rmom", "Rener", "Dmass", "Dmom", "Dener")
TOL = 1e-12
THREADS = min(16, os.cpu_count() or 1)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu_tet(tp, c_rho=1.0, c_ener=1.0):
    """Both TET kernels on the GPU; returns (Rmass, Rmom, Rener, Dmass, Dmom, Dener)."""
    R, D = tp.outputs(), tp.outputs()
    ins = tp.device_inputs()
    m.copyin(*ins)
    m.create(*R, *D)
    try:
        m.set_diffusion_coefficients(c_rho, c_ener)
        m.fem_convec(*tp.convec_args(*R))
        m.fem_diffu(*tp.diffu_args(*D))
        m.update_host(*R, *D)
    finally:
        m.set_diffusion_coefficients(1.0, 1.0)
        m.delete(*ins, *R, *D)
    return R + D


def oracle_tet(tp, c_rho=1.0, c_ener=1.0):
    o = Oracle()
    return o.fem_convec(tp) + o.fem_diffu(tp, c_rho, c_ener)


@pytest.mark.parametrize("name", TET_CASES)
def test_matches_reference_golden(gpu, name):
    from golden.make_golden import TET_CASES as B, tet_input_digests
    say(f"[gpu] TET golden case {name}")
    z = np.load(os.path.join(GOLD, name + ".npz"))
    meta = json.loads(str(z["meta"]))
    build, c_rho, c_ener = B[name]
    tp = build()
    assert tet_input_digests(tp) == meta["inputs"]
    for key, got in zip(KEYS, gpu_tet(tp, c_rho, c_ener)):
        err = rel_l2(got, z[key])
        assert err <= TOL, (key, err)


@pytest.mark.parametrize("nelem", [1, 63, 64, 65, 1027])
def test_matches_oracle_ragged(gpu, nelem):
    say(f"[gpu] TET ragged nelem={nelem}")
    # a box whose node count gives this many tetrahedra, with 1..3 nodes left over
    base = m.box_problem(4, 4, jitter=0.05, rho_jump=0.1, seeds=(41, 42, 43, 44))  # 4913 nodes
    extra = 1 + nelem % 3
    keep = 4 * nelem + extra
    assert keep <= base.npoin
    for k in ("u", "q"):
        setattr(base, k, np.asfortranarray(getattr(base, k)[:keep]))
    for k in ("rho", "pr", "E", "Tem", "mu_fluid"):
        setattr(base, k, np.asfortranarray(getattr(base, k)[:keep]))
    base.npoin = keep
    tp = m.tet_problem(base, jitter=0.3, random_mu=True, seed=nelem, strict=False)
    assert tp.nelem == nelem
    out = gpu_tet(tp, 0.8, 1.6)
    want = oracle_tet(tp, 0.8, 1.6)
    for key, got, w in zip(KEYS, out, want):
        assert rel_l2(got, w) <= TOL, (key, rel_l2(got, w))
        assert np.all(np.asarray(got)[4 * nelem:] == 0.0), key


def test_shared_node_mesh(gpu):
    """Tetrahedra sharing nodes (random connectivity): the element residuals
    meet at the nodes through atomics, and nodes no element uses stay 0."""
    say("[gpu] TET mesh with shared nodes")
    rng = np.random.default_rng(17)
    nelem, npoin = 3000, 2000
    connec = np.asfortranarray(np.array([rng.choice(npoin - 50, 4, replace=False) + 1
                                         for _ in range(nelem)], dtype=np.int32))
    coord = np.asfortranarray(rng.random((npoin, 3)))
    t = m.tet_tables()
    He = np.zeros((3, 3, 4, nelem), order="F")
    gp = np.zeros((1, 4, nelem), order="F")
    m.host_lib().msodh_elem_jacobian(nelem, npoin, 4, connec.ctypes.data, coord.ctypes.data,
                                     t.dNgp.ctypes.data, t.wgp.ctypes.data, gp.ctypes.data,
                                     He.ctypes.data)
    f = lambda *sh: np.asfortranarray(rng.random(sh))
    tp = m.TetProblem(t, nelem, npoin, connec, coord, He, gp, f(npoin, 3), f(npoin, 3),
                      1 + f(npoin), 1 + f(npoin), 2 + f(npoin), 1 + f(npoin), f(npoin),
                      f(nelem, 4), f(nelem, 4))
    out = gpu_tet(tp, 1.1, 0.9)
    want = oracle_tet(tp, 1.1, 0.9)
    for key, got, w in zip(KEYS, out, want):
        assert rel_l2(got, w) <= TOL, (key, rel_l2(got, w))
        assert np.all(np.asarray(got)[npoin - 50:] == 0.0), key


def _tet_on(rng, connec, npoin, coord, fields):
    """TetProblem of a given connectivity: He, gpvol from the host routine."""
    nelem = connec.shape[0]
    t = m.tet_tables()
    He = np.zeros((3, 3, 4, nelem), order="F")
    gp = np.zeros((1, 4, nelem), order="F")
    m.host_lib().msodh_elem_jacobian(nelem, npoin, 4, connec.ctypes.data, coord.ctypes.data,
                                     t.dNgp.ctypes.data, t.wgp.ctypes.data, gp.ctypes.data,
                                     He.ctypes.data)
    return m.TetProblem(t, nelem, npoin, connec, coord, He, gp, *fields)


def test_connectivity_rewritten_in_place(gpu):
    """The mesh analysis is cached per connectivity (plain stores when no
    node is shared).  A disjoint connectivity rewritten in place into one
    with shared nodes, then msod_data_update_device: calls through the host
    address and through the device address (msod_data_deviceptr) both see
    the new mesh and match the oracle."""
    say("[gpu] TET connectivity rewritten in place")
    rng = np.random.default_rng(23)
    nelem, npoin = 500, 2000
    coord = np.asfortranarray(rng.random((npoin, 3)))
    f = lambda *sh: np.asfortranarray(rng.random(sh))
    fields = (f(npoin, 3), f(npoin, 3), 1 + f(npoin), 1 + f(npoin), 2 + f(npoin), 1 + f(npoin),
              f(npoin), f(nelem, 4), f(nelem, 4))
    disjoint = np.asfortranarray(np.arange(1, 4 * nelem + 1, dtype=np.int32).reshape(nelem, 4))
    shared = np.asfortranarray(np.array([rng.choice(npoin - 50, 4, replace=False) + 1
                                         for _ in range(nelem)], dtype=np.int32))
    A = _tet_on(rng, disjoint, npoin, coord, fields)
    B = _tet_on(rng, shared, npoin, coord, fields)
    # one set of host arrays, made present once; A's contents first
    tp = _tet_on(rng, disjoint.copy(order="F"), npoin, coord, fields)
    R, D = tp.outputs(), tp.outputs()
    ins = tp.device_inputs()
    m.copyin(*ins)
    m.create(*R, *D)
    try:
        dc = m.hip_lib().msod_data_deviceptr(tp.connec.ctypes.data)
        assert dc
        for want_p, label in ((A, "disjoint"), (B, "shared")):
            if want_p is B:
                tp.connec[...] = B.connec
                tp.He[...] = B.He
                tp.gpvol[...] = B.gpvol
                m.update_device(tp.connec, tp.He, tp.gpvol)
            want = oracle_tet(want_p)
            for via, conn in (("host address", tp.connec), ("device address", dc)):
                args_c = (nelem, npoin, conn) + tp.convec_args(*R)[3:]
                args_d = (nelem, npoin, conn) + tp.diffu_args(*D)[3:]
                m.fem_convec(*args_c)
                m.fem_diffu(*args_d)
                m.update_host(*R, *D)
                for key, got, w in zip(KEYS, R + D, want):
                    assert rel_l2(got, w) <= TOL, (label, via, key, rel_l2(got, w))
        # invalidating the device address keeps the calls right
        m._check(m.hip_lib().msod_data_invalidate(dc))
        m.fem_convec(*((nelem, npoin, dc) + tp.convec_args(*R)[3:]))
        m.update_host(*R)
        for key, got, w in zip(KEYS, R, oracle_tet(B)[:3]):
            assert rel_l2(got, w) <= TOL, ("after invalidate", key, rel_l2(got, w))
    finally:
        m.delete(*ins, *R, *D)


def test_tet_jacobian_matches_golden_and_host(gpu):
    from golden.make_golden import digest
    say("[gpu] TET geometry vs the reference's tet_jacobian")
    z = np.load(os.path.join(GOLD, "tet_jac.npz"))
    tp = m.tet_problem(m.box_problem(3, 2), jitter=0.4, strict=False)
    assert digest(tp.coord) == str(z["coord_sha"])
    He = np.zeros_like(tp.He)
    gp = np.zeros_like(tp.gpvol)
    m.copyin(tp.connec, tp.coord)
    m.create(He, gp)
    try:
        nbad = m.tet_jacobian(tp.nelem, tp.npoin, tp.connec, tp.coord, tp.tables.dNgp, tp.tables.wgp,
                              gp, He)
        m.update_host(He, gp)
    finally:
        m.delete(tp.connec, tp.coord, He, gp)
    assert nbad == 0
    assert rel_l2(He, z["He"]) <= 1e-15 and rel_l2(gp, z["gpvol"]) <= 1e-15
    assert rel_l2(He, tp.He) <= 1e-15 and rel_l2(gp, tp.gpvol) <= 1e-15


def test_tet_jacobian_counts_inverted_elements(gpu):
    say("[gpu] TET geometry: mirrored elements")
    tp = m.tet_problem(m.box_problem(2, 2), strict=False)  # 125 nodes, 31 tetrahedra
    coord = tp.coord.copy(order="F")
    coord[:8, 0] *= -1.0  # mirror the first two elements: det < 0 at their 8 points
    He = np.zeros_like(tp.He)
    gp = np.zeros_like(tp.gpvol)
    m.copyin(tp.connec, coord)
    m.create(He, gp)
    try:
        nbad = m.tet_jacobian(tp.nelem, tp.npoin, tp.connec, coord, tp.tables.dNgp, tp.tables.wgp,
                              gp, He)
        m.update_host(gp)
    finally:
        m.delete(tp.connec, coord, He, gp)
    assert nbad == 8
    assert np.all(gp[0, :, :2] < 0) and np.all(gp[0, :, 2:] > 0)


def test_bad_connectivity_is_rejected(gpu):
    say("[gpu] TET error paths")
    tp = m.tet_problem(m.reference_problem(2, 3))  # 128 nodes, 32 tetrahedra
    R = tp.outputs()
    ins = tp.device_inputs()
    m.copyin(*ins)
    m.create(*R)
    try:
        tp.connec[5, 2] = tp.npoin + 1
        m.update_device(tp.connec)
        with pytest.raises(m.MsodError) as ei:
            m.fem_convec(*tp.convec_args(*R))
        assert ei.value.status == m.MSOD_ERR_MESH
        with pytest.raises(m.MsodError) as ei:
            m.fem_diffu(*tp.diffu_args(*R))
        assert ei.value.status == m.MSOD_ERR_MESH
        tp.connec[5, 2] = 0
        m.update_device(tp.connec)
        with pytest.raises(m.MsodError):
            m.fem_convec(*tp.convec_args(*R))
        # repaired: the new version of connec is checked again and accepted
        tp.connec[5, 2] = 4 * 5 + 3
        m.update_device(tp.connec)
        m.fem_convec(*tp.convec_args(*R))
        L = m.hip_lib()
        assert L.msod_fem_convec(0, tp.npoin, *[m._ptr(a) for a in (
            tp.connec, tp.tables.Ngp, tp.tables.dNgp, tp.He, tp.gpvol, tp.u, tp.q, tp.rho, tp.pr,
            tp.E, *R)]) == m.MSOD_ERR_ARG
        # Pr = 0, a missing table, a missing input
        assert L.msod_fem_diffu(tp.nelem, tp.npoin, *[m._ptr(a) for a in (
            tp.connec, tp.tables.Ngp, tp.tables.dNgp, tp.He, tp.gpvol)], 1.0, 0.0, *[m._ptr(a) for a in (
                tp.rho, tp.u, tp.Tem, tp.mu_fluid, tp.mu_e, tp.mu_sgs, *R)]) == m.MSOD_ERR_ARG
        assert L.msod_fem_convec(tp.nelem, tp.npoin, *[m._ptr(a) for a in (
            tp.connec, None, tp.tables.dNgp, tp.He, tp.gpvol, tp.u, tp.q, tp.rho, tp.pr,
            tp.E, *R)]) == m.MSOD_ERR_ARG
        assert L.msod_fem_convec(tp.nelem, tp.npoin, *[m._ptr(a) for a in (
            tp.connec, tp.tables.Ngp, tp.tables.dNgp, tp.He, tp.gpvol, tp.u, tp.q, tp.rho, None,
            tp.E, *R)]) == m.MSOD_ERR_ARG
    finally:
        m.delete(*ins, *R)


def test_full_size_sampled_parity_and_uniform_state(gpu):
    say("[gpu] TET full size: the 128^3 p=4 box's nodal fields")
    prob = m.box_problem(128, 4, threads=THREADS)
    tp = m.tet_problem(prob, jitter=0.2, random_mu=True, strict=False)
    say(f"[gpu] TET full size: {tp.nelem} tetrahedra over {tp.npoin} nodes")
    out = gpu_tet(tp)
    rng = np.random.default_rng(5)
    els = np.sort(rng.choice(tp.nelem, size=2000, replace=False))
    # the mesh is disjoint: the sampled elements' nodes receive nothing else
    f = np.asfortranarray
    nodes = (tp.connec[els] - 1).ravel()
    sub = m.TetProblem(tp.tables, len(els), 4 * len(els),
                       f(np.arange(1, 4 * len(els) + 1, dtype=np.int32).reshape(len(els), 4)),
                       f(tp.coord[nodes]), f(tp.He[:, :, :, els]), f(tp.gpvol[:, :, els]),
                       *[f(getattr(tp, k)[nodes]) for k in ("u", "q", "rho", "pr", "E", "Tem", "mu_fluid")],
                       f(tp.mu_e[els]), f(tp.mu_sgs[els]), Cp=tp.Cp, Pr=tp.Pr)
    say("[gpu] TET full size: oracle on 2000 sampled elements")
    want = oracle_tet(sub)
    for key, got, w in zip(KEYS, out, want):
        err = rel_l2(np.asarray(got)[nodes], w)
        assert err <= TOL, (key, err)
    assert all(np.all(np.asarray(r)[4 * tp.nelem:] == 0.0) for r in out)
    scale = [float(np.sqrt(np.mean(np.square(r)))) for r in out]
    del out
    say("[gpu] TET full size: uniform state")
    rho, u, p = 1.2, np.array([0.3, -0.2, 0.1]), 0.9
    tp.rho = np.full(tp.npoin, rho)
    tp.u = np.asfortranarray(np.tile(u, (tp.npoin, 1)))
    tp.q = np.asfortranarray(rho * tp.u)
    tp.pr = np.full(tp.npoin, p)
    tp.E = np.full(tp.npoin, p / 0.4 + 0.5 * rho * float(u @ u))
    tp.Tem = np.full(tp.npoin, p / (rho * 287.0))
    for key, r, s in zip(KEYS, gpu_tet(tp), scale):
        rms = float(np.sqrt(np.mean(np.square(r))))
        assert rms <= 1e-10 * s, (key, rms, s)

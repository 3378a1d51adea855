"""GPU parity: the HIP kernels (through the C ABI) against the reference's
outputs (golden fixtures) and the CPU oracle.  Bar: relative L2 per residual
component <= 1e-12 for the FP64 kernels (atomic summation order is the only
expected difference).  The FP32-storage forms compute in FP64: <= 1e-12
against the oracle on the same rounded inputs; against the oracle on the FP64
inputs, <= 1e-6 for FP32 element arrays (geo32) and the FP32 rounding of the
inputs (<= 4e-6 at these sizes) for FP32 storage of everything (mixed)."""
import os
import re
import sys

import numpy as np
import pytest

import minisod2d_amd as m
from oracle_lib import Oracle, rel_l2

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = sorted(f[:-4] for f in os.listdir(GOLD) if f.startswith("case_"))
KEYS = ("Rmass", "Rmom", "Rener", "Dmass", "Dmom", "Dener")
TOL64 = 1e-12
TOL_MIXED = 1e-6


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu_residuals(prob, mixed=False, fp32=False):
    """Both kernels on the GPU; returns (Rmass, Rmom, Rener, Dmass, Dmom, Dener)."""
    R = prob.outputs()
    D = prob.outputs()
    t = prob.tables
    if mixed:
        f = prob.fp32()
        ins = (prob.connec, f["He"], f["gpvol"], f["u"], f["q"], f["rho"], f["E"], f["Tem"],
               f["mu_fluid"], f["mu_e"], f["mu_sgs"])
    else:
        ins = prob.device_inputs()
    m.copyin(*ins)
    m.create(*R, *D)
    try:
        if mixed:
            conv = m.full_convec_ijk_fp32 if fp32 else m.full_convec_ijk_mixed
            diff = m.full_diffusion_ijk_fp32 if fp32 else m.full_diffusion_ijk_mixed
            conv(prob.nelem, prob.npoin, prob.connec, f["He"], f["gpvol"], t.dlxigp_ip, t.invAtoIJK,
                 t.AtoI, t.AtoJ, t.AtoK, f["u"], f["q"], f["rho"], f["E"], prob.gamma, *R)
            diff(prob.nelem, prob.npoin, prob.connec, f["He"], f["gpvol"], t.dlxigp_ip, t.invAtoIJK,
                 t.AtoI, t.AtoJ, t.AtoK, prob.Cp, prob.Pr, f["rho"], f["u"], f["Tem"],
                 f["mu_fluid"], f["mu_e"], f["mu_sgs"], *D)
        else:
            m.full_convec_ijk(*prob.convec_args(*R))
            m.full_diffusion_ijk(*prob.diffu_args(*D))
        m.update_host(*R, *D)
    finally:
        m.delete(*ins, *R, *D)
    return R + D


def load_case(name):
    import json
    z = np.load(os.path.join(GOLD, name + ".npz"))
    return {k: z[k] for k in z.files if k != "meta"}, json.loads(str(z["meta"]))


@pytest.mark.parametrize("name", CASES)
def test_fp64_matches_reference_golden(gpu, name):
    from golden.make_golden import CASES as BUILDERS, input_digests
    say(f"[gpu] golden case {name}")
    gold, meta = load_case(name)
    prob = BUILDERS[name]()
    assert input_digests(prob) == meta["inputs"]
    out = gpu_residuals(prob)
    for key, got in zip(KEYS, out):
        err = rel_l2(got, gold[key])
        assert err <= TOL64, (key, err)


@pytest.mark.parametrize("p", range(1, 9))
def test_fp64_matches_oracle_all_porders(gpu, p):
    say(f"[gpu] porder {p} jittered 3x2x2 box vs oracle")
    prob = m.box_problem(3, p, ny=2, nz=2, jitter=0.05, rho_jump=0.1, seeds=(21, 22, 23, 24))
    o = Oracle()
    ref = o.convec(prob) + o.diffu(prob)
    out = gpu_residuals(prob)
    for key, got, want in zip(KEYS, out, ref):
        assert rel_l2(got, want) <= TOL64, (key, rel_l2(got, want))


@pytest.mark.parametrize("p", range(1, 9))
def test_fp64_ragged_last_batch_all_porders(gpu, p):
    """15 elements: the last batch is short at every porder (p=4: three
    elements, its second pass holding one; p=3: seven, the second pass three
    of four; p>=5: the convection's last two-element batch holds one)."""
    say(f"[gpu] porder {p} jittered 5x3x1 box (ragged last batch) vs oracle")
    prob = m.box_problem(5, p, ny=3, nz=1, jitter=0.05, rho_jump=0.1, seeds=(31, 32, 33, 34))
    assert prob.nelem == 15
    o = Oracle()
    ref = o.convec(prob) + o.diffu(prob)
    out = gpu_residuals(prob)
    for key, got, want in zip(KEYS, out, ref):
        assert rel_l2(got, want) <= TOL64, (key, rel_l2(got, want))


def test_fp64_ragged_workgroup_tail(gpu):
    """37x31x17 elements at p=4: 4875 batches, the last of three elements;
    the convection takes two batches per workgroup there (at least 2048
    workgroups), so its last workgroup holds one batch."""
    say("[gpu] 37x31x17 p=4 jittered (ragged workgroup tail) vs oracle")
    prob = m.box_problem(37, 4, ny=31, nz=17, jitter=0.05, threads=min(16, os.cpu_count() or 1))
    assert prob.nelem == 19499 and -(-prob.nelem // 4) == 4875
    o = Oracle()
    ref = o.convec(prob) + o.diffu(prob)
    out = gpu_residuals(prob)
    for key, got, want in zip(KEYS, out, ref):
        assert rel_l2(got, want) <= TOL64, (key, rel_l2(got, want))


def test_fp64_config1_64cubed_p4_jittered(gpu):
    """64^3 elements, porder 4, jittered interior nodes (~17M nodes)."""
    say("[gpu] 64^3 p=4 jittered: generating")
    prob = m.box_problem(64, 4, jitter=0.05, threads=min(16, os.cpu_count() or 1))
    say("[gpu] 64^3 p=4: GPU kernels")
    out = gpu_residuals(prob)
    say("[gpu] 64^3 p=4: oracle")
    o = Oracle()
    nt = min(16, os.cpu_count() or 1)
    ref = o.convec(prob, nt) + o.diffu(prob, nt)
    for key, got, want in zip(KEYS, out, ref):
        err = rel_l2(got, want)
        say(f"    {key}: rel L2 {err:.3e}")
        assert err <= TOL64, (key, err)
    # size-independent property: the weak divergence of tau sums to zero
    for r in range(3):
        assert abs(out[4][:, r].sum()) <= 1e-12 * np.abs(out[4][:, r]).sum()
    # the reference's own residuals of the same inputs (digest: 4096 seeded nodes, norms)
    from test_gpu_full_configs import check_reference_digest
    check_reference_digest("digest_cfg1_64cubed_p4", prob, out)


def rounded_to_fp32(prob):
    """The same problem with geometry and state rounded to FP32 (widened back)."""
    import copy
    q = copy.copy(prob)
    for k in ("He", "gpvol", "u", "q", "rho", "E", "Tem", "mu_fluid", "mu_e", "mu_sgs"):
        setattr(q, k, np.asfortranarray(getattr(prob, k).astype(np.float32).astype(np.float64)))
    # the mixed convective kernel recovers p from E, rho, u
    ke = 0.5 * (q.u ** 2).sum(axis=1)
    q.pr = (prob.gamma - 1.0) * (q.E - q.rho * ke)
    return q


@pytest.mark.parametrize("p,n", [(3, 8), (4, 6), (4, 5), (6, 3), (7, 3)])
def test_mixed_precision_error(gpu, p, n):
    """FP32 storage of every input, FP64 arithmetic: equal to the FP64 oracle
    on the same FP32-rounded inputs (1e-12), and within the FP32 rounding of
    the inputs of the oracle on the FP64 inputs."""
    say(f"[gpu] mixed precision p={p} n={n}")
    prob = m.box_problem(n, p, jitter=0.05)
    o = Oracle()
    ref = o.convec(prob, 4) + o.diffu(prob, 4)
    q32 = rounded_to_fp32(prob)
    ref32 = o.convec(q32, 4) + o.diffu(q32, 4)
    out = gpu_residuals(prob, mixed=True)
    e2e = {k: rel_l2(g, w) for k, g, w in zip(KEYS, out, ref)}
    kern = {k: rel_l2(g, w) for k, g, w in zip(KEYS, out, ref32)}
    say("    vs FP64 oracle on FP64 inputs: " + " ".join(f"{k}={v:.2e}" for k, v in e2e.items()))
    say("    vs FP64 oracle on the FP32 inputs: " + " ".join(f"{k}={v:.2e}" for k, v in kern.items()))
    for k in KEYS:
        assert kern[k] <= TOL64, (k, kern[k])
        assert e2e[k] <= 4 * TOL_MIXED, (k, e2e[k])


# Bar of the FP32 arithmetic itself (gathers, fluxes, sweeps in FP32), against
# the FP64 oracle on the same FP32-rounded inputs; measured up to 1.3e-6 at
# these sizes.  Against the oracle on the FP64 inputs these forms also carry
# the FP32 rounding of the velocity, which exceeds the 1e-6 bar of the
# precision split on large meshes (tests/test_gpu_full_configs.py, xfail);
# the f32split form below meets it.
TOL_FP32 = 5e-6


@pytest.mark.parametrize("p,n", [(3, 8), (4, 6), (4, 5), (6, 3), (7, 3)])
def test_fp32_arithmetic_error(gpu, p, n):
    """FP32 storage and FP32 gathers, fluxes and derivative sweeps, FP64 node
    sums and assembly: the error against the FP64 oracle on the FP32-rounded
    inputs stays within TOL_FP32 (printed with the error on the FP64
    inputs); it differs from the FP64-arithmetic form by more than rounding
    of the sums."""
    say(f"[gpu] FP32 arithmetic p={p} n={n}")
    prob = m.box_problem(n, p, jitter=0.05)
    o = Oracle()
    ref = o.convec(prob, 4) + o.diffu(prob, 4)
    q32 = rounded_to_fp32(prob)
    ref32 = o.convec(q32, 4) + o.diffu(q32, 4)
    out = gpu_residuals(prob, mixed=True, fp32=True)
    e2e = {k: rel_l2(g, w) for k, g, w in zip(KEYS, out, ref)}
    kern = {k: rel_l2(g, w) for k, g, w in zip(KEYS, out, ref32)}
    say("    vs FP64 oracle on FP64 inputs: " + " ".join(f"{k}={v:.2e}" for k, v in e2e.items()))
    say("    vs FP64 oracle on the FP32 inputs: " + " ".join(f"{k}={v:.2e}" for k, v in kern.items()))
    for k in KEYS:
        assert kern[k] <= TOL_FP32, (k, kern[k])
    assert max(kern.values()) > 1e-10  # the sweeps really ran in FP32


TOL_SPLIT = 1e-6  # accuracy bar of the FP32 split against the FP64 oracle (relative L2)


@pytest.mark.parametrize("p,n", [(3, 8), (4, 6), (4, 5), (6, 3), (7, 3)])
def test_f32split_error(gpu, p, n):
    """FP32 He, gpvol, mu_e, mu_sgs, q, rho, E, Tem, mu_fluid; FP64 u and pr;
    FP64 arithmetic: equal to the FP64 oracle on the same rounded inputs
    (1e-12) and within TOL_SPLIT of the oracle on the FP64 inputs."""
    import copy
    say(f"[gpu] f32split p={p} n={n}")
    prob = m.box_problem(n, p, jitter=0.05)
    o = Oracle()
    ref = o.convec(prob, 4) + o.diffu(prob, 4)
    q = copy.copy(prob)
    for k in ("He", "gpvol", "q", "rho", "E", "Tem", "mu_fluid", "mu_e", "mu_sgs"):
        setattr(q, k, np.asfortranarray(getattr(prob, k).astype(np.float32).astype(np.float64)))
    refs = o.convec(q, 4) + o.diffu(q, 4)
    t = prob.tables
    f = prob.fp32()
    R, D = prob.outputs(), prob.outputs()
    ins = (prob.connec, f["He"], f["gpvol"], prob.u, f["q"], f["rho"], prob.pr, f["E"], f["Tem"],
           f["mu_fluid"], f["mu_e"], f["mu_sgs"])
    m.copyin(*ins)
    m.create(*R, *D)
    try:
        m.full_convec_ijk_f32split(prob.nelem, prob.npoin, prob.connec, f["He"], f["gpvol"], t.dlxigp_ip,
                                   t.invAtoIJK, t.AtoI, t.AtoJ, t.AtoK, prob.u, f["q"], f["rho"], prob.pr,
                                   f["E"], *R)
        m.full_diffusion_ijk_f32split(prob.nelem, prob.npoin, prob.connec, f["He"], f["gpvol"], t.dlxigp_ip,
                                      t.invAtoIJK, t.AtoI, t.AtoJ, t.AtoK, prob.Cp, prob.Pr, f["rho"], prob.u,
                                      f["Tem"], f["mu_fluid"], f["mu_e"], f["mu_sgs"], *D)
        m.update_host(*R, *D)
    finally:
        m.delete(*ins, *R, *D)
    out = R + D
    e2e = {k: rel_l2(g, w) for k, g, w in zip(KEYS, out, ref)}
    kern = {k: rel_l2(g, w) for k, g, w in zip(KEYS, out, refs)}
    say("    vs FP64 oracle on FP64 inputs: " + " ".join(f"{k}={v:.2e}" for k, v in e2e.items()))
    for k in KEYS:
        assert kern[k] <= TOL64, (k, kern[k])
        assert e2e[k] <= TOL_SPLIT, (k, e2e[k])


def gpu_residuals_geo32(prob):
    R, D = prob.outputs(), prob.outputs()
    t = prob.tables
    f = prob.fp32()
    ins = (prob.connec, f["He"], f["gpvol"], prob.u, prob.q, prob.rho, prob.pr, prob.E, prob.Tem,
           prob.mu_fluid, f["mu_e"], f["mu_sgs"])
    m.copyin(*ins)
    m.create(*R, *D)
    try:
        m.full_convec_ijk_geo32(prob.nelem, prob.npoin, prob.connec, f["He"], f["gpvol"],
                                t.dlxigp_ip, t.invAtoIJK, t.AtoI, t.AtoJ, t.AtoK, prob.u, prob.q,
                                prob.rho, prob.pr, prob.E, *R)
        m.full_diffusion_ijk_geo32(prob.nelem, prob.npoin, prob.connec, f["He"], f["gpvol"],
                                   t.dlxigp_ip, t.invAtoIJK, t.AtoI, t.AtoJ, t.AtoK, prob.Cp,
                                   prob.Pr, prob.rho, prob.u, prob.Tem, prob.mu_fluid, f["mu_e"],
                                   f["mu_sgs"], *D)
        m.update_host(*R, *D)
    finally:
        m.delete(*ins, *R, *D)
    return R + D


@pytest.mark.parametrize("p,n", [(3, 8), (4, 6), (4, 5), (6, 3), (7, 3)])
def test_geo32_precision_error(gpu, p, n):
    """FP32 element arrays (He, gpvol, mu_e, mu_sgs), FP64 nodal fields and
    arithmetic: within 1e-6 of the FP64 oracle on the FP64 inputs, and equal to
    the oracle on the same rounded element arrays (1e-12)."""
    import copy
    say(f"[gpu] geo32 precision p={p} n={n}")
    prob = m.box_problem(n, p, jitter=0.05)
    o = Oracle()
    ref = o.convec(prob, 4) + o.diffu(prob, 4)
    q = copy.copy(prob)
    for k in ("He", "gpvol", "mu_e", "mu_sgs"):
        setattr(q, k, np.asfortranarray(getattr(prob, k).astype(np.float32).astype(np.float64)))
    refg = o.convec(q, 4) + o.diffu(q, 4)
    out = gpu_residuals_geo32(prob)
    e2e = {k: rel_l2(g, w) for k, g, w in zip(KEYS, out, ref)}
    kern = {k: rel_l2(g, w) for k, g, w in zip(KEYS, out, refg)}
    say("    vs FP64 oracle on FP64 inputs: " + " ".join(f"{k}={v:.2e}" for k, v in e2e.items()))
    for k in KEYS:
        assert kern[k] <= TOL64, (k, kern[k])
        assert e2e[k] <= TOL_MIXED, (k, e2e[k])


def test_repeat_runs_differ_only_by_summation_order(gpu):
    say("[gpu] determinism")
    prob = m.box_problem(6, 4, jitter=0.05)
    a = gpu_residuals(prob)
    b = gpu_residuals(prob)
    for x, y in zip(a, b):
        assert rel_l2(x, y) <= 1e-14


def test_async_and_timing(gpu):
    say("[gpu] async + timing")
    prob = m.box_problem(4, 4)
    R = prob.outputs()
    ins = prob.device_inputs()
    m.copyin(*ins)
    m.create(*R)
    try:
        m.timing_reset()
        m.timing_enable(True)
        m.set_async(True)
        for _ in range(3):
            m.full_convec_ijk(*prob.convec_args(*R))
        m.synchronize()
        m.set_async(False)
        ms, n = m.timing_get(m.TIMER_CONVEC)
        assert n == 3 and ms > 0
        zms, zn = m.timing_get(m.TIMER_ZERO)
        assert zn == 3
        m.update_host(*R)
        want = Oracle().convec(prob)
        for g, w in zip(R, want):
            assert rel_l2(g, w) <= TOL64
    finally:
        m.timing_enable(False)
        m.set_async(False)
        m.delete(*ins, *R)


def test_error_paths(gpu):
    say("[gpu] error paths")
    prob = m.box_problem(2, 3)
    R = prob.outputs()
    # nothing present yet
    with pytest.raises(m.MsodError) as e:
        m.full_convec_ijk(*prob.convec_args(*R))
    assert e.value.status == m.MSOD_ERR_NOT_PRESENT
    ins = prob.device_inputs()
    m.copyin(*ins)
    m.create(*R)
    try:
        # inconsistent 1-D derivative table
        args = list(prob.convec_args(*R))
        bad_dl = prob.tables.dlxigp_ip.copy(order="F")
        bad_dl[5, 1, 2] += 1.0
        args[7] = bad_dl
        with pytest.raises(m.MsodError) as e:
            m.full_convec_ijk(*args)
        assert e.value.status == m.MSOD_ERR_TABLE
        # porder without a kernel
        L = m.hip_lib()
        st = L.msod_full_convec_ijk(9, prob.nelem, prob.npoin, *[m._ptr(a) for a in prob.convec_args(*R)[2:]])
        assert st == m.MSOD_ERR_PORDER
        # connec entry out of range: rejected before any kernel reads through it
        bad_c = prob.connec.copy(order="F")
        bad_c[1, 3] = prob.npoin + 5
        m.copyin(bad_c)
        try:
            args = list(prob.convec_args(*R))
            args[2] = bad_c
            with pytest.raises(m.MsodError) as e:
                m.full_convec_ijk(*args)
            assert e.value.status == m.MSOD_ERR_MESH
        finally:
            m.delete(bad_c)
        # output array too small for npoin
        small = np.zeros(prob.npoin - 1)
        m.create(small)
        try:
            with pytest.raises(m.MsodError) as e:
                m.full_convec_ijk(*prob.convec_args(small, R[1], R[2]))
            assert e.value.status == m.MSOD_ERR_MESH
        finally:
            m.delete(small)
        # gamma <= 1 for the forms that recover the pressure
        f = prob.fp32()
        t = prob.tables
        for fn in (L.msod_full_convec_ijk_mixed, L.msod_full_convec_ijk_fp32):
            st = fn(3, prob.nelem, prob.npoin, *[m._ptr(x) for x in (
                prob.connec, f["He"], f["gpvol"], t.dlxigp_ip, t.invAtoIJK, t.AtoI, t.AtoJ, t.AtoK,
                f["u"], f["q"], f["rho"], f["E"])], 1.0, *[m._ptr(x) for x in R])
            assert st == m.MSOD_ERR_ARG
        # the valid call still works after the failures
        m.full_convec_ijk(*prob.convec_args(*R))
        m.update_host(*R)
        for g, w in zip(R, Oracle().convec(prob)):
            assert rel_l2(g, w) <= TOL64
    finally:
        m.delete(*ins, *R)


@pytest.mark.parametrize("mode", [m.ASSEMBLY_COLORED, m.ASSEMBLY_ATOMIC])
def test_assembly_modes(gpu, mode):
    say(f"[gpu] assembly mode {mode}")
    m.set_assembly(mode)
    try:
        for name in ("case_box8_p3", "case_box2_p7_jit_jump", "case_refdriver_p4"):
            from golden.make_golden import CASES as BUILDERS
            gold, _ = load_case(name)
            out = gpu_residuals(BUILDERS[name]())
            for key, got in zip(KEYS, out):
                assert rel_l2(got, gold[key]) <= TOL64, (name, key)
    finally:
        m.set_assembly(m.ASSEMBLY_DEFAULT)


def test_colored_assembly_is_bitwise_reproducible(gpu):
    say("[gpu] bitwise reproducibility of the coloured assembly")
    prob = m.box_problem(5, 4, jitter=0.05)
    m.set_assembly(m.ASSEMBLY_COLORED)
    try:
        a = gpu_residuals(prob)
        b = gpu_residuals(prob)
    finally:
        m.set_assembly(m.ASSEMBLY_DEFAULT)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_shuffled_elements_and_unreferenced_nodes(gpu):
    """Any element order (batches then share few nodes and need more colours)
    and nodes that no element references (their residual is zero)."""
    say("[gpu] shuffled element order + unreferenced nodes")
    prob = m.box_problem(4, 3, jitter=0.05)
    rng = np.random.default_rng(5)
    perm = rng.permutation(prob.nelem)
    prob.connec = np.asfortranarray(prob.connec[perm])
    prob.He = np.asfortranarray(prob.He[:, :, :, perm])
    prob.gpvol = np.asfortranarray(prob.gpvol[:, :, perm])
    prob.mu_e = np.asfortranarray(prob.mu_e[perm])
    prob.mu_sgs = np.asfortranarray(prob.mu_sgs[perm])
    # append 7 unreferenced nodes
    extra = 7
    for k in ("u", "q"):
        setattr(prob, k, np.asfortranarray(np.vstack([getattr(prob, k), np.ones((extra, 3))])))
    for k in ("rho", "pr", "E", "Tem", "mu_fluid"):
        setattr(prob, k, np.concatenate([getattr(prob, k), np.ones(extra)]))
    prob.npoin += extra
    o = Oracle()
    ref = o.convec(prob) + o.diffu(prob)
    for mode in (m.ASSEMBLY_COLORED, m.ASSEMBLY_ATOMIC):
        m.set_assembly(mode)
        try:
            out = gpu_residuals(prob)
        finally:
            m.set_assembly(m.ASSEMBLY_DEFAULT)
        for key, got, want in zip(KEYS, out, ref):
            assert rel_l2(got, want) <= TOL64, (mode, key)
            assert np.all(np.asarray(got)[-extra:] == 0.0)


@pytest.mark.parametrize("args,case", [
    (["--nelem", "6", "--porder", "4"], "case_refdriver_p4"),
    (["--n", "8", "--porder", "3"], "case_box8_p3"),
    (["--n", "8", "--porder", "3", "--gpu-geometry"], "case_box8_p3"),
    (["--nelem", "5", "--porder", "3"], "case_refdriver_p3"),
])
def test_fortran_driver(gpu, tmp_path, args, case):
    """The Fortran timer driver (drop-in elem_convec/elem_diffu modules over the
    C ABI) reproduces the reference's residuals and writes timers.dat."""
    import subprocess
    exe = os.path.join(m.LIB_DIR, "miniSOD")
    say(f"[gpu] Fortran driver {' '.join(args)}")
    r = subprocess.run([exe, *args, "--nruns", "3"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    gold, _ = load_case(case)
    rc = np.loadtxt(tmp_path / "results_convec.dat")
    rd = np.loadtxt(tmp_path / "results_diffu.dat")
    assert rel_l2(rc[:, 1], gold["Rmass"]) <= TOL64
    assert rel_l2(rc[:, 2:5], gold["Rmom"]) <= TOL64
    assert rel_l2(rc[:, 5], gold["Rener"]) <= TOL64
    assert rel_l2(rd[:, 1], gold["Dmass"]) <= TOL64
    assert rel_l2(rd[:, 2:5], gold["Dmom"]) <= TOL64
    assert rel_l2(rd[:, 5], gold["Dener"]) <= TOL64
    # timers.dat records: format (i3,1x,4(f16.8,1x)) of main.f90:495,499
    lines = (tmp_path / "timers.dat").read_text().splitlines()
    assert len(lines) == 3
    for i, line in enumerate(lines, 1):
        assert re.fullmatch(r"[ \d]{2}\d( [ \d-]{6}\d\.\d{8}){4} ?", line), line
        assert int(line[:3]) == i
    timers = np.loadtxt(tmp_path / "timers.dat")
    assert timers.shape == (3, 5) and np.all(timers[:, 1] > 0) and np.all(timers[:, 3] > 0)
    assert "GLL-node updates/s" in r.stdout
    # the timing lines of main.f90:504-534, the split form's included
    for name in ("Avg. convective time", "Avg. split convective time", "Max. split convective time",
                 "Min. split convective time", "Variation split convec.", "Avg. diffusive time"):
        assert re.search(rf"{re.escape(name)}\s*=\s*[-+0-9.Ee]+", r.stdout), name
    # the summary block of main.f90:544-566 quotes the extremes of the dumps
    summary = r.stdout[r.stdout.index("Basic results:"):].split("Basic TET results:")[0]
    cols = {"Rmass": (rc, 1), "Rmom(:,1)": (rc, 2), "Rmom(:,2)": (rc, 3), "Rmom(:,3)": (rc, 4),
            "Rener": (rc, 5), "Dmass": (rd, 1), "Dmom(:,1)": (rd, 2), "Dmom(:,2)": (rd, 3),
            "Dmom(:,3)": (rd, 4), "Dener": (rd, 5)}
    num = r"([-+0-9.Ee]+)"
    for name, (arr, j) in cols.items():
        hit = re.search(rf"Max {re.escape(name)}\s*=\s*{num}\s*Min {re.escape(name)}\s*=\s*{num}", summary)
        assert hit, name
        assert float(hit.group(1)) == pytest.approx(arr[:, j].max(), rel=1e-14, abs=1e-300)
        assert float(hit.group(2)) == pytest.approx(arr[:, j].min(), rel=1e-14, abs=1e-300)
    # TET04 part (main.f90:364-407, 465-491, 567-581): runs when 4 divides npoin
    from golden.make_golden import CASES as BUILDERS
    prob = BUILDERS[case]()
    if prob.npoin % 4:
        assert "Basic TET results:" not in r.stdout and np.all(timers[:, [2, 4]] == 0)
        return
    assert np.all(timers[:, 2] > 0) and np.all(timers[:, 4] > 0)
    tp = m.tet_problem(prob)
    o = Oracle()
    want = dict(zip(KEYS, o.fem_convec(tp) + o.fem_diffu(tp)))
    tsum = r.stdout[r.stdout.index("Basic TET results:"):]
    cols = {"Rmass": ("Rmass", None), "Rmom(:,1)": ("Rmom", 0), "Rmom(:,2)": ("Rmom", 1),
            "Rmom(:,3)": ("Rmom", 2), "Rener": ("Rener", None), "Dmass": ("Dmass", None),
            "Dmom(:,1)": ("Dmom", 0), "Dmom(:,2)": ("Dmom", 1), "Dmom(:,3)": ("Dmom", 2),
            "Dener": ("Dener", None)}
    for name, (key, j) in cols.items():
        arr = want[key] if j is None else want[key][:, j]
        hit = re.search(rf"Max {re.escape(name)}\s*=\s*{num}\s*Min {re.escape(name)}\s*=\s*{num}", tsum)
        assert hit, name
        scale = np.abs(arr).max()
        assert float(hit.group(1)) == pytest.approx(arr.max(), rel=1e-12, abs=1e-12 * scale), name
        assert float(hit.group(2)) == pytest.approx(arr.min(), rel=1e-12, abs=1e-12 * scale), name

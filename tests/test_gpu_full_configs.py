"""GPU parity at the full sizes of the benchmark configurations, where the
CPU oracle cannot run the whole mesh in seconds:

  128^3 porder 4 (the benchmark workload), 48^3 porder 7 with the 10%
  density jump, 96^3 porder 3 (the closure configuration).

Checks per mesh:
  * sampled nodes: the residuals at 2000 random nodes against the oracle run
    on exactly the elements that touch them (so each sampled node receives
    all of its contributions), relative L2 <= 1e-12;
  * a uniform state (constant rho, u, p, T on the same jittered mesh): every
    residual vanishes, to rounding, relative to the residuals of the
    non-uniform state;
  * where the reference's own residuals were digested (tests/golden/digest_*:
    48^3 p=7, 96^3 p=3), the values at its 4096 seeded nodes and the
    per-component L2 norms, relative error <= 1e-12.
"""
import os
import sys

import numpy as np
import pytest

import minisod2d_amd as m
from oracle_lib import Oracle, rel_l2

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)
KEYS = ("Rmass", "Rmom", "Rener", "Dmass", "Dmom", "Dener")
CONFIGS = [
    (128, 4, 0.0, 0.0),   # benchmark workload
    (48, 7, 0.05, 0.1),   # porder 7, density jump
    (96, 3, 0.05, 0.0),   # closure configuration
]
DIGEST = {(48, 7): "digest_cfg3_48cubed_p7", (96, 3): "digest_cfg4_96cubed_p3"}
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def check_reference_digest(name, prob, out):
    """GPU residuals against the reference's digest of the same inputs."""
    import json
    from golden.make_golden import DIGEST_GROUPS, input_digests, residual_digest
    z = np.load(os.path.join(GOLD, name + ".npz"))
    meta = json.loads(str(z["meta"]))
    assert input_digests(prob) == meta["inputs"], "generator output changed"
    dg = residual_digest(out, prob.npoin)
    assert np.array_equal(dg["nodes"], z["nodes"])
    # per residual array, like every parity check (a component such as Rmom3 of
    # these fields is mostly cancellation, so its own relative error measures
    # the summation order, not the residual)
    for key, cols in DIGEST_GROUPS:
        err = rel_l2(dg["values"][:, cols], z["values"][:, cols])
        assert err <= 1e-12, (key, err)
        ng, nz = np.linalg.norm(dg["norms"][list(cols)]), np.linalg.norm(z["norms"][list(cols)])
        assert abs(ng - nz) <= 1e-12 * nz, key


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def run_pair(prob):
    R, D = prob.outputs(), prob.outputs()
    ins = prob.device_inputs()
    m.copyin(*ins)
    m.create(*R, *D)
    try:
        m.full_convec_ijk(*prob.convec_args(*R))
        m.full_diffusion_ijk(*prob.diffu_args(*D))
        m.update_host(*R, *D)
    finally:
        m.delete(*ins, *R, *D)
    return R + D


def sampled_oracle(prob, nodes):
    """Oracle residuals at the 1-based `nodes`, from the elements touching them."""
    sub, pos = sampled_problem(prob, nodes)
    o = Oracle()
    want = o.convec(sub, 8) + o.diffu(sub, 8)
    return [w[pos] for w in want]


def sampled_problem(prob, nodes):
    """The elements touching the 1-based `nodes` as a problem of their own, and
    the positions of `nodes` in its node numbering."""
    mark = np.zeros(prob.npoin + 1, dtype=bool)
    mark[nodes] = True
    hit = mark[prob.connec].any(axis=1)
    els = np.nonzero(hit)[0]
    sub_conn = prob.connec[els]
    uniq, inv = np.unique(sub_conn, return_inverse=True)
    f = np.asfortranarray
    sub = m.Problem(prob.tables, len(els), len(uniq),
                    f((inv.reshape(sub_conn.shape) + 1).astype(np.int32)), prob.coord[uniq - 1],
                    f(prob.He[:, :, :, els]), f(prob.gpvol[:, :, els]),
                    *[f(getattr(prob, k)[uniq - 1]) for k in
                      ("u", "q", "rho", "pr", "E", "Tem", "mu_fluid")],
                    f(prob.mu_e[els]), f(prob.mu_sgs[els]), Cp=prob.Cp, Pr=prob.Pr, gamma=prob.gamma)
    return sub, np.searchsorted(uniq, nodes)


def uniform_state(prob):
    """The same mesh and viscosities with constant rho, u, p, T."""
    g = prob.gamma
    rho, u, p = 1.2, np.array([0.3, -0.2, 0.1]), 0.9
    n = prob.npoin
    prob.rho[:] = rho
    prob.u[:] = u[None, :]
    prob.q[:] = rho * u[None, :]
    prob.pr[:] = p
    prob.E[:] = p / (g - 1.0) + 0.5 * rho * float(u @ u)
    prob.Tem[:] = p / (rho * 287.0)
    assert prob.rho.shape == (n,)
    return prob


@pytest.mark.parametrize("n,p,jitter,jump", CONFIGS)
def test_full_size_sampled_parity_and_uniform_state(gpu, n, p, jitter, jump):
    say(f"[gpu] full size {n}^3 p={p}: building the mesh")
    prob = m.box_problem(n, p, jitter=jitter, rho_jump=jump, threads=THREADS)
    say(f"[gpu] full size {n}^3 p={p}: {prob.npoin} nodes, kernels")
    out = run_pair(prob)
    rng = np.random.default_rng(11 + p)
    nodes = rng.choice(prob.npoin, size=2000, replace=False) + 1
    say(f"[gpu] full size {n}^3 p={p}: oracle on the elements of 2000 nodes")
    want = sampled_oracle(prob, nodes)
    for key, got, w in zip(KEYS, out, want):
        err = rel_l2(got[nodes - 1], w)
        assert err <= 1e-12, (key, err)
    if (n, p) in DIGEST:
        say(f"[gpu] full size {n}^3 p={p}: against the reference's digest")
        check_reference_digest(DIGEST[(n, p)], prob, out)
    scale = [float(np.sqrt(np.mean(np.square(r)))) for r in out]
    del out
    say(f"[gpu] full size {n}^3 p={p}: uniform state")
    flat = run_pair(uniform_state(prob))
    for key, r, s in zip(KEYS, flat, scale):
        rms = float(np.sqrt(np.mean(np.square(r))))
        assert rms <= 1e-10 * s, (key, rms, s)


# ---------------------------------------------------------------------------
# precision forms at full size
# ---------------------------------------------------------------------------
SPLIT_F32 = ("He", "gpvol", "q", "rho", "E", "Tem", "mu_fluid", "mu_e", "mu_sgs")   # u, pr stay FP64
ALL_F32 = SPLIT_F32 + ("u", "pr")
BAR = 1e-6   # accuracy bar of the FP32 split against the FP64 oracle (relative L2)


def rounded(prob, fields):
    """A copy of prob with `fields` rounded to FP32 (widened back to FP64)."""
    import copy
    q = copy.copy(prob)
    for k in fields:
        setattr(q, k, np.asfortranarray(getattr(prob, k).astype(np.float32).astype(np.float64)))
    return q


def run_form(prob, form):
    t = prob.tables
    f = prob.fp32()
    R, D = prob.outputs(), prob.outputs()
    if form == "f32split":
        ins = (prob.connec, f["He"], f["gpvol"], prob.u, f["q"], f["rho"], prob.pr, f["E"], f["Tem"],
               f["mu_fluid"], f["mu_e"], f["mu_sgs"])
    else:
        ins = (prob.connec, f["He"], f["gpvol"], f["u"], f["q"], f["rho"], f["E"], f["Tem"], f["mu_fluid"],
               f["mu_e"], f["mu_sgs"])
    m.copyin(*ins)
    m.create(*R, *D)
    try:
        if form == "f32split":
            m.full_convec_ijk_f32split(prob.nelem, prob.npoin, prob.connec, f["He"], f["gpvol"], t.dlxigp_ip,
                                       t.invAtoIJK, t.AtoI, t.AtoJ, t.AtoK, prob.u, f["q"], f["rho"], prob.pr,
                                       f["E"], *R)
            m.full_diffusion_ijk_f32split(prob.nelem, prob.npoin, prob.connec, f["He"], f["gpvol"],
                                          t.dlxigp_ip, t.invAtoIJK, t.AtoI, t.AtoJ, t.AtoK, prob.Cp, prob.Pr,
                                          f["rho"], prob.u, f["Tem"], f["mu_fluid"], f["mu_e"], f["mu_sgs"], *D)
        else:
            conv = m.full_convec_ijk_fp32 if form == "fp32" else m.full_convec_ijk_mixed
            diff = m.full_diffusion_ijk_fp32 if form == "fp32" else m.full_diffusion_ijk_mixed
            conv(prob.nelem, prob.npoin, prob.connec, f["He"], f["gpvol"], t.dlxigp_ip, t.invAtoIJK, t.AtoI,
                 t.AtoJ, t.AtoK, f["u"], f["q"], f["rho"], f["E"], prob.gamma, *R)
            diff(prob.nelem, prob.npoin, prob.connec, f["He"], f["gpvol"], t.dlxigp_ip, t.invAtoIJK, t.AtoI,
                 t.AtoJ, t.AtoK, prob.Cp, prob.Pr, f["rho"], f["u"], f["Tem"], f["mu_fluid"], f["mu_e"],
                 f["mu_sgs"], *D)
        m.update_host(*R, *D)
    finally:
        m.delete(*ins, *R, *D)
    return R + D


@pytest.fixture(scope="module", params=[(128, 4, 0.0, 0.0), (48, 7, 0.05, 0.1)],
                ids=["128cubed_p4", "48cubed_p7_jump"])
def precision_case(request):
    n, p, jitter, jump = request.param
    say(f"[gpu] precision forms {n}^3 p={p}: building the mesh")
    prob = m.box_problem(n, p, jitter=jitter, rho_jump=jump, threads=THREADS)
    nodes = np.random.default_rng(23 + p).choice(prob.npoin, size=2000, replace=False) + 1
    say(f"[gpu] precision forms {n}^3 p={p}: FP64 oracle at 2000 nodes")
    want = sampled_oracle(prob, nodes)
    out = {}
    for form in ("f32split", "mixed", "fp32"):
        say(f"[gpu] precision forms {n}^3 p={p}: {form} kernels")
        out[form] = [r[nodes - 1] for r in run_form(prob, form)]
    yield dict(n=n, p=p, prob=prob, nodes=nodes, want=want, out=out)


def test_f32split_within_bar_at_full_size(gpu, precision_case):
    """The FP32 split (FP64 u and pr) against the FP64 oracle on the FP64
    inputs: every residual within BAR = 1e-6; against the oracle on the same
    rounded inputs, 1e-12 (the only error is the storage of the inputs)."""
    c = precision_case
    got = c["out"]["f32split"]
    e2e = {k: rel_l2(g, w) for k, g, w in zip(KEYS, got, c["want"])}
    say(f"    f32split {c['n']}^3 p={c['p']} vs FP64 oracle: " + " ".join(f"{k}={v:.2e}" for k, v in e2e.items()))
    want_r = sampled_oracle(rounded(c["prob"], SPLIT_F32), c["nodes"])
    for k, g, w in zip(KEYS, got, want_r):
        assert rel_l2(g, w) <= 1e-12, (k, rel_l2(g, w))
    for k, v in e2e.items():
        assert v <= BAR, (k, v)


def test_mixed_equals_oracle_on_its_rounded_inputs_at_full_size(gpu, precision_case):
    """Every input in FP32, FP64 arithmetic: equal to the FP64 oracle on the
    same rounded inputs (1e-12); the distance to the FP64 inputs' oracle is
    printed (it is the input rounding, see the xfail below)."""
    c = precision_case
    got = c["out"]["mixed"]
    e2e = {k: rel_l2(g, w) for k, g, w in zip(KEYS, got, c["want"])}
    say(f"    mixed {c['n']}^3 p={c['p']} vs FP64 oracle: " + " ".join(f"{k}={v:.2e}" for k, v in e2e.items()))
    q = rounded(c["prob"], ALL_F32)
    q.pr = None   # the mixed form recovers the pressure from the rounded state
    want_r = sampled_oracle_recovered_p(q, c["nodes"])
    for k, g, w in zip(KEYS, got, want_r):
        assert rel_l2(g, w) <= 1e-12, (k, rel_l2(g, w))


def test_fp32_arithmetic_error_at_full_size(gpu, precision_case):
    """FP32 gathers, fluxes and sweeps: against the FP64 oracle on the same
    rounded inputs the FP32 arithmetic adds at most 5e-5 (printed per
    residual; measured 2.7e-5 in the diffusive momentum residual at both
    sizes, whose two derivative sweeps cancel, and <= 4e-6 elsewhere)."""
    c = precision_case
    got = c["out"]["fp32"]
    q = rounded(c["prob"], ALL_F32)
    q.pr = None
    want_r = sampled_oracle_recovered_p(q, c["nodes"])
    kern = {k: rel_l2(g, w) for k, g, w in zip(KEYS, got, want_r)}
    e2e = {k: rel_l2(g, w) for k, g, w in zip(KEYS, got, c["want"])}
    say(f"    fp32 {c['n']}^3 p={c['p']} vs oracle on its inputs: " + " ".join(f"{k}={v:.2e}" for k, v in kern.items()))
    say(f"    fp32 {c['n']}^3 p={c['p']} vs FP64 oracle: " + " ".join(f"{k}={v:.2e}" for k, v in e2e.items()))
    for k, v in kern.items():
        assert v <= 5e-5, (k, v)


@pytest.mark.xfail(reason="FP32 storage of the velocity: the diffusive momentum residual carries "
                          "4e-5 (128^3 p=4) and 2.5e-5 (48^3 p=7) against the FP64 oracle; "
                          "the f32split form keeps u and pr in FP64 and meets the bar",
                   strict=False)
@pytest.mark.parametrize("form", ["mixed", "fp32"])
def test_all_fp32_storage_forms_against_bar(gpu, precision_case, form):
    c = precision_case
    e2e = {k: rel_l2(g, w) for k, g, w in zip(KEYS, c["out"][form], c["want"])}
    assert max(e2e.values()) <= BAR, e2e


def sampled_oracle_recovered_p(q, nodes):
    """sampled_oracle for the forms that recover the pressure from the rounded
    conservative state, p = (gamma-1)(E - rho|u|^2/2), in FP64."""
    g = q.gamma
    q.pr = np.asfortranarray((g - 1.0) * (q.E - 0.5 * q.rho * np.sum(q.u * q.u, axis=1)))
    return sampled_oracle(q, nodes)


# ---------------------------------------------------------------------------
# learned closure at the closure configuration's full size
# ---------------------------------------------------------------------------
def test_fused_closure_at_full_size(gpu):
    """96^3 p=3: full_diffusion_ijk with mu_sgs predicted inside the kernel
    (bf16 MFMA), at 2000 sampled nodes, against the oracle diffusion fed the
    oracle MLP's mu_sgs on the features of the same elements (1e-4, the bound
    of the bf16 network's rounding, as at small sizes); weights with a scale
    that makes mu_sgs matter (up to half of mu_fluid)."""
    from minisod2d_amd import closure as C
    say("[gpu] closure 96^3 p=3: building the mesh")
    prob = m.box_problem(96, 3, jitter=0.05, threads=THREADS)
    t = prob.tables
    nodes = np.random.default_rng(31).choice(prob.npoin, size=2000, replace=False) + 1
    sub, pos = sampled_problem(prob, nodes)
    # features of the sampled elements (they depend on the element's own nodes only)
    F = np.zeros((sub.nelem * t.nnode, 12), np.float32)
    ins = [sub.connec, sub.He, sub.gpvol, sub.rho, sub.u]
    m.copyin(*ins)
    m.create(F)
    try:
        m.closure_features(sub.nelem, sub.npoin, sub.connec, sub.He, sub.gpvol, t.dlxigp_ip, t.invAtoIJK,
                           t.AtoI, t.AtoJ, t.AtoK, sub.rho, sub.u, F)
        m.update_host(F)
    finally:
        m.delete(*ins, F)
    rng = np.random.default_rng(5)
    P = np.zeros(m.CLOSURE_NPARAM, np.float32)
    P[C.OW1:C.OB1] = rng.normal(0, 0.5, 12 * 32)
    P[C.OB1:C.OW2] = rng.normal(0, 0.2, 32)
    P[C.OW2:C.OB2] = rng.normal(0, 0.3, 32 * 32)
    P[C.OB2:C.OW3] = rng.normal(0, 0.2, 32)
    P[C.OW3:C.OB3] = rng.normal(0, 0.5, 32)
    P[C.OB3] = -0.3
    f64 = F.astype(np.float64)
    P[C.OMEAN:C.OISTD] = f64.mean(0)
    sd = f64.std(0)
    P[C.OISTD:C.OSCALE] = np.where(sd > 0, 1.0 / np.where(sd > 0, sd, 1.0), 1.0)
    P[C.OSCALE] = 5e-4
    say("[gpu] closure 96^3 p=3: fused kernel on the whole mesh")
    R = prob.outputs()
    ins = [prob.connec, prob.He, prob.gpvol, prob.rho, prob.u, prob.Tem, prob.mu_fluid, prob.mu_e]
    m.copyin(*ins)
    m.create(*R)
    try:
        used = m.full_diffusion_ijk_closure(prob.nelem, prob.npoin, prob.connec, t.Ngp, t.dNgp, prob.He,
                                            prob.gpvol, t.dlxigp_ip, t.xgp, t.invAtoIJK, t.AtoI, t.AtoJ,
                                            t.AtoK, prob.Cp, prob.Pr, prob.rho, prob.u, prob.Tem, prob.mu_fluid,
                                            prob.mu_e, P, *R)
        m.update_host(*R)
    finally:
        m.delete(*ins, *R)
    assert used
    say("[gpu] closure 96^3 p=3: oracle MLP and diffusion at the sampled nodes")
    o = Oracle()
    _, muo = o.mlp_forward(P, F, quant=1)
    l2g = t.invAtoIJK.ravel(order="F") - 1
    mu = np.zeros((sub.nelem, t.nnode), order="F")
    mu[:, l2g] = muo.reshape(sub.nelem, t.nnode)
    assert float(mu.max()) > 0.1 * float(prob.mu_fluid.max())  # the closure matters here
    want = o.diffu(sub, 8, mu_sgs=mu)
    zero = o.diffu(sub, 8, mu_sgs=np.zeros_like(mu))
    for key, got, w, z in zip(("Dmass", "Dmom", "Dener"), R, want, zero):
        err = rel_l2(got[nodes - 1], w[pos])
        dz = rel_l2(z[pos], w[pos])
        say(f"    {key}: fused vs oracle {err:.2e} (mu_sgs = 0 differs by {dz:.2e})")
        assert err <= 1e-4, (key, err)
        if key != "Dmass":  # the mass diffusion does not depend on mu_sgs (elem_diffu.f90:143,183)
            assert dz > 100 * err, (key, dz, err)

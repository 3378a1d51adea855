"""GPU tests of the learned closure, through the C ABI, against the closure
oracle (oracle/msod_oracle_mlp.c), the diffusion oracle (oracle/msod_oracle.c)
and closed forms.

Tolerances (stated per test):
* features: float32 storage, 1e-5 relative to the field's scale;
* MLP forward: bf16 matrix-core inputs on both sides; the GPU sums in FP32
  and uses single-precision exp, so mu_sgs agrees to 1e-3 relative L2;
* parameter gradient: the GPU also rounds the back-propagated deltas to bf16
  for its matrix products, 2e-2 relative L2 (cosine >= 0.999);
* sensitivity d<r, R_diffu>/d mu_sgs: the residual is affine in mu_sgs, so the
  FP64 oracle's finite difference is exact; the GPU stores float32: 1e-5;
* fused inference vs explicit mu_sgs: the same MLP code in both kernels,
  1e-10 relative L2.
"""
import sys

import numpy as np
import pytest

import minisod2d_amd as m
from minisod2d_amd import closure as C
from oracle_lib import Oracle, rel_l2

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def orc():
    return Oracle()


def l2g(t):
    return t.invAtoIJK.ravel(order="F") - 1


def random_params(seed=0, scale=1e-4, features=None):
    rng = np.random.default_rng(seed)
    P = np.zeros(m.CLOSURE_NPARAM, np.float32)
    P[C.OW1:C.OB1] = rng.normal(0, 0.5, 12 * 32)
    P[C.OB1:C.OW2] = rng.normal(0, 0.2, 32)
    P[C.OW2:C.OB2] = rng.normal(0, 0.3, 32 * 32)
    P[C.OB2:C.OW3] = rng.normal(0, 0.2, 32)
    P[C.OW3:C.OB3] = rng.normal(0, 0.5, 32)
    P[C.OB3] = -0.3
    if features is None:
        P[C.OISTD:C.OSCALE] = 1.0
    else:
        f = features.astype(np.float64)
        P[C.OMEAN:C.OISTD] = f.mean(0)
        s = f.std(0)
        P[C.OISTD:C.OSCALE] = np.where(s > 0, 1.0 / np.where(s > 0, s, 1.0), 1.0)
    P[C.OSCALE] = scale
    return P


def gpu_features(prob):
    t = prob.tables
    F = np.zeros((prob.nelem * t.nnode, 12), np.float32)
    ins = [prob.connec, prob.He, prob.gpvol, prob.rho, prob.u]
    m.copyin(*ins)
    m.create(F)
    try:
        m.closure_features(prob.nelem, prob.npoin, prob.connec, prob.He, prob.gpvol, t.dlxigp_ip,
                           t.invAtoIJK, t.AtoI, t.AtoJ, t.AtoK, prob.rho, prob.u, F)
        m.update_host(F)
    finally:
        m.delete(*ins, F)
    return F


def test_features_closed_form(gpu):
    """u = A x + c is represented exactly by every isoparametric element, so
    G = A at every element node, also on jittered meshes; h = element edge on
    an affine uniform mesh."""
    say("[gpu] closure features, closed form")
    A = np.array([[0.3, -1.2, 0.7], [2.0, 0.1, -0.4], [-0.5, 0.9, 1.1]])
    for jitter, n in ((0.05, 3), (0.0, 4)):
        prob = m.box_problem(n, 3, jitter=jitter)
        prob.u[:] = prob.coord @ A.T + np.array([0.2, -0.1, 0.3])
        F = gpu_features(prob)
        conn = prob.connec[:, l2g(prob.tables)] - 1           # (nelem, nn) lexicographic
        g = conn.ravel()
        np.testing.assert_allclose(F[:, :9], np.broadcast_to(A.ravel(), (len(F), 9)),
                                   atol=1e-5 * np.abs(A).max())
        np.testing.assert_allclose(F[:, 9], prob.rho[g], rtol=1e-6)
        np.testing.assert_allclose(F[:, 10], np.linalg.norm(prob.u[g], axis=1), rtol=1e-6)
        if jitter == 0.0:
            np.testing.assert_allclose(F[:, 11], 1.0 / n, rtol=1e-6)


def test_predict_matches_oracle(gpu, orc):
    say("[gpu] closure predict vs oracle")
    prob = m.box_problem(3, 3, jitter=0.05)
    F = gpu_features(prob)
    P = random_params(1, features=F)
    t = prob.tables
    mu = np.zeros((prob.nelem, t.nnode), order="F")
    m.copyin(F)
    m.create(mu)
    try:
        m.closure_predict(prob.nelem, t.invAtoIJK, P, F, mu)
        m.update_host(mu)
    finally:
        m.delete(F, mu)
    _, ref = orc.mlp_forward(P, F, quant=1)
    got = mu[:, l2g(t)].ravel()
    assert rel_l2(got, ref) <= 1e-3
    assert np.max(np.abs(got - ref) / np.abs(ref)) <= 2e-2
    assert np.all(got >= 0)


def _sensitivity(prob, rerr):
    t = prob.tables
    gnu = np.zeros(prob.nelem * t.nnode, np.float32)
    ins = [prob.connec, prob.He, prob.gpvol, prob.rho, prob.u, prob.Tem, rerr]
    m.copyin(*ins)
    m.create(gnu)
    try:
        m.closure_sensitivity(prob.nelem, prob.npoin, prob.connec, prob.He, prob.gpvol,
                              t.dlxigp_ip, t.invAtoIJK, t.AtoI, t.AtoJ, t.AtoK, prob.Cp, prob.rho,
                              prob.u, prob.Tem, rerr, gnu)
        m.update_host(gnu)
    finally:
        m.delete(*ins, gnu)
    return gnu.reshape(prob.nelem, t.nnode)[:, np.argsort(l2g(t))]   # -> reference order (e, a)


def _dot(prob, orc, rerr, mu):
    _, Dmom, Dener = orc.diffu(prob, mu_sgs=np.asfortranarray(mu))
    return float((rerr[:3] * Dmom.T).sum() + (rerr[3] * Dener).sum())


@pytest.mark.parametrize("p,n,jitter", [(2, 1, 0.05), (3, 2, 0.05)])
def test_sensitivity_is_exact_derivative(gpu, orc, p, n, jitter):
    say(f"[gpu] closure sensitivity p={p} n={n}")
    prob = m.box_problem(n, p, jitter=jitter)
    rng = np.random.default_rng(p)
    rerr = np.ascontiguousarray(rng.normal(size=(4, prob.npoin)))
    rerr[3] *= 1e-3                       # energy residuals are larger; keep terms comparable
    g = _sensitivity(prob, rerr)
    mu0 = np.asfortranarray(prob.mu_sgs)
    base = _dot(prob, orc, rerr, mu0)
    if prob.nelem == 1:                   # every component
        ref = np.zeros_like(g)
        for a in range(g.shape[1]):
            d = mu0.copy()
            d[0, a] += 1.0
            ref[0, a] = _dot(prob, orc, rerr, d) - base
        np.testing.assert_allclose(g, ref, atol=1e-5 * np.abs(ref).max(), rtol=1e-5)
    for k in range(3):                    # random directions
        dm = rng.normal(size=mu0.shape)
        lhs = float((g * dm).sum())
        rhs = _dot(prob, orc, rerr, mu0 + dm) - base
        assert abs(lhs - rhs) <= 1e-5 * np.abs(g * dm).sum(), (lhs, rhs)


@pytest.mark.parametrize("nsamp", [37, 8000])
def test_gradient_matches_oracle(gpu, orc, nsamp):
    say(f"[gpu] closure gradient nsamp={nsamp}")
    prob = m.box_problem(6, 3, jitter=0.05)            # 13824 element nodes
    F = gpu_features(prob)[:nsamp].copy()
    P = random_params(2, features=F)
    gnu = np.random.default_rng(3).uniform(0.5, 1.5, nsamp).astype(np.float32)
    grad = np.zeros(m.CLOSURE_NTRAIN, np.float32)
    m.copyin(F, gnu)
    m.create(grad)
    try:
        m.closure_gradient(nsamp, P, F, gnu, grad)
        m.update_host(grad)
    finally:
        m.delete(F, gnu, grad)
    ref = orc.mlp_gradient(P, F, gnu.astype(np.float64), quant=1)
    cos = float(grad @ ref / (np.linalg.norm(grad) * np.linalg.norm(ref)))
    say(f"    rel L2 {rel_l2(grad, ref):.2e}, cosine {cos:.6f}")
    assert rel_l2(grad, ref) <= 2e-2
    assert cos >= 0.999


def test_adam_matches_oracle(gpu, orc):
    say("[gpu] closure adam")
    rng = np.random.default_rng(4)
    P = random_params(5)
    Pg, m1, v1 = P.copy(), np.zeros(m.CLOSURE_NTRAIN, np.float32), np.zeros(m.CLOSURE_NTRAIN, np.float32)
    Po, m2, v2 = P.copy(), m1.copy(), v1.copy()
    grad = np.zeros(m.CLOSURE_NTRAIN, np.float32)
    m.copyin(Pg, m1, v1, grad)
    try:
        for step in range(1, 4):
            grad[:] = rng.normal(size=grad.shape)
            m.update_device(grad)
            m.closure_adam(Pg, grad, m1, v1, step, 1e-2)
            orc.adam(Po, grad, m2, v2, step, 1e-2)
        m.update_host(Pg, m1, v1)
    finally:
        m.delete(Pg, m1, v1, grad)
    # float32 moment updates on the GPU, double then float in the oracle:
    # equal to float rounding (a few ulp of the largest entries)
    np.testing.assert_allclose(Pg, Po, rtol=1e-5, atol=1e-6 * np.abs(Po).max())
    np.testing.assert_allclose(m1, m2, rtol=1e-5, atol=1e-6 * np.abs(m2).max())
    np.testing.assert_allclose(v1, v2, rtol=1e-5, atol=1e-6 * np.abs(v2).max())


def test_loss_matches_numpy(gpu):
    say("[gpu] closure loss")
    rng = np.random.default_rng(6)
    n = 70001
    Rmom = np.asfortranarray(rng.normal(size=(n, 3)))
    Rener = rng.normal(size=n)
    tgt = np.ascontiguousarray(rng.normal(size=(4, n)))
    w = np.array([1.0, 2.0, 0.5, 1e-3])
    rerr = np.zeros((4, n))
    m.copyin(Rmom, Rener, tgt)
    m.create(rerr)
    try:
        L = m.closure_loss(n, Rmom, Rener, tgt, w, rerr)
        m.update_host(rerr)
    finally:
        m.delete(Rmom, Rener, tgt, rerr)
    R = np.vstack([Rmom.T, Rener[None]])
    d = R - tgt
    assert abs(L - float((w[:, None] * d * d).sum())) <= 1e-12 * abs(L)
    np.testing.assert_allclose(rerr, 2 * w[:, None] * d, rtol=1e-15)


def _closure_diffu(prob, params):
    t = prob.tables
    R = prob.outputs()
    ins = [prob.connec, prob.He, prob.gpvol, prob.rho, prob.u, prob.Tem, prob.mu_fluid, prob.mu_e]
    m.copyin(*ins)
    m.create(*R)
    try:
        used = m.full_diffusion_ijk_closure(prob.nelem, prob.npoin, prob.connec, t.Ngp, t.dNgp,
                                            prob.He, prob.gpvol, t.dlxigp_ip, t.xgp, t.invAtoIJK,
                                            t.AtoI, t.AtoJ, t.AtoK, prob.Cp, prob.Pr, prob.rho,
                                            prob.u, prob.Tem, prob.mu_fluid, prob.mu_e, params, *R)
        m.update_host(*R)
    finally:
        m.delete(*ins, *R)
    return used, R


def _plain_diffu(prob, mu):
    R = prob.outputs()
    ins = list(prob.device_inputs()) + [mu]
    m.copyin(*ins)
    m.create(*R)
    try:
        m.full_diffusion_ijk(*prob.diffu_args(*R, mu_sgs=mu))
        m.update_host(*R)
    finally:
        m.delete(*ins, *R)
    return R


@pytest.mark.parametrize("p", [3, 4, 7])
def test_fused_inference_equals_explicit(gpu, orc, p):
    say(f"[gpu] fused closure diffusion p={p}")
    prob = m.box_problem(2, p, jitter=0.05)
    F = gpu_features(prob)
    P = random_params(7, scale=5e-4, features=F)
    t = prob.tables
    mu = np.zeros((prob.nelem, t.nnode), order="F")
    m.copyin(F)
    m.create(mu)
    try:
        m.closure_predict(prob.nelem, t.invAtoIJK, P, F, mu)
        m.update_host(mu)
    finally:
        m.delete(F, mu)
    used, Rf = _closure_diffu(prob, P)
    assert used
    Re = _plain_diffu(prob, mu)
    for a, b in zip(Rf, Re):
        assert rel_l2(a, b) <= 1e-10
    # end to end against the oracles: oracle MLP (bf16 points) + oracle diffusion
    _, muo = orc.mlp_forward(P, F, quant=1)
    mref = np.zeros_like(mu)
    mref[:, l2g(t)] = muo.reshape(prob.nelem, t.nnode)
    Ro = orc.diffu(prob, mu_sgs=mref)
    for a, b in zip(Rf, Ro):
        assert rel_l2(a, b) <= 1e-4


def test_fallback_to_zero_viscosity(gpu):
    say("[gpu] closure fallback")
    prob = m.box_problem(2, 3, jitter=0.05)
    zero = np.zeros_like(prob.mu_sgs)
    m.set_assembly(m.ASSEMBLY_COLORED)  # fixed summation order: bitwise comparisons
    try:
        _fallback_cases(prob, zero)
    finally:
        m.set_assembly(m.ASSEMBLY_DEFAULT)


def _fallback_cases(prob, zero):
    R0 = _plain_diffu(prob, zero)
    bad_nan = random_params(8)
    bad_nan[100] = np.nan
    bad_scale = random_params(8)
    bad_scale[C.OSCALE] = -1.0
    for params in (None, bad_nan, bad_scale):
        used, R = _closure_diffu(prob, params)
        assert not used
        for a, b in zip(R, R0):
            assert np.array_equal(a, b)
    F = np.zeros((prob.nelem * prob.tables.nnode, 12), np.float32)
    with pytest.raises(m.MsodError) as e:
        m.closure_predict(prob.nelem, prob.tables.invAtoIJK, bad_nan, F,
                          np.zeros((prob.nelem, prob.tables.nnode), order="F"))
    assert e.value.status == m.MSOD_ERR_MODEL


def test_training_learns_a_planted_closure(gpu):
    """End to end on the GPU (features, MLP, diffusion, loss, sensitivity,
    back-propagation, Adam): the target residuals come from mu_sgs of a
    known teacher network; a student from another initialisation recovers
    most of the gap on its training pairs and on a held-out pair."""
    say("[gpu] closure training on a planted closure")
    rng = np.random.default_rng(11)

    def problem(n, seed):
        prob = m.box_problem(n, 3, jitter=0.05, seeds=C.pair_seeds(seed))
        return prob

    F = np.concatenate([gpu_features(problem(4, 0)), gpu_features(problem(5, 1))])
    teacher = C.init_params(F, scale=2e-4, b3=1.0, seed=5)
    teacher[C.OW3:C.OB3] = rng.normal(0, 0.3, 32)       # a visibly non-constant closure
    pairs = [C.planted_pair(problem(4, 0), teacher), C.planted_pair(problem(5, 1), teacher)]
    held = C.planted_pair(problem(5, 2), teacher)
    res = C.train(pairs, steps=400, lr=1e-2, seed=1, log=say)
    h = res["history"]
    say(f"    training loss {h[0]:.4f} -> {h[-1]:.4f} (mu_sgs = 0 scores {4 * len(pairs)})")
    assert h[-1] < 0.2 * 4 * len(pairs)
    ev = C.evaluate(held, res["params"])
    say(f"    held-out loss {ev['loss']:.4f} (mu_sgs = 0 scores 4), gain {ev['gain']:.3f}")
    assert ev["used_model"] and ev["gain"] > 0.8


def test_trained_closure_on_held_out_pairs(gpu):
    """The committed weights (trained by ``python -m minisod2d_amd.closure
    train`` on pairs generated by the method) on held-out pairs of other seeds
    and sizes.  Reports the gain over mu_sgs = 0; on this generator the
    least-squares constant mu_sgs is negative (the fine fields' commutation
    error is anti-diffusive on average) while softplus keeps mu_sgs >= 0, so
    the bar is that the closure does no harm: held-out loss within 0.1% of
    mu_sgs = 0 or better."""
    say("[gpu] trained closure on held-out pairs")
    w = C.load_weights()
    for n, seed in ((6, 31), (8, 32)):
        held = C.make_pair(n, seed=seed, log=say)
        ev = C.evaluate(held, w["params"])
        say(f"    n={n} seed={seed}: loss {ev['loss']:.6f} vs 4 for mu_sgs = 0, gain {ev['gain']:+.5f}")
        assert ev["used_model"]
        assert ev["loss"] <= 4.0 * (1 + 1e-3)


def test_constant_fit_on_noise_free_pairs(gpu):
    """The least-squares constant mu_sgs of a pair (closure.constant_fit, from
    Rd(0) and Rd(1) only) against the loss evaluated through the diffusion
    kernel at mu_sgs = c*, 0.5 c* and 1.5 c*: L(c) = 4 - 2c num + c^2 den is a
    parabola with minimum 4 (1 - gain) at c*.  On the noise-free Taylor-Green
    pairs the optimum is negative (anti-diffusion), so a closure, whose
    mu_sgs >= 0, cannot gain there: its best is mu_sgs = 0."""
    from minisod2d_amd import closure as cl
    say("[gpu] closure: constant fit on a noise-free 8^3 pair (k = 2)")
    pair = cl.make_pair(8, seed=0, smooth_k=2)
    c, gain = cl.constant_fit(pair)
    p = pair.coarse
    nn = p.tables.nnode

    def loss(cv):
        mu = np.full((p.nelem, nn), cv, order="F")
        _, Rd = cl.gpu_residuals(p, mu_sgs=mu)
        r = cl._mom_ener(Rd[1], Rd[2])
        return float((pair.weights * ((r - pair.target) ** 2).sum(axis=1)).sum())

    L0, Lc, Llo, Lhi = loss(0.0), loss(c), loss(0.5 * c), loss(1.5 * c)
    say(f"    c* = {c:.4e}, gain {gain:.4e}; L(0) = {L0:.12f}, L(c*) = {Lc:.12f}")
    assert abs(L0 - 4.0) <= 1e-9
    assert abs(Lc - 4.0 * (1.0 - gain)) <= 1e-9
    # parabola through the three points: L(c/2) = L(3c/2) = L(c*) + gain
    assert abs(Llo - Lc - gain) <= 1e-9 and abs(Lhi - Lc - gain) <= 1e-9
    assert c < 0.0 and 0.0 <= gain < 1e-3

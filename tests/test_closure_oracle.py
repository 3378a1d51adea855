"""Pins of the closure oracle (oracle/msod_oracle_mlp.c) and of the host side
of the closure (order transfer, weight files).  CPU only.

The network has no reference counterpart, so the oracle is pinned by
independent routines: torch's float64 autograd for the forward pass and the
parameter gradient, central finite differences, torch.optim.Adam for the
optimiser step, and torch's bfloat16 conversion for the rounding points of the
quantised forward.  A planted mistake (a dropped bias, a transposed weight)
must fail these checks; test_planted_mistakes_fail shows that it does.
"""
import json
import os

import numpy as np
import pytest
import torch

from oracle_lib import Oracle

NIN, NH = 12, 32
OW1, OB1 = 0, 384
OW2, OB2 = 416, 1440
OW3, OB3 = 1472, 1504
OMEAN, OISTD, OSCALE = 1505, 1517, 1529


@pytest.fixture(scope="module")
def orc():
    return Oracle()


def random_params(seed=0, scale=2.5e-4):
    rng = np.random.default_rng(seed)
    P = np.zeros(1530, np.float32)
    P[OW1:OB1] = rng.normal(0, 0.5, NIN * NH)
    P[OB1:OW2] = rng.normal(0, 0.2, NH)
    P[OW2:OB2] = rng.normal(0, 0.3, NH * NH)
    P[OB2:OW3] = rng.normal(0, 0.2, NH)
    P[OW3:OB3] = rng.normal(0, 0.5, NH)
    P[OB3] = -0.3
    P[OMEAN:OISTD] = rng.normal(0, 0.1, NIN)
    P[OISTD:OSCALE] = rng.uniform(0.5, 2.0, NIN)
    P[OSCALE] = scale
    return P


def random_features(n, seed=1):
    rng = np.random.default_rng(seed)
    return rng.normal(0, 1, (n, NIN)).astype(np.float32)


def torch_net(P, quant):
    """The network written with torch modules' arithmetic, float64."""
    t = lambda a: torch.tensor(np.asarray(a, np.float64), dtype=torch.float64, requires_grad=True)
    W1 = t(P[OW1:OB1].reshape(NIN, NH))
    b1 = t(P[OB1:OW2])
    W2 = t(P[OW2:OB2].reshape(NH, NH))
    b2 = t(P[OB2:OW3])
    w3 = t(P[OW3:OB3])
    b3 = t(P[OB3:OB3 + 1])
    mean = torch.tensor(P[OMEAN:OISTD], dtype=torch.float32)
    istd = torch.tensor(P[OISTD:OSCALE], dtype=torch.float32)
    scale = float(P[OSCALE])

    def q(x):  # bf16 rounding in the forward, identity in the backward
        if not quant:
            return x
        r = x.detach().to(torch.float32).to(torch.bfloat16).to(torch.float64)
        return x + (r - x).detach()

    def f(feat):
        x = ((torch.tensor(feat) - mean) * istd).to(torch.float64)
        x = q(x)
        h1 = q(torch.nn.functional.silu(x @ q(W1) + b1))   # q rounds through float32
        z3 = torch.nn.functional.silu(h1 @ q(W2) + b2) @ w3 + b3
        return z3, scale * torch.nn.functional.softplus(z3)

    return f, [W1, b1, W2, b2, w3, b3]


def test_forward_matches_torch(orc):
    P = random_params()
    F = random_features(2000)
    z3, mu = orc.mlp_forward(P, F, quant=0)
    f, _ = torch_net(P, quant=False)
    tz, tmu = f(F)
    np.testing.assert_allclose(z3, tz.detach().numpy(), rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(mu, tmu.detach().numpy(), rtol=1e-12, atol=0)


def test_quantised_forward_matches_torch_bf16(orc):
    P = random_params(3)
    F = random_features(2000, 4)
    z3, _ = orc.mlp_forward(P, F, quant=1)
    f, _ = torch_net(P, quant=True)
    tz, _ = f(F)
    # both round float32(value) to bf16 at the same three points
    np.testing.assert_allclose(z3, tz.detach().numpy(), rtol=1e-9, atol=1e-9)
    z0, _ = orc.mlp_forward(P, F, quant=0)
    d = np.abs(z3 - z0)
    assert 1e-5 < np.median(d) < 5e-2, "bf16 rounding must change z3 by about 2^-8 relative"


def test_quant_is_identity_on_bf16_values(orc):
    """With weights, hidden values and inputs exactly representable in bf16,
    quant = 1 and quant = 0 agree."""
    P = np.zeros(1530, np.float32)
    P[OW1 + 0 * NH + 0] = 0.5          # x0 -> z1[0]
    P[OW2 + 0 * NH + 0] = 1.0          # h1[0] -> z2[0]
    P[OW3 + 0] = 2.0
    P[OB1:OW2] = 0.0
    P[OISTD:OSCALE] = 1.0
    P[OSCALE] = 1.0
    F = np.zeros((1, NIN), np.float32)   # x = 0: silu(0) = 0 exactly
    a, _ = orc.mlp_forward(P, F, 1)
    b, _ = orc.mlp_forward(P, F, 0)
    assert a[0] == b[0] == 0.0


def test_gradient_matches_torch_autograd(orc):
    P = random_params(5)
    F = random_features(500, 6)
    gnu = np.random.default_rng(7).normal(0, 1, 500)
    for quant in (0, 1):
        g = orc.mlp_gradient(P, F, gnu, quant=quant)
        f, params = torch_net(P, quant=bool(quant))
        _, mu = f(F)
        (mu * torch.tensor(gnu)).sum().backward()
        tg = np.concatenate([p.grad.numpy().ravel() for p in params])
        np.testing.assert_allclose(g, tg, rtol=1e-9, atol=1e-10 * np.abs(tg).max())


def test_gradient_matches_finite_differences(orc):
    P = random_params(8)
    F = random_features(64, 9)
    gnu = np.random.default_rng(10).normal(0, 1, 64)
    g = orc.mlp_gradient(P, F, gnu, quant=0)
    L = lambda Q: float(orc.mlp_forward(Q, F, quant=0)[1] @ gnu)
    rng = np.random.default_rng(11)
    for j in list(rng.choice(1505, 40, replace=False)) + [OB3, OW1, OB1, OW2, OW3]:
        eps = 1e-2 * max(1.0, abs(float(P[j])))
        Pp, Pm = P.copy(), P.copy()
        Pp[j] += eps
        Pm[j] -= eps
        h = float(Pp[j]) - float(Pm[j])          # the step actually taken in float32
        fd = (L(Pp) - L(Pm)) / h
        assert abs(fd - g[j]) <= 1e-4 * np.abs(g).max() + 1e-3 * abs(g[j]), (j, fd, g[j])


def test_adam_matches_torch(orc):
    rng = np.random.default_rng(12)
    P = random_params(13)
    m = np.zeros(1505, np.float32)
    v = np.zeros(1505, np.float32)
    tp = torch.nn.Parameter(torch.tensor(P[:1505].copy()))
    opt = torch.optim.Adam([tp], lr=3e-3, betas=(0.9, 0.999), eps=1e-8)
    for step in range(1, 6):
        g = rng.normal(0, 1, 1505).astype(np.float32)
        orc.adam(P, g, m, v, step, 3e-3)
        tp.grad = torch.tensor(g)
        opt.step()
        np.testing.assert_allclose(P[:1505], tp.detach().numpy(), rtol=2e-6, atol=2e-7)
    assert np.all(P[1505:] == random_params(13)[1505:]), "only the trained part moves"


def test_planted_mistakes_fail(orc):
    """A dropped bias or a transposed W2 in the network changes the output far
    beyond the tolerances used above."""
    P = random_params(14)
    F = random_features(300, 15)
    z, _ = orc.mlp_forward(P, F, 0)
    Q = P.copy()
    Q[OB1:OW2] = 0.0
    assert np.abs(orc.mlp_forward(Q, F, 0)[0] - z).max() > 1e-3
    Q = P.copy()
    Q[OW2:OB2] = P[OW2:OB2].reshape(NH, NH).T.ravel()
    assert np.abs(orc.mlp_forward(Q, F, 0)[0] - z).max() > 1e-3


# ---------------------------------------------------------------------------
# host side: order transfer (fortran/msod_interp.f90) and weight files
# ---------------------------------------------------------------------------
def test_interp_1d_reproduces_polynomials():
    from minisod2d_amd.closure import interp_1d
    for pf, pt in ((7, 3), (4, 2), (3, 7)):
        M = interp_1d(pf, pt)
        assert M.shape == (pt + 1, pf + 1)
        np.testing.assert_allclose(M.sum(axis=1), 1.0, atol=1e-13)   # partition of unity
        if pt <= pf:
            # interpolating a degree-pt polynomial sampled at the fine nodes is exact
            from oracle_lib import Oracle as O
            xf, _ = O().gll(pf)
            xt, _ = O().gll(pt)
            sod = lambda x: np.concatenate([[x[0], x[-1]], x[1:-1]])
            xf, xt = sod(xf), sod(xt)
            for k in range(pt + 1):
                np.testing.assert_allclose(M @ xf ** k, xt ** k, atol=1e-12)


def _pair_meshes(n, pf, pt, jitter=0.0):
    from minisod2d_amd import box_problem
    return box_problem(n, pf, jitter=jitter), box_problem(n, pt, jitter=jitter)


def test_prolong_exact_for_polynomials():
    from minisod2d_amd.closure import prolong_fields
    fine, coarse = _pair_meshes(2, 5, 2)
    x = fine.coord
    f = np.column_stack([1 + x[:, 0] ** 2 * x[:, 1], x[:, 2] ** 5 - x[:, 0], np.ones(len(x))])
    ft = prolong_fields(fine, coarse, f)
    xc = coarse.coord
    exact = np.column_stack([1 + xc[:, 0] ** 2 * xc[:, 1], xc[:, 2] ** 5 - xc[:, 0], np.ones(len(xc))])
    np.testing.assert_allclose(ft, exact, atol=1e-12)


def test_restriction_of_lumped_mass():
    """R_f(i) = integral of phi^pf_i (the fine lumped mass) restricts to the
    integral of phi^pt_j, which the order-pt GLL rule gives exactly on affine
    elements: the restricted fine mass equals the coarse lumped mass."""
    from minisod2d_amd.closure import restrict_residual
    fine, coarse = _pair_meshes(3, 7, 3)

    def lumped(p):
        m = np.zeros(p.npoin)
        np.add.at(m, p.connec.ravel(order="F") - 1, p.gpvol[0].T.ravel(order="F"))
        return m

    rt = restrict_residual(fine, coarse, lumped(fine)[:, None])
    # exact up to the binary32 rounding of the GLL nodes and weights (as in
    # the reference's tables): about 1e-7 relative
    np.testing.assert_allclose(rt[:, 0], lumped(coarse), rtol=1e-6)
    # and the total integral is the unit cube's volume
    assert abs(rt.sum() - 1.0) < 1e-6


def test_restriction_is_prolongation_transpose():
    """<P^T r, v> = <r, P v> for a coarse field v (P: exact interpolation of
    order-pt fields on the fine nodes)."""
    from minisod2d_amd.closure import prolong_fields, restrict_residual
    fine, coarse = _pair_meshes(2, 7, 3, jitter=0.0)
    rng = np.random.default_rng(20)
    r = rng.normal(size=(fine.npoin, 1))
    v = rng.normal(size=(coarse.npoin, 1))
    Pv = prolong_fields(coarse, fine, v)   # coarse -> fine (exact embedding)
    lhs = float(restrict_residual(fine, coarse, r)[:, 0] @ v[:, 0])
    rhs = float(r[:, 0] @ Pv[:, 0])
    assert abs(lhs - rhs) <= 1e-11 * (abs(lhs) + np.abs(r).sum())


def test_weight_file_roundtrip_and_corruption(tmp_path):
    from minisod2d_amd.closure import ClosureModelError, load_weights, save_weights
    P = random_params(21)
    path = str(tmp_path / "w.npz")
    sha = save_weights(path, P, step=5, meta=dict(note="test"))
    w = load_weights(path)
    assert np.array_equal(w["params"], P) and w["step"] == 5
    assert w["manifest"]["sha256"] == sha
    with open(path, "r+b") as fh:          # flip one byte
        fh.seek(100)
        b = fh.read(1)
        fh.seek(100)
        fh.write(bytes([b[0] ^ 0xFF]))
    with pytest.raises(ClosureModelError):
        load_weights(path)
    with pytest.raises(ClosureModelError):
        load_weights(str(tmp_path / "missing.npz"))
    man = json.load(open(path + ".json"))
    man["format"] = "other"
    json.dump(man, open(path + ".json", "w"))
    with pytest.raises(ClosureModelError):
        load_weights(path)

"""Host-side Fortran library (libmsod_host.so): tables, conforming mesh,
Jacobian and fields, checked against the reference's tables, closed forms
and invariants."""
import os

import numpy as np
import pytest

import minisod2d_amd as m
from oracle_lib import Oracle, rel_l2

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.mark.parametrize("p", range(2, 9))
def test_tables_bit_exact_with_reference(p):
    t = m.hex_tables(p)
    z = np.load(os.path.join(GOLD, f"tables_p{p}.npz"))
    for k in ("invAtoIJK", "AtoIJK", "AtoI", "AtoJ", "AtoK", "xgp", "wgp", "dlxigp_ip"):
        assert np.array_equal(getattr(t, k), z[k]), k
    from golden.make_golden import digest
    assert digest(t.Ngp) == str(z["Ngp_sha"]) and digest(t.dNgp) == str(z["dNgp_sha"])


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_tables_match_oracle(p):
    t = m.hex_tables(p)
    o = Oracle().tables(p)
    for k in ("invAtoIJK", "AtoIJK", "AtoI", "AtoJ", "AtoK", "xgp", "wgp", "Ngp", "dNgp", "dlxigp_ip"):
        assert np.array_equal(getattr(t, k), o[k]), k


@pytest.mark.parametrize("n,p", [(2, 2), (3, 3), (2, 4), (1, 7)])
def test_box_mesh_is_conforming(n, p):
    prob = m.box_problem(n, p)
    nn = (p + 1) ** 3
    assert prob.npoin == (n * p + 1) ** 3 and prob.nelem == n ** 3
    c = prob.connec - 1
    assert c.min() == 0 and c.max() == prob.npoin - 1
    assert len(np.unique(c)) == prob.npoin  # every node referenced
    mult = np.bincount(c.ravel(), minlength=prob.npoin)
    # lattice position of every node -> multiplicity from the count of elements touching it
    n1 = n * p + 1
    idx = np.arange(prob.npoin)
    pos = np.stack([idx % n1, (idx // n1) % n1, idx // (n1 * n1)])
    on_grid = (pos % p == 0)
    inner = (pos > 0) & (pos < n1 - 1)
    expect = np.prod(np.where(on_grid & inner, 2, 1), axis=0)
    assert np.array_equal(mult, expect)
    # nodes sit on the GLL lattice of their element
    t = prob.tables
    h = 1.0 / n
    for e in (0, prob.nelem - 1):
        ex, ey, ez = e % n, (e // n) % n, e // (n * n)
        X = prob.coord[c[e]]
        np.testing.assert_allclose(X, h * (np.array([ex, ey, ez]) + 0.5 * (t.xgp + 1)), atol=1e-15)
    assert nn == prob.connec.shape[1]


def test_affine_jacobian_closed_form():
    n, p = 2, 4
    prob = m.box_problem(n, p)
    h = 1.0 / n
    He = prob.He
    np.testing.assert_allclose(He[0, 0], 2 / h, rtol=1e-13)
    np.testing.assert_allclose(He[1, 1], 2 / h, rtol=1e-13)
    np.testing.assert_allclose(He[2, 2], 2 / h, rtol=1e-13)
    off = He.copy()
    for i in range(3):
        off[i, i] = 0
    assert np.abs(off).max() < 1e-12
    np.testing.assert_allclose(prob.gpvol[0], np.broadcast_to(prob.tables.wgp[:, None] * (h / 2) ** 3, prob.gpvol[0].shape), rtol=1e-13)
    # binary32 weights: total volume exact to float rounding only
    assert abs(prob.gpvol.sum() - 1.0) < 2e-6


def test_jacobian_matches_reference_golden():
    prob = m.box_problem(2, 4, jitter=0.05)
    z = np.load(os.path.join(GOLD, "jac_p4.npz"))
    from golden.make_golden import digest
    assert digest(prob.coord) == str(z["coord_sha"])
    assert rel_l2(prob.He, z["He"]) < 1e-15 and rel_l2(prob.gpvol, z["gpvol"]) < 1e-15


def test_jacobian_matches_oracle_on_jittered_mesh():
    p = 5
    prob = m.box_problem(2, p, ny=3, nz=1, jitter=0.05)
    He, gp = Oracle().jacobian(p, prob.nelem, prob.npoin, prob.connec, prob.coord,
                               prob.tables.dNgp, prob.tables.wgp)
    # the C oracle may contract to FMA: equal to rounding, not bitwise
    assert rel_l2(He, prob.He) < 1e-14 and rel_l2(gp, prob.gpvol) < 1e-14


def test_jitter_moves_only_interior_nodes_within_bound():
    p, n = 4, 2
    a = m.box_problem(n, p)
    b = m.box_problem(n, p, jitter=0.05)
    moved = np.any(a.coord != b.coord, axis=1)
    mult = np.bincount((a.connec - 1).ravel(), minlength=a.npoin)
    assert moved.sum() == n ** 3 * (p - 1) ** 3
    assert np.all(mult[moved] == 1)
    xi, _ = Oracle().gll(p)
    gap = np.diff(xi).max() * 0.5 / n  # largest local spacing bounds every move
    assert np.abs(a.coord - b.coord).max() <= 0.05 * gap
    assert np.all(b.gpvol > 0)
    # He is no longer affine
    assert np.abs(b.He[0, 1]).max() > 1e-4


def test_fields_relations_and_determinism():
    prob = m.box_problem(3, 3, threads=1)
    g = m.GAMMA
    np.testing.assert_allclose(prob.q, prob.rho[:, None] * prob.u, rtol=1e-15)
    ke = 0.5 * prob.rho * (prob.u ** 2).sum(axis=1)
    np.testing.assert_allclose(prob.E, prob.pr / (g - 1) + ke, rtol=1e-14)
    np.testing.assert_allclose(prob.Tem, prob.pr / (prob.rho * m.RGAS), rtol=1e-15)
    x = prob.coord[:, 0]
    np.testing.assert_allclose(prob.pr, 1 / g + 0.1 * np.cos(4 * np.pi * x), atol=1e-15)
    assert abs(prob.rho.mean() - 1.0) < 0.01 and abs(prob.rho.std() - 0.05) < 0.01
    assert prob.mu_e.min() >= 0 and prob.mu_e.max() < 1e-3 and prob.mu_sgs.max() < 1e-4
    other = m.box_problem(3, 3, threads=8)
    for k in ("rho", "E", "mu_e", "mu_sgs", "He"):
        assert np.array_equal(getattr(prob, k), getattr(other, k)), k


def test_rho_jump():
    a = m.box_problem(2, 3)
    b = m.box_problem(2, 3, rho_jump=0.1)
    right = a.coord[:, 0] > 0.5
    np.testing.assert_allclose(b.rho[right], 1.1 * a.rho[right], rtol=1e-15)
    assert np.array_equal(b.rho[~right], a.rho[~right])


def test_reference_driver_case():
    prob = m.reference_problem(3, 4)
    assert prob.npoin == 3 * 125
    assert np.array_equal(prob.connec.ravel(order="F") - 1,
                          np.arange(3 * 125).reshape(3, 125).ravel(order="F"))
    np.testing.assert_allclose(prob.He[0, 0], 1.0, rtol=1e-14)  # copies of [-1,1]^3: He = I
    np.testing.assert_allclose(prob.gpvol[0, :, 0], prob.tables.wgp, rtol=1e-14)

"""The TET04 part of the CPU oracle (msod_oracle_tet_*, msod_oracle_fem_*),
pinned before it is trusted: the reference's own outputs (golden fixtures
from tests/golden/make_golden.py, and the live reference libraries where
built), closed forms of P1 elements, and planted mistakes that must fail.
Also the host side of the product: the TET04 tables and the tetrahedral mesh
of minisod2d_amd.tet_problem against the reference's."""
import json
import os

import numpy as np
import pytest

import minisod2d_amd as m
from oracle_lib import Oracle, RefTet, ref_available, rel_l2

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TET_CASES = sorted(f[:-4] for f in os.listdir(GOLD) if f.startswith("tet_case_"))
KEYS = ("Rmass", "Rmom", "Rener", "Dmass", "Dmom", "Dener")


@pytest.fixture(scope="module")
def oracle():
    return Oracle()


def load(name):
    z = np.load(os.path.join(GOLD, name + ".npz"))
    meta = json.loads(str(z["meta"])) if "meta" in z.files else {}
    return {k: z[k] for k in z.files if k != "meta"}, meta


def build(name):
    from golden.make_golden import TET_CASES as B, tet_input_digests
    fn, c_rho, c_ener = B[name]
    tp = fn()
    return tp, c_rho, c_ener, tet_input_digests(tp)


def residuals(oracle, tp, c_rho=1.0, c_ener=1.0):
    return oracle.fem_convec(tp) + oracle.fem_diffu(tp, c_rho, c_ener)


# ---------------------------------------------------------------------------
# tables and mesh
# ---------------------------------------------------------------------------
def test_tet_tables_match_reference(oracle):
    gold, _ = load("tet_tables")
    o = oracle.tet_tables()
    h = m.tet_tables()
    for k in ("xgp", "wgp", "Ngp", "dNgp"):
        assert np.array_equal(o[k], gold[k]), k
        assert np.array_equal(getattr(h, k), gold[k]), k


def test_tet_tables_closed_form():
    t = m.tet_tables()
    # P1 basis: partition of unity at every point, constant gradients summing to zero
    assert np.allclose(t.Ngp.sum(axis=1), 1.0, atol=1e-15)
    assert np.array_equal(t.dNgp[:, 0, 0], [-1, -1, -1]) and np.all(t.dNgp.sum(axis=1) == 0)
    # the 4-point rule integrates linear functions exactly over the reference tet (weights sum 1)
    assert t.wgp.sum() == 1.0
    assert np.allclose(t.wgp @ t.xgp, 0.25, atol=1e-15)
    # rows of Ngp are its columns permuted, to rounding (1 - a - 2b is b only
    # in exact arithmetic): with the constant He of a P1 element, weighting
    # node n by Ngp(n,g), as fem_diffu does, permutes the points' terms
    perm = [2, 3, 0, 1]
    assert np.allclose(t.Ngp, t.Ngp.T[perm], rtol=0, atol=1e-15)


def test_tet_mesh_matches_reference_generator():
    from golden.make_golden import digest
    gold, _ = load("tet_mesh")
    nelem, npoin = int(gold["nelem"]), int(gold["npoin"])
    H = m.host_lib()
    connec = np.zeros((nelem, 4), np.int32, order="F")
    coord = np.zeros((npoin, 3), order="F")
    H.msodh_disjoint_mesh(nelem, npoin, 4, np.asfortranarray(m.TET_BASE).ctypes.data,
                          connec.ctypes.data, coord.ctypes.data)
    assert digest(connec) == str(gold["connec_sha"])
    assert digest(coord[:4 * nelem]) == str(gold["coord_sha"])


def test_tet_problem_requires_divisible_npoin():
    p = m.box_problem(3, 2)  # 343 nodes
    with pytest.raises(ValueError):
        m.tet_problem(p)
    tp = m.tet_problem(p, strict=False)
    assert tp.nelem == 85 and tp.connec.max() == 340


# ---------------------------------------------------------------------------
# geometry
# ---------------------------------------------------------------------------
def test_tet_jacobian_matches_reference_golden(oracle):
    from golden.make_golden import digest
    gold, _ = load("tet_jac")
    tp = m.tet_problem(m.box_problem(3, 2), jitter=0.4, strict=False)
    assert digest(tp.coord) == str(gold["coord_sha"]) and digest(tp.connec) == str(gold["connec_sha"])
    t = tp.tables
    He, gp = oracle.tet_jacobian(tp.nelem, tp.npoin, tp.connec, tp.coord, t.dNgp, t.wgp)
    assert rel_l2(He, gold["He"]) < 1e-15 and rel_l2(gp, gold["gpvol"]) < 1e-15
    # the host library's routine (the product's host path) as well
    assert rel_l2(tp.He, gold["He"]) < 1e-15 and rel_l2(tp.gpvol, gold["gpvol"]) < 1e-15


def test_tet_jacobian_closed_forms(oracle):
    t = m.tet_tables()
    # an affine map x = x0 + A xi: He = A^{-T}... in the reference's convention
    # He(i,k) = d xi_k / d x_i, gpvol = det(A) / 4 at every point
    rng = np.random.default_rng(3)
    nelem = 20
    A = np.eye(3) + 0.3 * rng.standard_normal((nelem, 3, 3))
    A[np.linalg.det(A) < 0] *= -1
    x0 = rng.standard_normal((nelem, 3))
    coord = np.zeros((4 * nelem, 3), order="F")
    for e in range(nelem):
        for a, xi in enumerate(m.TET_BASE):
            coord[4 * e + a] = x0[e] + A[e] @ xi
    connec = np.asfortranarray(np.arange(1, 4 * nelem + 1, dtype=np.int32).reshape(nelem, 4))
    He, gp = oracle.tet_jacobian(nelem, 4 * nelem, connec, coord, t.dNgp, t.wgp)
    for e in range(nelem):
        Ainv = np.linalg.inv(A[e])
        for g in range(4):
            # d xi_k / d x_i = Ainv[k, i]
            assert np.allclose(He[:, :, g, e], Ainv.T, rtol=1e-12, atol=1e-13)
            assert gp[0, g, e] == pytest.approx(np.linalg.det(A[e]) / 4, rel=1e-13)


# ---------------------------------------------------------------------------
# kernels against the reference's outputs
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("name", TET_CASES)
def test_oracle_matches_reference_golden(oracle, name):
    gold, meta = load(name)
    tp, c_rho, c_ener, dig = build(name)
    assert dig == meta["inputs"], "generator output changed: fixtures no longer describe these inputs"
    for key, got in zip(KEYS, residuals(oracle, tp, c_rho, c_ener)):
        assert rel_l2(got, gold[key]) < 1e-14, key


def test_oracle_matches_live_reference(oracle):
    if not ref_available(4):
        pytest.skip("reference library not built (make -C oracle ref needs /root/reference)")
    ref = RefTet(4)
    tp = m.tet_problem(m.box_problem(4, 2, jitter=0.05, rho_jump=0.2, seeds=(31, 32, 33, 34)),
                       jitter=0.25, random_mu=True, seed=35, strict=False)
    want = ref.fem_convec(tp) + ref.fem_diffu(tp, 1.7, 0.4)
    for key, got, w in zip(KEYS, residuals(oracle, tp, 1.7, 0.4), want):
        assert rel_l2(got, w) < 1e-14, key


# ---------------------------------------------------------------------------
# closed forms of P1 elements
# ---------------------------------------------------------------------------
def fields_problem(nelem, rho, u, Tem=None, jitter=0.3, seed=7):
    """A disjoint tetrahedral mesh with nodal fields rho(x), u(x) given as functions."""
    npoin = 4 * nelem
    base = m.box_problem(1, 1)  # tables, placeholders
    H = m.host_lib()
    t = m.tet_tables()
    connec = np.zeros((nelem, 4), np.int32, order="F")
    coord = np.zeros((npoin, 3), order="F")
    H.msodh_disjoint_mesh(nelem, npoin, 4, np.asfortranarray(m.TET_BASE).ctypes.data,
                          connec.ctypes.data, coord.ctypes.data)
    coord += jitter * (np.random.default_rng(seed).random((npoin, 3)) - 0.5)
    He = np.zeros((3, 3, 4, nelem), order="F")
    gpvol = np.zeros((1, 4, nelem), order="F")
    assert H.msodh_elem_jacobian(nelem, npoin, 4, connec.ctypes.data, coord.ctypes.data,
                                 t.dNgp.ctypes.data, t.wgp.ctypes.data, gpvol.ctypes.data,
                                 He.ctypes.data) == 0
    r = np.asfortranarray(rho(coord))
    uu = np.asfortranarray(u(coord))
    q = np.asfortranarray(r[:, None] * uu)
    pr = np.ones(npoin)
    E = 2.5 * pr + 0.5 * (q * uu).sum(axis=1)
    T = np.ones(npoin) if Tem is None else np.asfortranarray(Tem(coord))
    mf = np.ones(npoin)
    zero = np.zeros((nelem, 4), order="F")
    return m.TetProblem(t, nelem, npoin, connec, coord, He, gpvol, uu, q, r, pr, E, T, mf,
                        zero.copy(), zero.copy(), Cp=base.Cp, Pr=base.Pr)


def test_uniform_state_has_zero_residuals(oracle):
    tp = fields_problem(30, lambda x: np.full(len(x), 1.3), lambda x: np.tile([0.2, -0.1, 0.4], (len(x), 1)))
    # rounding scales with the largest coefficient: kappa = Cp/Pr in Dener
    for key, got in zip(KEYS, residuals(oracle, tp)):
        scale = tp.Cp / tp.Pr if key == "Dener" else 1.0
        assert np.abs(got).max() < 1e-13 * scale, key


def test_linear_density_convection_closed_form(oracle):
    # rho = 1 + g.x, u constant: gradients are exact, divU = 0, q = rho u, and
    # Rmass(n) = sum_g gpvol Ngp(g,n) (u.g) = (u.g) det(J)/4 per node
    # (gpvol = det(J)/4 at each point, sum_g Ngp(g,n) = 1)
    gvec = np.array([0.3, -0.2, 0.5])
    uc = np.array([0.7, 0.1, -0.4])
    tp = fields_problem(25, lambda x: 1.0 + x @ gvec, lambda x: np.tile(uc, (len(x), 1)))
    Rm = oracle.fem_convec(tp)[0]
    det = np.repeat(tp.gpvol[0].sum(axis=0), 4)  # det(J) per element, per node
    assert np.allclose(Rm, (uc @ gvec) * det / 4, rtol=1e-12, atol=1e-15)


def test_linear_velocity_diffusion_closed_form(oracle):
    # u = B x, mu_fluid = 1, mu_e = mu_sgs = 0: tau = B + B^T - 2/3 tr(B) I (the
    # reference's gradU(i,j) = d u_j / d x_i, so tau is symmetric either way),
    # Dmom(n,i) = sum_g gpvol sum_j dN_n/dx_j tau(i,j) = det(J) (tau grad N_n)_i
    B = np.array([[0.2, -0.3, 0.1], [0.4, 0.05, -0.2], [0.0, 0.3, -0.1]])
    tp = fields_problem(15, lambda x: np.full(len(x), 1.1), lambda x: x @ B.T)
    Dmo = oracle.fem_diffu(tp)[1]
    tau = B + B.T - 2.0 / 3.0 * np.trace(B) * np.eye(3)
    for e in range(tp.nelem):
        det = tp.gpvol[0, :, e].sum()
        Hm = tp.He[:, :, 0, e]  # He(i,k) = d xi_k / d x_i
        for n in range(4):
            gradN = Hm @ tp.tables.dNgp[:, n, 0]
            assert np.allclose(Dmo[4 * e + n], det * tau @ gradN, rtol=1e-11, atol=1e-13)
    # and each element's momentum residuals sum to zero (sum_n grad N_n = 0)
    assert np.abs(Dmo.reshape(tp.nelem, 4, 3).sum(axis=1)).max() < 1e-13


def test_linear_temperature_diffusion_closed_form(oracle):
    # T = t.x, u = 0, mu_fluid = 1: Dener(n) = det(J) kappa grad N_n . t with
    # kappa = Cp/Pr
    tv = np.array([1.5, -0.5, 0.25])
    tp = fields_problem(15, lambda x: np.full(len(x), 1.0), lambda x: np.zeros((len(x), 3)),
                        Tem=lambda x: x @ tv)
    De = oracle.fem_diffu(tp)[2]
    kappa = tp.Cp / tp.Pr
    for e in range(tp.nelem):
        det = tp.gpvol[0, :, e].sum()
        Hm = tp.He[:, :, 0, e]
        for n in range(4):
            gradN = Hm @ tp.tables.dNgp[:, n, 0]
            assert De[4 * e + n] == pytest.approx(det * kappa * gradN @ tv, rel=1e-11, abs=1e-12)


# ---------------------------------------------------------------------------
# planted mistakes: plausible errors must fail the checks above
# ---------------------------------------------------------------------------
def test_planted_mistakes_are_caught(oracle):
    name = "tet_case_box3_p2_jit"
    gold, _ = load(name)
    tp, c_rho, c_ener, _ = build(name)
    # transposed operand: He(i,k) read as He(k,i)
    bad = build(name)[0]
    bad.He = np.asfortranarray(np.swapaxes(tp.He, 0, 1))
    assert rel_l2(oracle.fem_convec(bad)[1], gold["Rmom"]) > 1e-3
    assert rel_l2(oracle.fem_diffu(bad)[1], gold["Dmom"]) > 1e-3
    # dropped term: c_rho = 0 removes the mass diffusion, c_ener = 0 part of kappa
    assert rel_l2(oracle.fem_diffu(tp, 0.0, c_ener)[0], gold["Dmass"]) > 0.5
    assert rel_l2(oracle.fem_diffu(tp, c_rho, 0.0)[2], gold["Dener"]) > 1e-6
    # wrong index: nodes 2 and 3 of every element swapped in connec
    bad2 = build(name)[0]
    bad2.connec = np.asfortranarray(tp.connec[:, [0, 2, 1, 3]])
    assert rel_l2(oracle.fem_convec(bad2)[0], gold["Rmass"]) > 1e-2
    # wrong sign of the pressure gradient: p -> -p changes the momentum residual
    bad3 = build(name)[0]
    bad3.pr = -tp.pr
    bad3.E = tp.E + 2 * tp.pr  # keep E + p
    assert rel_l2(oracle.fem_convec(bad3)[1], gold["Rmom"]) > 1e-3
    # and the unmodified oracle still matches
    assert rel_l2(oracle.fem_convec(tp)[0], gold["Rmass"]) < 1e-14

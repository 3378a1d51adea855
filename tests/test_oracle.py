"""The CPU oracle (oracle/msod_oracle.c) pinned before it is trusted:
closed forms, polynomial exactness, conservation, the reference's own
kernels (golden fixtures from tests/golden/make_golden.py and, where built,
the live reference libraries), and planted mistakes that must be caught."""
import json
import os

import numpy as np
import pytest

import minisod2d_amd as m
from oracle_lib import Oracle, RefLib, ref_available, rel_l2

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = sorted(f[:-4] for f in os.listdir(GOLD) if f.startswith("case_"))


@pytest.fixture(scope="module")
def oracle():
    return Oracle()


def build_case(name):
    from golden.make_golden import CASES as BUILDERS
    return BUILDERS[name]()


def load_case(name):
    z = np.load(os.path.join(GOLD, name + ".npz"))
    return {k: z[k] for k in z.files if k != "meta"}, json.loads(str(z["meta"]))


def check_inputs(prob, meta):
    from golden.make_golden import input_digests
    got = input_digests(prob)
    bad = [k for k, v in meta["inputs"].items() if got[k] != v]
    assert not bad, f"generator output changed for {bad}: fixtures no longer describe these inputs"


# ---------------------------------------------------------------------------
# GLL nodes/weights and the 1-D derivative matrix
# ---------------------------------------------------------------------------
def f32(x):
    return float(np.float32(x))


@pytest.mark.parametrize("p,nodes,weights", [
    (2, [-1, 0, 1], [1 / 3, 4 / 3, 1 / 3]),
    (3, [-1, -1 / np.sqrt(5), 1 / np.sqrt(5), 1], [1 / 6, 5 / 6, 5 / 6, 1 / 6]),
    (4, [-1, -np.sqrt(3 / 7), 0, np.sqrt(3 / 7), 1], [1 / 10, 49 / 90, 32 / 45, 49 / 90, 1 / 10]),
])
def test_gll_closed_forms(oracle, p, nodes, weights):
    xi, w = oracle.gll(p)
    # the reference rounds through binary32 (mod_maths.f90:25,39-40)
    assert list(xi) == [f32(v) for v in nodes]
    assert list(w) == [f32(v) for v in weights]


@pytest.mark.parametrize("p", range(1, 9))
def test_gll_against_legendre(oracle, p):
    xi, w = oracle.gll(p, exact=True)
    # interior nodes: roots of P_p'; weights 2/(p(p+1) P_p(x)^2)
    Pp = np.polynomial.legendre.Legendre.basis(p)
    if p > 1:
        np.testing.assert_allclose(xi[1:-1], np.sort(Pp.deriv().roots()), atol=1e-14)
    np.testing.assert_allclose(w, 2.0 / (p * (p + 1) * Pp(xi) ** 2), rtol=1e-13)
    assert abs(w.sum() - 2.0) < 1e-13
    xr, wr = oracle.gll(p)
    assert np.array_equal(xr, xi.astype(np.float32).astype(np.float64))
    assert np.array_equal(wr, w.astype(np.float32).astype(np.float64))


@pytest.mark.parametrize("p", range(2, 9))
def test_oracle_tables_match_reference_golden(oracle, p):
    t = oracle.tables(p)
    z = np.load(os.path.join(GOLD, f"tables_p{p}.npz"))
    for k in ("gmshIJK", "invAtoIJK", "AtoIJK", "AtoI", "AtoJ", "AtoK", "xgp", "wgp", "dlxigp_ip"):
        assert np.array_equal(t[k], z[k]), k


@pytest.mark.parametrize("p", range(1, 9))
def test_derivative_matrix_exact_on_polynomials(oracle, p):
    t = oracle.tables(p)
    q = p + 1
    inv, dl = t["invAtoIJK"], t["dlxigp_ip"]
    # D[i][n] = dlxigp_ip(node on the xi-line (i,1,1), 1, n); nodes x_i = xgp of that node
    nodes = [inv[i, 0, 0] - 1 for i in range(q)]
    D = np.array([[dl[a, 0, n] for n in range(q)] for a in nodes])
    x = np.array([t["xgp"][a, 0] for a in nodes])
    for mdeg in range(p + 1):
        np.testing.assert_allclose(D @ x ** mdeg, mdeg * x ** max(mdeg - 1, 0) if mdeg else 0 * x,
                                   atol=1e-11 * (1 + mdeg))
    # every node and direction carries the row of its own 1-D index
    for a in range(q ** 3):
        idx = (t["AtoI"][a] - 1, t["AtoJ"][a] - 1, t["AtoK"][a] - 1)
        for d in range(3):
            assert np.array_equal(dl[a, d, :], D[idx[d]])


@pytest.mark.parametrize("p", range(1, 9))
def test_index_maps_are_consistent(oracle, p):
    t = oracle.tables(p)
    q = p + 1
    inv = t["invAtoIJK"]
    assert sorted(inv.ravel(order="F")) == list(range(1, q ** 3 + 1))
    for a in range(q ** 3):
        assert inv[t["AtoI"][a] - 1, t["AtoJ"][a] - 1, t["AtoK"][a] - 1] == a + 1
    # Gmsh corner order (mod_gmsh_indices.f90:19-30, SOD 1-D index 1 <-> xi=-1, 2 <-> +1)
    corners = [(1, 1, 1), (2, 1, 1), (2, 2, 1), (1, 2, 1), (1, 1, 2), (2, 1, 2), (2, 2, 2), (1, 2, 2)]
    for a, c in enumerate(corners):
        assert (t["AtoI"][a], t["AtoJ"][a], t["AtoK"][a]) == c


# ---------------------------------------------------------------------------
# kernels against the reference's outputs
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("name", CASES)
def test_oracle_matches_reference_golden(oracle, name):
    gold, meta = load_case(name)
    prob = build_case(name)
    check_inputs(prob, meta)
    Rm, Rmo, Re = oracle.convec(prob)
    Dm, Dmo, De = oracle.diffu(prob)
    for got, key in ((Rm, "Rmass"), (Rmo, "Rmom"), (Re, "Rener"), (Dm, "Dmass"), (Dmo, "Dmom"), (De, "Dener")):
        err = rel_l2(got, gold[key])
        assert err < 1e-13, (key, err)


@pytest.mark.parametrize("p", range(2, 9))
def test_oracle_matches_live_reference(oracle, p):
    if not ref_available(p):
        pytest.skip("reference library not built (make -C oracle ref needs /root/reference)")
    ref = RefLib(p)
    prob = m.box_problem(2, p, ny=1, nz=2, jitter=0.05, seeds=(11, 12, 13, 14))
    Rm, Rmo, Re, split = ref.convec(prob)
    Dm, Dmo, De = ref.diffu(prob, c_rho=0.7, c_ener=1.3)
    om, omo, oe = oracle.convec(prob)
    dm, dmo, de = oracle.diffu(prob, c_rho=0.7, c_ener=1.3)
    for a, b in ((om, Rm), (omo, Rmo), (oe, Re), (dm, Dm), (dmo, Dmo), (de, De)):
        assert rel_l2(a, b) < 1e-13
    # fused and split reference forms agree on Rmass/Rmom
    assert rel_l2(split[0], Rm) < 1e-13 and rel_l2(split[1], Rmo) < 1e-13


def test_oracle_threaded_equals_sequential(oracle):
    prob = m.box_problem(4, 3, jitter=0.05)
    a = oracle.convec(prob, 1) + oracle.diffu(prob, 1)
    b = oracle.convec(prob, 4) + oracle.diffu(prob, 4)
    for x, y in zip(a, b):
        assert rel_l2(x, y) < 1e-14


# ---------------------------------------------------------------------------
# properties fixed by the mathematics
# ---------------------------------------------------------------------------
def constant_state(prob):
    prob.u[:] = [0.3, -0.2, 0.1]
    prob.rho[:] = 1.2
    prob.pr[:] = 0.8
    prob.q[:] = prob.rho[:, None] * prob.u
    prob.E[:] = 2.5
    prob.Tem[:] = 0.01
    return prob


@pytest.mark.parametrize("p", [2, 4, 7])
def test_constant_state_has_zero_residual(oracle, p):
    prob = constant_state(m.box_problem(2, p, jitter=0.05))
    for r in oracle.convec(prob) + oracle.diffu(prob):
        assert np.abs(r).max() < 1e-11


@pytest.mark.parametrize("p", [2, 3, 5])
def test_linear_velocity_mass_residual_is_rho_divu(oracle, p):
    # affine mesh, constant rho and p, linear u: Rmass = gpvol*rho*div(u) exactly
    prob = m.box_problem(3, p, ny=2, nz=2)
    x, y, z = prob.coord.T
    prob.u[:, 0], prob.u[:, 1], prob.u[:, 2] = 0.5 + 2.0 * x - y, 0.3 * z + 1.5 * y, -0.7 * z + x
    prob.rho[:] = 1.3
    prob.q[:] = prob.rho[:, None] * prob.u
    prob.pr[:] = 1.0
    divu = 2.0 + 1.5 - 0.7
    Rm, _, _ = oracle.convec(prob)
    # nodal sum over elements of gpvol: assemble gpvol the same way
    vol = np.zeros(prob.npoin)
    np.add.at(vol, (prob.connec - 1).ravel(order="F"), prob.gpvol[0].T.ravel(order="F"))
    np.testing.assert_allclose(Rm, 1.3 * divu * vol, rtol=1e-11, atol=1e-14)


@pytest.mark.parametrize("p", [2, 4, 7])
def test_weak_divergence_sums_to_zero(oracle, p):
    # sum over all nodes of the transposed sweep vanishes (columns of D sum to 0)
    prob = m.box_problem(2, p, jitter=0.05)
    _, Dmo, _ = oracle.diffu(prob)
    for r in range(3):
        assert abs(Dmo[:, r].sum()) < 1e-12 * np.abs(Dmo[:, r]).sum()


# ---------------------------------------------------------------------------
# planted mistakes: plausible errors must fail the checks above
# ---------------------------------------------------------------------------
def test_planted_mistakes_are_caught(oracle):
    name = "case_box3x2x2_p4_jit"
    gold, meta = load_case(name)
    prob = build_case(name)
    # transposed operand: He(i,k) read as He(k,i)
    bad = build_case(name)
    bad.He[:] = np.asfortranarray(np.swapaxes(bad.He, 0, 1))
    assert rel_l2(oracle.convec(bad)[1], gold["Rmom"]) > 1e-3
    assert rel_l2(oracle.diffu(bad)[1], gold["Dmom"]) > 1e-3
    # dropped term: c_rho = 0 removes the mass diffusion
    assert rel_l2(oracle.diffu(prob, c_rho=0.0)[0], gold["Dmass"]) > 0.5
    # wrong index: a permuted 1-D derivative row (swap two SOD nodes)
    bad2 = build_case(name)
    dl = bad2.tables.dlxigp_ip
    dl[:, :, [0, 1]] = dl[:, :, [1, 0]]
    assert rel_l2(oracle.convec(bad2)[0], gold["Rmass"]) > 1e-2
    # the oracle itself still matches on the unmodified case
    assert rel_l2(oracle.convec(prob)[0], gold["Rmass"]) < 1e-13

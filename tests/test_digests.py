"""The oracle against the reference's own residuals at the full-size
configurations (64^3 p=4 jittered, 48^3 p=7 with the density jump, 96^3 p=3
jittered), through digest fixtures written by tests/golden/make_golden.py
from the compiled reference: the values at 4096 seeded nodes, where the
oracle runs on exactly the elements touching those nodes."""
import json
import os

import numpy as np
import pytest

import minisod2d_amd as m
from oracle_lib import Oracle, rel_l2

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DIGESTS = sorted(f[:-4] for f in os.listdir(GOLD) if f.startswith("digest_"))


def load_digest(name):
    z = np.load(os.path.join(GOLD, name + ".npz"))
    return {k: z[k] for k in z.files if k != "meta"}, json.loads(str(z["meta"]))


def sampled_oracle(prob, nodes0):
    """Oracle residuals (10 components) at the 0-based nodes, from the elements touching them."""
    mark = np.zeros(prob.npoin + 1, dtype=bool)
    mark[nodes0 + 1] = True
    els = np.nonzero(mark[prob.connec].any(axis=1))[0]
    sub_conn = prob.connec[els]
    uniq, inv = np.unique(sub_conn, return_inverse=True)
    f = np.asfortranarray
    sub = m.Problem(prob.tables, len(els), len(uniq),
                    f((inv.reshape(sub_conn.shape) + 1).astype(np.int32)), prob.coord[uniq - 1],
                    f(prob.He[:, :, :, els]), f(prob.gpvol[:, :, els]),
                    *[f(getattr(prob, k)[uniq - 1]) for k in
                      ("u", "q", "rho", "pr", "E", "Tem", "mu_fluid")],
                    f(prob.mu_e[els]), f(prob.mu_sgs[els]), Cp=prob.Cp, Pr=prob.Pr, gamma=prob.gamma)
    o = Oracle()
    out = o.convec(sub, 8) + o.diffu(sub, 8)
    pos = np.searchsorted(uniq, nodes0 + 1)
    cols = []
    for a in out:
        a = np.asarray(a).reshape(len(uniq), -1, order="F")
        cols.extend(a[pos, j] for j in range(a.shape[1]))
    return np.stack(cols, axis=1)


@pytest.mark.parametrize("name", DIGESTS)
def test_oracle_matches_reference_digest(name):
    from golden.make_golden import DIGEST_CASES, DIGEST_GROUPS, digest_nodes, input_digests
    dg, meta = load_digest(name)
    prob = DIGEST_CASES[name]()
    assert input_digests(prob) == meta["inputs"], "generator output changed"
    assert np.array_equal(dg["nodes"], digest_nodes(prob.npoin))
    got = sampled_oracle(prob, dg["nodes"])
    for key, cols in DIGEST_GROUPS:
        err = rel_l2(got[:, cols], dg["values"][:, cols])
        assert err <= 1e-12, (key, err)

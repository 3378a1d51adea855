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
    o = Oracle()
    want = o.convec(sub, 8) + o.diffu(sub, 8)
    pos = np.searchsorted(uniq, nodes)
    return [w[pos] for w in want]


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

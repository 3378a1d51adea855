"""GPU test of msod_closure_gradient_noalloc, the entry point that allocates
no device memory (its per-workgroup partial sums live in a fixed library
array), against msod_closure_gradient, which grows a library buffer, through
the C ABI.

Both run the same kernels over the same number of workgroups, and every
workgroup sums its samples in a fixed order, so the gradients are compared
bit for bit; the first call is also checked against the closure oracle
(bf16 rounding of the back-propagated deltas: 2e-2 relative L2, cosine
>= 0.999, as for msod_closure_gradient).  Two calls in a row, the second
with fewer samples (fewer workgroups: rows left by the first call stay in
the array and must not be read).
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


def params(rng, F):
    P = np.zeros(m.CLOSURE_NPARAM, np.float32)
    P[C.OW1:C.OB1] = rng.normal(0, 0.5, 12 * 32)
    P[C.OB1:C.OW2] = rng.normal(0, 0.2, 32)
    P[C.OW2:C.OB2] = rng.normal(0, 0.3, 32 * 32)
    P[C.OB2:C.OW3] = rng.normal(0, 0.2, 32)
    P[C.OW3:C.OB3] = rng.normal(0, 0.5, 32)
    P[C.OB3] = -0.3
    P[C.OMEAN:C.OISTD] = F.mean(0)
    P[C.OISTD:C.OSCALE] = 1.0 / F.std(0)
    P[C.OSCALE] = 1e-4
    return P


def both(L, nsamp, P, F, gnu):
    """Gradients from the allocating and the non-allocating entry point."""
    g_ref = np.zeros(m.CLOSURE_NTRAIN, np.float32)
    g_new = np.full(m.CLOSURE_NTRAIN, np.nan, np.float32)
    m.copyin(P, F, gnu)
    m.create(g_ref, g_new)
    try:
        st = L.msod_closure_gradient(nsamp, P.ctypes.data, F.ctypes.data, gnu.ctypes.data,
                                     g_ref.ctypes.data)
        assert st == 0, L.msod_last_error()
        st = L.msod_closure_gradient_noalloc(nsamp, P.ctypes.data, F.ctypes.data, gnu.ctypes.data,
                                             g_new.ctypes.data)
        assert st == 0, L.msod_last_error()
        m.update_host(g_ref, g_new)
    finally:
        m.delete(P, F, gnu, g_ref, g_new)
    return g_ref, g_new


def test_noalloc_gradient_equals_allocating_entry_point(gpu):
    L = m.hip_lib()
    rng = np.random.default_rng(21)
    # 100000 samples: 1563 chunks of 64 over the 1024 workgroups of the array
    # (the cap), then 3000 samples: 47 workgroups
    F1 = rng.normal(0.0, 1.0, (100000, 12)).astype(np.float32)
    P = params(rng, F1)
    gnu1 = rng.uniform(0.5, 1.5, 100000).astype(np.float32)
    say("[gpu] closure gradient noalloc: call 1, nsamples=100000")
    r1, n1 = both(L, 100000, P, F1, gnu1)
    assert np.array_equal(r1, n1), (rel_l2(n1, r1), int(np.sum(r1 != n1)))
    ref = Oracle().mlp_gradient(P, F1, gnu1.astype(np.float64), quant=1)
    cos = float(n1 @ ref / (np.linalg.norm(n1) * np.linalg.norm(ref)))
    say(f"    equal bit for bit; against the oracle: rel L2 {rel_l2(n1, ref):.2e}, cosine {cos:.6f}")
    assert rel_l2(n1, ref) <= 2e-2 and cos >= 0.999

    F2 = rng.normal(0.3, 2.0, (3000, 12)).astype(np.float32)
    gnu2 = rng.uniform(-1.0, 1.0, 3000).astype(np.float32)
    say("[gpu] closure gradient noalloc: call 2, nsamples=3000")
    r2, n2 = both(L, 3000, P, F2, gnu2)
    assert np.array_equal(r2, n2), (rel_l2(n2, r2), int(np.sum(r2 != n2)))
    assert not np.allclose(n2, n1)  # a different gradient, not the first call's rows
    say("    equal bit for bit")


def test_noalloc_gradient_refuses_host_weights(gpu):
    """A weight block in plain host memory would need a device copy, which
    this entry point cannot make: it is refused, nothing runs."""
    L = m.hip_lib()
    rng = np.random.default_rng(1)
    F = rng.normal(0.0, 1.0, (64, 12)).astype(np.float32)
    P = params(rng, F)
    gnu = np.ones(64, np.float32)
    grad = np.zeros(m.CLOSURE_NTRAIN, np.float32)
    m.copyin(F, gnu)
    m.create(grad)
    try:
        st = L.msod_closure_gradient_noalloc(64, P.ctypes.data, F.ctypes.data, gnu.ctypes.data,
                                             grad.ctypes.data)
    finally:
        m.delete(F, gnu, grad)
    assert st == m.MSOD_ERR_NOT_PRESENT, (st, L.msod_last_error())
    assert b"present or device memory" in L.msod_last_error()

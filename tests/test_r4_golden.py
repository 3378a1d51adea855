"""CPU pins of the rp=4 fixtures (tests/golden/r4_*.npz): the reference
compiled at its default real kind rp=4 stays within float32 rounding of the
FP64 oracle on the same inputs (2e-6 relative L2 per residual, measured
2e-7..7e-7), and, where oracle/_ref/r4 is built, the live library
reproduces the fixtures bit for bit.  The bar of the GPU drop-in
(tests/test_gpu_r4.py) is stated from these distances."""
import numpy as np
import pytest

from oracle_lib import RefLib4
from r4_common import OUTS, R4_CASES, distances, load_case, oracle64
import os


@pytest.mark.parametrize("name", list(R4_CASES))
def test_r4_fixture_within_float_rounding_of_fp64_oracle(name):
    c, ref, prob = load_case(name)
    want = oracle64(c, prob)
    for k, (_, dref, _) in zip(OUTS, distances(ref, ref, want)):
        assert dref <= 2e-6, (name, k, dref)
        assert dref >= 1e-8, (name, k, dref)  # a float32 computation, not a copy of the oracle


@pytest.mark.parametrize("name", list(R4_CASES))
def test_live_rp4_reference_reproduces_fixture(name):
    c, ref, prob = load_case(name)
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref",
                        "r4", f"libsodref_r4_p{prob.porder}.so")
    if not os.path.exists(path):
        pytest.skip("rp=4 reference not built (make -C oracle ref, needs the reference sources)")
    L = RefLib4(prob.porder)
    t = L.tables()
    assert np.array_equal(t["dlxigp_ip"].astype(np.float32), c["dlxigp_ip"])
    cc = dict(c, Ngp=np.asfortranarray(t["Ngp"], np.float32), dNgp=np.asfortranarray(t["dNgp"], np.float32))
    got = list(L.convec(cc)) + list(L.diffu(cc))
    for k, a, b in zip(OUTS, got, ref):
        assert np.array_equal(a, b), (name, k)

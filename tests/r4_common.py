"""Shared helpers of the rp=4 drop-in tests (tests/test_r4_golden.py on the
CPU, tests/test_gpu_r4.py on the GPU).

The fixtures tests/golden/r4_*.npz hold the residuals of the reference
compiled at its default real kind rp=4 (oracle/_ref/r4/, written by
tests/golden/make_golden.py --r4) for this repository's generator arrays
rounded to float32, with the reference's own rp=4 tables."""
import hashlib
import json
import os

import numpy as np

import minisod2d_amd as m
from oracle_lib import R4_FIELDS, Oracle, rel_l2

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")

# (fixture, builder) -- the inputs of tests/golden/make_golden.py R4_CASES
R4_CASES = {
    "r4_box8_p3": lambda: m.box_problem(8, 3),
    "r4_box3x2x2_p4_jit": lambda: m.box_problem(3, 4, ny=2, nz=2, jitter=0.05),
    "r4_box2_p7_jit_jump": lambda: m.box_problem(2, 7, jitter=0.05, rho_jump=0.1),
}
OUTS = ("Rmass", "Rmom", "Rener", "Dmass", "Dmom", "Dener")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(np.asfortranarray(a).ravel(order="F")).tobytes()).hexdigest()


def load_case(name):
    """(inputs, golden outputs): the float32 arrays a caller built at rp=4
    passes, checked against the digests kept in the fixture."""
    g = np.load(os.path.join(GOLDEN, name + ".npz"))
    meta = json.loads(str(g["meta"]))
    prob = R4_CASES[name]()
    f32 = lambda a: np.asfortranarray(a, dtype=np.float32)
    t = prob.tables
    c = dict(nelem=prob.nelem, npoin=prob.npoin, porder=t.porder, Cp=prob.Cp, Pr=prob.Pr,
             connec=prob.connec, dlxigp_ip=f32(g["dlxigp_ip"]), xgp=f32(g["xgp"]), Ngp=f32(t.Ngp),
             dNgp=f32(t.dNgp), invAtoIJK=t.invAtoIJK, AtoI=t.AtoI, AtoJ=t.AtoJ, AtoK=t.AtoK)
    for k in R4_FIELDS:
        c[k] = f32(getattr(prob, k))
    for k, want in meta["inputs"].items():
        assert sha(c[k]) == want, f"{name}: input {k} differs from the fixture's"
    return c, [g[k] for k in OUTS], prob


def oracle64(c, prob):
    """The FP64 oracle on the float32 inputs widened to double, with the
    rp=4 derivative table."""
    for k in R4_FIELDS:
        setattr(prob, k, np.asfortranarray(c[k], dtype=np.float64))
    prob.tables.dlxigp_ip = np.asfortranarray(c["dlxigp_ip"], dtype=np.float64)
    o = Oracle()
    return list(o.convec(prob) + o.diffu(prob))


def distances(got, ref, want):
    """Per residual: (got vs FP64 oracle, rp=4 reference vs FP64 oracle, got vs reference)."""
    return [(rel_l2(np.asarray(a, np.float64), w), rel_l2(np.asarray(r, np.float64), w),
             rel_l2(np.asarray(a, np.float64), np.asarray(r, np.float64)))
            for a, r, w in zip(got, ref, want)]

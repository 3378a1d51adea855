"""Writes the golden fixtures of tests/golden/ from the reference's own kernels.

Run in a container where /root/reference exists, after `make -C oracle ref`
(which compiles the reference's Fortran sources, rp=8, one library per
porder, into oracle/_ref/):

    python tests/golden/make_golden.py

Each fixture holds the reference's OUTPUTS for inputs produced by this
repository's own deterministic host generator (minisod2d_amd.box_problem /
reference_problem), plus a SHA-256 of every input array so that a test can
tell a generator change from a kernel change.  Fixtures:

  tables_pP.npz   reference element tables for porder P (main.f90:110-255)
  jac_p4.npz      He, gpvol of elem_jacobian on a jittered 2x2x2 p=4 box
  case_*.npz      Rmass/Rmom of full_convec_ijk, Rener of full_convec_ijk_split,
                  Dmass/Dmom/Dener of full_diffusion_ijk
  tet_tables.npz  TET04 tables of the reference driver (main.f90:374-383)
  tet_mesh.npz    digests of gen_tet_mesh's connec and coord (mod_mesh.f90:71-125)
  tet_jac.npz     He, gpvol of tet_jacobian on a jittered tetrahedral mesh
  tet_case_*.npz  Rmass/Rmom/Rener of fem_convec, Dmass/Dmom/Dener of fem_diffu

  digest_*.npz    the reference's residuals at configurations too large to
                  commit whole (64^3 p=4, 48^3 p=7, 96^3 p=3): per-component
                  L2 norms and sums, and the values at 4096 seeded nodes

  r4_*.npz        the reference compiled at its default real kind rp=4
                  (oracle/_ref/r4/): its own tables (float32 values) and its
                  residuals for the case's inputs rounded to float32 (SHA-256
                  of each rounded array kept)

`python tests/golden/make_golden.py --tet` writes the TET fixtures only,
`--digests` the digests only, `--r4` the rp=4 fixtures only.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import minisod2d_amd as m  # noqa: E402
from oracle_lib import R4_FIELDS, RefLib, RefLib4, RefTet, r4_inputs  # noqa: E402

# (name, builder) -- small enough for the reference to finish in seconds
CASES = {
    "case_box8_p3": lambda: m.box_problem(8, 3),                                  # 8^3 elements, porder 3
    "case_box2_p2_jit": lambda: m.box_problem(2, 2, jitter=0.05),
    "case_box2_p3_jit": lambda: m.box_problem(2, 3, jitter=0.05),
    "case_box3x2x2_p4_jit": lambda: m.box_problem(3, 4, ny=2, nz=2, jitter=0.05),
    "case_box2_p5_jit": lambda: m.box_problem(2, 5, jitter=0.05),
    "case_box2_p6_jit": lambda: m.box_problem(2, 6, jitter=0.05),
    "case_box2_p7_jit_jump": lambda: m.box_problem(2, 7, jitter=0.05, rho_jump=0.1),
    "case_box2_p8_jit": lambda: m.box_problem(2, 8, jitter=0.05),
    "case_refdriver_p4": lambda: m.reference_problem(6, 4),
    "case_refdriver_p3": lambda: m.reference_problem(5, 3),
}

# (name, builder, c_rho, c_ener) -- the TET04 kernels on the reference driver's
# disjoint tetrahedral mesh over the nodal fields of a hex problem
TET_CASES = {
    "tet_case_refdriver_p3": (lambda: m.tet_problem(m.reference_problem(4, 3)), 1.0, 1.0),
    "tet_case_box3_p2_jit": (lambda: m.tet_problem(m.box_problem(3, 2, jitter=0.05, rho_jump=0.1),
                                                   jitter=0.3, random_mu=True, strict=False), 1.0, 1.0),
    "tet_case_box4_p3_coef": (lambda: m.tet_problem(m.box_problem(4, 3), jitter=0.2, random_mu=True,
                                                    strict=False), 0.7, 1.3),
}
TET_INPUT_FIELDS = ("connec", "coord", "He", "gpvol", "u", "q", "rho", "pr", "E", "Tem", "mu_fluid",
                    "mu_e", "mu_sgs")

# (name, builder) -- the full-size configurations, as the GPU tests build them
DIGEST_CASES = {
    "digest_cfg1_64cubed_p4": lambda: m.box_problem(64, 4, jitter=0.05, threads=8),
    "digest_cfg3_48cubed_p7": lambda: m.box_problem(48, 7, jitter=0.05, rho_jump=0.1, threads=8),
    "digest_cfg4_96cubed_p3": lambda: m.box_problem(96, 3, jitter=0.05, threads=8),
}
DIGEST_NODES = 4096
DIGEST_SEED = 4096


def digest_nodes(npoin):
    """The 0-based nodes whose values a digest keeps."""
    return np.sort(np.random.default_rng(DIGEST_SEED).choice(npoin, size=DIGEST_NODES, replace=False))


# digest columns of each residual array (Rmass, Rmom, Rener, Dmass, Dmom, Dener):
# the parity bar is the relative L2 error of each array
DIGEST_GROUPS = (("Rmass", (0,)), ("Rmom", (1, 2, 3)), ("Rener", (4,)), ("Dmass", (5,)),
                 ("Dmom", (6, 7, 8)), ("Dener", (9,)))


def residual_digest(outputs, npoin):
    """outputs: Rmass, Rmom, Rener, Dmass, Dmom, Dener -> columns of 10 components."""
    cols = []
    for a in outputs:
        a = np.asarray(a).reshape(npoin, -1, order="F")
        cols.extend(a[:, j] for j in range(a.shape[1]))
    idx = digest_nodes(npoin)
    return dict(nodes=idx, values=np.stack([c[idx] for c in cols], axis=1),
                norms=np.array([np.linalg.norm(c) for c in cols]),
                sums=np.array([c.sum() for c in cols]))


INPUT_FIELDS = ("connec", "He", "gpvol", "u", "q", "rho", "pr", "E", "Tem", "mu_fluid", "mu_e", "mu_sgs")


def digest(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(np.asfortranarray(a).ravel(order="F")).tobytes()).hexdigest()


def input_digests(prob) -> dict:
    return {k: digest(getattr(prob, k)) for k in INPUT_FIELDS}


def tet_input_digests(tp) -> dict:
    return {k: digest(getattr(tp, k)) for k in TET_INPUT_FIELDS}


def write_tet():
    ref = RefTet(4)
    t = ref.tables()
    np.savez_compressed(os.path.join(HERE, "tet_tables.npz"), **t)
    nelem = 50
    connec, xyz = ref.mesh(nelem, 4 * nelem + 3, m.TET_BASE)
    np.savez_compressed(os.path.join(HERE, "tet_mesh.npz"), nelem=nelem, npoin=4 * nelem + 3,
                        connec_sha=digest(connec), coord_sha=digest(xyz[:4 * nelem]))
    tp = m.tet_problem(m.box_problem(3, 2), jitter=0.4, strict=False)
    He, gp = ref.jacobian(tp.nelem, tp.npoin, tp.connec, tp.coord, t["dNgp"], t["wgp"])
    np.savez_compressed(os.path.join(HERE, "tet_jac.npz"), He=He, gpvol=gp, coord_sha=digest(tp.coord),
                        connec_sha=digest(tp.connec))
    for name, (build, c_rho, c_ener) in TET_CASES.items():
        tp = build()
        Rmass, Rmom, Rener = ref.fem_convec(tp)
        Dmass, Dmom, Dener = ref.fem_diffu(tp, c_rho, c_ener)
        meta = dict(tp.meta, nelem=tp.nelem, npoin=tp.npoin, Cp=tp.Cp, Pr=tp.Pr, c_rho=c_rho,
                    c_ener=c_ener, inputs=tet_input_digests(tp))
        np.savez_compressed(os.path.join(HERE, name + ".npz"), Rmass=Rmass, Rmom=Rmom, Rener=Rener,
                            Dmass=Dmass, Dmom=Dmom, Dener=Dener, meta=json.dumps(meta))
        print(name, tp.nelem, tp.npoin, "ok")


def write_digests():
    for name, build in DIGEST_CASES.items():
        prob = build()
        ref = RefLib(prob.porder)
        Rmass, Rmom, Rener, _ = ref.convec(prob)
        Dmass, Dmom, Dener = ref.diffu(prob)
        dg = residual_digest((Rmass, Rmom, Rener, Dmass, Dmom, Dener), prob.npoin)
        meta = dict(prob.meta, nelem=prob.nelem, npoin=prob.npoin, inputs=input_digests(prob),
                    components="Rmass Rmom1 Rmom2 Rmom3 Rener Dmass Dmom1 Dmom2 Dmom3 Dener")
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **dg, meta=json.dumps(meta))
        print(name, prob.nelem, prob.npoin, "ok", flush=True)


# (name, builder) -- cases of the rp=4 drop-in
R4_CASES = {
    "r4_box8_p3": lambda: m.box_problem(8, 3),
    "r4_box3x2x2_p4_jit": lambda: m.box_problem(3, 4, ny=2, nz=2, jitter=0.05),
    "r4_box2_p7_jit_jump": lambda: m.box_problem(2, 7, jitter=0.05, rho_jump=0.1),
}


def write_r4():
    for name, build in R4_CASES.items():
        prob = build()
        ref = RefLib4(prob.porder)
        t = ref.tables()
        c = r4_inputs(prob, t)
        Rmass, Rmom, Rener = ref.convec(c)
        Dmass, Dmom, Dener = ref.diffu(c)
        meta = dict(prob.meta, nelem=prob.nelem, npoin=prob.npoin, Cp=prob.Cp, Pr=prob.Pr,
                    inputs={k: digest(c[k]) for k in R4_FIELDS + ("connec",)})
        np.savez_compressed(os.path.join(HERE, name + ".npz"), Rmass=Rmass, Rmom=Rmom, Rener=Rener,
                            Dmass=Dmass, Dmom=Dmom, Dener=Dener, dlxigp_ip=t["dlxigp_ip"].astype(np.float32),
                            xgp=t["xgp"].astype(np.float32), meta=json.dumps(meta))
        print(name, prob.nelem, prob.npoin, "ok", flush=True)


def main():
    if "--r4" in sys.argv[1:]:
        write_r4()
        return
    if "--tet" in sys.argv[1:]:
        write_tet()
        return
    if "--digests" in sys.argv[1:]:
        write_digests()
        return
    write_tet()
    for p in range(2, 9):
        t = RefLib(p).tables()
        np.savez_compressed(os.path.join(HERE, f"tables_p{p}.npz"),
                            **{k: v for k, v in t.items() if k not in ("Ngp", "dNgp")},
                            Ngp_sha=digest(t["Ngp"]), dNgp_sha=digest(t["dNgp"]))
    prob = m.box_problem(2, 4, jitter=0.05)
    ref = RefLib(4)
    He, gp = ref.jacobian(prob.nelem, prob.npoin, prob.connec, prob.coord, prob.tables.dNgp, prob.tables.wgp)
    np.savez_compressed(os.path.join(HERE, "jac_p4.npz"), He=He, gpvol=gp, coord_sha=digest(prob.coord),
                        connec_sha=digest(prob.connec))
    for name, build in CASES.items():
        prob = build()
        ref = RefLib(prob.porder)
        Rmass, Rmom, Rener, _ = ref.convec(prob)
        Dmass, Dmom, Dener = ref.diffu(prob)
        meta = dict(prob.meta, nelem=prob.nelem, npoin=prob.npoin, Cp=prob.Cp, Pr=prob.Pr,
                    inputs=input_digests(prob))
        np.savez_compressed(os.path.join(HERE, name + ".npz"), Rmass=Rmass, Rmom=Rmom, Rener=Rener,
                            Dmass=Dmass, Dmom=Dmom, Dener=Dener, meta=json.dumps(meta))
        print(name, prob.nelem, prob.npoin, "ok")


if __name__ == "__main__":
    main()

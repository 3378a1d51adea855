"""Benchmark of the convective + diffusive residual pair on one MI355X.

Workload (default): unit cube of 128x128x128 hexahedra of order 4
(135,005,697 global GLL nodes, 262,144,000 element nodes), FP64, synthetic
fields from the host generator.  A step is one call of the convective and one
of the diffusive residual over the whole mesh (zero-fill + kernel each, the
reference's timed region, src/app/sources/main.f90:446-478), with every input
already resident in HBM.

Prints one JSON line: GLL-node updates/s = npoin / (t_convec + t_diffu), the
roofline of the dominant kernel from HIP-event durations and the canonical
algorithmic bytes, the FP32-storage variants beside it, a sampled-node
parity check against the CPU oracle, and the reference's own CPU kernels timed
on a bounded sample of the same workload.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import minisod2d_amd as m  # noqa: E402

HBM_PEAK = 8.0e12  # MI355X, bytes/s
FP64_PEAK = 78.6e12  # MI355X vector FP64, flop/s
CALIB_BYTES = 2 << 30  # device-to-device copy that scales the HBM byte counters
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_convec_diffu_128cubed_p4.md")
TRAFFIC_TAG = "traffic-json: "
KERNEL_RE = {"full_convec_ijk": r"k_convec<(\d+), msod::ConvecArgs<double, double, double, double, double>, double>",
             "full_diffusion_ijk": r"k_diffu<(\d+), msod::DiffuArgs<double, double, double, double>, false, double>",
             "fem_convec": r"k_fem_convec<true>", "fem_diffu": r"k_fem_diffu<true>",
             "zero_fill": r"msod::k_zero\("}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def canonical_bytes(E, N, wg, wf, we=None):
    """Algorithmic bytes of one convective and one diffusive call: each input
    element read once, each output written once (FP64).  wg/wf/we: bytes per
    geometry word, nodal-field word and element-field word (mu_e, mu_sgs)."""
    we = wf if we is None else we
    bc = E * (4 + 10 * wg) + N * (9 * wf + 40)
    bd = E * (4 + 10 * wg + 2 * we) + N * (6 * wf + 40)
    return bc, bd


def calibrate_copy():
    """One device-to-device copy of CALIB_BYTES through the HIP runtime the
    kernel library uses (a blit kernel: CALIB_BYTES read, CALIB_BYTES
    written), so that a counter run sees a dispatch of known HBM traffic."""
    hip = ctypes.CDLL("libamdhip64.so.7")
    src, dst = ctypes.c_void_p(), ctypes.c_void_p()
    for ptr in (src, dst):
        if hip.hipMalloc(ctypes.byref(ptr), ctypes.c_size_t(CALIB_BYTES)) != 0:
            raise RuntimeError("hipMalloc of the calibration buffers failed")
    try:
        if hip.hipMemset(src, 1, ctypes.c_size_t(CALIB_BYTES)) != 0:
            raise RuntimeError("hipMemset failed")
        if hip.hipMemcpy(dst, src, ctypes.c_size_t(CALIB_BYTES), 3) != 0:  # device to device
            raise RuntimeError("hipMemcpy failed")
        if hip.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipDeviceSynchronize failed")
    finally:
        hip.hipFree(src)
        hip.hipFree(dst)


def summarize_pmc(fetch_csv, write_csv, n, porder):
    """HBM bytes per dispatch of the residual kernels from one FETCH_SIZE and
    one WRITE_SIZE counter run (rocprofv3 --pmc ... --output-format csv),
    each scaled by what the counter read for the calibration copy (the
    largest device-to-device blit, CALIB_BYTES each way)."""
    import csv
    import re

    def per_dispatch(path, counter):
        rows = {}
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] != counter:
                continue
            key = (r["Dispatch_Id"], r["Kernel_Name"])
            rows[key] = rows.get(key, 0.0) + float(r["Counter_Value"]) * 1024.0  # KiB
        return rows

    out = {"calibration": {}, "kernels": {}}
    for counter, path in (("FETCH_SIZE", fetch_csv), ("WRITE_SIZE", write_csv)):
        rows = per_dispatch(path, counter)
        copies = [v for (d, k), v in rows.items() if "copyBuffer" in k]
        if not copies:
            raise SystemExit(f"{path}: no device-to-device copy dispatch (run with --calibrate-copy)")
        factor = max(copies) / CALIB_BYTES
        if not 0.4 <= factor <= 2.5:
            raise SystemExit(f"{path}: largest copy reads {max(copies):.4g} bytes of {counter}, "
                             f"not the {CALIB_BYTES}-byte calibration copy")
        out["calibration"][counter] = {"copy_bytes": CALIB_BYTES, "counter_bytes": max(copies),
                                       "factor": factor}
        for name, pat in KERNEL_RE.items():
            vals = [v for (d, k), v in rows.items() if re.search(pat, k)]
            if not vals:
                continue
            ent = out["kernels"].setdefault(name, {})
            ent["dispatches"] = len(vals)
            ent[counter.lower() + "_raw_bytes"] = sum(vals) / len(vals)
            ent[counter.lower() + "_bytes"] = sum(vals) / len(vals) / factor
    prob_n = (n * porder + 1) ** 3
    E = n ** 3 * (porder + 1) ** 3
    bc, bd = canonical_bytes(E, prob_n, 8, 8)
    tc, td = tet_bytes(prob_n // 4)
    algo = {"full_convec_ijk": bc, "full_diffusion_ijk": bd, "fem_convec": tc, "fem_diffu": td}
    for name, ent in out["kernels"].items():
        if "fetch_size_bytes" in ent and "write_size_bytes" in ent:
            ent["traffic_bytes"] = ent["fetch_size_bytes"] + ent["write_size_bytes"]
            if name in algo:
                ent["algorithmic_bytes"] = algo[name]
                ent["traffic_over_algorithmic"] = ent["traffic_bytes"] / algo[name]
    out["workload"] = {"n": n, "porder": porder, "npoin": prob_n, "elem_nodes": E}
    return out


def short_kernel_name(k):
    """A kernel's demangled name without its argument list, storage types
    shortened (d = double, f = float)."""
    import re
    k = k.replace("(anonymous namespace)::", "")
    k = re.sub(r"^void ", "", k.split("(")[0]).replace("msod::", "")
    k = re.sub(r"(Convec|Diffu)Args<([^>]*)>", lambda mo: mo.group(1) + "<" + "".join(
        t.strip()[0] for t in mo.group(2).split(",")) + ">", k)
    return k


def summarize_counters(paths, min_ms=None):
    """Per kernel, the mean of every counter per dispatch over one or more
    rocprofv3 --pmc csv files (one counter set per run), as a markdown table
    with derived ratios where their counters are present."""
    import csv
    acc = {}
    for path in paths:
        for r in csv.DictReader(open(path)):
            key = short_kernel_name(r["Kernel_Name"])
            d = acc.setdefault(key, {})
            c = d.setdefault(r["Counter_Name"], {})
            c[r["Dispatch_Id"]] = c.get(r["Dispatch_Id"], 0.0) + float(r["Counter_Value"])
    names = sorted({c for d in acc.values() for c in d})
    ratios = [("wait/wave", "SQ_WAIT_ANY", "SQ_WAVE_CYCLES"),
              ("wait_issue/wave", "SQ_WAIT_INST_ANY", "SQ_WAVE_CYCLES"),
              ("valu/wave", "SQ_ACTIVE_INST_VALU", "SQ_WAVE_CYCLES"),
              ("lds/wave", "SQ_ACTIVE_INST_LDS", "SQ_WAVE_CYCLES"),
              ("bank conflict/LDS inst", "SQ_LDS_BANK_CONFLICT", "SQ_INSTS_LDS"),
              ("L2 hit rate", "TCC_HIT_sum", None)]
    rows = []
    for k, d in acc.items():
        mean = {c: sum(v.values()) / len(v) for c, v in d.items()}
        ndisp = max(len(v) for v in d.values())
        row = [k, str(ndisp)] + [f"{mean[c]:.4g}" if c in mean else "" for c in names]
        for _, a, b in ratios:
            if b is None:
                row.append(f"{mean[a] / (mean[a] + mean['TCC_MISS_sum']):.3f}"
                           if a in mean and "TCC_MISS_sum" in mean else "")
            else:
                row.append(f"{mean[a] / mean[b]:.3f}" if a in mean and b in mean and mean[b] else "")
        rows.append(row)
    head = ["kernel", "dispatches"] + names + [r[0] for r in ratios]
    out = ["| " + " | ".join(head) + " |", "|" + "---|" * len(head)]
    out += ["| " + " | ".join(r) + " |" for r in sorted(rows)]
    return "\n".join(out)


def load_traffic(n, porder, kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary of
    the 128^3 porder-4 workload, or None for any other workload."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith(TRAFFIC_TAG):
                    rec = json.loads(line[len(TRAFFIC_TAG):])
                    w = rec["workload"]
                    if (w["n"], w["porder"]) != (n, porder):
                        return None
                    return rec["kernels"].get(kernel, {}).get("traffic_bytes")
    except (OSError, ValueError, KeyError):
        return None
    return None


def minimal_flops(E, q):
    """Flops of the minimal formulation of the two residuals (per element
    node: convection 252q + 506, diffusion 66q + 290), for reporting the
    arithmetic side of the co-limitation next to the bytes."""
    return E * (252 * q + 506), E * (66 * q + 290)


TIMERS = {"fp64": (m.TIMER_CONVEC, m.TIMER_DIFFU),
          "mixed": (m.TIMER_CONVEC_MIXED, m.TIMER_DIFFU_MIXED),
          "geo32": (m.TIMER_CONVEC_GEO32, m.TIMER_DIFFU_GEO32),
          "fp32": (m.TIMER_CONVEC_FP32, m.TIMER_DIFFU_FP32),
          "f32split": (m.TIMER_CONVEC_F32SPLIT, m.TIMER_DIFFU_F32SPLIT)}


def split_bytes(E, N):
    """Canonical bytes of the FP32 split (FP32 He, gpvol, mu_e, mu_sgs, q,
    rho, E, Tem, mu_fluid; FP64 u and pr; FP64 outputs)."""
    return E * (4 + 10 * 4) + N * (3 * 8 + 8 + 3 * 4 + 4 + 4 + 40), E * (4 + 10 * 4 + 2 * 4) + N * (3 * 8 + 3 * 4 + 40)


def run_pair(prob, R, D, variant, f32=None):
    t = prob.tables
    if variant in ("mixed", "fp32"):
        conv = m.full_convec_ijk_fp32 if variant == "fp32" else m.full_convec_ijk_mixed
        diff = m.full_diffusion_ijk_fp32 if variant == "fp32" else m.full_diffusion_ijk_mixed
        conv(prob.nelem, prob.npoin, prob.connec, f32["He"], f32["gpvol"], t.dlxigp_ip,
             t.invAtoIJK, t.AtoI, t.AtoJ, t.AtoK, f32["u"], f32["q"], f32["rho"], f32["E"],
             prob.gamma, *R)
        diff(prob.nelem, prob.npoin, prob.connec, f32["He"], f32["gpvol"], t.dlxigp_ip,
             t.invAtoIJK, t.AtoI, t.AtoJ, t.AtoK, prob.Cp, prob.Pr, f32["rho"], f32["u"],
             f32["Tem"], f32["mu_fluid"], f32["mu_e"], f32["mu_sgs"], *D)
    elif variant == "f32split":
        m.full_convec_ijk_f32split(prob.nelem, prob.npoin, prob.connec, f32["He"], f32["gpvol"],
                                   t.dlxigp_ip, t.invAtoIJK, t.AtoI, t.AtoJ, t.AtoK, prob.u, f32["q"],
                                   f32["rho"], prob.pr, f32["E"], *R)
        m.full_diffusion_ijk_f32split(prob.nelem, prob.npoin, prob.connec, f32["He"], f32["gpvol"],
                                      t.dlxigp_ip, t.invAtoIJK, t.AtoI, t.AtoJ, t.AtoK, prob.Cp, prob.Pr,
                                      f32["rho"], prob.u, f32["Tem"], f32["mu_fluid"], f32["mu_e"],
                                      f32["mu_sgs"], *D)
    elif variant == "geo32":
        m.full_convec_ijk_geo32(prob.nelem, prob.npoin, prob.connec, f32["He"], f32["gpvol"],
                                t.dlxigp_ip, t.invAtoIJK, t.AtoI, t.AtoJ, t.AtoK, prob.u, prob.q,
                                prob.rho, prob.pr, prob.E, *R)
        m.full_diffusion_ijk_geo32(prob.nelem, prob.npoin, prob.connec, f32["He"], f32["gpvol"],
                                   t.dlxigp_ip, t.invAtoIJK, t.AtoI, t.AtoJ, t.AtoK, prob.Cp,
                                   prob.Pr, prob.rho, prob.u, prob.Tem, prob.mu_fluid,
                                   f32["mu_e"], f32["mu_sgs"], *D)
    else:
        m.full_convec_ijk(*prob.convec_args(*R))
        m.full_diffusion_ijk(*prob.diffu_args(*D))


def timed(prob, R, D, steps, warmup, variant, f32=None):
    """Wall time per pair over `steps` pairs with the per-launch timers off
    (their events are measurement, not work), then the kernels' HIP-event
    times over as many pairs again."""
    m.set_async(True)
    for _ in range(warmup):
        run_pair(prob, R, D, variant, f32)
    m.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        run_pair(prob, R, D, variant, f32)
    m.synchronize()
    t1 = time.perf_counter()
    m.timing_reset()
    m.timing_enable(True)
    for _ in range(steps):
        run_pair(prob, R, D, variant, f32)
    m.synchronize()
    m.timing_enable(False)
    m.set_async(False)
    kc, kd = TIMERS[variant]
    zc = m.timing_get(m.TIMER_ZERO)
    c = m.timing_get(kc)
    d = m.timing_get(kd)
    return dict(ms_per_step=(t1 - t0) * 1e3 / steps,
                convec_ms=c[0] / max(c[1], 1), diffu_ms=d[0] / max(d[1], 1),
                zero_ms=zc[0] / max(zc[1], 1), launches_per_call=c[1] / max(steps, 1))


def flops_block(E, q, r):
    fc, fd = minimal_flops(E, q)
    tc, td = r["convec_ms"] * 1e-3, r["diffu_ms"] * 1e-3
    return {"per_pair_gflop": (fc + fd) / 1e9, "convec_tflops": fc / tc / 1e12,
            "diffu_tflops": fd / td / 1e12, "pair_fp64_peak_frac": (fc + fd) / (tc + td) / FP64_PEAK,
            "note": "minimal-formulation flops (252q+506, 66q+290 per element node) over kernel time"}


def time_geometry(prob, threads):
    """elem_jacobian on the GPU (msod_elem_jacobian, outputs in device-only
    mirrors) against the host Fortran routine on `threads` cores."""
    t = prob.tables
    nn = t.nnode
    He = np.empty(prob.He.shape, order="F")
    gp = np.empty(prob.gpvol.shape, order="F")
    m.copyin(prob.coord, t.dNgp, t.wgp)
    m.create(He, gp)
    try:
        m.timing_reset()
        m.timing_enable(True)
        for _ in range(2):  # the first call builds nothing lasting: same work each time
            nbad = m.elem_jacobian(t.porder, prob.nelem, prob.npoin, prob.connec, prob.coord,
                                   t.dNgp, t.wgp, gp, He)
        m.timing_enable(False)
        ms, n = m.timing_get(m.TIMER_GEOMETRY)
    finally:
        m.delete(prob.coord, t.dNgp, t.wgp, He, gp)
    H = m.host_lib()
    H.msodh_set_threads(threads)
    He2 = np.empty(prob.He.shape, order="F")
    gp2 = np.empty(prob.gpvol.shape, order="F")
    t0 = time.perf_counter()
    H.msodh_elem_jacobian(prob.nelem, prob.npoin, nn, prob.connec.ctypes.data, prob.coord.ctypes.data,
                          t.dNgp.ctypes.data, t.wgp.ctypes.data, gp2.ctypes.data, He2.ctypes.data)
    host_ms = (time.perf_counter() - t0) * 1e3
    del He2, gp2
    E = prob.nelem * nn
    nbytes = E * (4 + 80) + prob.npoin * 24
    dev_ms = ms / max(n, 1)
    return {"device_ms": dev_ms, "host_ms": host_ms, "host_threads": threads, "nbad": nbad,
            "bytes": nbytes, "device_GBps": nbytes / (dev_ms * 1e-3) / 1e9,
            "note": "elem_jacobian: He, gpvol from coord on the GPU vs the host Fortran routine"}


def time_generator(n, p, jitter, host_s, threads):
    """Mesh, He/gpvol and fields of the workload generated on the GPU
    (device_box_problem), against the host generator's wall time."""
    m.timing_reset()
    m.timing_enable(True)
    m.synchronize()
    t0 = time.perf_counter()
    d = m.device_box_problem(n, p, jitter=jitter)
    m.synchronize()
    wall = time.perf_counter() - t0
    m.timing_enable(False)
    gen_ms, _ = m.timing_get(m.TIMER_GENERATOR)
    geo_ms, _ = m.timing_get(m.TIMER_GEOMETRY)
    m.delete(*d.device_inputs(), d.coord)
    return {"device_kernels_ms": gen_ms + geo_ms, "device_wall_s": wall, "host_s": host_s,
            "host_threads": threads,
            "note": "box mesh, He/gpvol and fields: device kernels vs the host generator"}


def tet_bytes(nelem):
    """Algorithmic bytes of fem_convec and fem_diffu on a disjoint mesh of
    nelem tetrahedra (4 nodes each, every input read once, outputs written
    once): per element connec 16, He 288, gpvol 32 (+ mu_e, mu_sgs 64 in the
    diffusion); per node u,q,rho,pr,E 72 (rho,u,Tem,mu_fluid 48) and 40 of
    outputs."""
    return nelem * (336 + 4 * (72 + 40)), nelem * (400 + 4 * (48 + 40))


def tet_cpu_baseline(tp, nsub=1 << 20, nruns=3):
    """The reference's own fem_convec and fem_diffu (oracle/_ref, compiled
    from its sources, one core) on the first nsub tetrahedra of the mesh."""
    from oracle_lib import RefTet, ref_available
    if not ref_available(4):
        return None
    ref = RefTet(4)
    n = min(nsub, tp.nelem)
    f = np.asfortranarray
    sub = m.TetProblem(tp.tables, n, 4 * n, f(tp.connec[:n]), f(tp.coord[:4 * n]),
                       f(tp.He[:, :, :, :n]), f(tp.gpvol[:, :, :n]),
                       *[f(getattr(tp, k)[:4 * n]) for k in ("u", "q", "rho", "pr", "E", "Tem", "mu_fluid")],
                       f(tp.mu_e[:n]), f(tp.mu_sgs[:n]), Cp=tp.Cp, Pr=tp.Pr)
    tc, td = [], []
    for _ in range(nruns):
        t0 = time.perf_counter()
        ref.fem_convec(sub)
        t1 = time.perf_counter()
        ref.fem_diffu(sub)
        t2 = time.perf_counter()
        tc.append(t1 - t0)
        td.append(t2 - t1)
    tcm, tdm = float(np.median(tc)), float(np.median(td))
    return dict(value=n / (tcm + tdm), unit="tetrahedra/s (fem_convec + fem_diffu)", cores=1,
                kind="reference",
                sample=f"first {n} tetrahedra of the same mesh, reference fem_convec + fem_diffu "
                       f"(amdflang -O3, rp=8), median of {nruns}: convec {tcm*1e3:.1f} ms, "
                       f"diffu {tdm*1e3:.1f} ms")


def time_tet(prob, R, D, steps, warmup, workload, nsample=2000, with_cpu=True):
    """fem_convec + fem_diffu on the reference driver's TET04 data over the
    workload's nodal fields (npoin/4 disjoint tetrahedra, tet_jacobian on the
    device), outputs in the workload's R, D; sampled elements against the oracle."""
    from oracle_lib import Oracle, rel_l2
    tp = m.tet_problem(prob, jitter=0.2, random_mu=True, strict=False)
    t = tp.tables
    He, gp = np.empty(tp.He.shape, order="F"), np.empty(tp.gpvol.shape, order="F")
    els_in = (tp.connec, tp.coord, tp.mu_e, tp.mu_sgs)
    m.copyin(*els_in)
    m.create(He, gp)
    try:
        m.timing_reset()
        m.timing_enable(True)
        nbad = m.tet_jacobian(tp.nelem, tp.npoin, tp.connec, tp.coord, t.dNgp, t.wgp, gp, He)
        m.timing_enable(False)
        jac_ms, _ = m.timing_get(m.TIMER_TET_GEOMETRY)
        tp.He, tp.gpvol = He, gp  # the device-computed geometry, as the reference driver uses
        m.set_async(True)
        for _ in range(warmup):
            m.fem_convec(*tp.convec_args(*R))
            m.fem_diffu(*tp.diffu_args(*D))
        m.synchronize()
        m.timing_reset()
        m.timing_enable(True)
        t0 = time.perf_counter()
        for _ in range(steps):
            m.fem_convec(*tp.convec_args(*R))
            m.fem_diffu(*tp.diffu_args(*D))
        m.synchronize()
        t1 = time.perf_counter()
        m.timing_enable(False)
        m.set_async(False)
        c, d, z = (m.timing_get(k) for k in (m.TIMER_FEM_CONVEC, m.TIMER_FEM_DIFFU, m.TIMER_ZERO))
        m.update_host(*R, *D, He, gp)
    finally:
        m.delete(*els_in, He, gp)
    cms, dms, zms = c[0] / max(c[1], 1), d[0] / max(d[1], 1), z[0] / max(z[1], 1)
    bc, bd = tet_bytes(tp.nelem)
    rng = np.random.default_rng(9)
    els = np.sort(rng.choice(tp.nelem, size=min(nsample, tp.nelem), replace=False))
    nodes = (tp.connec[els] - 1).ravel()
    f = np.asfortranarray
    sub = m.TetProblem(t, len(els), 4 * len(els),
                       f(np.arange(1, 4 * len(els) + 1, dtype=np.int32).reshape(len(els), 4)),
                       f(tp.coord[nodes]), f(He[:, :, :, els]), f(gp[:, :, els]),
                       *[f(getattr(tp, k)[nodes]) for k in ("u", "q", "rho", "pr", "E", "Tem", "mu_fluid")],
                       f(tp.mu_e[els]), f(tp.mu_sgs[els]), Cp=tp.Cp, Pr=tp.Pr)
    o = Oracle()
    want = o.fem_convec(sub) + o.fem_diffu(sub)
    errs = {k: rel_l2(g[nodes], w) for k, g, w in
            zip(("Rmass", "Rmom", "Rener", "Dmass", "Dmom", "Dener"), R + D, want)}
    cpu = tet_cpu_baseline(tp) if with_cpu else None
    return {"nelem": tp.nelem, "npoin": tp.npoin, "nbad": nbad, "tet_jacobian_ms": jac_ms,
            "ms_per_step": (t1 - t0) * 1e3 / steps,
            "tetrahedra_per_s": tp.nelem / ((cms + dms) * 1e-3),
            "kernels_ms": {"fem_convec": cms, "fem_diffu": dms, "zero_fill": zms},
            "bytes": {"fem_convec": bc, "fem_diffu": bd},
            "hbm_frac": {"fem_convec": bc / (cms * 1e-3) / HBM_PEAK,
                         "fem_diffu": bd / (dms * 1e-3) / HBM_PEAK},
            "traffic": {k: load_traffic(*workload, k) for k in ("fem_convec", "fem_diffu")},
            "sampled_rel_l2": errs,
            "cpu_baseline": cpu,
            "note": "TET04 kernels on npoin/4 disjoint tetrahedra over the workload's nodal "
                    "fields (main.f90:364-407), jittered vertices, random mu_e, mu_sgs"}


def sampled_parity(prob, out, nsample=3000, seed=7):
    """Residuals at randomly chosen nodes against the oracle run on exactly
    the elements that touch those nodes (each sampled node then receives all
    of its contributions)."""
    from oracle_lib import Oracle, rel_l2
    rng = np.random.default_rng(seed)
    nodes = rng.choice(prob.npoin, size=min(nsample, prob.npoin), replace=False) + 1
    hit = np.isin(prob.connec, nodes).any(axis=1)
    els = np.nonzero(hit)[0]
    sub_conn = prob.connec[els]
    uniq, inv = np.unique(sub_conn, return_inverse=True)
    sub = m.Problem(prob.tables, len(els), len(uniq),
                    np.asfortranarray((inv.reshape(sub_conn.shape) + 1).astype(np.int32)),
                    prob.coord[uniq - 1], np.asfortranarray(prob.He[:, :, :, els]),
                    np.asfortranarray(prob.gpvol[:, :, els]),
                    *[np.asfortranarray(getattr(prob, k)[uniq - 1]) for k in
                      ("u", "q", "rho", "pr", "E", "Tem", "mu_fluid")],
                    np.asfortranarray(prob.mu_e[els]), np.asfortranarray(prob.mu_sgs[els]),
                    Cp=prob.Cp, Pr=prob.Pr, gamma=prob.gamma)
    o = Oracle()
    want = o.convec(sub, 8) + o.diffu(sub, 8)
    pos = np.searchsorted(uniq, nodes)
    errs = {}
    for key, g, w in zip(("Rmass", "Rmom", "Rener", "Dmass", "Dmom", "Dener"), out, want):
        errs[key] = rel_l2(g[nodes - 1], w[pos])
    return errs


SPLIT_NOTES = {
    "f32split": "FP32 He, gpvol, mu_e, mu_sgs, q, rho, E, Tem, mu_fluid; FP64 u and pr "
                "(chosen field by field for <= 1e-6); FP64 arithmetic and atomic assembly",
    "mixed": "every input stored in FP32, FP64 arithmetic and assembly; error = FP32 "
             "rounding of the inputs, amplified by the derivative sweeps",
    "geo32": "He, gpvol, mu_e, mu_sgs stored in FP32, nodal fields FP64, FP64 "
             "arithmetic and assembly",
    "fp32": "every input stored in FP32, FP32 gathers, fluxes and derivative sweeps; "
            "FP64 node sums, atomic assembly and pressure recovery"}


def split_record(E, N, r64, r, variant):
    """One precision form against the FP64 run of the same mesh."""
    wb = (4, 8, 4) if variant == "geo32" else (4, 4, 4)
    bcm, bdm = split_bytes(E, N) if variant == "f32split" else canonical_bytes(E, N, *wb)
    errs = r.get("sampled_rel_l2")
    return {"value": N / (r["ms_per_step"] * 1e-3), "ms_per_step": r["ms_per_step"],
            "speedup_vs_fp64": r64["ms_per_step"] / r["ms_per_step"],
            "kernel_speedup_vs_fp64": (r64["convec_ms"] + r64["diffu_ms"]) / (r["convec_ms"] + r["diffu_ms"]),
            "kernels_ms": {"convec": r["convec_ms"], "diffu": r["diffu_ms"]},
            "hbm_frac_convec": bcm / (r["convec_ms"] * 1e-3) / HBM_PEAK,
            "hbm_frac_diffu": bdm / (r["diffu_ms"] * 1e-3) / HBM_PEAK,
            "rel_l2_vs_fp64_oracle": errs,
            "max_rel_l2": max(errs.values()) if errs else None,
            "storage": SPLIT_NOTES[variant]}


def config_record(n, p, jitter, jump, steps, warmup, threads, variants=("f32split", "mixed", "fp32"),
                  parity=True):
    """A second mesh of the benchmark set, timed in FP64 and in the precision
    forms, with each form's sampled error against the FP64 oracle."""
    log(f"config {n}^3 porder {p} jitter {jitter} density jump {jump}: generating")
    prob = m.box_problem(n, p, jitter=jitter, rho_jump=jump, threads=threads)
    N, E = prob.npoin, prob.nelem * prob.tables.nnode
    R, D = prob.outputs(), prob.outputs()
    ins = prob.device_inputs()
    m.copyin(*ins)
    m.create(*R, *D)
    f32 = prob.fp32()
    f32_ins = [f32[k] for k in ("He", "gpvol", "u", "q", "rho", "E", "Tem", "mu_fluid", "mu_e", "mu_sgs")]
    try:
        r64 = timed(prob, R, D, steps, warmup, "fp64")
        log(f"config {n}^3 p={p} FP64 {r64}")
        errs = None
        if parity:
            m.update_host(*R, *D)
            errs = sampled_parity(prob, R + D, nsample=2000)
        bc, bd = canonical_bytes(E, N, 8, 8)
        rec = {"workload": f"unit cube {n}x{n}x{n} hex elements, porder {p}, {N} GLL nodes, "
                           f"{E} element nodes, jitter {jitter}, density jump {jump}",
               "npoin": N, "nelem": prob.nelem, "porder": p,
               "fp64": {"value": N / (r64["ms_per_step"] * 1e-3), "ms_per_step": r64["ms_per_step"],
                        "kernels_ms": {"convec": r64["convec_ms"], "diffu": r64["diffu_ms"],
                                       "zero_fill": r64["zero_ms"]},
                        "hbm_frac_convec": bc / (r64["convec_ms"] * 1e-3) / HBM_PEAK,
                        "hbm_frac_diffu": bd / (r64["diffu_ms"] * 1e-3) / HBM_PEAK,
                        "sampled_rel_l2_vs_oracle": errs}}
        m.copyin(*f32_ins)
        for variant in variants:
            r = timed(prob, R, D, steps, warmup, variant, f32=f32)
            if parity:
                m.update_host(*R, *D)
                r["sampled_rel_l2"] = sampled_parity(prob, R + D, nsample=2000)
            rec[variant] = split_record(E, N, r64, r, variant)
            log(f"config {n}^3 p={p} {variant} {rec[variant]['ms_per_step']:.3f} ms, "
                f"max err {rec[variant]['max_rel_l2']}")
    finally:
        m.delete(*ins, *R, *D, *f32_ins)
    return rec


def closure_record(n=96, seed=50, repeats=10):
    """The learned closure at porder 3 on an n^3 box: the fused inference +
    diffusion kernel against the plain diffusion (mu_sgs read from memory), and
    the error of the order-3 residual against the order-7 residual of the same
    fields projected to order 3, with the closure and with mu_sgs = 0."""
    from minisod2d_amd import closure as cl
    try:
        w = cl.load_weights()
    except cl.ClosureModelError as e:
        return {"error": str(e)}
    log(f"closure: order-7/order-3 pair on {n}^3 elements")
    t0 = time.time()
    pair = cl.make_pair(n, seed=seed)
    rec = cl.report(pair, w["params"], repeats=repeats)
    rec["weights_sha256"] = w["manifest"].get("sha256")
    rec["pair_build_s"] = round(time.time() - t0, 1)
    rec["note"] = ("closure_ms: mu_sgs inferred per element node by the MLP (bf16 MFMA) inside the "
                   "diffusion kernel; plain_ms: full_diffusion_ijk with mu_sgs from memory; errors per "
                   "component (momentum x, y, z, energy) relative to the projected order-7 residual")
    del pair
    # noise-free pairs (Taylor-Green velocity of wavenumber 2, smooth density
    # and pressure): the least-squares constant mu_sgs and its gain, and a
    # network trained on 12^3 and 16^3 pairs, evaluated on a held-out 20^3 pair
    log("closure: noise-free pairs, constant fit and a trained network")
    t0 = time.time()
    d = cl.diagnose(2, [12, 16], [10], [20], steps=300)
    rec["noise_free_k2"] = {"constant_fit": d["constant"], "model_gain": d["model"],
                            "best_step": d["best_step"], "seconds": round(time.time() - t0, 1)}
    return rec


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), for the baseline record."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(porder, n=32, nruns=3):
    """The reference's own Fortran kernels (oracle/_ref, compiled from the
    reference sources, rp=8, single thread as the reference is) on an
    n^3-element box of the same generator."""
    from oracle_lib import RefLib, ref_available
    if not ref_available(porder):
        return None
    ref = RefLib(porder)
    prob = m.box_problem(n, porder, jitter=0.0)
    t = prob.tables
    R = prob.outputs()
    D = prob.outputs()
    p_ = lambda a: a.ctypes.data
    cargs = (prob.nelem, prob.npoin, p_(prob.connec), p_(t.Ngp), p_(t.dNgp), p_(prob.He),
             p_(prob.gpvol), p_(t.dlxigp_ip), p_(t.xgp), p_(t.invAtoIJK), p_(t.AtoI), p_(t.AtoJ),
             p_(t.AtoK), p_(prob.u), p_(prob.q), p_(prob.rho), p_(prob.pr), p_(prob.E),
             *map(p_, R))
    tc, td = [], []
    for _ in range(nruns):
        t0 = time.perf_counter()
        ref.L.sodref_full_convec_ijk(*cargs)
        t1 = time.perf_counter()
        ref.diffu(prob)
        t2 = time.perf_counter()
        tc.append(t1 - t0)
        td.append(t2 - t1)
    tcm, tdm = float(np.median(tc)), float(np.median(td))
    return dict(value=prob.npoin / (tcm + tdm), unit="GLL-node updates/s", cores=1,
                kind="reference", cpu_model=cpu_model(), nproc=os.cpu_count(),
                sample=f"{n}^3 elements porder {porder} ({prob.npoin} nodes), reference "
                       f"full_convec_ijk + full_diffusion_ijk (amdflang -O3, rp=8), median of "
                       f"{nruns}: convec {tcm*1e3:.1f} ms, diffu {tdm*1e3:.1f} ms")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=128, help="elements per axis")
    ap.add_argument("--porder", type=int, default=4)
    ap.add_argument("--jitter", type=float, default=0.0)
    ap.add_argument("--rho-jump", type=float, default=0.0, help="density jump at x > 0.5 (relative)")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-parity", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-geometry", action="store_true", help="skip the elem_jacobian timing")
    ap.add_argument("--no-tet", action="store_true", help="skip the TET04 kernels")
    ap.add_argument("--threads", type=int, default=16)
    ap.add_argument("--assembly", choices=("colored", "atomic"), default="atomic",
                    help="node assembly of the reported run")
    ap.add_argument("--compare-assembly", action="store_true",
                    help="also time the other assembly mode (FP64)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run one 2 GiB device-to-device copy first (HBM counter calibration)")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the 48^3 porder 7 precision record and the 96^3 porder 3 closure record")
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="print HBM bytes per dispatch from two counter runs and exit")
    ap.add_argument("--summarize-counters", nargs="+", metavar="CSV",
                    help="print every counter's mean per dispatch and kernel (markdown) and exit")
    args = ap.parse_args()
    if args.summarize_counters:
        print(summarize_counters(args.summarize_counters))
        return
    if args.summarize_pmc:
        print(json.dumps(summarize_pmc(*args.summarize_pmc, args.n, args.porder)))
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported here "
              f"(the element-partitioned multi-GPU form is a design, not a build)", file=sys.stderr)
        sys.exit(2)

    n, p = args.n, args.porder
    log(f"generating {n}^3 porder {p} (host, {args.threads} threads)")
    t0 = time.time()
    prob = m.box_problem(n, p, jitter=args.jitter, rho_jump=args.rho_jump,
                         threads=min(args.threads, os.cpu_count() or 1))
    host_gen_s = time.time() - t0
    N, E = prob.npoin, prob.nelem * prob.tables.nnode
    log(f"generated npoin={N} elem-nodes={E} in {host_gen_s:.1f}s")

    m.init(0)
    if args.calibrate_copy:
        calibrate_copy()
    R, D = prob.outputs(), prob.outputs()
    ins = prob.device_inputs()
    t0 = time.time()
    m.copyin(*ins)
    m.create(*R, *D)
    log(f"uploaded in {time.time()-t0:.1f}s; timing FP64")
    mode = m.ASSEMBLY_ATOMIC if args.assembly == "atomic" else m.ASSEMBLY_COLORED
    other = None
    if args.compare_assembly:
        m.set_assembly(m.ASSEMBLY_COLORED if mode == m.ASSEMBLY_ATOMIC else m.ASSEMBLY_ATOMIC)
        other = timed(prob, R, D, args.steps, args.warmup, "fp64")
        log(f"FP64 other assembly {other}")
    m.set_assembly(mode)
    r64 = timed(prob, R, D, args.steps, args.warmup, "fp64")
    log(f"FP64 {r64}")
    parity = None
    if not args.no_parity:
        m.update_host(*R, *D)
        parity = sampled_parity(prob, R + D)
        log(f"sampled parity {parity}")

    geometry = None
    if not args.no_geometry:
        log("timing the device geometry (elem_jacobian)")
        geometry = time_geometry(prob, min(args.threads, os.cpu_count() or 1))
        log(f"geometry {geometry}")
        log("timing the device input generator")
        geometry["generator"] = time_generator(n, p, args.jitter, host_gen_s,
                                               min(args.threads, os.cpu_count() or 1))
        log(f"generator {geometry['generator']}")

    tet = None
    if not args.no_tet:
        log("timing the TET04 kernels")
        tet = time_tet(prob, R, D, args.steps, args.warmup, (n, p), with_cpu=not args.no_cpu_baseline)
        log(f"tet {tet}")

    split = {}
    if not args.no_mixed:
        f32 = prob.fp32()
        m.copyin(*[f32[k] for k in ("He", "gpvol", "u", "q", "rho", "E", "Tem", "mu_fluid", "mu_e", "mu_sgs")])
        for variant in ("f32split", "mixed", "geo32", "fp32"):
            log(f"timing {variant}")
            r = timed(prob, R, D, args.steps, args.warmup, variant, f32=f32)
            log(f"{variant} {r}")
            if parity is not None:
                m.update_host(*R, *D)
                r["sampled_rel_l2"] = sampled_parity(prob, R + D)
            split[variant] = r

    bc, bd = canonical_bytes(E, N, 8, 8)
    dom, dom_ms, dom_b = ("full_convec_ijk", r64["convec_ms"], bc)
    if r64["diffu_ms"] > r64["convec_ms"]:
        dom, dom_ms, dom_b = ("full_diffusion_ijk", r64["diffu_ms"], bd)
    achieved = dom_b / (dom_ms * 1e-3)
    value = N / (r64["ms_per_step"] * 1e-3)
    rec = {
        "metric": "GLL-node updates/s (convec+diffu), FP64 vs FP32-split at stated err; HBM GB/s",
        "value": value,
        "unit": "GLL-node updates/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": r64["ms_per_step"],
        "elem_node_updates_per_s": E / (r64["ms_per_step"] * 1e-3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"unit cube {n}x{n}x{n} hex elements, porder {p}, "
                               f"{N} GLL nodes, {E} element nodes, jitter {args.jitter}"
                               + (f", density jump {args.rho_jump}" if args.rho_jump else ""),
                   "npoin": N, "nelem": prob.nelem, "porder": p},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved / 1e9,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": achieved / HBM_PEAK,
                     "traffic": load_traffic(n, p, dom),
                     "traffic_source": os.path.relpath(TRAFFIC_FILE, ROOT)},
        "kernels_ms": {"convec": r64["convec_ms"], "diffu": r64["diffu_ms"],
                       "zero_fill": r64["zero_ms"]},
        "flops": flops_block(E, p + 1, r64),
        "pair_hbm_frac": (bc + bd) / ((r64["convec_ms"] + r64["diffu_ms"] + 2 * r64["zero_ms"]) * 1e-3) / HBM_PEAK,
        "parity_sampled_rel_l2": parity,
        "geometry": geometry,
        "tet": tet,
        "assembly": args.assembly,
    }
    if other is not None:
        rec["other_assembly"] = {"mode": "atomic" if args.assembly == "colored" else "colored",
                                 "ms_per_step": other["ms_per_step"],
                                 "kernels_ms": {"convec": other["convec_ms"], "diffu": other["diffu_ms"],
                                                "zero_fill": other["zero_ms"]}}
    for variant, r in split.items():
        rec[variant] = split_record(E, N, r64, r, variant)
    m.delete(*ins, *R, *D)
    if split:
        m.delete(*[f32[k] for k in ("He", "gpvol", "u", "q", "rho", "E", "Tem", "mu_fluid", "mu_e", "mu_sgs")])
        del f32
    del prob, R, D, ins
    if not args.no_configs:
        threads = min(args.threads, os.cpu_count() or 1)
        rec["cfg3"] = config_record(48, 7, 0.05, 0.1, args.steps, args.warmup, threads,
                                    parity=not args.no_parity)
        rec["closure"] = closure_record()
        log(f"closure {rec['closure']}")
    m.finalize()
    if not args.no_cpu_baseline:
        log("timing the reference CPU kernels on a bounded sample")
        rec["cpu_baseline"] = cpu_baseline(p)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()

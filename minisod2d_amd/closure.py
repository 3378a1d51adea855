"""Learned subgrid-viscosity closure: training pairs, GPU trainer, weight files.

The reference has no subgrid model: ``mu_sgs`` is an input of
full_diffusion_ijk (elem_diffu.f90:70,146) that its driver sets to 1
(main.f90:334).  Here a per-node MLP (12 -> 32 -> 32 -> 1, SiLU, softplus;
csrc/msod_mlp.h) predicts it from the local velocity gradient, rho, |u| and
the element size, and is trained on data the method generates itself:

* a fine run: the order-``pf`` (7) residuals Rc_f + Rd_f of the synthetic
  fields on a box mesh of n^3 affine hexahedra, mu_sgs = 0;
* the same fields interpolated to the order-``pt`` (3) nodes of the same
  hexahedra, and the fine residual restricted to the order-``pt`` test
  functions, P^T (Rc_f + Rd_f) (exact for weak-form residuals, since the
  order-pt basis lies in the order-pf space);
* the target of the order-pt diffusion residual,
      target = P^T (Rc_f + Rd_f) - Rc_t,
  over momentum (3) and energy; the closure should make Rd_t(mu_sgs) reach it.

Loss: sum_c w_c |Rd_t(mu_sgs)_c - target_c|^2 with w_c = 1/|Rd_t(0)_c - target_c|^2,
so that the method without the closure (mu_sgs = 0) scores 1 per component.
The gradient goes through the diffusion operator, which is affine in mu_sgs:
msod_closure_sensitivity gives d loss / d mu_sgs at every element node and
msod_closure_gradient back-propagates it through the MLP (bf16 MFMA).  All of
it runs on the GPU; the host only sees the loss.

Weight files: ``.npz`` (no pickled objects) with the parameter block and the
Adam state, and a ``.json`` manifest with its sha256, so that training can
resume and a corrupt or foreign file is refused (the kernels then fall back to
mu_sgs = 0).
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import sys
import time
from dataclasses import dataclass, field

import numpy as np

from . import (CLOSURE_NFEAT, CLOSURE_NPARAM, CLOSURE_NTRAIN, GAMMA, RGAS, SEED_JITTER, SEED_MUE,
               SEED_MUSGS, SEED_RHO, Problem, box_problem, closure_adam, closure_features, closure_gradient,
               closure_loss, closure_predict, closure_sensitivity, copyin, create, delete,
               full_convec_ijk, full_diffusion_ijk, full_diffusion_ijk_closure, host_lib,
               update_device, update_host)

NIN, NH = 12, 32
# parameter block offsets (csrc/msod_mlp.h)
OW1 = 0
OB1 = OW1 + NIN * NH
OW2 = OB1 + NH
OB2 = OW2 + NH * NH
OW3 = OB2 + NH
OB3 = OW3 + NH
OMEAN = OB3 + 1
OISTD = OMEAN + NIN
OSCALE = OISTD + NIN
assert OMEAN == CLOSURE_NTRAIN and OSCALE + 1 == CLOSURE_NPARAM and NIN == CLOSURE_NFEAT

FORMAT = "msod-closure-v1"
DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data")
DEFAULT_WEIGHTS = os.path.join(DATA_DIR, "closure_p3.npz")


class ClosureModelError(RuntimeError):
    """A weight file that is missing, corrupt or of another format."""


# ---------------------------------------------------------------------------
# order transfer (Fortran host library, fortran/msod_interp.f90)
# ---------------------------------------------------------------------------
def interp_1d(pf: int, pt: int) -> np.ndarray:
    """M[i, n] = l^pf_n(xi^pt_i) in the SOD 1-D node order."""
    M = np.zeros((pt + 1, pf + 1), order="F")
    host_lib().msodh_interp_1d(pf, pt, M.ctypes.data)
    return M


def prolong_fields(fine: Problem, coarse: Problem, ff: np.ndarray) -> np.ndarray:
    """Nodal fields ff(npoin_f, ncomp) of the fine order at the coarse nodes."""
    ff = np.asfortranarray(ff, dtype=np.float64)
    nc = ff.shape[1]
    ft = np.zeros((coarse.npoin, nc), order="F")
    host_lib().msodh_prolong_fields(fine.porder, coarse.porder, fine.nelem, fine.npoin,
                                    coarse.npoin, nc, fine.connec.ctypes.data,
                                    coarse.connec.ctypes.data, ff.ctypes.data, ft.ctypes.data)
    return ft


def restrict_residual(fine: Problem, coarse: Problem, rf: np.ndarray) -> np.ndarray:
    """Weak-form residuals rf(npoin_f, ncomp) moved to the coarse test functions."""
    rf = np.asfortranarray(rf, dtype=np.float64)
    nc = rf.shape[1]
    rt = np.zeros((coarse.npoin, nc), order="F")
    host_lib().msodh_restrict_residual(fine.porder, coarse.porder, fine.nelem, fine.npoin,
                                       coarse.npoin, nc, fine.connec.ctypes.data,
                                       coarse.connec.ctypes.data, rf.ctypes.data, rt.ctypes.data)
    return rt


def prolong_element_field(fine: Problem, coarse: Problem, mf: np.ndarray) -> np.ndarray:
    """Element-node field mf(nelem, nnode_f) (reference layout) interpolated
    inside each element to the coarse element nodes."""
    tf, tc = fine.tables, coarse.tables
    qf, qc = tf.porder + 1, tc.porder + 1
    M = interp_1d(tf.porder, tc.porder)
    F = np.zeros((fine.nelem, qf, qf, qf))
    F[:, tf.AtoI - 1, tf.AtoJ - 1, tf.AtoK - 1] = mf
    T = np.einsum("ai,eijk->eajk", M, F)
    T = np.einsum("bj,eajk->eabk", M, T)
    T = np.einsum("ck,eabk->eabc", M, T)
    return np.asfortranarray(T[:, tc.AtoI - 1, tc.AtoJ - 1, tc.AtoK - 1])


# ---------------------------------------------------------------------------
# training pairs
# ---------------------------------------------------------------------------
def gpu_residuals(prob: Problem, mu_sgs: np.ndarray | None = None):
    """(Rmass, Rmom, Rener) of convection and of diffusion, FP64 on the GPU."""
    mu = prob.mu_sgs if mu_sgs is None else mu_sgs
    ins = [prob.connec, prob.He, prob.gpvol, prob.u, prob.q, prob.rho, prob.pr, prob.E,
           prob.Tem, prob.mu_fluid, prob.mu_e, mu]
    Rc, Rd = prob.outputs(), prob.outputs()
    copyin(*ins)
    create(*Rc, *Rd)
    try:
        full_convec_ijk(*prob.convec_args(*Rc))
        full_diffusion_ijk(*prob.diffu_args(*Rd, mu_sgs=mu))
        update_host(*Rc, *Rd)
    finally:
        delete(*ins, *Rc, *Rd)
    return Rc, Rd


def _mom_ener(Rmom: np.ndarray, Rener: np.ndarray) -> np.ndarray:
    """[4][npoin] (C order): momentum x, y, z, energy."""
    return np.ascontiguousarray(np.vstack([Rmom.T, Rener[None, :]]))


@dataclass
class ClosurePair:
    """An order-pt problem, its diffusion target and the mu_sgs = 0 residual."""
    coarse: Problem
    target: np.ndarray     # [4][npoin]
    r0: np.ndarray         # Rd_t(mu_sgs = 0), [4][npoin]
    weights: np.ndarray    # [4]
    meta: dict = field(default_factory=dict)

    @property
    def nsamples(self) -> int:
        return self.coarse.nelem * self.coarse.tables.nnode


def smooth_fields(prob: Problem, k: int = 2, mu_e: float = 5e-4) -> Problem:
    """Replaces the synthetic fields by ones without per-node noise: the
    Taylor-Green velocity of wavenumber k,
        u = (sin a x cos a y cos a z, -cos a x sin a y cos a z, 0),  a = 2 pi k,
    its pressure p = 1/gamma + (cos 2ax + cos 2ay)(cos 2az + 2)/16, a smooth
    density rho = 1 + 0.05 sin a x sin a y sin a z, q = rho u,
    E = p/(gamma-1) + rho|u|^2/2, T = p/(rho R), and a constant mu_e.  With the
    noise gone, what the order-pt residual misses of the order-pf one is the
    truncation error of under-resolved smooth flow, which an eddy viscosity
    may or may not represent."""
    x, y, z = (prob.coord[:, i] for i in range(3))
    a = 2.0 * np.pi * k
    u0 = np.sin(a * x) * np.cos(a * y) * np.cos(a * z)
    u1 = -np.cos(a * x) * np.sin(a * y) * np.cos(a * z)
    prob.u[:, 0] = u0
    prob.u[:, 1] = u1
    prob.u[:, 2] = 0.0
    prob.rho[:] = 1.0 + 0.05 * np.sin(a * x) * np.sin(a * y) * np.sin(a * z)
    prob.pr[:] = 1.0 / GAMMA + (np.cos(2 * a * x) + np.cos(2 * a * y)) * (np.cos(2 * a * z) + 2.0) / 16.0
    prob.q[:] = prob.rho[:, None] * prob.u
    prob.E[:] = prob.pr / (GAMMA - 1.0) + 0.5 * prob.rho * (u0 * u0 + u1 * u1)
    prob.Tem[:] = prob.pr / (prob.rho * RGAS)
    prob.mu_e[:] = mu_e
    return prob


def pair_seeds(seed: int):
    return (SEED_RHO + 17 * seed, SEED_MUE + 17 * seed, SEED_MUSGS + 17 * seed, SEED_JITTER)


def make_pair(n: int, seed: int = 0, pf: int = 7, pt: int = 3, rho_jump: float = 0.0,
              log=None, smooth_k: int = 0, smooth_mu_e: float = 5e-4) -> ClosurePair:
    """Training pair on n^3 affine hexahedra (see the module docstring);
    smooth_k > 0: the noise-free fields of smooth_fields at wavenumber
    smooth_k (with mu_e = smooth_mu_e) instead of the synthetic ones."""
    t0 = time.time()
    say = log or (lambda *_: None)
    seeds = pair_seeds(seed)
    fine = box_problem(n, pf, jitter=0.0, rho_jump=rho_jump, seeds=seeds)
    if smooth_k:
        smooth_fields(fine, smooth_k, mu_e=smooth_mu_e)
    say(f"  n={n}: order-{pf} problem built ({time.time() - t0:.1f} s)")
    fine.mu_sgs[:] = 0.0
    Rcf, Rdf = gpu_residuals(fine)
    say(f"  n={n}: order-{pf} residuals ({time.time() - t0:.1f} s)")
    coarse = box_problem(n, pt, jitter=0.0, rho_jump=rho_jump, seeds=seeds)
    nodal = np.column_stack([fine.u, fine.q, fine.rho, fine.pr, fine.E, fine.Tem, fine.mu_fluid])
    ft = prolong_fields(fine, coarse, nodal)
    coarse.u[:] = ft[:, 0:3]
    coarse.q[:] = ft[:, 3:6]
    coarse.rho[:] = ft[:, 6]
    coarse.pr[:] = ft[:, 7]
    coarse.E[:] = ft[:, 8]
    coarse.Tem[:] = ft[:, 9]
    coarse.mu_fluid[:] = ft[:, 10]
    coarse.mu_e[:] = prolong_element_field(fine, coarse, fine.mu_e)
    coarse.mu_sgs[:] = 0.0
    say(f"  n={n}: fields moved to order {pt} ({time.time() - t0:.1f} s)")
    rt = restrict_residual(fine, coarse, np.column_stack([Rcf[1] + Rdf[1], Rcf[2] + Rdf[2]]))
    say(f"  n={n}: residual restricted ({time.time() - t0:.1f} s)")
    del fine, Rcf, Rdf, nodal, ft
    Rct, Rd0 = gpu_residuals(coarse)                    # mu_sgs = 0
    target = _mom_ener(rt[:, 0:3], rt[:, 3]) - _mom_ener(Rct[1], Rct[2])
    r0 = _mom_ener(Rd0[1], Rd0[2])
    gap = ((r0 - target) ** 2).sum(axis=1)
    weights = 1.0 / np.maximum(gap, np.finfo(float).tiny)
    meta = dict(n=n, seed=seed, seeds=list(seeds), pf=pf, pt=pt, rho_jump=rho_jump, smooth_k=smooth_k,
                smooth_mu_e=smooth_mu_e if smooth_k else None,
                nelem=coarse.nelem, npoin=coarse.npoin, build_s=round(time.time() - t0, 2),
                fine_norm=np.linalg.norm(_mom_ener(rt[:, 0:3], rt[:, 3]), axis=1).tolist())
    if log:
        log(f"pair n={n} seed={seed}: {coarse.nelem * coarse.tables.nnode} samples, "
            f"{meta['build_s']} s")
    return ClosurePair(coarse, target, r0, weights, meta)


def planted_pair(prob: Problem, teacher: np.ndarray) -> ClosurePair:
    """A pair whose target is the diffusion residual with mu_sgs given by a
    known ("teacher") network: a closure that can be learnt exactly, to check
    the trainer end to end."""
    t = prob.tables
    prob.mu_sgs[:] = 0.0
    _, Rd0 = gpu_residuals(prob)
    feats = np.zeros((prob.nelem * t.nnode, NIN), np.float32)
    mu = np.zeros((prob.nelem, t.nnode), order="F")
    ins = [prob.connec, prob.He, prob.gpvol, prob.rho, prob.u]
    copyin(*ins)
    create(feats, mu)
    try:
        closure_features(prob.nelem, prob.npoin, prob.connec, prob.He, prob.gpvol, t.dlxigp_ip,
                         t.invAtoIJK, t.AtoI, t.AtoJ, t.AtoK, prob.rho, prob.u, feats)
        closure_predict(prob.nelem, t.invAtoIJK, teacher, feats, mu)
        update_host(mu)
    finally:
        delete(*ins, feats, mu)
    _, Rdt = gpu_residuals(prob, mu_sgs=mu)
    target = _mom_ener(Rdt[1], Rdt[2])
    r0 = _mom_ener(Rd0[1], Rd0[2])
    gap = ((r0 - target) ** 2).sum(axis=1)
    return ClosurePair(prob, target, r0, 1.0 / np.maximum(gap, np.finfo(float).tiny),
                       dict(planted=True, nelem=prob.nelem, npoin=prob.npoin))


# ---------------------------------------------------------------------------
# parameters and weight files
# ---------------------------------------------------------------------------
def init_params(features: np.ndarray, scale: float, b3: float = 0.0, seed: int = 0) -> np.ndarray:
    """Parameter block: standardisation from the training features, small
    random hidden layers and an output layer near the constant mu_sgs =
    scale * softplus(b3)."""
    rng = np.random.default_rng(seed)
    P = np.zeros(CLOSURE_NPARAM, dtype=np.float32)
    P[OW1:OB1] = rng.normal(0.0, 1.0 / np.sqrt(NIN), NIN * NH)
    P[OW2:OB2] = rng.normal(0.0, 1.0 / np.sqrt(NH), NH * NH)
    P[OW3:OB3] = rng.normal(0.0, 0.1 / np.sqrt(NH), NH)
    P[OB3] = b3
    f = features.astype(np.float64)
    mean = f.mean(axis=0)
    std = f.std(axis=0)
    P[OMEAN:OISTD] = mean
    P[OISTD:OSCALE] = np.where(std > 1e-12 * np.maximum(1.0, np.abs(mean)), 1.0 / np.maximum(std, 1e-300), 1.0)
    P[OSCALE] = scale
    return P


def _sha256(path: str) -> str:
    h = hashlib.sha256()
    with open(path, "rb") as fh:
        for chunk in iter(lambda: fh.read(1 << 20), b""):
            h.update(chunk)
    return h.hexdigest()


def save_weights(path: str, params: np.ndarray, m: np.ndarray | None = None,
                 v: np.ndarray | None = None, step: int = 0, meta: dict | None = None) -> str:
    """Writes path (.npz) and path + '.json' (format, sha256, meta)."""
    params = np.asarray(params, dtype=np.float32)
    if params.shape != (CLOSURE_NPARAM,):
        raise ValueError(f"params: {CLOSURE_NPARAM} floats expected")
    z = np.zeros(CLOSURE_NTRAIN, np.float32)
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    with open(path, "wb") as fh:
        np.savez(fh, params=params, adam_m=z if m is None else np.asarray(m, np.float32),
                 adam_v=z if v is None else np.asarray(v, np.float32), step=np.int64(step))
    man = dict(format=FORMAT, file=os.path.basename(path), sha256=_sha256(path),
               nparam=CLOSURE_NPARAM, ntrain=CLOSURE_NTRAIN, step=int(step), meta=meta or {})
    with open(path + ".json", "w") as fh:
        json.dump(man, fh, indent=1, sort_keys=True)
        fh.write("\n")
    return man["sha256"]


def load_weights(path: str = DEFAULT_WEIGHTS) -> dict:
    """Parameter block, Adam state and manifest of a weight file; raises
    ClosureModelError if the manifest is missing, the checksum or format
    differs, or the arrays are malformed."""
    try:
        with open(path + ".json") as fh:
            man = json.load(fh)
    except (OSError, ValueError) as e:
        raise ClosureModelError(f"{path}.json: {e}") from None
    if man.get("format") != FORMAT or man.get("nparam") != CLOSURE_NPARAM:
        raise ClosureModelError(f"{path}: not a {FORMAT} weight file")
    try:
        sha = _sha256(path)
    except OSError as e:
        raise ClosureModelError(str(e)) from None
    if sha != man.get("sha256"):
        raise ClosureModelError(f"{path}: sha256 {sha} differs from the manifest")
    try:
        with np.load(path, allow_pickle=False) as z:
            out = dict(params=z["params"].astype(np.float32), adam_m=z["adam_m"].astype(np.float32),
                       adam_v=z["adam_v"].astype(np.float32), step=int(z["step"]), manifest=man)
    except (OSError, ValueError, KeyError) as e:
        raise ClosureModelError(f"{path}: {e}") from None
    if out["params"].shape != (CLOSURE_NPARAM,) or out["adam_m"].shape != (CLOSURE_NTRAIN,):
        raise ClosureModelError(f"{path}: wrong array shapes")
    return out


# ---------------------------------------------------------------------------
# GPU trainer
# ---------------------------------------------------------------------------
class _Device:
    """Device-resident arrays of one pair (present table of the C ABI)."""

    def __init__(self, pair: ClosurePair):
        p = pair.coarse
        t = p.tables
        self.pair = pair
        self.nn = t.nnode
        self.E = p.nelem * self.nn
        self.features = np.zeros((self.E, NIN), dtype=np.float32)
        self.mu = np.zeros((p.nelem, self.nn), order="F")
        self.gnu = np.zeros(self.E, dtype=np.float32)
        self.rerr = np.zeros((4, p.npoin))
        self.R = p.outputs()
        self.ins = [p.connec, p.He, p.gpvol, p.u, p.rho, p.Tem, p.mu_fluid, p.mu_e, pair.target]
        copyin(*self.ins)
        create(self.features, self.mu, self.gnu, self.rerr, *self.R)
        closure_features(p.nelem, p.npoin, p.connec, p.He, p.gpvol, t.dlxigp_ip, t.invAtoIJK,
                         t.AtoI, t.AtoJ, t.AtoK, p.rho, p.u, self.features)
        update_host(self.features)

    def close(self):
        delete(*self.ins, self.features, self.mu, self.gnu, self.rerr, *self.R)

    def loss(self, params, with_grad: bool, grad_acc=None, grad_tmp=None) -> float:
        """Loss of the pair; with_grad: adds d loss / d params into grad_acc."""
        p = self.pair.coarse
        t = p.tables
        closure_predict(p.nelem, t.invAtoIJK, params, self.features, self.mu)
        full_diffusion_ijk(*p.diffu_args(*self.R, mu_sgs=self.mu))
        L = closure_loss(p.npoin, self.R[1], self.R[2], self.pair.target, self.pair.weights,
                         self.rerr)
        if with_grad:
            closure_sensitivity(p.nelem, p.npoin, p.connec, p.He, p.gpvol, t.dlxigp_ip,
                                t.invAtoIJK, t.AtoI, t.AtoJ, t.AtoK, p.Cp, p.rho, p.u, p.Tem,
                                self.rerr, self.gnu)
            closure_gradient(self.E, params, self.features, self.gnu, grad_tmp)
        return L

    def constant_optimum(self) -> float:
        """Least-squares constant mu_sgs of the pair (the loss is quadratic in
        a constant mu_sgs since the residual is affine in it)."""
        return self.constant_fit()[0]

    def constant_fit(self) -> tuple[float, float]:
        """(c*, gain at c*): with L1 = Rd(mu_sgs = 1) - Rd(0), the loss of a
        constant c is 4 - 2c num + c^2 den, num = sum w (target - r0) L1,
        den = sum w L1^2; its minimum over c is at num/den with gain
        num^2 / (4 den).  A closure, whose mu_sgs is >= 0, can only use a
        positive c*."""
        p = self.pair.coarse
        one = np.ones((p.nelem, self.nn), order="F")
        copyin(one)
        try:
            full_diffusion_ijk(*p.diffu_args(*self.R, mu_sgs=one))
            update_host(*self.R)
        finally:
            delete(one)
        L1 = _mom_ener(self.R[1], self.R[2]) - self.pair.r0
        w = self.pair.weights[:, None]
        num = float((w * (self.pair.target - self.pair.r0) * L1).sum())
        den = float((w * L1 * L1).sum())
        if den <= 0:
            return 0.0, 0.0
        return num / den, num * num / (4.0 * den)


def train(pairs: list[ClosurePair], steps: int = 600, lr: float = 3e-3, seed: int = 0,
          init: dict | None = None, log=None, val_pairs: list[ClosurePair] | None = None,
          eval_every: int = 25) -> dict:
    """Adam on the sum of the pairs' losses, full batch, all on the GPU.

    With val_pairs, the parameters of the lowest validation loss are kept
    (early stopping); the mu_sgs = 0 model scores 4 per validation pair.
    Returns params, Adam state, the training loss history and the
    validation history [(step, loss)]."""
    devs = [_Device(p) for p in pairs]
    vdevs = []
    try:
        vdevs = [_Device(p) for p in (val_pairs or [])]
        cstar = None
        if init is not None:
            params = init["params"].copy()
            m, v, step0 = init["adam_m"].copy(), init["adam_v"].copy(), init["step"]
        else:
            feats = np.concatenate([d.features for d in devs])
            cstar = [d.constant_optimum() for d in devs]
            c = float(np.mean(cstar))
            # start near the best constant viscosity when one helps, near 0 otherwise
            scale = abs(c) / np.log(2.0) if c > 0 else max(abs(c), 1e-8)
            params = init_params(feats, scale=scale, b3=0.0 if c > 0 else -6.0, seed=seed)
            m = np.zeros(CLOSURE_NTRAIN, np.float32)
            v = np.zeros(CLOSURE_NTRAIN, np.float32)
            step0 = 0
            if log:
                log(f"constant optimum mu_sgs per pair: {cstar}; scale {scale:.3e}")
        grad = np.zeros(CLOSURE_NTRAIN, np.float32)
        gtmp = np.zeros(CLOSURE_NTRAIN, np.float32)
        copyin(params, m, v, grad, gtmp)
        hist, vhist = [], []
        best = None
        t0 = time.time()

        def validate(step):
            nonlocal best
            Lv = float(sum(d.loss(params, False) for d in vdevs))
            vhist.append((step, Lv))
            if best is None or Lv < best[0]:
                update_host(params)
                best = (Lv, params.copy(), step)
            return Lv

        try:
            if vdevs:
                validate(step0)
            for it in range(1, steps + 1):
                L = 0.0
                if len(devs) == 1:
                    L = devs[0].loss(params, True, grad_tmp=grad)
                else:
                    acc = np.zeros(CLOSURE_NTRAIN, np.float64)
                    for d in devs:
                        L += d.loss(params, True, grad_tmp=gtmp)
                        update_host(gtmp)
                        acc += gtmp
                    grad[:] = acc.astype(np.float32)
                    update_device(grad)
                closure_adam(params, grad, m, v, step0 + it, lr)
                hist.append(L)
                Lv = validate(step0 + it) if vdevs and (it % eval_every == 0 or it == steps) else None
                if log and (it == 1 or it % 50 == 0 or it == steps):
                    extra = f", validation {Lv:.6f}" if Lv is not None else ""
                    log(f"step {step0 + it}: loss {L:.6f}{extra} ({time.time() - t0:.1f} s)")
            update_host(params, m, v)
        finally:
            delete(params, m, v, grad, gtmp)
        out = dict(params=params, adam_m=m, adam_v=v, step=step0 + steps, history=hist,
                   val_history=vhist, last_params=params.copy(), constant_optimum=cstar)
        if best is not None:
            out["params"] = best[1]
            out["best_step"] = best[2]
            out["best_val"] = best[0]
        return out
    finally:
        for d in devs + vdevs:
            d.close()


def evaluate(pair: ClosurePair, params: np.ndarray | None) -> dict:
    """Loss of the pair with the model (or mu_sgs = 0 for params None), per
    component relative to mu_sgs = 0, through the fused closure kernel."""
    p = pair.coarse
    t = p.tables
    R = p.outputs()
    ins = [p.connec, p.He, p.gpvol, p.rho, p.u, p.Tem, p.mu_fluid, p.mu_e]
    copyin(*ins)
    create(*R)
    try:
        used = full_diffusion_ijk_closure(p.nelem, p.npoin, p.connec, t.Ngp, t.dNgp, p.He, p.gpvol,
                                          t.dlxigp_ip, t.xgp, t.invAtoIJK, t.AtoI, t.AtoJ, t.AtoK,
                                          p.Cp, p.Pr, p.rho, p.u, p.Tem, p.mu_fluid, p.mu_e,
                                          params, *R)
        update_host(*R)
    finally:
        delete(*ins, *R)
    r = _mom_ener(R[1], R[2])
    per = (pair.weights * ((r - pair.target) ** 2).sum(axis=1))
    return dict(used_model=used, loss=float(per.sum()), per_component=per.tolist(),
                gain=float(1.0 - per.sum() / 4.0))


def constant_fit(pair: ClosurePair) -> tuple[float, float]:
    """Least-squares constant mu_sgs of a pair and its gain (_Device.constant_fit)."""
    d = _Device(pair)
    try:
        return d.constant_fit()
    finally:
        d.close()


def diagnose(k: int, sizes, val_sizes, held_out, steps: int = 300, lr: float = 3e-3, seed: int = 0,
             log=None, mu_e: float = 5e-4) -> dict:
    """Noise-free pairs (smooth_fields at wavenumber k): the least-squares
    constant mu_sgs and its gain on every pair, then a network trained on the
    pairs of `sizes` (early stopping on `val_sizes`) and its gain on the
    held-out sizes, which it never saw."""
    mk = lambda ns, off: [make_pair(n, seed=seed + off + i, smooth_k=k, smooth_mu_e=mu_e, log=log)
                          for i, n in enumerate(ns)]
    groups = {"train": mk(sizes, 0), "val": mk(val_sizes, 100), "held_out": mk(held_out, 200)}
    const = {g: [dict(n=p.meta["n"], c=c, gain=gn) for p in ps for c, gn in [constant_fit(p)]]
             for g, ps in groups.items()}
    res = train(groups["train"], steps=steps, lr=lr, seed=seed, log=log, val_pairs=groups["val"] or None)
    gain = {g: [dict(n=p.meta["n"], gain=evaluate(p, res["params"])["gain"]) for p in ps]
            for g, ps in groups.items()}
    return dict(k=k, mu_e=mu_e, constant=const, model=gain, best_step=res.get("best_step"),
                loss_first=res["history"][0], loss_last=res["history"][-1], params=res["params"],
                adam_m=res["adam_m"], adam_v=res["adam_v"], step=res["step"])


def report(pair: ClosurePair, params: np.ndarray, repeats: int = 10) -> dict:
    """Error against the projected fine residual with the closure and with
    mu_sgs = 0, and the cost of the fused closure kernel against the plain
    diffusion kernel (mu_sgs read from memory) on the same order-pt problem."""
    from . import (TIMER_DIFFU, TIMER_DIFFU_MLP, set_async, synchronize, timing_enable,
                   timing_get, timing_reset)
    ev = evaluate(pair, params)
    ev0 = evaluate(pair, None)
    p = pair.coarse
    t = p.tables
    R = p.outputs()
    ins = [p.connec, p.He, p.gpvol, p.rho, p.u, p.Tem, p.mu_fluid, p.mu_e, p.mu_sgs]
    copyin(*ins, params)
    create(*R)
    try:
        def run(fused):
            if fused:
                full_diffusion_ijk_closure(p.nelem, p.npoin, p.connec, t.Ngp, t.dNgp, p.He, p.gpvol,
                                           t.dlxigp_ip, t.xgp, t.invAtoIJK, t.AtoI, t.AtoJ, t.AtoK,
                                           p.Cp, p.Pr, p.rho, p.u, p.Tem, p.mu_fluid, p.mu_e, params,
                                           *R)
            else:
                full_diffusion_ijk(*p.diffu_args(*R))
        cost = {}
        for fused, timer in ((False, TIMER_DIFFU), (True, TIMER_DIFFU_MLP)):
            run(fused)
            synchronize()
            timing_reset()
            timing_enable(True)
            for _ in range(repeats):
                run(fused)
            synchronize()
            timing_enable(False)
            ms, nl = timing_get(timer)
            cost["closure_ms" if fused else "plain_ms"] = ms / max(repeats, 1)
    finally:
        delete(*ins, params, *R)
    # relative L2 error of the order-pt residual against the projected fine
    # residual, per component (momentum x, y, z, energy)
    fine = np.asarray(pair.meta.get("fine_norm", [np.nan] * 4))
    err = lambda per: (np.sqrt(np.asarray(per) / pair.weights) / fine).tolist()
    return dict(n=pair.meta.get("n"), nelem=p.nelem, npoin=p.npoin,
                loss_model=ev["loss"], loss_zero=ev0["loss"], gain=ev["gain"],
                gain_per_component=[1 - x for x in ev["per_component"]],
                rel_err_vs_projected_fine_model=err(ev["per_component"]),
                rel_err_vs_projected_fine_zero=err(ev0["per_component"]),
                **cost, cost_ratio=cost["closure_ms"] / cost["plain_ms"])


# ---------------------------------------------------------------------------
# command line: python -m minisod2d_amd.closure train|evaluate|report ...
# ---------------------------------------------------------------------------
def _main(argv=None) -> int:
    from . import init as dev_init
    ap = argparse.ArgumentParser(prog="python -m minisod2d_amd.closure")
    sub = ap.add_subparsers(dest="cmd", required=True)
    tr = sub.add_parser("train")
    tr.add_argument("--sizes", type=int, nargs="+", default=[24, 32, 40])
    tr.add_argument("--seed", type=int, default=0)
    tr.add_argument("--steps", type=int, default=600)
    tr.add_argument("--lr", type=float, default=3e-3)
    tr.add_argument("--val-sizes", type=int, nargs="*", default=[16, 20])
    tr.add_argument("--out", default=DEFAULT_WEIGHTS)
    tr.add_argument("--resume", action="store_true")
    rp = sub.add_parser("report", help="gain and cost of the closure on one n^3 pair")
    rp.add_argument("--weights", default=DEFAULT_WEIGHTS)
    rp.add_argument("--n", type=int, default=96)
    rp.add_argument("--seed", type=int, default=50)
    rp.add_argument("--repeats", type=int, default=10)
    dg = sub.add_parser("diagnose", help="noise-free pairs: constant optimum and trained gain")
    dg.add_argument("--k", type=int, nargs="+", default=[1, 2, 3])
    dg.add_argument("--sizes", type=int, nargs="+", default=[12, 16])
    dg.add_argument("--val-sizes", type=int, nargs="*", default=[10])
    dg.add_argument("--held-out", type=int, nargs="+", default=[20])
    dg.add_argument("--steps", type=int, default=300)
    dg.add_argument("--lr", type=float, default=3e-3)
    dg.add_argument("--seed", type=int, default=0)
    dg.add_argument("--mu-e", type=float, default=5e-4, help="constant mu_e of the noise-free fields")
    dg.add_argument("--out", default=None, help="weight file for the model of the last k")
    ev = sub.add_parser("evaluate")
    ev.add_argument("--weights", default=DEFAULT_WEIGHTS)
    ev.add_argument("--sizes", type=int, nargs="+", default=[28])
    ev.add_argument("--seed", type=int, default=7)
    a = ap.parse_args(argv)
    log = lambda s: print(s, flush=True)
    dev_init(0)
    if a.cmd == "train":
        pairs = [make_pair(n, seed=a.seed + i, log=log) for i, n in enumerate(a.sizes)]
        val = [make_pair(n, seed=a.seed + 100 + i, log=log) for i, n in enumerate(a.val_sizes)]
        init = load_weights(a.out) if a.resume else None
        res = train(pairs, steps=a.steps, lr=a.lr, seed=a.seed, init=init, log=log,
                    val_pairs=val or None)
        evals = [evaluate(p, res["params"]) for p in pairs]
        vevals = [evaluate(p, res["params"]) for p in val]
        meta = dict(sizes=a.sizes, seed=a.seed, steps=res["step"], lr=a.lr,
                    samples=int(sum(p.nsamples for p in pairs)),
                    loss_first=res["history"][0], loss_last=res["history"][-1],
                    best_step=res.get("best_step"), val_sizes=a.val_sizes,
                    train_gain=[e["gain"] for e in evals], val_gain=[e["gain"] for e in vevals],
                    constant_optimum=res.get("constant_optimum"))
        sha = save_weights(a.out, res["params"], res["adam_m"], res["adam_v"], res["step"], meta)
        print(json.dumps(dict(out=a.out, sha256=sha, **meta)), flush=True)
    elif a.cmd == "diagnose":
        for k in a.k:
            r = diagnose(k, a.sizes, a.val_sizes, a.held_out, steps=a.steps, lr=a.lr, seed=a.seed, log=log,
                         mu_e=a.mu_e)
            if a.out and k == a.k[-1]:
                meta = dict(smooth_k=k, sizes=a.sizes, val_sizes=a.val_sizes, held_out=a.held_out,
                            steps=r["step"], lr=a.lr, constant=r["constant"], model=r["model"])
                r["sha256"] = save_weights(a.out, r["params"], r["adam_m"], r["adam_v"], r["step"], meta)
            for key in ("params", "adam_m", "adam_v"):
                r.pop(key)
            print(json.dumps(r), flush=True)
    elif a.cmd == "report":
        w = load_weights(a.weights)
        pair = make_pair(a.n, seed=a.seed, log=log)
        print(json.dumps(report(pair, w["params"], a.repeats)), flush=True)
    else:
        w = load_weights(a.weights)
        for i, n in enumerate(a.sizes):
            pair = make_pair(n, seed=a.seed + i, log=log)
            print(json.dumps(dict(n=n, seed=a.seed + i, model=evaluate(pair, w["params"]),
                                  zero=evaluate(pair, None))), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(_main())

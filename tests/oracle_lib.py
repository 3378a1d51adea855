"""ctypes bindings of the checkers under oracle/ (test infrastructure).

* ``Oracle``: the C restatement of the reference algorithm
  (oracle/msod_oracle.c, built into oracle/_build/libmsod_oracle.so).
* ``RefLib``: the reference's own Fortran kernels compiled by oracle/Makefile
  into oracle/_ref/libsodref_pP.so (only where they have been built).
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "_build", "libmsod_oracle.so")
REF_DIR = os.path.join(ROOT, "oracle", "_ref")

_vp = ctypes.c_void_p
_d = ctypes.c_double


def _p(a):
    return None if a is None else a.ctypes.data


def _f(shape, dtype=np.float64):
    return np.zeros(shape, dtype=dtype, order="F")


class Oracle:
    def __init__(self):
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle not built: {ORACLE_SO} (make -C oracle oracle)")
        L = ctypes.CDLL(ORACLE_SO, mode=ctypes.RTLD_LOCAL)
        L.msod_oracle_convec.argtypes = [ctypes.c_int] * 3 + [_vp] * 16 + [ctypes.c_int]
        L.msod_oracle_diffu.argtypes = [ctypes.c_int] * 3 + [_vp] * 8 + [_d] * 4 + [_vp] * 9 + [ctypes.c_int]
        L.msod_oracle_mlp_forward.argtypes = [_vp, _vp, ctypes.c_int64, ctypes.c_int, _vp, _vp]
        L.msod_oracle_mlp_gradient.argtypes = [_vp, _vp, _vp, ctypes.c_int64, ctypes.c_int, _vp]
        L.msod_oracle_adam.argtypes = [_vp] * 4 + [ctypes.c_int, _d, _d, _d, _d]
        L.msod_oracle_tet_tables.argtypes = [_vp] * 4
        L.msod_oracle_tet_jacobian.argtypes = [ctypes.c_int] * 2 + [_vp] * 6
        L.msod_oracle_fem_convec.argtypes = [ctypes.c_int] * 2 + [_vp] * 13
        L.msod_oracle_fem_diffu.argtypes = [ctypes.c_int] * 2 + [_vp] * 5 + [_d] * 4 + [_vp] * 9
        self.L = L

    def gll(self, p, exact=False):
        xi, w = np.zeros(p + 1), np.zeros(p + 1)
        fn = self.L.msod_oracle_gll_double if exact else self.L.msod_oracle_gll
        fn(p, _vp(xi.ctypes.data), _vp(w.ctypes.data))
        return xi, w

    def tables(self, p):
        q = p + 1
        nn = q ** 3
        t = [_f((nn, 3), np.int32), _f((q, q, q), np.int32), _f(nn, np.int32), _f(nn, np.int32),
             _f(nn, np.int32), _f(nn, np.int32), _f((nn, 3)), _f(nn), _f((nn, nn)),
             _f((3, nn, nn)), _f((nn, 3, q))]
        self.L.msod_oracle_tables(p, *[_vp(x.ctypes.data) for x in t])
        return dict(zip("gmshIJK invAtoIJK AtoIJK AtoI AtoJ AtoK xgp wgp Ngp dNgp dlxigp_ip".split(), t))

    def jacobian(self, p, nelem, npoin, connec, coord, dNgp, wgp):
        nn = (p + 1) ** 3
        He, gp = _f((3, 3, nn, nelem)), _f((1, nn, nelem))
        self.L.msod_oracle_elem_jacobian(p, nelem, npoin, *[_vp(_p(a)) for a in (connec, coord, dNgp, wgp, gp, He)])
        return He, gp

    def convec(self, prob, nthreads=1):
        t = prob.tables
        R = prob.outputs()
        self.L.msod_oracle_convec(prob.porder, prob.nelem, prob.npoin, *map(_p, (
            prob.connec, prob.He, prob.gpvol, t.dlxigp_ip, t.invAtoIJK, t.AtoI, t.AtoJ, t.AtoK,
            prob.u, prob.q, prob.rho, prob.pr, prob.E)), *map(_p, R), nthreads)
        return R

    def diffu(self, prob, nthreads=1, mu_sgs=None, c_rho=1.0, c_ener=1.0):
        t = prob.tables
        R = prob.outputs()
        self.L.msod_oracle_diffu(prob.porder, prob.nelem, prob.npoin, *map(_p, (
            prob.connec, prob.He, prob.gpvol, t.dlxigp_ip, t.invAtoIJK, t.AtoI, t.AtoJ, t.AtoK)),
            prob.Cp, prob.Pr, c_rho, c_ener, *map(_p, (
                prob.rho, prob.u, prob.Tem, prob.mu_fluid, prob.mu_e,
                prob.mu_sgs if mu_sgs is None else mu_sgs)), *map(_p, R), nthreads)
        return R


    # ---- linear tetrahedra ----------------------------------------------------
    def tet_tables(self):
        t = [_f((4, 3)), _f(4), _f((4, 4)), _f((3, 4, 4))]
        self.L.msod_oracle_tet_tables(*[_vp(x.ctypes.data) for x in t])
        return dict(zip("xgp wgp Ngp dNgp".split(), t))

    def tet_jacobian(self, nelem, npoin, connec, coord, dNgp, wgp):
        He, gp = _f((3, 3, 4, nelem)), _f((1, 4, nelem))
        self.L.msod_oracle_tet_jacobian(nelem, npoin, *map(_p, (connec, coord, dNgp, wgp, gp, He)))
        return He, gp

    def fem_convec(self, tp):
        R = tp.outputs()
        self.L.msod_oracle_fem_convec(tp.nelem, tp.npoin, *map(_p, (
            tp.connec, tp.tables.Ngp, tp.tables.dNgp, tp.He, tp.gpvol, tp.u, tp.q, tp.rho, tp.pr,
            tp.E)), *map(_p, R))
        return R

    def fem_diffu(self, tp, c_rho=1.0, c_ener=1.0, Ngp=None):
        R = tp.outputs()
        self.L.msod_oracle_fem_diffu(tp.nelem, tp.npoin, *map(_p, (
            tp.connec, tp.tables.Ngp if Ngp is None else Ngp, tp.tables.dNgp, tp.He, tp.gpvol)),
            tp.Cp, tp.Pr, c_rho, c_ener, *map(_p, (
                tp.rho, tp.u, tp.Tem, tp.mu_fluid, tp.mu_e, tp.mu_sgs)), *map(_p, R))
        return R

    # ---- learned closure (oracle/msod_oracle_mlp.c) ---------------------------
    def mlp_forward(self, P, feat, quant=1):
        """(z3, mu_sgs) per sample."""
        P = np.ascontiguousarray(P, np.float32)
        feat = np.ascontiguousarray(feat, np.float32)
        n = feat.shape[0]
        z3, mu = np.zeros(n), np.zeros(n)
        self.L.msod_oracle_mlp_forward(P.ctypes.data, feat.ctypes.data, n, quant, z3.ctypes.data,
                                       mu.ctypes.data)
        return z3, mu

    def mlp_gradient(self, P, feat, gnu, quant=1):
        P = np.ascontiguousarray(P, np.float32)
        feat = np.ascontiguousarray(feat, np.float32)
        gnu = np.ascontiguousarray(gnu, np.float64)
        g = np.zeros(1505)
        self.L.msod_oracle_mlp_gradient(P.ctypes.data, feat.ctypes.data, gnu.ctypes.data,
                                        feat.shape[0], quant, g.ctypes.data)
        return g

    def adam(self, P, grad, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
        """In place on float32 arrays P, m, v."""
        grad = np.ascontiguousarray(grad, np.float32)
        self.L.msod_oracle_adam(P.ctypes.data, grad.ctypes.data, m.ctypes.data, v.ctypes.data,
                                step, lr, b1, b2, eps)


def ref_available(p):
    return os.path.exists(os.path.join(REF_DIR, f"libsodref_p{p}.so"))


class RefLib:
    """The reference's Fortran kernels (rp=8) for one porder."""

    def __init__(self, p):
        path = os.path.join(REF_DIR, f"libsodref_p{p}.so")
        if not os.path.exists(path):
            raise RuntimeError(f"reference library not built: {path}")
        L = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
        assert L.sodref_porder() == p and L.sodref_rp() == 8
        L.sodref_full_convec_ijk.argtypes = [ctypes.c_int] * 2 + [_vp] * 19
        L.sodref_full_convec_ijk_split.argtypes = [ctypes.c_int] * 2 + [_vp] * 19
        L.sodref_full_diffusion_ijk.argtypes = [ctypes.c_int] * 2 + [_vp] * 11 + [_d] * 4 + [_vp] * 9
        L.sodref_elem_jacobian.argtypes = [ctypes.c_int] * 2 + [_vp] * 6
        L.sodref_create_mesh.argtypes = [ctypes.c_int] * 2 + [_vp] * 3
        L.sodref_tet_tables.argtypes = [_vp] * 4
        L.sodref_gen_tet_mesh.argtypes = [ctypes.c_int] * 2 + [_vp] * 3
        L.sodref_tet_jacobian.argtypes = [ctypes.c_int] * 2 + [_vp] * 6
        L.sodref_fem_convec.argtypes = [ctypes.c_int] * 2 + [_vp] * 13
        L.sodref_fem_diffu.argtypes = [ctypes.c_int] * 2 + [_vp] * 5 + [_d] * 4 + [_vp] * 9
        self.L = L
        self.p = p

    def tables(self):
        q = self.p + 1
        nn = q ** 3
        t = [_f((nn, 3), np.int32), _f((q, q, q), np.int32), _f(nn, np.int32), _f(nn, np.int32),
             _f(nn, np.int32), _f(nn, np.int32), _f((nn, 3)), _f(nn), _f((nn, nn)),
             _f((3, nn, nn)), _f((nn, 3, q))]
        self.L.sodref_tables(*[_vp(x.ctypes.data) for x in t])
        return dict(zip("gmshIJK invAtoIJK AtoIJK AtoI AtoJ AtoK xgp wgp Ngp dNgp dlxigp_ip".split(), t))

    def jacobian(self, nelem, npoin, connec, coord, dNgp, wgp):
        nn = (self.p + 1) ** 3
        He, gp = _f((3, 3, nn, nelem)), _f((1, nn, nelem))
        self.L.sodref_elem_jacobian(nelem, npoin, *map(_p, (connec, coord, dNgp, wgp, gp, He)))
        return He, gp

    def _cargs(self, prob):
        t = prob.tables
        return (prob.nelem, prob.npoin, _p(prob.connec), _p(t.Ngp), _p(t.dNgp), _p(prob.He),
                _p(prob.gpvol), _p(t.dlxigp_ip), _p(t.xgp), _p(t.invAtoIJK), _p(t.AtoI),
                _p(t.AtoJ), _p(t.AtoK), _p(prob.u), _p(prob.q), _p(prob.rho), _p(prob.pr),
                _p(prob.E))

    def convec(self, prob):
        """Rmass, Rmom of full_convec_ijk; Rener of full_convec_ijk_split."""
        Rf = prob.outputs()
        Rs = prob.outputs()
        self.L.sodref_full_convec_ijk(*self._cargs(prob), *map(_p, Rf))
        self.L.sodref_full_convec_ijk_split(*self._cargs(prob), *map(_p, Rs))
        return Rf[0], Rf[1], Rs[2], Rs

    def diffu(self, prob, c_rho=1.0, c_ener=1.0):
        t = prob.tables
        R = prob.outputs()
        self.L.sodref_full_diffusion_ijk(
            prob.nelem, prob.npoin, *map(_p, (prob.connec, t.Ngp, t.dNgp, prob.He, prob.gpvol,
                                              t.dlxigp_ip, t.xgp, t.invAtoIJK, t.AtoI, t.AtoJ, t.AtoK)),
            prob.Cp, prob.Pr, c_rho, c_ener,
            *map(_p, (prob.rho, prob.u, prob.Tem, prob.mu_fluid, prob.mu_e, prob.mu_sgs)), *map(_p, R))
        return R


class RefLib4:
    """The reference's Fortran kernels at its default real kind rp=4
    (mod_constants.f90:5): every real array is float32; the tables come back
    in float64 holding float32 values."""

    def __init__(self, p):
        path = os.path.join(REF_DIR, "r4", f"libsodref_r4_p{p}.so")
        if not os.path.exists(path):
            raise RuntimeError(f"reference library not built: {path}")
        L = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
        assert L.sodref_porder() == p and L.sodref_rp() == 4
        L.sodref_full_convec_ijk.argtypes = [ctypes.c_int] * 2 + [_vp] * 19
        L.sodref_full_convec_ijk_split.argtypes = [ctypes.c_int] * 2 + [_vp] * 19
        L.sodref_full_diffusion_ijk.argtypes = [ctypes.c_int] * 2 + [_vp] * 11 + [_d] * 4 + [_vp] * 9
        self.L = L
        self.p = p

    def tables(self):
        return RefLib.tables(self)

    @staticmethod
    def _f32(a):
        assert a.dtype == np.float32 and a.flags.f_contiguous, a.dtype
        return _p(a)

    def convec(self, c):
        """c: dict of float32/int32 Fortran arrays (names of the reference's
        arguments).  Rmass, Rmom of full_convec_ijk, Rener of the split kernel."""
        n = c["npoin"]
        out = lambda shape: np.zeros(shape, np.float32, order="F")
        Rf = [out(n), out((n, 3)), out(n)]
        Rs = [out(n), out((n, 3)), out(n)]
        args = (c["nelem"], n, _p(c["connec"]), *map(self._f32, (c["Ngp"], c["dNgp"], c["He"], c["gpvol"],
                                                                  c["dlxigp_ip"], c["xgp"])),
                _p(c["invAtoIJK"]), _p(c["AtoI"]), _p(c["AtoJ"]), _p(c["AtoK"]),
                *map(self._f32, (c["u"], c["q"], c["rho"], c["pr"], c["E"])))
        self.L.sodref_full_convec_ijk(*args, *map(_p, Rf))
        self.L.sodref_full_convec_ijk_split(*args, *map(_p, Rs))
        return Rf[0], Rf[1], Rs[2]

    def diffu(self, c, c_rho=1.0, c_ener=1.0):
        n = c["npoin"]
        R = [np.zeros(n, np.float32, order="F"), np.zeros((n, 3), np.float32, order="F"),
             np.zeros(n, np.float32, order="F")]
        self.L.sodref_full_diffusion_ijk(
            c["nelem"], n, _p(c["connec"]), *map(self._f32, (c["Ngp"], c["dNgp"], c["He"], c["gpvol"],
                                                             c["dlxigp_ip"], c["xgp"])),
            _p(c["invAtoIJK"]), _p(c["AtoI"]), _p(c["AtoJ"]), _p(c["AtoK"]), c["Cp"], c["Pr"], c_rho, c_ener,
            *map(self._f32, (c["rho"], c["u"], c["Tem"], c["mu_fluid"], c["mu_e"], c["mu_sgs"])), *map(_p, R))
        return R


R4_FIELDS = ("He", "gpvol", "u", "q", "rho", "pr", "E", "Tem", "mu_fluid", "mu_e", "mu_sgs")
R4_TABLES = ("Ngp", "dNgp", "dlxigp_ip", "xgp")


def r4_inputs(prob, tables):
    """The arguments a caller built at rp=4 passes: the problem's arrays
    rounded to float32, the rp=4 reference's own tables (float32 values)."""
    f32 = lambda a: np.asfortranarray(a, dtype=np.float32)
    c = dict(nelem=prob.nelem, npoin=prob.npoin, Cp=prob.Cp, Pr=prob.Pr, connec=prob.connec)
    for k in R4_FIELDS:
        c[k] = f32(getattr(prob, k))
    for k in R4_TABLES:
        c[k] = f32(tables[k])
    for k in ("invAtoIJK", "AtoI", "AtoJ", "AtoK"):
        c[k] = np.asfortranarray(tables[k], dtype=np.int32)
    return c


class RefTet:
    """The reference's TET04 routines (any porder's library holds them)."""

    def __init__(self, p=4):
        self.R = RefLib(p)
        self.L = self.R.L

    def tables(self):
        t = [_f((4, 3)), _f(4), _f((4, 4)), _f((3, 4, 4))]
        self.L.sodref_tet_tables(*[_vp(x.ctypes.data) for x in t])
        return dict(zip("xgp wgp Ngp dNgp".split(), t))

    def mesh(self, nelem, npoin, base):
        connec, xyz = _f((nelem, 4), np.int32), _f((npoin, 3))
        self.L.sodref_gen_tet_mesh(nelem, npoin, _p(np.asfortranarray(base)), _p(connec), _p(xyz))
        return connec, xyz

    def jacobian(self, nelem, npoin, connec, coord, dNgp, wgp):
        He, gp = _f((3, 3, 4, nelem)), _f((1, 4, nelem))
        self.L.sodref_tet_jacobian(nelem, npoin, *map(_p, (connec, coord, dNgp, wgp, gp, He)))
        return He, gp

    def fem_convec(self, tp):
        R = tp.outputs()
        self.L.sodref_fem_convec(tp.nelem, tp.npoin, *map(_p, (
            tp.connec, tp.tables.Ngp, tp.tables.dNgp, tp.He, tp.gpvol, tp.u, tp.q, tp.rho, tp.pr,
            tp.E)), *map(_p, R))
        return R

    def fem_diffu(self, tp, c_rho=1.0, c_ener=1.0):
        R = tp.outputs()
        self.L.sodref_fem_diffu(tp.nelem, tp.npoin, *map(_p, (
            tp.connec, tp.tables.Ngp, tp.tables.dNgp, tp.He, tp.gpvol)), tp.Cp, tp.Pr, c_rho, c_ener,
            *map(_p, (tp.rho, tp.u, tp.Tem, tp.mu_fluid, tp.mu_e, tp.mu_sgs)), *map(_p, R))
        return R


def rel_l2(a, b):
    nb = np.linalg.norm(np.ravel(b))
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / (nb if nb > 0 else 1.0))


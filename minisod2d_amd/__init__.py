"""MI355X-native spectral-element residual kernels of miniSOD2D.

Python host mirror of the reference's kernel interface: the functions below
take the same arguments, in the same order and Fortran layouts, as
``full_convec_ijk`` (src/kernels/sources/elem_convec.f90:198) and
``full_diffusion_ijk`` (src/kernels/sources/elem_diffu.f90:10), and run the
hand-written HIP kernels of ``libmsod2d_hip.so`` through its C ABI
(include/msod2d.h).  There is no CPU fallback: without the built library or
a GPU every call raises.

Arrays are numpy arrays in Fortran order.  Like the reference, which expects
every array to be `present` on the device (OpenACC ``enter data`` in the
driver, src/app/sources/main.f90:197-425), arrays must be mirrored first with
:func:`copyin` / :func:`create`; results come back with :func:`update_host`.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(_HERE, "lib")
# MSOD2D_HIP_LIB: another build of the kernel library (same ABI), e.g. for tuning
HIP_LIB = os.environ.get("MSOD2D_HIP_LIB") or os.path.join(LIB_DIR, "libmsod2d_hip.so")
HOST_LIB = os.path.join(LIB_DIR, "libmsod_host.so")

# status codes (include/msod2d.h)
MSOD_OK = 0
MSOD_ERR_ARG = 1
MSOD_ERR_NOT_PRESENT = 2
MSOD_ERR_PORDER = 3
MSOD_ERR_TABLE = 4
MSOD_ERR_HIP = 5
MSOD_ERR_STATE = 6
MSOD_ERR_MESH = 7
MSOD_ERR_MODEL = 8

# timer ids of msod_timing_get
(TIMER_ZERO, TIMER_CONVEC, TIMER_DIFFU, TIMER_CONVEC_MIXED, TIMER_DIFFU_MIXED, TIMER_DIFFU_MLP,
 TIMER_CLOSURE_TRAIN, TIMER_CONVEC_GEO32, TIMER_DIFFU_GEO32, TIMER_GEOMETRY,
 TIMER_GENERATOR, TIMER_FEM_CONVEC, TIMER_FEM_DIFFU, TIMER_TET_GEOMETRY, TIMER_CONVEC_FP32,
 TIMER_DIFFU_FP32, TIMER_CONVEC_R4, TIMER_DIFFU_R4, TIMER_CONVEC_F32SPLIT,
 TIMER_DIFFU_F32SPLIT) = range(20)

# the C-ABI entry points declared in include/msod2d.h (checked by the tests)
ABI_SYMBOLS = (
    "msod_init", "msod_finalize", "msod_status_string", "msod_last_error",
    "msod_data_copyin", "msod_data_create", "msod_data_update_device",
    "msod_data_update_host", "msod_data_delete", "msod_data_present",
    "msod_data_deviceptr", "msod_data_invalidate", "msod_set_assembly", "msod_set_async",
    "msod_synchronize", "msod_timing_enable", "msod_timing_get", "msod_timing_reset",
    "msod_full_convec_ijk", "msod_full_diffusion_ijk", "msod_set_diffusion_coefficients",
    "msod_full_convec_ijk_mixed", "msod_full_diffusion_ijk_mixed",
    "msod_full_convec_ijk_geo32", "msod_full_diffusion_ijk_geo32",
    "msod_full_diffusion_ijk_closure", "msod_closure_features", "msod_closure_predict",
    "msod_closure_sensitivity", "msod_closure_gradient", "msod_closure_adam",
    "msod_closure_loss", "msod_elem_jacobian", "msod_box_mesh", "msod_flow_fields",
    "msod_viscosity_fields", "msod_tet_jacobian", "msod_fem_convec", "msod_fem_diffu",
    "msod_full_convec_ijk_fp32", "msod_full_diffusion_ijk_fp32", "msod_data_register",
    "msod_closure_gradient_noalloc", "msod_full_convec_ijk_r4", "msod_full_diffusion_ijk_r4",
    "msod_full_convec_ijk_f32split", "msod_full_diffusion_ijk_f32split",
)


class MsodError(RuntimeError):
    """A non-zero status from the C ABI."""

    def __init__(self, status: int, message: str):
        super().__init__(f"msod status {status}: {message}")
        self.status = status


_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double

_hip = None
_host = None


def _load(path: str, what: str) -> ctypes.CDLL:
    if not os.path.exists(path):
        raise RuntimeError(f"{what} not built: {path} is missing (run __graft_entry__.build() "
                           f"or `make -C minisod2d_amd`)")
    return ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)


def hip_lib() -> ctypes.CDLL:
    """The HIP kernel library (raises if it is not built)."""
    global _hip
    if _hip is None:
        L = _load(HIP_LIB, "HIP kernel library")
        L.msod_status_string.restype = ctypes.c_char_p
        L.msod_last_error.restype = ctypes.c_char_p
        L.msod_data_deviceptr.restype = _vp
        L.msod_data_deviceptr.argtypes = [_vp]
        for n in ("msod_data_copyin", "msod_data_create", "msod_data_update_device",
                  "msod_data_update_host", "msod_data_register"):
            getattr(L, n).argtypes = [_vp, ctypes.c_size_t]
        for n in ("msod_data_delete", "msod_data_present", "msod_data_invalidate"):
            getattr(L, n).argtypes = [_vp]
        L.msod_timing_get.argtypes = [_i, ctypes.POINTER(_d), ctypes.POINTER(ctypes.c_int64)]
        L.msod_set_diffusion_coefficients.argtypes = [_d, _d]
        L.msod_full_convec_ijk.argtypes = [_i, _i, _i] + [_vp] * 19
        L.msod_full_diffusion_ijk.argtypes = [_i, _i, _i] + [_vp] * 11 + [_d, _d] + [_vp] * 9
        L.msod_full_convec_ijk_mixed.argtypes = [_i, _i, _i] + [_vp] * 12 + [_d] + [_vp] * 3
        L.msod_full_diffusion_ijk_mixed.argtypes = [_i, _i, _i] + [_vp] * 8 + [_d, _d] + [_vp] * 9
        L.msod_full_convec_ijk_geo32.argtypes = [_i, _i, _i] + [_vp] * 16
        L.msod_full_convec_ijk_fp32.argtypes = [_i, _i, _i] + [_vp] * 12 + [_d] + [_vp] * 3
        L.msod_full_diffusion_ijk_fp32.argtypes = [_i, _i, _i] + [_vp] * 8 + [_d, _d] + [_vp] * 9
        L.msod_full_diffusion_ijk_geo32.argtypes = [_i, _i, _i] + [_vp] * 8 + [_d, _d] + [_vp] * 9
        L.msod_full_convec_ijk_f32split.argtypes = [_i, _i, _i] + [_vp] * 16
        L.msod_full_diffusion_ijk_f32split.argtypes = [_i, _i, _i] + [_vp] * 8 + [_d, _d] + [_vp] * 9
        L.msod_full_convec_ijk_r4.argtypes = [_i, _i, _i] + [_vp] * 19
        L.msod_full_diffusion_ijk_r4.argtypes = ([_i, _i, _i] + [_vp] * 11 + [ctypes.c_float] * 2 +
                                                 [_vp] * 9)
        L.msod_full_diffusion_ijk_closure.argtypes = ([_i, _i, _i] + [_vp] * 11 + [_d, _d] +
                                                      [_vp] * 10)
        L.msod_closure_features.argtypes = [_i, _i, _i] + [_vp] * 11
        L.msod_closure_predict.argtypes = [_i, _i] + [_vp] * 4
        L.msod_closure_sensitivity.argtypes = [_i, _i, _i] + [_vp] * 8 + [_d] + [_vp] * 5
        L.msod_closure_gradient.argtypes = [ctypes.c_int64] + [_vp] * 4
        L.msod_closure_gradient_noalloc.argtypes = [ctypes.c_int64] + [_vp] * 4
        L.msod_closure_adam.argtypes = [_vp] * 4 + [_i, _d, _d, _d, _d]
        L.msod_closure_loss.argtypes = [_i] + [_vp] * 6
        L.msod_elem_jacobian.argtypes = [_i, _i, _i] + [_vp] * 7
        L.msod_box_mesh.argtypes = [_i, _i, _i, _i, _d, ctypes.c_int64] + [_vp] * 6
        L.msod_flow_fields.argtypes = [ctypes.c_int64, _vp, _d, _d, ctypes.c_int64, _d] + [_vp] * 7
        L.msod_viscosity_fields.argtypes = [ctypes.c_int64, _i, ctypes.c_int64, ctypes.c_int64,
                                            _d, _d, _vp, _vp]
        L.msod_tet_jacobian.argtypes = [_i, _i] + [_vp] * 7
        L.msod_fem_convec.argtypes = [_i, _i] + [_vp] * 13
        L.msod_fem_diffu.argtypes = [_i, _i] + [_vp] * 5 + [_d, _d] + [_vp] * 9
        _hip = L
    return _hip


def host_lib() -> ctypes.CDLL:
    """The Fortran host library (tables, meshes, geometry, fields)."""
    global _host
    if _host is None:
        L = _load(HOST_LIB, "Fortran host library")
        L.msodh_box_mesh.argtypes = [_i, _i, _i, _i, _d, ctypes.c_int64, _vp, _vp]
        L.msodh_elem_jacobian.argtypes = [_i, _i, _i] + [_vp] * 6
        L.msodh_elem_jacobian.restype = _i
        L.msodh_flow_fields.argtypes = [ctypes.c_int64, _vp, _d, _d, ctypes.c_int64, _d] + [_vp] * 7
        L.msodh_viscosity_fields.argtypes = [ctypes.c_int64, _i, ctypes.c_int64, ctypes.c_int64,
                                             _d, _d, _vp, _vp]
        L.msodh_disjoint_mesh.argtypes = [_i, _i, _i, _vp, _vp, _vp]
        L.msodh_tet_tables.argtypes = [_vp] * 4
        L.msodh_interp_1d.argtypes = [_i, _i, _vp]
        L.msodh_prolong_fields.argtypes = [_i] * 6 + [_vp] * 4
        L.msodh_restrict_residual.argtypes = [_i] * 6 + [_vp] * 4
        _host = L
    return _host


def _check(status: int) -> None:
    if status != MSOD_OK:
        msg = hip_lib().msod_last_error().decode(errors="replace")
        raise MsodError(status, msg)


def _ptr(a) -> int | None:
    if a is None:
        return None
    if isinstance(a, int):
        return a
    return a.ctypes.data


# ---------------------------------------------------------------------------
# device state and data region
# ---------------------------------------------------------------------------
def init(device: int = 0) -> None:
    _check(hip_lib().msod_init(device))


def finalize() -> None:
    _check(hip_lib().msod_finalize())


def copyin(*arrays: np.ndarray) -> None:
    """``!$acc enter data copyin`` for each array."""
    for a in arrays:
        _check(hip_lib().msod_data_copyin(a.ctypes.data, a.nbytes))


def create(*arrays: np.ndarray) -> None:
    """``!$acc enter data create`` for each array (device copy zeroed)."""
    for a in arrays:
        _check(hip_lib().msod_data_create(a.ctypes.data, a.nbytes))


def update_device(*arrays: np.ndarray) -> None:
    for a in arrays:
        _check(hip_lib().msod_data_update_device(a.ctypes.data, a.nbytes))


def update_host(*arrays: np.ndarray) -> None:
    for a in arrays:
        _check(hip_lib().msod_data_update_host(a.ctypes.data, a.nbytes))


def delete(*arrays: np.ndarray) -> None:
    for a in arrays:
        _check(hip_lib().msod_data_delete(a.ctypes.data))


def present(a: np.ndarray) -> bool:
    return bool(hip_lib().msod_data_present(a.ctypes.data))


ASSEMBLY_COLORED = 0
ASSEMBLY_ATOMIC = 1
ASSEMBLY_DEFAULT = ASSEMBLY_ATOMIC


def set_assembly(mode: int) -> None:
    """ASSEMBLY_ATOMIC (default: one launch, FP64 atomics into zeroed outputs,
    the fastest) or ASSEMBLY_COLORED (one launch per batch colour, no atomics,
    bitwise reproducible)."""
    _check(hip_lib().msod_set_assembly(mode))


def set_async(flag: bool) -> None:
    _check(hip_lib().msod_set_async(1 if flag else 0))


def synchronize() -> None:
    _check(hip_lib().msod_synchronize())


def timing_enable(flag: bool) -> None:
    _check(hip_lib().msod_timing_enable(1 if flag else 0))


def timing_reset() -> None:
    _check(hip_lib().msod_timing_reset())


def timing_get(kernel: int) -> tuple[float, int]:
    ms = ctypes.c_double()
    n = ctypes.c_int64()
    _check(hip_lib().msod_timing_get(kernel, ctypes.byref(ms), ctypes.byref(n)))
    return ms.value, n.value


def set_diffusion_coefficients(c_rho: float = 1.0, c_ener: float = 1.0) -> None:
    _check(hip_lib().msod_set_diffusion_coefficients(c_rho, c_ener))


# ---------------------------------------------------------------------------
# geometry, reference argument list
# ---------------------------------------------------------------------------
def elem_jacobian(porder, nelem, npoin, connec, coord, dNgp, wgp, gpvol, He) -> int:
    """elem_jacobian (jacobian_oper.f90:7-136) on the GPU: fills He(3,3,ngaus,nelem)
    and gpvol(1,ngaus,nelem) from coord(npoin,3); returns the number of element
    nodes with det(Je) <= 0.  Arrays: host-present (copyin/create) or device."""
    nbad = ctypes.c_int64(0)
    _check(hip_lib().msod_elem_jacobian(int(porder), int(nelem), int(npoin),
                                        *map(_ptr, (connec, coord, dNgp, wgp, gpvol, He)),
                                        ctypes.byref(nbad)))
    return int(nbad.value)


def tet_jacobian(nelem, npoin, connec, coord, dNgp, wgp, gpvol, He) -> int:
    """tet_jacobian (jacobian_oper.f90:138-267) on the GPU: He(3,3,4,nelem) and
    gpvol(1,4,nelem) of linear tetrahedra from coord(npoin,3); returns the
    number of points with det(Je) <= 0."""
    nbad = ctypes.c_int64(0)
    _check(hip_lib().msod_tet_jacobian(int(nelem), int(npoin),
                                       *map(_ptr, (connec, coord, dNgp, wgp, gpvol, He)),
                                       ctypes.byref(nbad)))
    return int(nbad.value)


def fem_convec(nelem, npoin, connec, Ngp, dNgp, He, gpvol, u, q, rho, pr, E, Rmass, Rmom, Rener):
    """Convective residual on linear tetrahedra (elem_convec.f90:839-948), FP64, on the GPU."""
    _check(hip_lib().msod_fem_convec(int(nelem), int(npoin), *map(_ptr, (
        connec, Ngp, dNgp, He, gpvol, u, q, rho, pr, E, Rmass, Rmom, Rener))))


def fem_diffu(nelem, npoin, connec, Ngp, dNgp, He, gpvol, Cp, Pr, rho, u, Tem, mu_fluid, mu_e,
              mu_sgs, Rmass, Rmom, Rener):
    """Diffusive residual on linear tetrahedra (elem_diffu.f90:199-317), FP64, on the GPU."""
    _check(hip_lib().msod_fem_diffu(int(nelem), int(npoin), *map(_ptr, (
        connec, Ngp, dNgp, He, gpvol)), float(Cp), float(Pr), *map(_ptr, (
            rho, u, Tem, mu_fluid, mu_e, mu_sgs, Rmass, Rmom, Rener))))


# ---------------------------------------------------------------------------
# the two kernels, reference argument lists
# ---------------------------------------------------------------------------
def _porder(invAtoIJK: np.ndarray) -> int:
    return int(invAtoIJK.shape[0]) - 1


def full_convec_ijk(nelem, npoin, connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp, invAtoIJK,
                    gmshAtoI, gmshAtoJ, gmshAtoK, u, q, rho, pr, E, Rmass, Rmom, Rener):
    """Convective residual (elem_convec.f90:198-431), FP64, on the GPU."""
    _check(hip_lib().msod_full_convec_ijk(
        _porder(invAtoIJK), nelem, npoin, *map(_ptr, (
            connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp, invAtoIJK, gmshAtoI, gmshAtoJ,
            gmshAtoK, u, q, rho, pr, E, Rmass, Rmom, Rener))))


def full_diffusion_ijk(nelem, npoin, connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp, invAtoIJK,
                       gmshAtoI, gmshAtoJ, gmshAtoK, Cp, Pr, rho, u, Tem, mu_fluid, mu_e, mu_sgs,
                       Rmass, Rmom, Rener):
    """Diffusive residual (elem_diffu.f90:10-197), FP64, on the GPU."""
    L = hip_lib()
    p = list(map(_ptr, (connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp, invAtoIJK, gmshAtoI,
                        gmshAtoJ, gmshAtoK)))
    r = list(map(_ptr, (rho, u, Tem, mu_fluid, mu_e, mu_sgs, Rmass, Rmom, Rener)))
    _check(L.msod_full_diffusion_ijk(_porder(invAtoIJK), nelem, npoin, *p, float(Cp), float(Pr), *r))


def full_convec_ijk_mixed(nelem, npoin, connec, He32, gpvol32, dlxigp_ip, invAtoIJK, gmshAtoI,
                          gmshAtoJ, gmshAtoK, u32, q32, rho32, E32, gamma_gas, Rmass, Rmom, Rener):
    """Convective residual from FP32 state and geometry (FP64 arithmetic and assembly)."""
    _check(hip_lib().msod_full_convec_ijk_mixed(
        _porder(invAtoIJK), nelem, npoin,
        *map(_ptr, (connec, He32, gpvol32, dlxigp_ip, invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK,
                    u32, q32, rho32, E32)), float(gamma_gas), *map(_ptr, (Rmass, Rmom, Rener))))


def full_convec_ijk_f32split(nelem, npoin, connec, He32, gpvol32, dlxigp_ip, invAtoIJK, gmshAtoI,
                             gmshAtoJ, gmshAtoK, u, q32, rho32, pr, E32, Rmass, Rmom, Rener):
    """FP32 split: FP32 He, gpvol, q, rho, E; FP64 u and pr; FP64 outputs."""
    _check(hip_lib().msod_full_convec_ijk_f32split(
        _porder(invAtoIJK), nelem, npoin, *map(_ptr, (
            connec, He32, gpvol32, dlxigp_ip, invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, u, q32, rho32,
            pr, E32, Rmass, Rmom, Rener))))


def full_diffusion_ijk_f32split(nelem, npoin, connec, He32, gpvol32, dlxigp_ip, invAtoIJK, gmshAtoI,
                                gmshAtoJ, gmshAtoK, Cp, Pr, rho32, u, Tem32, mu_fluid32, mu_e32,
                                mu_sgs32, Rmass, Rmom, Rener):
    """FP32 split: FP32 He, gpvol, rho, Tem, mu_fluid, mu_e, mu_sgs; FP64 u; FP64 outputs."""
    _check(hip_lib().msod_full_diffusion_ijk_f32split(
        _porder(invAtoIJK), nelem, npoin, *map(_ptr, (
            connec, He32, gpvol32, dlxigp_ip, invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK)),
        float(Cp), float(Pr), *map(_ptr, (rho32, u, Tem32, mu_fluid32, mu_e32, mu_sgs32, Rmass, Rmom,
                                          Rener))))


def full_convec_ijk_r4(nelem, npoin, connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp, invAtoIJK,
                       gmshAtoI, gmshAtoJ, gmshAtoK, u, q, rho, pr, E, Rmass, Rmom, Rener):
    """full_convec_ijk for a caller built at rp=4: float32 arrays and outputs."""
    _check(hip_lib().msod_full_convec_ijk_r4(
        _porder(invAtoIJK), nelem, npoin, *map(_ptr, (
            connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp, invAtoIJK, gmshAtoI, gmshAtoJ,
            gmshAtoK, u, q, rho, pr, E, Rmass, Rmom, Rener))))


def full_diffusion_ijk_r4(nelem, npoin, connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp, invAtoIJK,
                          gmshAtoI, gmshAtoJ, gmshAtoK, Cp, Pr, rho, u, Tem, mu_fluid, mu_e, mu_sgs,
                          Rmass, Rmom, Rener):
    """full_diffusion_ijk for a caller built at rp=4: float32 arrays and outputs."""
    _check(hip_lib().msod_full_diffusion_ijk_r4(
        _porder(invAtoIJK), nelem, npoin, *map(_ptr, (
            connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp, invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK)),
        float(Cp), float(Pr), *map(_ptr, (rho, u, Tem, mu_fluid, mu_e, mu_sgs, Rmass, Rmom, Rener))))


def full_diffusion_ijk_mixed(nelem, npoin, connec, He32, gpvol32, dlxigp_ip, invAtoIJK, gmshAtoI,
                             gmshAtoJ, gmshAtoK, Cp, Pr, rho32, u32, Tem32, mu_fluid32, mu_e32,
                             mu_sgs32, Rmass, Rmom, Rener):
    """Diffusive residual from FP32 state and geometry (FP64 arithmetic and assembly)."""
    _check(hip_lib().msod_full_diffusion_ijk_mixed(
        _porder(invAtoIJK), nelem, npoin,
        *map(_ptr, (connec, He32, gpvol32, dlxigp_ip, invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK)),
        float(Cp), float(Pr),
        *map(_ptr, (rho32, u32, Tem32, mu_fluid32, mu_e32, mu_sgs32, Rmass, Rmom, Rener))))


def full_convec_ijk_fp32(nelem, npoin, connec, He32, gpvol32, dlxigp_ip, invAtoIJK, gmshAtoI,
                         gmshAtoJ, gmshAtoK, u32, q32, rho32, E32, gamma_gas, Rmass, Rmom, Rener):
    """Convective residual with FP32 storage and FP32 gathers, fluxes and sweeps
    (FP64 node sums, assembly and pressure recovery)."""
    _check(hip_lib().msod_full_convec_ijk_fp32(
        _porder(invAtoIJK), nelem, npoin,
        *map(_ptr, (connec, He32, gpvol32, dlxigp_ip, invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK,
                    u32, q32, rho32, E32)), float(gamma_gas), *map(_ptr, (Rmass, Rmom, Rener))))


def full_diffusion_ijk_fp32(nelem, npoin, connec, He32, gpvol32, dlxigp_ip, invAtoIJK, gmshAtoI,
                            gmshAtoJ, gmshAtoK, Cp, Pr, rho32, u32, Tem32, mu_fluid32, mu_e32,
                            mu_sgs32, Rmass, Rmom, Rener):
    """Diffusive residual with FP32 storage and FP32 arithmetic (FP64 node sums and assembly)."""
    _check(hip_lib().msod_full_diffusion_ijk_fp32(
        _porder(invAtoIJK), nelem, npoin,
        *map(_ptr, (connec, He32, gpvol32, dlxigp_ip, invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK)),
        float(Cp), float(Pr),
        *map(_ptr, (rho32, u32, Tem32, mu_fluid32, mu_e32, mu_sgs32, Rmass, Rmom, Rener))))


def full_convec_ijk_geo32(nelem, npoin, connec, He32, gpvol32, dlxigp_ip, invAtoIJK, gmshAtoI,
                          gmshAtoJ, gmshAtoK, u, q, rho, pr, E, Rmass, Rmom, Rener):
    """Convective residual with FP32 geometry and FP64 fields (FP64 arithmetic)."""
    _check(hip_lib().msod_full_convec_ijk_geo32(
        _porder(invAtoIJK), nelem, npoin,
        *map(_ptr, (connec, He32, gpvol32, dlxigp_ip, invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK,
                    u, q, rho, pr, E, Rmass, Rmom, Rener))))


def full_diffusion_ijk_geo32(nelem, npoin, connec, He32, gpvol32, dlxigp_ip, invAtoIJK, gmshAtoI,
                             gmshAtoJ, gmshAtoK, Cp, Pr, rho, u, Tem, mu_fluid, mu_e32, mu_sgs32,
                             Rmass, Rmom, Rener):
    """Diffusive residual with FP32 geometry, mu_e, mu_sgs and FP64 fields."""
    _check(hip_lib().msod_full_diffusion_ijk_geo32(
        _porder(invAtoIJK), nelem, npoin,
        *map(_ptr, (connec, He32, gpvol32, dlxigp_ip, invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK)),
        float(Cp), float(Pr),
        *map(_ptr, (rho, u, Tem, mu_fluid, mu_e32, mu_sgs32, Rmass, Rmom, Rener))))


# ---------------------------------------------------------------------------
# learned subgrid-viscosity closure
# ---------------------------------------------------------------------------
CLOSURE_NPARAM = 1530   # weight block (floats); layout in csrc/msod_mlp.h
CLOSURE_NTRAIN = 1505   # trained part: W1 b1 W2 b2 w3 b3
CLOSURE_NFEAT = 12


def full_diffusion_ijk_closure(nelem, npoin, connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp,
                               invAtoIJK, gmshAtoI, gmshAtoJ, gmshAtoK, Cp, Pr, rho, u, Tem,
                               mu_fluid, mu_e, mlp_params, Rmass, Rmom, Rener) -> bool:
    """full_diffusion_ijk with mu_sgs predicted by the MLP inside the kernel.
    Returns whether the model was used (False: invalid or missing weights,
    mu_sgs = 0)."""
    used = ctypes.c_int(0)
    p = list(map(_ptr, (connec, Ngp, dNgp, He, gpvol, dlxigp_ip, xgp, invAtoIJK, gmshAtoI,
                        gmshAtoJ, gmshAtoK)))
    r = list(map(_ptr, (rho, u, Tem, mu_fluid, mu_e, mlp_params)))
    _check(hip_lib().msod_full_diffusion_ijk_closure(
        _porder(invAtoIJK), nelem, npoin, *p, float(Cp), float(Pr), *r, ctypes.addressof(used),
        *map(_ptr, (Rmass, Rmom, Rener))))
    return bool(used.value)


def closure_features(nelem, npoin, connec, He, gpvol, dlxigp_ip, invAtoIJK, AtoI, AtoJ, AtoK,
                     rho, u, features):
    """features[nelem*nnode, 12] (element-major, lexicographic nodes)."""
    _check(hip_lib().msod_closure_features(
        _porder(invAtoIJK), nelem, npoin,
        *map(_ptr, (connec, He, gpvol, dlxigp_ip, invAtoIJK, AtoI, AtoJ, AtoK, rho, u,
                    features))))


def closure_predict(nelem, invAtoIJK, mlp_params, features, mu_sgs):
    """mu_sgs(nelem, nnode) (reference layout) from the MLP."""
    _check(hip_lib().msod_closure_predict(_porder(invAtoIJK), nelem,
                                          *map(_ptr, (invAtoIJK, mlp_params, features, mu_sgs))))


def closure_sensitivity(nelem, npoin, connec, He, gpvol, dlxigp_ip, invAtoIJK, AtoI, AtoJ, AtoK,
                        Cp, rho, u, Tem, rerr, gnu):
    """gnu[e*nnode + l] = d<rerr, (Rmom, Rener)>/d mu_sgs at element node (e, l)."""
    _check(hip_lib().msod_closure_sensitivity(
        _porder(invAtoIJK), nelem, npoin,
        *map(_ptr, (connec, He, gpvol, dlxigp_ip, invAtoIJK, AtoI, AtoJ, AtoK)), float(Cp),
        *map(_ptr, (rho, u, Tem, rerr, gnu))))


def closure_gradient(nsamples, mlp_params, features, gnu, grad):
    _check(hip_lib().msod_closure_gradient(int(nsamples),
                                           *map(_ptr, (mlp_params, features, gnu, grad))))


def closure_adam(mlp_params, grad, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    _check(hip_lib().msod_closure_adam(*map(_ptr, (mlp_params, grad, m, v)), int(step),
                                       float(lr), float(beta1), float(beta2), float(eps)))


def closure_loss(npoin, Rmom, Rener, target, weights, rerr) -> float:
    """sum_c w_c |R_c - target_c|^2 over momentum (3) and energy; writes
    rerr = 2 w_c (R_c - target_c)."""
    w = np.ascontiguousarray(weights, dtype=np.float64)
    if w.shape != (4,):
        raise ValueError("weights: 4 values (momentum x, y, z, energy)")
    loss = ctypes.c_double()
    _check(hip_lib().msod_closure_loss(npoin, *map(_ptr, (Rmom, Rener, target, w, rerr)),
                                       ctypes.addressof(loss)))
    return loss.value


# ---------------------------------------------------------------------------
# synthetic problems (Fortran host library)
# ---------------------------------------------------------------------------
GAMMA = 1.4
RGAS = 287.0
CP = 1004.5
PR = 0.71
SEED_RHO, SEED_MUE, SEED_MUSGS, SEED_JITTER = 1001, 1002, 1003, 1004


def _f(shape, dtype=np.float64):
    return np.zeros(shape, dtype=dtype, order="F")


@dataclass
class Tables:
    porder: int
    invAtoIJK: np.ndarray
    AtoIJK: np.ndarray
    AtoI: np.ndarray
    AtoJ: np.ndarray
    AtoK: np.ndarray
    xgp: np.ndarray
    wgp: np.ndarray
    Ngp: np.ndarray
    dNgp: np.ndarray
    dlxigp_ip: np.ndarray

    @property
    def nnode(self) -> int:
        return (self.porder + 1) ** 3


def hex_tables(porder: int) -> Tables:
    q = porder + 1
    nn = q ** 3
    t = [_f((q, q, q), np.int32), _f(nn, np.int32), _f(nn, np.int32), _f(nn, np.int32),
         _f(nn, np.int32), _f((nn, 3)), _f(nn), _f((nn, nn)), _f((3, nn, nn)), _f((nn, 3, q))]
    host_lib().msodh_hex_tables(porder, *[x.ctypes.data_as(_vp) for x in t])
    return Tables(porder, *t)


@dataclass
class Problem:
    """Inputs of both kernels for one mesh, in the reference layouts."""
    tables: Tables
    nelem: int
    npoin: int
    connec: np.ndarray
    coord: np.ndarray
    He: np.ndarray
    gpvol: np.ndarray
    u: np.ndarray
    q: np.ndarray
    rho: np.ndarray
    pr: np.ndarray
    E: np.ndarray
    Tem: np.ndarray
    mu_fluid: np.ndarray
    mu_e: np.ndarray
    mu_sgs: np.ndarray
    Cp: float = CP
    Pr: float = PR
    gamma: float = GAMMA
    meta: dict = field(default_factory=dict)

    @property
    def porder(self) -> int:
        return self.tables.porder

    def convec_args(self, Rmass, Rmom, Rener):
        t = self.tables
        return (self.nelem, self.npoin, self.connec, t.Ngp, t.dNgp, self.He, self.gpvol,
                t.dlxigp_ip, t.xgp, t.invAtoIJK, t.AtoI, t.AtoJ, t.AtoK, self.u, self.q,
                self.rho, self.pr, self.E, Rmass, Rmom, Rener)

    def diffu_args(self, Rmass, Rmom, Rener, mu_sgs=None):
        t = self.tables
        return (self.nelem, self.npoin, self.connec, t.Ngp, t.dNgp, self.He, self.gpvol,
                t.dlxigp_ip, t.xgp, t.invAtoIJK, t.AtoI, t.AtoJ, t.AtoK, self.Cp, self.Pr,
                self.rho, self.u, self.Tem, self.mu_fluid, self.mu_e,
                self.mu_sgs if mu_sgs is None else mu_sgs, Rmass, Rmom, Rener)

    def outputs(self):
        return _f(self.npoin), _f((self.npoin, 3)), _f(self.npoin)

    def device_inputs(self):
        t = self.tables
        return (self.connec, self.He, self.gpvol, self.u, self.q, self.rho, self.pr, self.E,
                self.Tem, self.mu_fluid, self.mu_e, self.mu_sgs)

    def fp32(self) -> dict:
        """FP32 copies of geometry and state for the mixed-precision kernels."""
        c = lambda a: np.asfortranarray(a, dtype=np.float32)
        return dict(He=c(self.He), gpvol=c(self.gpvol), u=c(self.u), q=c(self.q),
                    rho=c(self.rho), E=c(self.E), Tem=c(self.Tem), mu_fluid=c(self.mu_fluid),
                    mu_e=c(self.mu_e), mu_sgs=c(self.mu_sgs))


def box_problem(nx: int, porder: int, ny: int | None = None, nz: int | None = None,
                jitter: float = 0.0, rho_jump: float = 0.0, threads: int = 0,
                seeds=(SEED_RHO, SEED_MUE, SEED_MUSGS, SEED_JITTER)) -> Problem:
    """Conforming [0,1]^3 hex mesh of nx*ny*nz elements of order porder with the
    synthetic fields (see minisod2d_amd/fortran/msod_fields.f90)."""
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    H = host_lib()
    if threads:
        H.msodh_set_threads(threads)
    t = hex_tables(porder)
    ne = ctypes.c_int64()
    npn = ctypes.c_int64()
    H.msodh_box_sizes(nx, ny, nz, porder, ctypes.byref(ne), ctypes.byref(npn))
    nelem, npoin = ne.value, npn.value
    if npoin >= 2 ** 31 or nelem * t.nnode >= 2 ** 31:
        raise ValueError("mesh too large for the int32 index format of the reference")
    nn = t.nnode
    connec = _f((nelem, nn), np.int32)
    coord = _f((npoin, 3))
    H.msodh_box_mesh(nx, ny, nz, porder, jitter, seeds[3], connec.ctypes.data, coord.ctypes.data)
    He = _f((3, 3, nn, nelem))
    gpvol = _f((1, nn, nelem))
    nbad = H.msodh_elem_jacobian(nelem, npoin, nn, connec.ctypes.data, coord.ctypes.data,
                                 t.dNgp.ctypes.data, t.wgp.ctypes.data, gpvol.ctypes.data,
                                 He.ctypes.data)
    if nbad:
        raise ValueError(f"{nbad} quadrature points with non-positive Jacobian")
    u, q = _f((npoin, 3)), _f((npoin, 3))
    rho, pr, E, Tem, muf = (_f(npoin) for _ in range(5))
    H.msodh_flow_fields(npoin, coord.ctypes.data, GAMMA, RGAS, seeds[0], rho_jump,
                        *[a.ctypes.data for a in (u, q, rho, pr, E, Tem, muf)])
    mu_e, mu_sgs = _f((nelem, nn)), _f((nelem, nn))
    H.msodh_viscosity_fields(nelem, nn, seeds[1], seeds[2], 1e-3, 1e-4, mu_e.ctypes.data,
                             mu_sgs.ctypes.data)
    return Problem(t, nelem, npoin, connec, coord, He, gpvol, u, q, rho, pr, E, Tem, muf, mu_e,
                   mu_sgs, meta=dict(nx=nx, ny=ny, nz=nz, porder=porder, jitter=jitter,
                                     rho_jump=rho_jump, seeds=list(seeds)))


def device_box_problem(nx: int, porder: int, ny: int | None = None, nz: int | None = None,
                       jitter: float = 0.0, rho_jump: float = 0.0,
                       seeds=(SEED_RHO, SEED_MUE, SEED_MUSGS, SEED_JITTER)) -> "Problem":
    """box_problem's mesh and fields generated on the GPU (msod_box_mesh,
    msod_flow_fields, msod_viscosity_fields), then He and gpvol
    (msod_elem_jacobian).  The arrays are allocated on the host (untouched,
    no resident pages until read) and created on the device: call
    update_host to read them, delete when done."""
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    t = hex_tables(porder)
    nn = t.nnode
    nelem = nx * ny * nz
    npoin = (nx * porder + 1) * (ny * porder + 1) * (nz * porder + 1)
    e = lambda shape, dt=np.float64: np.empty(shape, dtype=dt, order="F")
    connec, coord = e((nelem, nn), np.int32), e((npoin, 3))
    He, gpvol = e((3, 3, nn, nelem)), e((1, nn, nelem))
    u, q = e((npoin, 3)), e((npoin, 3))
    rho, pr, E, Tem, muf = (e(npoin) for _ in range(5))
    mu_e, mu_sgs = e((nelem, nn)), e((nelem, nn))
    arrays = (connec, coord, He, gpvol, u, q, rho, pr, E, Tem, muf, mu_e, mu_sgs)
    create(*arrays)
    L = hip_lib()
    _check(L.msod_box_mesh(nx, ny, nz, porder, float(jitter), int(seeds[3]), *map(_ptr, (
        t.xgp, t.AtoI, t.AtoJ, t.AtoK, connec, coord))))
    nbad = elem_jacobian(porder, nelem, npoin, connec, coord, t.dNgp, t.wgp, gpvol, He)
    if nbad:
        delete(*arrays)
        raise ValueError(f"{nbad} quadrature points with non-positive Jacobian")
    _check(L.msod_flow_fields(npoin, _ptr(coord), GAMMA, RGAS, int(seeds[0]), float(rho_jump),
                              *map(_ptr, (u, q, rho, pr, E, Tem, muf))))
    _check(L.msod_viscosity_fields(nelem, nn, int(seeds[1]), int(seeds[2]), 1e-3, 1e-4,
                                   _ptr(mu_e), _ptr(mu_sgs)))
    return Problem(t, nelem, npoin, connec, coord, He, gpvol, u, q, rho, pr, E, Tem, muf, mu_e,
                   mu_sgs, meta=dict(nx=nx, ny=ny, nz=nz, porder=porder, jitter=jitter,
                                     rho_jump=rho_jump, seeds=list(seeds), device=True))


@dataclass
class TetTables:
    xgp: np.ndarray   # (4,3)
    wgp: np.ndarray   # (4,)
    Ngp: np.ndarray   # (4,4): (point, node)
    dNgp: np.ndarray  # (3,4,4)


def tet_tables() -> TetTables:
    """The TET04 element of the reference driver (main.f90:374-383)."""
    t = [_f((4, 3)), _f(4), _f((4, 4)), _f((3, 4, 4))]
    host_lib().msodh_tet_tables(*[x.ctypes.data_as(_vp) for x in t])
    return TetTables(*t)


TET_BASE = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
SEED_TET = 1005


@dataclass
class TetProblem:
    """Inputs of fem_convec / fem_diffu: the reference driver's disjoint
    tetrahedral mesh (main.f90:367-407) over the nodal fields of a hex problem."""
    tables: TetTables
    nelem: int
    npoin: int
    connec: np.ndarray
    coord: np.ndarray
    He: np.ndarray
    gpvol: np.ndarray
    u: np.ndarray
    q: np.ndarray
    rho: np.ndarray
    pr: np.ndarray
    E: np.ndarray
    Tem: np.ndarray
    mu_fluid: np.ndarray
    mu_e: np.ndarray
    mu_sgs: np.ndarray
    Cp: float = CP
    Pr: float = PR
    meta: dict = field(default_factory=dict)

    def convec_args(self, Rmass, Rmom, Rener):
        t = self.tables
        return (self.nelem, self.npoin, self.connec, t.Ngp, t.dNgp, self.He, self.gpvol, self.u,
                self.q, self.rho, self.pr, self.E, Rmass, Rmom, Rener)

    def diffu_args(self, Rmass, Rmom, Rener):
        t = self.tables
        return (self.nelem, self.npoin, self.connec, t.Ngp, t.dNgp, self.He, self.gpvol, self.Cp,
                self.Pr, self.rho, self.u, self.Tem, self.mu_fluid, self.mu_e, self.mu_sgs, Rmass,
                Rmom, Rener)

    def outputs(self):
        return _f(self.npoin), _f((self.npoin, 3)), _f(self.npoin)

    def device_inputs(self):
        return (self.connec, self.He, self.gpvol, self.u, self.q, self.rho, self.pr, self.E,
                self.Tem, self.mu_fluid, self.mu_e, self.mu_sgs)


def tet_problem(prob: Problem, jitter: float = 0.0, random_mu: bool = False,
                seed: int = SEED_TET, strict: bool = True) -> TetProblem:
    """The reference driver's TET04 data over prob's nodal fields: nelem =
    npoin/4 disjoint unit tetrahedra, each shifted by 3 along the diagonal
    (mod_mesh.f90:71-125), He and gpvol from tet_jacobian, mu_e = mu_sgs = 1
    (main.f90:403-407).  The reference stops unless 4 divides npoin; with
    strict=False the last npoin mod 4 nodes are left out of the mesh instead.
    jitter moves every vertex by up to jitter/2 per axis (seeded), so that He
    is not the identity; random_mu draws mu_e, mu_sgs in [0,1) instead of ones."""
    npoin = prob.npoin
    if npoin % 4 and strict:
        raise ValueError(f"npoin={npoin} is not divisible by 4 (main.f90:369-372)")
    nelem = npoin // 4
    H = host_lib()
    t = tet_tables()
    connec = _f((nelem, 4), np.int32)
    coord = _f((npoin, 3))
    H.msodh_disjoint_mesh(nelem, npoin, 4, np.asfortranarray(TET_BASE).ctypes.data,
                          connec.ctypes.data, coord.ctypes.data)
    rng = np.random.default_rng(seed)
    if jitter:
        coord[:4 * nelem] += jitter * (rng.random((4 * nelem, 3)) - 0.5)
    He = _f((3, 3, 4, nelem))
    gpvol = _f((1, 4, nelem))
    nbad = H.msodh_elem_jacobian(nelem, npoin, 4, connec.ctypes.data, coord.ctypes.data,
                                 t.dNgp.ctypes.data, t.wgp.ctypes.data, gpvol.ctypes.data,
                                 He.ctypes.data)
    if nbad:
        raise ValueError(f"{nbad} tetrahedron points with non-positive Jacobian")
    if random_mu:
        mu_e, mu_sgs = np.asfortranarray(rng.random((nelem, 4))), np.asfortranarray(rng.random((nelem, 4)))
    else:
        mu_e, mu_sgs = np.ones((nelem, 4), order="F"), np.ones((nelem, 4), order="F")
    return TetProblem(t, nelem, npoin, connec, coord, He, gpvol, prob.u, prob.q, prob.rho, prob.pr,
                      prob.E, prob.Tem, prob.mu_fluid, mu_e, mu_sgs, Cp=prob.Cp, Pr=prob.Pr,
                      meta=dict(prob.meta, tet=True, tet_jitter=jitter, tet_random_mu=random_mu,
                                tet_seed=seed))


def reference_problem(nelem: int, porder: int) -> Problem:
    """The reference driver's own case (src/app/sources/main.f90:257-413): the
    disjoint mesh of mod_mesh.f90, fields of ones with spikes at local nodes
    1 and 4 of every element, E and Tem from the driver's relations, Cp=Pr=1."""
    H = host_lib()
    t = hex_tables(porder)
    nn = t.nnode
    npoin = nelem * nn
    connec = _f((nelem, nn), np.int32)
    coord = _f((npoin, 3))
    H.msodh_disjoint_mesh(nelem, npoin, nn, t.xgp.ctypes.data, connec.ctypes.data,
                          coord.ctypes.data)
    He = _f((3, 3, nn, nelem))
    gpvol = _f((1, nn, nelem))
    H.msodh_elem_jacobian(nelem, npoin, nn, connec.ctypes.data, coord.ctypes.data,
                          t.dNgp.ctypes.data, t.wgp.ctypes.data, gpvol.ctypes.data, He.ctypes.data)
    u = np.ones((npoin, 3), order="F")
    rho = np.ones(npoin)
    pr = np.ones(npoin)
    i1 = connec[:, 0] - 1
    i4 = connec[:, 3] - 1
    u[i1, :] = 10.0
    u[i4, :] = 20.0
    rho[i1] = 10.0
    rho[i4] = 20.0
    pr[i1] = 100.0
    pr[i4] = 200.0
    q = np.asfortranarray(rho[:, None] * u)
    fmag = (q * u).sum(axis=1)
    eint = pr / ((GAMMA - 1.0) * rho)
    E = rho * eint + 0.5 * fmag
    Tem = pr / (rho * RGAS)
    muf = np.ones(npoin)
    mu_e = np.ones((nelem, nn), order="F")
    mu_sgs = np.ones((nelem, nn), order="F")
    return Problem(t, nelem, npoin, connec, coord, He, gpvol, u, q, rho, pr, E, Tem, muf, mu_e,
                   mu_sgs, Cp=1.0, Pr=1.0, meta=dict(reference_case=True, nelem=nelem,
                                                     porder=porder))

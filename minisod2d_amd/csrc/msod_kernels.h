// Internal interface between the runtime (msod_runtime.hip) and the kernel
// translation units.  All pointers are device pointers in the kernels'
// element-major layouts:
//
//   conn   [nelem][nn]       uint32: bits 0..30 = 0-based global node,
//                            bit 31 = node referenced by this element only
//   geo    [nelem][9][nn]    He(i,k) at component c = i + 3k (SoA per element)
//   gvol   [nelem][nn]       gpvol
// where the element-local node index is lexicographic in the reference's
// 1-D SOD order: l = I + q*(J + q*K), with (I,J,K) = (AtoI,AtoJ,AtoK)-1.
// tab.D[q*q] is the 1-D derivative D[i*q+n] = l_n'(xi_i) and tab.l2g[nn]
// maps l -> 0-based reference (Gmsh) node number, used to read arrays kept
// in the reference layout (mu_e, mu_sgs).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace msod {

constexpr int kMaxP = 8;

template <int P>
struct Shape {
    static constexpr int Q = P + 1;
    static constexpr int NN = Q * Q * Q;
    // one node per thread; several elements per block while they fit 256
    static constexpr int EPB = NN <= 128 ? 256 / NN : 1;
    static constexpr int NT = NN <= 128 ? 256 : ((NN + 63) / 64) * 64;
};

struct Tables {
    const double *D;      // [q*q]
    const int32_t *l2g;   // [nn]
};

template <typename T>
struct ConvecArgs {
    int64_t nelem, npoin;
    const uint32_t *conn;
    const T *geo, *gvol;
    const T *u, *q, *rho, *pr, *E;   // pr == nullptr: recover pressure from E
    double gamma_m1;                 // gamma-1, used when pr == nullptr
    double *Rmass, *Rmom, *Rener;
    Tables tab;
};

template <typename T>
struct DiffuArgs {
    int64_t nelem, npoin;
    const uint32_t *conn;
    const T *geo, *gvol;
    const T *rho, *u, *Tem, *mu_fluid, *mu_e, *mu_sgs;
    double Cp, Pr, c_rho, c_ener;
    double *Rmass, *Rmom, *Rener;
    Tables tab;
};

// Launchers (return hipError_t of the launch).  T = double: FP64 form;
// T = float: mixed form (FP32 sweeps, FP64 assembly).
template <typename T>
hipError_t launch_convec(int porder, const ConvecArgs<T> &a, hipStream_t s);
template <typename T>
hipError_t launch_diffu(int porder, const DiffuArgs<T> &a, hipStream_t s);

// Zero Rmass (n), Rmom (3n), Rener (n).
hipError_t launch_zero_outputs(double *Rmass, double *Rmom, double *Rener, int64_t n,
                               hipStream_t s);

// Element-major layouts from the reference layouts (see top of file).
// *nbad receives the number of connec entries outside 1..npoin (counted in
// the device int *bad); conn is written only when that number is zero.
hipError_t launch_build_conn(const int32_t *connec, int64_t nelem, int nn, int64_t npoin,
                             const int32_t *l2g, uint32_t *conn, uint32_t *mult, int *bad,
                             int *nbad, hipStream_t s);
template <typename T>
hipError_t launch_build_geo(const T *He, const T *gpvol, int64_t nelem, int nn,
                            const int32_t *l2g, T *geo, T *gvol, hipStream_t s);

// XCD-aware mapping: blocks are dispatched round-robin over the 8 XCDs of
// the chip; give each XCD a contiguous range of element groups so that
// neighbouring elements (which share nodes) meet in the same L2.
constexpr int kNumXcd = 8;
__device__ inline int64_t xcd_group(int64_t ngroups)
{
    const int64_t b = blockIdx.x;
    const int64_t per = gridDim.x / kNumXcd;
    const int64_t g = (b % kNumXcd) * per + b / kNumXcd;
    return g < ngroups ? g : -1;
}
inline int64_t xcd_grid(int64_t ngroups)
{
    return (ngroups + kNumXcd - 1) / kNumXcd * kNumXcd;
}

}  // namespace msod

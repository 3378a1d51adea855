// Internal interface between the runtime (msod_runtime.hip) and the kernel
// translation units.  All pointers are device pointers.
//
// Element data are element-major, with the element-local node index l
// lexicographic in the reference's 1-D SOD order: l = I + q*(J + q*K),
// (I,J,K) = (AtoI,AtoJ,AtoK) - 1.
//   geo   [nelem][5][nn][2] He(i,k) (component c = i + 3k) and gpvol of each
//                          element node as five pairs (c = 0,1 | 2,3 | 4,5 |
//                          6,7 | 8 and gpvol): a node's ten values in five
//                          16-byte (FP64) or 8-byte (FP32) loads, each wave's
//                          load contiguous (GeoNode)
// tab.D[q*q] is the 1-D derivative D[i*q+n] = l_n'(xi_i); tab.l2g[nn] maps l
// to the 0-based reference (Gmsh) node number, used for arrays kept in the
// reference layout (mu_e, mu_sgs).
//
// Assembly plan (built once per mesh by the runtime).  Elements are grouped
// into batches of Shape<P>::BATCH consecutive elements; one workgroup owns a
// batch, sums the contributions of its elements to every node of the batch in
// LDS, and writes each of the batch's distinct nodes once:
//   conn  [nelem][nn]   0-based global node of each element node
//   lidx  [nelem][nn]   position of that node in its batch's node list
//   uoff  [nbatch+1]    node list of batch b: unode[uoff[b] .. uoff[b+1])
//   unode [..]          sorted distinct global nodes of each batch; bit 31 set
//                       when this batch is the first writer of the node
//   batches[nb]         the batches processed by one launch; a workgroup takes
//                       kpb consecutive ones, one after the other, with the
//                       loads of the next pass always in flight.  Null: the
//                       batches 0 .. nb-1 in mesh order, with no list to load
// Within a batch the element-node values are summed in a fixed order
// (add_pass).  Two ways to combine the batches' partial sums at nodes shared
// between batches (mode):
//   kColored: batches are coloured so that batches of one colour share no
//     node; one launch per colour, each node written with a plain store by its
//     first batch and read-add-stored by later ones.  No atomics, no zero-fill,
//     and the summation order is fixed (bitwise reproducible).
//   kAtomic: one launch over all batches, FP64 atomics into zeroed outputs.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace msod {

constexpr int kMaxP = 8;
constexpr int kMaxKpb = 16;           // batches per workgroup, at most (LDS list of the block's batches)
constexpr int kBatchesPerBlock = 8;   // batches per workgroup (convection: measured best of 1..16 at p=4)
static_assert(kBatchesPerBlock <= kMaxKpb, "the block's batch list holds kpb entries");
enum AssemblyMode { kColored = 0, kAtomic = 1 };

template <int P>
struct Shape {
    static constexpr int Q = P + 1;
    static constexpr int NN = Q * Q * Q;
    // one node per thread; several elements per pass while they fit 256 threads
    static constexpr int EPB = NN <= 128 ? 256 / NN : 1;
    static constexpr int NT = NN <= 128 ? 256 : ((NN + 63) / 64) * 64;
    // elements per batch: about 512 element nodes, a whole number of passes
    static constexpr int BATCH_RAW = 512 / NN < 1 ? 1 : (512 / NN > 64 ? 64 : 512 / NN);
    static constexpr int BATCH = BATCH_RAW < EPB ? EPB : (BATCH_RAW / EPB) * EPB;
    static constexpr int NPASS = BATCH / EPB;
    static constexpr int UMAX = BATCH * NN;
};

// The convection's batches: Shape<P>'s, except from porder 6 on, where one
// element fills a workgroup; there two elements per batch (two passes) share
// a face in the node sums and halve the write phases per element (48^3 p=7:
// 4.75 -> 4.01 ms).  The diffusion measured slower with two (3.58 -> 3.81 ms):
// each kernel has its own plan where the sizes differ.
template <int P>
struct CShape {
    static constexpr int Q = Shape<P>::Q;
    static constexpr int NN = Shape<P>::NN;
    static constexpr int EPB = Shape<P>::EPB;
    static constexpr int NT = Shape<P>::NT;
    static constexpr int BATCH = NN >= 343 ? 2 * EPB : Shape<P>::BATCH;
    static constexpr int NPASS = BATCH / EPB;
    static constexpr int UMAX = BATCH * NN;
};

// Host-side copies of Shape<P>::BATCH (diffusion and closure kernels) and
// CShape<P>::BATCH (convection) for run-time porder.
int batch_size(int porder);
int convec_batch_size(int porder);

struct Tables {
    const double *D;      // [q*q]
    const int32_t *l2g;   // [nn]
};

struct Plan {
    const uint32_t *conn;
    const uint16_t *lidx;
    const uint32_t *uoff;
    const uint32_t *unode;
    const uint32_t *batches;
    int64_t nb;
    int mode;
    int kpb;  // consecutive batches of the list per workgroup
};

// Workgroups of a launch: each takes kpb consecutive batches of the list.
inline unsigned plan_blocks(const Plan &p) { return (unsigned)((p.nb + p.kpb - 1) / p.kpb); }

// Kernel arguments.  Storage types: TG for the element-sized arrays (geo;
// in the diffusion also mu_e, mu_sgs), TF for the nodal fields.  The
// kernels compute in FP64 whatever the storage: on gfx950 an FP64 FMA issues
// at the rate of an unpacked FP32 one, so FP32 storage buys bytes, not time
// in the ALUs, and FP64 arithmetic keeps rounding out of the derivative
// sweeps (whose cancellation would amplify FP32 rounding by about 1/h).
//   <double, double>  FP64 form (the reference's operands)
//   <float,  float>   FP32 storage of every input
//   <float,  double>  FP32 element arrays, FP64 nodal fields
// TO: type of the outputs, FP64, or FP32 for callers built at the
// reference's default real kind (rp = 4); the sums before the output stay
// FP64.  TU, TP: storage of the velocity and of the pressure where it differs
// from the other nodal fields (the FP32 split keeps both in FP64: the
// velocity is differentiated twice in the diffusive momentum residual, the
// pressure gradient dominates the convective one).  The kernels take the
// argument type as a template parameter and read the storage types from it.
template <typename TG, typename TF, typename TO = double, typename TU = TF, typename TP = TF>
struct ConvecArgs {
    using G = TG;
    using F = TF;
    using O = TO;
    using U = TU;
    using Pt = TP;
    int64_t nelem, npoin;
    Plan plan;
    const TG *geo;
    const TU *u;
    const TF *q, *rho, *E;
    const TP *pr;                    // nullptr: recover pressure from E
    double gamma_m1;                 // gamma-1, used when pr == nullptr
    TO *Rmass, *Rmom, *Rener;
    Tables tab;
};

template <typename TG, typename TF, typename TO = double, typename TU = TF>
struct DiffuArgs {
    using G = TG;
    using F = TF;
    using O = TO;
    using U = TU;
    int64_t nelem, npoin;
    Plan plan;
    const TG *geo;
    const TF *rho, *Tem, *mu_fluid;
    const TU *u;
    const TG *mu_e, *mu_sgs;
    double Cp, Pr, c_rho, c_ener;
    TO *Rmass, *Rmom, *Rener;
    Tables tab;
};

// The FP32 split: FP32 geometry and nodal fields, FP64 velocity (and
// pressure), FP64 arithmetic and outputs.
using ConvecSplitArgs = ConvecArgs<float, float, double, double, double>;
using DiffuSplitArgs = DiffuArgs<float, float, double, double>;

// Launchers (return hipError_t of the launch); A is a ConvecArgs or
// DiffuArgs instantiation.
template <class A>
hipError_t launch_convec(int porder, const A &a, hipStream_t s);
template <class A>
hipError_t launch_diffu(int porder, const A &a, hipStream_t s);
// FP32 storage and FP32 gathers, fluxes and derivative sweeps; FP64 node sums,
// assembly and (convection) pressure recovery.
hipError_t launch_convec_fp32(int porder, const ConvecArgs<float, float> &a, hipStream_t s);
hipError_t launch_diffu_fp32(int porder, const DiffuArgs<float, float> &a, hipStream_t s);

// Zero Rmass (n), Rmom (3n), Rener (n).
hipError_t launch_zero_outputs(double *Rmass, double *Rmom, double *Rener, int64_t n,
                               hipStream_t s);
hipError_t launch_zero_outputs(float *Rmass, float *Rmom, float *Rener, int64_t n, hipStream_t s);
// Zero the listed nodes of the three outputs.
template <typename TO>
hipError_t launch_zero_nodes(TO *Rmass, TO *Rmom, TO *Rener, int64_t npoin, const uint32_t *nodes,
                             int64_t count, hipStream_t s);

// Element-major geometry from the reference layouts (see top of file).
template <typename T>
hipError_t launch_build_geo(const T *He, const T *gpvol, int64_t nelem, int nn,
                            const int32_t *l2g, T *geo, hipStream_t s);

// He and gpvol of one element node from geo: He(i,k) = he(i + 3k), with the
// component index a compile-time constant (a register of the loaded pair, no
// copy: the values stay in flight until used).
template <typename T>
using GeoPair = T __attribute__((ext_vector_type(2)));

template <typename T>
struct GeoNode {
    GeoPair<T> v[5];
    __device__ __forceinline__ T he(int c) const { return (c & 1) ? v[c >> 1].y : v[c >> 1].x; }
    __device__ __forceinline__ T gvol() const { return v[4].y; }
};

// NT: non-temporal loads (the geometry is read once per call).
template <int NN, bool NT, typename T>
__device__ __forceinline__ GeoNode<T> load_geo(const T *geo, int64_t e, int l)
{
    const GeoPair<T> *p = reinterpret_cast<const GeoPair<T> *>(geo) + e * 5 * NN + l;
    GeoNode<T> g;
#pragma unroll
    for (int k = 0; k < 5; ++k) g.v[k] = NT ? __builtin_nontemporal_load(&p[k * NN]) : p[k * NN];
    return g;
}

// Entry i of a launch's batch list (Plan::batches); a null list is mesh order.
__device__ __forceinline__ uint32_t plan_batch(const uint32_t *batches, uint32_t i)
{
    return batches ? batches[i] : i;
}

// Blocks of a launch are dealt to the chip's 8 XCDs round robin (block i to
// XCD i % 8), each XCD with its own L2.  xcd_slot gives every XCD a
// contiguous range of the launch's batch list, so that batches adjacent in
// the list (which share nodes and cache lines) meet in the same L2.
// Bijective on [0, n) for any n.
__device__ __forceinline__ uint32_t xcd_slot(uint32_t i, uint32_t n)
{
    constexpr uint32_t kXcd = 8;
    const uint32_t x = i % kXcd, k = i / kXcd, q = n / kXcd, r = n % kXcd;
    return x * q + (x < r ? x : r) + k;
}

// Write phase shared by the kernels: the batch's node sums sR[c][i]
// (c = Rmass, Rmom x3, Rener) go to the outputs.  Thread t holds entries t,
// t + NT, ... of the batch's node list in registers (wn), loaded when the
// batch started.  Atomic mode: every entry is added atomically onto the
// zero-filled outputs (agent-scope FP64 atomics, performed past the L2; a
// plain store of the nodes no other batch holds was tried and is much
// slower, since those lines also take atomics).  Coloured mode: the first
// writer stores, the others add.
template <int UMAX, int NT, typename TO>
__device__ __forceinline__ void write_batch_nodes(int mode, const uint32_t (&wn)[(UMAX + NT - 1) / NT],
                                                  uint32_t U, const double (*sR)[UMAX], TO *Rmass,
                                                  TO *Rmom, TO *Rener, int64_t npoin)
{
#pragma unroll
    for (int j = 0; j < (UMAX + NT - 1) / NT; ++j) {
        const uint32_t i = threadIdx.x + j * NT;
        if (i >= U) break;
        const uint32_t w = wn[j];
        const int64_t g = w & 0x7fffffffu;
        TO *dst[5] = {Rmass + g, Rmom + g, Rmom + npoin + g, Rmom + 2 * npoin + g, Rener + g};
        if (mode == kAtomic) {
#pragma unroll
            for (int c = 0; c < 5; ++c) unsafeAtomicAdd(dst[c], (TO)sR[c][i]);
        } else if (w >> 31) {
#pragma unroll
            for (int c = 0; c < 5; ++c) *dst[c] = (TO)sR[c][i];
        } else {
#pragma unroll
            for (int c = 0; c < 5; ++c) *dst[c] += (TO)sR[c][i];
        }
    }
}

// Adds the five residual values of every element node of a pass into the
// batch's node sums, one element of the pass after the other: the nodes of
// one element are distinct, so no two threads of a round touch the same sum,
// and the summation order (pass, then element, within the batch) is fixed.
// Every thread of the block calls it (uniform control flow, barriers).
// In the atomic mode, whose global sums have no fixed order anyway, the
// batch sums use LDS atomics instead (no rounds); the caller synchronises
// before the batch is written.
template <int UMAX, int EPB>
__device__ __forceinline__ void add_pass(double (*sR)[UMAX], int mode, int le, bool active, uint32_t li,
                                         const double (&v)[5])
{
    if (mode == kAtomic) {
        if (active) {
#pragma unroll
            for (int c = 0; c < 5; ++c)
                __hip_atomic_fetch_add(&sR[c][li], v[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
        return;
    }
#pragma unroll 1
    for (int k = 0; k < EPB; ++k) {
        if (active && le == k) {
#pragma unroll
            for (int c = 0; c < 5; ++c) sR[c][li] += v[c];
        }
        __syncthreads();
    }
}

}  // namespace msod

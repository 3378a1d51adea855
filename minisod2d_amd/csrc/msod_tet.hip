// Linear-tetrahedron kernels: fem_convec (elem_convec.f90:839-948),
// fem_diffu (elem_diffu.f90:199-317) and tet_jacobian
// (jacobian_oper.f90:138-267).
//
// One thread per element, one 64-thread workgroup (one wave) per 64
// consecutive elements.  The element arrays are read once and streamed:
//  * He(3,3,4,e) is 288 contiguous bytes per element.  Read per thread, the
//    lanes of a wave would sit 288 bytes apart and each load would touch 64
//    cache lines; instead the wave copies its 64 elements' 18 KB with loads in
//    which neighbouring lanes take neighbouring doubles, into LDS padded to an
//    odd stride (37 doubles per element), and each thread reads its own
//    element from there.
//  * connec, gpvol, mu_e and mu_sgs have the element index fastest (or four
//    values per element contiguous): neighbouring lanes, neighbouring words.
// The four nodes' values are gathered into registers, the four Gauss points
// are evaluated in the reference's operation order, and the twenty element
// residuals are stored (mesh without shared nodes) or added with FP64
// atomics (any mesh).  Measured at 33.75 M disjoint tetrahedra (convection):
// 17.0 ms with atomics, 7.8 ms with stores (one atomic per lane, neighbouring
// lanes 32 bytes apart), 5.6 ms once the He copy issued its 36 loads at once;
// gathering before the He copy's LDS writes instead of after: 6.0 ms; loading
// the connectivity before the He copy: 5.8 ms.
//
// The kernels never read outside the arrays: connec entries outside
// 1..npoin are clamped (the runtime rejects such meshes before a launch, and
// the geometry kernel counts them).
#include "msod_tet.h"

namespace msod {

namespace {

constexpr int kWave = 64;
constexpr int kHeStride = 37;  // doubles per element in LDS (odd: spreads the elements over the banks)

__device__ __forceinline__ int64_t node_of(const int32_t *connec, int64_t nelem, int64_t npoin, int64_t e,
                                           int a)
{
    int64_t n = (int64_t)__builtin_nontemporal_load(&connec[e + nelem * a]) - 1;
    return n < 0 ? 0 : (n >= npoin ? npoin - 1 : n);
}

// Residual of element node n into the output: plain store or atomic add.
template <bool STORE>
__device__ __forceinline__ void put(double *p, double v)
{
    if (STORE)
        *p = v;
    else
        unsafeAtomicAdd(p, v);
}

// The wave's 64 elements' He into sHe[element][0..35]; returns after the
// copy is visible to every lane.  The 36 loads are unconditional (past the
// last element they re-read its last value, and the slots they fill belong
// to no element), so that all of them are in flight at once: guarded one by
// one, each load was waited for before the next was issued.
__device__ __forceinline__ void stage_he(const double *He, int64_t e0, int64_t nelem,
                                         double (*sHe)[kHeStride])
{
    const int lane = threadIdx.x;
    const int64_t ne = nelem - e0 < kWave ? nelem - e0 : kWave;
    const int64_t last = 36 * ne - 1;
    const double *src = He + 36 * e0;
    double v[36];
#pragma unroll
    for (int k = 0; k < 36; ++k) {
        const int64_t j = k * kWave + lane;
        v[k] = __builtin_nontemporal_load(&src[j < last ? j : last]);
    }
#pragma unroll
    for (int k = 0; k < 36; ++k) {
        const int j = k * kWave + lane;  // < 36 * 64: element j / 36 < 64
        sHe[j / 36][j % 36] = v[k];
    }
    __syncthreads();
}

// gpcar(n,i) = sum_k He(i,k,g) dNgp(k,n,g), as DOT_PRODUCT(He(i,:,g,e), dNgp(:,n,g))
__device__ __forceinline__ void shape_derivatives(const double *h, const TetTables &t, int g, double (&gc)[4][3])
{
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int n = 0; n < 4; ++n) {
            const double *dn = &t.dNgp[3 * (n + 4 * g)];
            gc[n][i] = h[i] * dn[0] + h[i + 3] * dn[1] + h[i + 6] * dn[2];
        }
}

template <bool STORE>
__global__ void __launch_bounds__(kWave) k_fem_convec(TetConvecArgs a)
{
    __shared__ double sHe[kWave][kHeStride];
    const int64_t e0 = (int64_t)blockIdx.x * kWave;
    const int64_t e = e0 + threadIdx.x;
    stage_he(a.He, e0, a.nelem, sHe);
    if (e >= a.nelem) return;
    const int64_t np = a.npoin;

    int64_t nd[4];
    double xu[4][3], xq[4][3], xr[4], xp[4], xEp[4];
#pragma unroll
    for (int n = 0; n < 4; ++n) nd[n] = node_of(a.connec, a.nelem, np, e, n);
#pragma unroll
    for (int n = 0; n < 4; ++n) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            xu[n][d] = a.u[nd[n] + np * d];
            xq[n][d] = a.q[nd[n] + np * d];
        }
        xr[n] = a.rho[nd[n]];
        xp[n] = a.pr[nd[n]];
        xEp[n] = a.E[nd[n]] + xp[n];  // E + p, as the reference forms it at every use
    }
    double gv[4];
#pragma unroll
    for (int g = 0; g < 4; ++g) gv[g] = __builtin_nontemporal_load(&a.gpvol[4 * e + g]);

    double Rm[4] = {0, 0, 0, 0}, Re[4] = {0, 0, 0, 0};
    double Rmo[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        double h[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) h[c] = sHe[threadIdx.x][9 * g + c];  // h[i + 3k] = He(i,k,g)
        double gc[4][3];
        shape_derivatives(h, a.t, g, gc);
        double divQ = 0, divFe = 0, gR[3] = {0, 0, 0}, gE[3] = {0, 0, 0}, gP[3] = {0, 0, 0};
        double gU[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, dF[3] = {0, 0, 0};
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int n = 0; n < 4; ++n) {
                divQ += gc[n][i] * xq[n][i];
                divFe += gc[n][i] * xEp[n] * xu[n][i];
                gR[i] += gc[n][i] * xr[n];
                gE[i] += gc[n][i] * xEp[n];
                gP[i] += gc[n][i] * xp[n];
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    gU[i][j] += gc[n][j] * xu[n][i];  // d u_i / d x_j
                    dF[i] += gc[n][j] * xq[n][i] * xu[n][j];
                }
            }
        const double divU = gU[0][0] + gU[1][1] + gU[2][2];
#pragma unroll
        for (int n = 0; n < 4; ++n) {
            double a1 = 0, a2 = 0;
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                a1 += gR[i] * xu[n][i];
                a2 += gE[i] * xu[n][i];
            }
            const double w = gv[g] * a.t.Ngp[g + 4 * n] * 0.5;  // gpvol Ngp(g,n) / 2
            Rm[n] += w * (divQ + xr[n] * divU + a1);
            Re[n] += w * (divFe + xEp[n] * divU + a2);
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                double a3 = 0, a4 = 0;
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    a3 += gU[i][j] * xq[n][j];
                    a4 += xu[n][i] * xu[n][j] * gR[j];
                }
                Rmo[n][i] += w * (dF[i] + xq[n][i] * divU + a3 + a4 + gP[i]);
            }
        }
    }
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        put<STORE>(&a.Rmass[nd[n]], Rm[n]);
        put<STORE>(&a.Rener[nd[n]], Re[n]);
#pragma unroll
        for (int i = 0; i < 3; ++i) put<STORE>(&a.Rmom[nd[n] + np * i], Rmo[n][i]);
    }
}

template <bool STORE>
__global__ void __launch_bounds__(kWave) k_fem_diffu(TetDiffuArgs a)
{
    __shared__ double sHe[kWave][kHeStride];
    const int64_t e0 = (int64_t)blockIdx.x * kWave;
    const int64_t e = e0 + threadIdx.x;
    stage_he(a.He, e0, a.nelem, sHe);
    if (e >= a.nelem) return;
    const int64_t np = a.npoin, ne = a.nelem;

    int64_t nd[4];
    double xr[4], xu[4][3], xT[4], xmf[4], me[4], ms[4];
#pragma unroll
    for (int n = 0; n < 4; ++n) nd[n] = node_of(a.connec, ne, np, e, n);
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        xr[n] = a.rho[nd[n]];
#pragma unroll
        for (int d = 0; d < 3; ++d) xu[n][d] = a.u[nd[n] + np * d];
        xT[n] = a.Tem[nd[n]];
        xmf[n] = a.mu_fluid[nd[n]];
        me[n] = __builtin_nontemporal_load(&a.mu_e[e + ne * n]);
        ms[n] = __builtin_nontemporal_load(&a.mu_sgs[e + ne * n]);
    }
    double gv[4];
#pragma unroll
    for (int g = 0; g < 4; ++g) gv[g] = __builtin_nontemporal_load(&a.gpvol[4 * e + g]);
    const double cpr = a.Cp / a.Pr;

    double Rm[4] = {0, 0, 0, 0}, Re[4] = {0, 0, 0, 0};
    double Rmo[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        // fluid properties at the point: the reference weights node n with
        // Ngp(n,g), the transpose of the table's (point, node) layout
        double aR = 0;
#pragma unroll
        for (int n = 0; n < 4; ++n) aR += a.t.Ngp[n + 4 * g] * xr[n];
        const double crr = a.c_rho / aR;
        double aNu = 0, aMu = 0, aK = 0;
#pragma unroll
        for (int n = 0; n < 4; ++n) {
            const double N = a.t.Ngp[n + 4 * g];
            aNu += N * me[n] * crr;
            aMu += N * (xmf[n] + aR * ms[n] + me[n]);
            aK += N * (xmf[n] * cpr + a.c_ener * me[n] / 0.4 + aR * ms[n] / 0.9);
        }
        double h[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) h[c] = sHe[threadIdx.x][9 * g + c];
        double gc[4][3];
        shape_derivatives(h, a.t, g, gc);
        double gR[3] = {0, 0, 0}, gT[3] = {0, 0, 0};
        double gU[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int n = 0; n < 4; ++n) {
                gR[i] += gc[n][i] * xr[n];
                gT[i] += gc[n][i] * xT[n];
#pragma unroll
                for (int j = 0; j < 3; ++j) gU[i][j] += gc[n][i] * xu[n][j];  // d u_j / d x_i
            }
        const double divU = gU[0][0] + gU[1][1] + gU[2][2];
        double tau[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            tau[i][i] = -(2.0 / 3.0) * divU;
#pragma unroll
            for (int j = 0; j < 3; ++j) tau[i][j] = tau[i][j] + gU[i][j] + gU[j][i];
        }
        double tauU[3] = {0, 0, 0};
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            double aU = 0;
#pragma unroll
            for (int n = 0; n < 4; ++n) aU += a.t.Ngp[n + 4 * g] * xu[n][j];
#pragma unroll
            for (int i = 0; i < 3; ++i) tauU[i] += tau[i][j] * aU;
        }
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int n = 0; n < 4; ++n) {
                Rm[n] += gv[g] * aNu * gc[n][i] * gR[i];
                Re[n] += gv[g] * gc[n][i] * (aMu * tauU[i] + aK * gT[i]);
#pragma unroll
                for (int j = 0; j < 3; ++j) Rmo[n][i] += gv[g] * gc[n][j] * aMu * tau[i][j];
            }
    }
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        put<STORE>(&a.Rmass[nd[n]], Rm[n]);
        put<STORE>(&a.Rener[nd[n]], Re[n]);
#pragma unroll
        for (int i = 0; i < 3; ++i) put<STORE>(&a.Rmom[nd[n] + np * i], Rmo[n][i]);
    }
}

// Je(i,j) = sum_a dNgp(i,a,g) coord(connec(e,a),j); det, cofactors and the
// inverse in the reference's operation order (jacobian_oper.f90:213-263).
__global__ void __launch_bounds__(256) k_tet_jacobian(TetJacobianArgs a)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= a.nelem) return;
    const int64_t np = a.npoin;
    double X[4][3];
    unsigned long long bad = 0;
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        int64_t v = (int64_t)a.connec[e + a.nelem * n] - 1;
        if (v < 0 || v >= np) {
            ++bad;
            v = 0;
        }
#pragma unroll
        for (int j = 0; j < 3; ++j) X[n][j] = a.coord[v + np * j];
    }
    unsigned long long nonpos = 0;
    double *He = a.He + 36 * e;
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        double J[3][3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                double s = a.t.dNgp[3 * (0 + 4 * g) + i] * X[0][j];
#pragma unroll
                for (int n = 1; n < 4; ++n) s += a.t.dNgp[3 * (n + 4 * g) + i] * X[n][j];
                J[i][j] = 0.0 + s;
            }
        const double det = J[0][0] * J[1][1] * J[2][2] + J[0][1] * J[1][2] * J[2][0] +
                           J[0][2] * J[1][0] * J[2][1] - J[2][0] * J[1][1] * J[0][2] -
                           J[2][1] * J[1][2] * J[0][0] - J[2][2] * J[1][0] * J[0][1];
        // minors, signs, transpose: He(i,j) = cofactor(j,i) / det
        const double m11 = J[1][1] * J[2][2] - J[2][1] * J[1][2];
        const double m12 = -(J[1][0] * J[2][2] - J[2][0] * J[1][2]);
        const double m13 = J[1][0] * J[2][1] - J[2][0] * J[1][1];
        const double m21 = -(J[0][1] * J[2][2] - J[2][1] * J[0][2]);
        const double m22 = J[0][0] * J[2][2] - J[2][0] * J[0][2];
        const double m23 = -(J[0][0] * J[2][1] - J[2][0] * J[0][1]);
        const double m31 = J[0][1] * J[1][2] - J[1][1] * J[0][2];
        const double m32 = -(J[0][0] * J[1][2] - J[1][0] * J[0][2]);
        const double m33 = J[0][0] * J[1][1] - J[1][0] * J[0][1];
        const double r = 1.0 / det;
        double *h = He + 9 * g;  // h[i + 3j] = He(i+1, j+1, g)
        h[0] = r * m11;
        h[3] = r * m21;
        h[6] = r * m31;
        h[1] = r * m12;
        h[4] = r * m22;
        h[7] = r * m32;
        h[2] = r * m13;
        h[5] = r * m23;
        h[8] = r * m33;
        a.gpvol[4 * e + g] = a.t.wgp[g] * det;
        nonpos += !(det > 0.0);
    }
    if (bad) atomicAdd(&a.counts[0], bad);
    if (nonpos) atomicAdd(&a.counts[1], nonpos);
}

__global__ void k_count_refs(const int32_t *connec, int64_t n, int64_t npoin, uint32_t *refs,
                             unsigned long long *stats)
{
    unsigned long long bad = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t v = connec[i];
        if (v < 1 || v > npoin)
            ++bad;
        else
            atomicAdd(&refs[v - 1], 1u);
    }
    if (bad) atomicAdd(&stats[0], bad);
}

__global__ void k_ref_summary(const uint32_t *refs, int64_t npoin, unsigned long long *stats)
{
    unsigned long long shared = 0, unref = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < npoin;
         i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t r = refs[i];
        shared += r > 1;
        unref += r == 0;
    }
    if (shared) atomicAdd(&stats[1], shared);
    if (unref) atomicAdd(&stats[2], unref);
}

__global__ void k_list_unreferenced(const uint32_t *refs, int64_t npoin, uint32_t *list,
                                    unsigned long long *cursor)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < npoin;
         i += (int64_t)gridDim.x * blockDim.x)
        if (refs[i] == 0) list[atomicAdd(cursor, 1ull)] = (uint32_t)i;
}

unsigned blocks_of(int64_t n, int per)
{
    return (unsigned)((n + per - 1) / per);
}

}  // namespace

hipError_t launch_fem_convec(const TetConvecArgs &a, hipStream_t s)
{
    if (a.nelem <= 0) return hipSuccess;
    if (a.store)
        hipLaunchKernelGGL(k_fem_convec<true>, dim3(blocks_of(a.nelem, kWave)), dim3(kWave), 0, s, a);
    else
        hipLaunchKernelGGL(k_fem_convec<false>, dim3(blocks_of(a.nelem, kWave)), dim3(kWave), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_fem_diffu(const TetDiffuArgs &a, hipStream_t s)
{
    if (a.nelem <= 0) return hipSuccess;
    if (a.store)
        hipLaunchKernelGGL(k_fem_diffu<true>, dim3(blocks_of(a.nelem, kWave)), dim3(kWave), 0, s, a);
    else
        hipLaunchKernelGGL(k_fem_diffu<false>, dim3(blocks_of(a.nelem, kWave)), dim3(kWave), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_tet_jacobian(const TetJacobianArgs &a, hipStream_t s)
{
    if (a.nelem <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_tet_jacobian, dim3(blocks_of(a.nelem, 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

static dim3 stride_grid(int64_t n)
{
    const int64_t b = blocks_of(n, 256);
    return dim3((unsigned)(b < 4096 ? (b < 1 ? 1 : b) : 4096));
}

hipError_t launch_count_refs(const int32_t *connec, int64_t n, int64_t npoin, uint32_t *refs,
                             unsigned long long *stats, hipStream_t s)
{
    hipLaunchKernelGGL(k_count_refs, stride_grid(n), dim3(256), 0, s, connec, n, npoin, refs, stats);
    return hipGetLastError();
}

hipError_t launch_ref_summary(const uint32_t *refs, int64_t npoin, unsigned long long *stats, hipStream_t s)
{
    hipLaunchKernelGGL(k_ref_summary, stride_grid(npoin), dim3(256), 0, s, refs, npoin, stats);
    return hipGetLastError();
}

hipError_t launch_list_unreferenced(const uint32_t *refs, int64_t npoin, uint32_t *list,
                                    unsigned long long *cursor, hipStream_t s)
{
    hipLaunchKernelGGL(k_list_unreferenced, stride_grid(npoin), dim3(256), 0, s, refs, npoin, list, cursor);
    return hipGetLastError();
}

}  // namespace msod

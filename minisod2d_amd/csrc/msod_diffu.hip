// Diffusive residual kernel (counterpart of full_diffusion_ijk,
// src/kernels/sources/elem_diffu.f90:10-197).
//
// One thread per element node, Shape<P>::EPB elements per block, two passes
// through LDS:
//  pass 1 (elem_diffu.f90:67-139): sweeps of rho, T, u along the node's three
//    grid lines; He of the node gives grad rho, grad T and G = grad u; then
//    the viscous stress tau and tau.u.  Instead of storing the six physical
//    flux vectors (grad rho, tau.u, grad T, tau rows) as the reference does
//    and re-reading He and gpvol of every source node in pass 2, each thread
//    stores their contravariant components
//        F_d = gpvol * sum_i He(i,d) Phi_i        (6 vectors x 3 = 18 values)
//    which is exactly the factor the reference's pass 2 forms at the source.
//  pass 2 (elem_diffu.f90:141-193): the transposed sweep
//        W[Phi](a) = sum_d sum_n D[n][I_d(a)] F_d(s_d(n))
//    then Rmass += nu_e W[grad rho], Rener += W[tau.u] + kappa_e W[grad T],
//    Rmom_r += W[tau_r], with nu_e and kappa_e of the target node (the
//    reference multiplies kappa_e inside the sum; same value, other rounding).
//
// T = double: FP64 form; T = float: mixed form (FP32 sweeps, FP64 assembly).
#include "msod_kernels.h"

namespace msod {

template <typename T>
__device__ __forceinline__ void scatter_d(double *dst, uint32_t w, T v)
{
    const uint32_t g = w & 0x7fffffffu;
    if (w >> 31)
        dst[g] = (double)v;
    else
        unsafeAtomicAdd(dst + g, (double)v);
}

template <int P, typename T>
__global__ void __launch_bounds__(Shape<P>::NT) k_diffu(DiffuArgs<T> a)
{
    using S = Shape<P>;
    constexpr int Q = S::Q, NN = S::NN, EPB = S::EPB;
    __shared__ T sD[Q * Q];
    __shared__ T sA[EPB][5][NN];    // rho, T, u0, u1, u2
    __shared__ T sG[EPB][18][NN];   // contravariant fluxes, index f*3 + d

    const int64_t grp = xcd_group((a.nelem + EPB - 1) / EPB);
    if (grp < 0) return;  // whole block
    const int t = threadIdx.x;
    for (int i = t; i < Q * Q; i += blockDim.x) sD[i] = (T)a.tab.D[i];

    const int le = t / NN;
    const int l = t - le * NN;
    const int64_t e = grp * EPB + le;
    const bool active = (le < EPB) && (e < a.nelem);

    uint32_t cw = 0;
    T r = 0, muf = 0, mue = 0, mus = 0, uu[3] = {0, 0, 0};
    if (active) {
        cw = a.conn[e * NN + l];
        const int64_t g = cw & 0x7fffffffu;
        const int64_t np = a.npoin;
        r = a.rho[g];
        muf = a.mu_fluid[g];
        const int64_t ge = e + a.nelem * (int64_t)a.tab.l2g[l];  // reference layout (nelem,ngaus)
        mue = a.mu_e[ge];
        mus = a.mu_sgs[ge];
        sA[le][0][l] = r;
        sA[le][1][l] = a.Tem[g];
        for (int d = 0; d < 3; ++d) {
            uu[d] = a.u[g + d * np];
            sA[le][2 + d][l] = uu[d];
        }
    }
    __syncthreads();

    const int I = l % Q, J = (l / Q) % Q, K = l / (Q * Q);
    if (active) {
        const T *geo = a.geo + e * 9 * NN + l;
        T h[3][3];  // h[i][d] = He(i,d)
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int i = 0; i < 3; ++i) h[i][d] = geo[(i + 3 * d) * NN];

        T gR[3] = {0, 0, 0}, gT[3] = {0, 0, 0}, G[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int Id = d == 0 ? I : (d == 1 ? J : K);
            const int stride = d == 0 ? 1 : (d == 1 ? Q : Q * Q);
            const int base = l - Id * stride;
            T aR = 0, aT = 0, aU0 = 0, aU1 = 0, aU2 = 0;
#pragma unroll 1
            for (int n = 0; n < Q; ++n) {
                const int s = base + n * stride;
                const T Dv = sD[Id * Q + n];
                aR += Dv * sA[le][0][s];
                aT += Dv * sA[le][1][s];
                aU0 += Dv * sA[le][2][s];
                aU1 += Dv * sA[le][3][s];
                aU2 += Dv * sA[le][4][s];
            }
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                gR[i] += h[i][d] * aR;
                gT[i] += h[i][d] * aT;
                // G[i][j] = d u_i / d x_j = sum_d He(j,d) D_d[u_i]
                G[0][i] += h[i][d] * aU0;
                G[1][i] += h[i][d] * aU1;
                G[2][i] += h[i][d] * aU2;
            }
        }
        const T divU = G[0][0] + G[1][1] + G[2][2];
        const T mu_f = muf + r * mus;
        const T mu_t = mu_f + mue;
        const T lam = mu_f * (T)(2.0 / 3.0) * divU;
        T tau[3][3], tauU[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            T acc = 0;
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                tau[i][j] = mu_t * (G[i][j] + G[j][i]);
                acc += tau[i][j] * uu[j];
            }
            tauU[i] = acc - lam * uu[i];
            tau[i][i] -= lam;
        }
        const T gv = a.gvol[e * NN + l];
        // contravariant fluxes F_d = gpvol * sum_i He(i,d) Phi_i
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const T c0 = gv * h[0][d], c1 = gv * h[1][d], c2 = gv * h[2][d];
            sG[le][0 * 3 + d][l] = c0 * gR[0] + c1 * gR[1] + c2 * gR[2];
            sG[le][1 * 3 + d][l] = c0 * tauU[0] + c1 * tauU[1] + c2 * tauU[2];
            sG[le][2 * 3 + d][l] = c0 * gT[0] + c1 * gT[1] + c2 * gT[2];
            sG[le][3 * 3 + d][l] = c0 * tau[0][0] + c1 * tau[0][1] + c2 * tau[0][2];
            sG[le][4 * 3 + d][l] = c0 * tau[1][0] + c1 * tau[1][1] + c2 * tau[1][2];
            sG[le][5 * 3 + d][l] = c0 * tau[2][0] + c1 * tau[2][1] + c2 * tau[2][2];
        }
    }
    __syncthreads();
    if (!active) return;

    T W[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int Id = d == 0 ? I : (d == 1 ? J : K);
        const int stride = d == 0 ? 1 : (d == 1 ? Q : Q * Q);
        const int base = l - Id * stride;
#pragma unroll 1
        for (int n = 0; n < Q; ++n) {
            const int s = base + n * stride;
            const T Dv = sD[n * Q + Id];  // transposed: l_{I_d(a)}'(xi_n)
#pragma unroll
            for (int f = 0; f < 6; ++f) W[f] += Dv * sG[le][f * 3 + d][s];
        }
    }
    const T Cp = (T)a.Cp, Pr = (T)a.Pr;
    const T nu_e = (T)a.c_rho * mue / r;
    const T kappa = muf * Cp / Pr + (T)a.c_ener * mue * Cp / Pr + Cp * r * mus / (T)0.9;
    scatter_d(a.Rmass, cw, nu_e * W[0]);
    scatter_d(a.Rener, cw, W[1] + kappa * W[2]);
    scatter_d(a.Rmom, cw, W[3]);
    scatter_d(a.Rmom + a.npoin, cw, W[4]);
    scatter_d(a.Rmom + 2 * a.npoin, cw, W[5]);
}

template <int P, typename T>
static hipError_t launch_p(const DiffuArgs<T> &a, hipStream_t s)
{
    using S = Shape<P>;
    const int64_t groups = (a.nelem + S::EPB - 1) / S::EPB;
    hipLaunchKernelGGL((k_diffu<P, T>), dim3((unsigned)xcd_grid(groups)), dim3(S::NT), 0, s, a);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_diffu(int porder, const DiffuArgs<T> &a, hipStream_t s)
{
    switch (porder) {
    case 1: return launch_p<1, T>(a, s);
    case 2: return launch_p<2, T>(a, s);
    case 3: return launch_p<3, T>(a, s);
    case 4: return launch_p<4, T>(a, s);
    case 5: return launch_p<5, T>(a, s);
    case 6: return launch_p<6, T>(a, s);
    case 7: return launch_p<7, T>(a, s);
    case 8: return launch_p<8, T>(a, s);
    default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_diffu<double>(int, const DiffuArgs<double> &, hipStream_t);
template hipError_t launch_diffu<float>(int, const DiffuArgs<float> &, hipStream_t);

}  // namespace msod

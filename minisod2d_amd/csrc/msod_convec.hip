// Convective residual kernel (counterpart of full_convec_ijk,
// src/kernels/sources/elem_convec.f90:198-431, with the energy fluxes of
// full_convec_ijk_split, :674-832).
//
// One workgroup per batch of Shape<P>::BATCH elements, processed in passes of
// Shape<P>::EPB elements with one thread per element node.  Per pass:
//  1. gather: each thread reads its global node (conn, element-major) and
//     writes 10 nodal values to LDS:
//       rho, p, H = E + k, rhoH = rho E + rho k, u(3), q(3)     (k = |u|^2/2)
//  2. sweeps: for each reference direction d the thread runs the 1-D
//     derivative D[I_d][n] along its grid line (q LDS reads per value).
//     Rather than differentiating the 39 flux components the reference forms,
//     the He column of the TARGET node is folded into the fluxes at the
//     source nodes:  U_d(s) = sum_i He(i,d)u_i(s), Q_d(s) = sum_i He(i,d)q_i(s).
//     Since He is taken at the target node,
//       div(q (x) u)_i = sum_d D_d[ q_i U_d ],  div u = sum_d D_d[U_d], ...
//     which needs 20 accumulators per direction and keeps only 10 values per
//     node in LDS; the node's own u, q, H then enter the non-divergence
//     terms through U_d(a), Q_d(a) (sum_i q_i dE/dx_i = sum_d Q_d(a) D_d[E]).
//     The result equals the reference's algebra exactly in real arithmetic;
//     rounding differs at the 1e-16 level.
//  3. residual x gpvol, added into the batch's per-node sums in LDS in a
//     fixed order (add_pass).
// After the passes the batch's distinct nodes are written once each
// (write_batch: plain store / read-add-store per colour, or FP64 atomics).
//
// T = double: the FP64 form.  T = float: the mixed form (all of the above in
// FP32, pressure recovered in FP64 from E, rho, u when pr == nullptr,
// node sums and assembly in FP64).
#include "msod_kernels.h"

namespace msod {

template <int P, typename T>
__global__ void __launch_bounds__(Shape<P>::NT) k_convec(ConvecArgs<T> a)
{
    using S = Shape<P>;
    constexpr int Q = S::Q, NN = S::NN, EPB = S::EPB, UMAX = S::UMAX;
    constexpr int NF = 10;
    __shared__ T sD[Q * Q];
    __shared__ T sF[EPB][NF][NN];
    __shared__ double sR[5][UMAX];

    const Plan &pl = a.plan;
    const int64_t b = pl.batches[blockIdx.x];
    const uint32_t u0 = pl.uoff[b], U = pl.uoff[b + 1] - u0;
    const int t = threadIdx.x;
    for (int i = t; i < Q * Q; i += blockDim.x) sD[i] = (T)a.tab.D[i];
    for (uint32_t i = t; i < U; i += blockDim.x)
#pragma unroll
        for (int c = 0; c < 5; ++c) sR[c][i] = 0.0;

    const int le = t / NN;
    const int l = t - le * NN;
    const int I = l % Q, J = (l / Q) % Q, K = l / (Q * Q);

    for (int pass = 0; pass < S::NPASS; ++pass) {
        const int64_t e = b * S::BATCH + pass * EPB + le;
        const bool active = (le < EPB) && (e < a.nelem);
        T r = 0, H = 0, rH = 0, uu[3] = {0, 0, 0}, qq[3] = {0, 0, 0};
        __syncthreads();  // previous pass done with sF (and sR zeroed)
        if (active) {
            const int64_t g = pl.conn[e * NN + l];
            const int64_t np = a.npoin;
            r = a.rho[g];
            const T Ev = a.E[g];
            for (int d = 0; d < 3; ++d) {
                uu[d] = a.u[g + d * np];
                qq[d] = a.q[g + d * np];
            }
            const T k = uu[0] * uu[0] * (T)0.5 + uu[1] * uu[1] * (T)0.5 + uu[2] * uu[2] * (T)0.5;
            T pv;
            if (a.pr) {
                pv = a.pr[g];
            } else {
                // p = (gamma-1)(E - rho|u|^2/2), evaluated in FP64
                const double kd = 0.5 * ((double)uu[0] * uu[0] + (double)uu[1] * uu[1] +
                                         (double)uu[2] * uu[2]);
                pv = (T)(a.gamma_m1 * ((double)Ev - (double)r * kd));
            }
            H = Ev + k;
            rH = r * Ev + r * k;
            sF[le][0][l] = r;
            sF[le][1][l] = pv;
            sF[le][2][l] = H;
            sF[le][3][l] = rH;
            sF[le][4][l] = uu[0];
            sF[le][5][l] = uu[1];
            sF[le][6][l] = uu[2];
            sF[le][7][l] = qq[0];
            sF[le][8][l] = qq[1];
            sF[le][9][l] = qq[2];
        }
        __syncthreads();
        double v[5] = {0, 0, 0, 0, 0};
        if (active) {
        const T *geo = a.geo + e * 9 * NN + l;
        // divergence-type accumulators (coefficient independent of d)
        T aQ = 0, aU = 0, aRHU = 0, aHU = 0, aQU[3] = {0, 0, 0}, aUU[3] = {0, 0, 0};
        // gradient-type terms, contracted per direction with the node's own values
        T sRho = 0, sH = 0, sRH = 0, gP[3] = {0, 0, 0}, qGu[3] = {0, 0, 0}, uGq[3] = {0, 0, 0};

#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const T h0 = geo[(0 + 3 * d) * NN], h1 = geo[(1 + 3 * d) * NN], h2 = geo[(2 + 3 * d) * NN];
            const T Ua = h0 * uu[0] + h1 * uu[1] + h2 * uu[2];
            const T Qa = h0 * qq[0] + h1 * qq[1] + h2 * qq[2];
            const int Id = d == 0 ? I : (d == 1 ? J : K);
            const int stride = d == 0 ? 1 : (d == 1 ? Q : Q * Q);
            const int base = l - Id * stride;
            T aRho = 0, aP = 0, aH = 0, aRH = 0, aUc[3] = {0, 0, 0}, aQc[3] = {0, 0, 0};
#pragma unroll 1
            for (int n = 0; n < Q; ++n) {
                const int s = base + n * stride;
                const T Dv = sD[Id * Q + n];
                const T rs = sF[le][0][s], ps = sF[le][1][s], Hs = sF[le][2][s], rHs = sF[le][3][s];
                const T us0 = sF[le][4][s], us1 = sF[le][5][s], us2 = sF[le][6][s];
                const T qs0 = sF[le][7][s], qs1 = sF[le][8][s], qs2 = sF[le][9][s];
                const T Us = h0 * us0 + h1 * us1 + h2 * us2;
                const T Qs = h0 * qs0 + h1 * qs1 + h2 * qs2;
                aRho += Dv * rs;
                aP += Dv * ps;
                aH += Dv * Hs;
                aRH += Dv * rHs;
                aUc[0] += Dv * us0; aUc[1] += Dv * us1; aUc[2] += Dv * us2;
                aQc[0] += Dv * qs0; aQc[1] += Dv * qs1; aQc[2] += Dv * qs2;
                aU += Dv * Us;
                aQ += Dv * Qs;
                const T DU = Dv * Us;
                aRHU += DU * rHs;
                aHU += DU * Hs;
                aQU[0] += DU * qs0; aQU[1] += DU * qs1; aQU[2] += DU * qs2;
                aUU[0] += DU * us0; aUU[1] += DU * us1; aUU[2] += DU * us2;
            }
            sRho += Ua * aRho;
            sH += Qa * aH;
            sRH += Ua * aRH;
            gP[0] += h0 * aP; gP[1] += h1 * aP; gP[2] += h2 * aP;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                qGu[c] += Qa * aUc[c];
                uGq[c] += Ua * aQc[c];
            }
        }

        const T gv = a.gvol[e * NN + l];
        const T Rm = (T)0.5 * (aQ + r * aU) + (T)0.5 * sRho;
        const T Re = (T)0.25 * (aRHU + H * aQ + rH * aU + r * aHU) + (T)0.25 * (sH + sRH + H * sRho);
        T Rmo[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
            Rmo[i] = (T)0.25 * (aQU[i] + uu[i] * aQ + qq[i] * aU + r * aUU[i] + uu[i] * sRho +
                                qGu[i] + uGq[i]) + gP[i];
        v[0] = (double)(gv * Rm);
        v[1] = (double)(gv * Rmo[0]);
        v[2] = (double)(gv * Rmo[1]);
        v[3] = (double)(gv * Rmo[2]);
        v[4] = (double)(gv * Re);
        }
        add_pass<UMAX, EPB>(sR, le, active, active ? pl.lidx[e * NN + l] : 0u, v);
    }
    write_batch<UMAX>(pl, u0, U, sR, a.Rmass, a.Rmom, a.Rener, a.npoin);
}

template <int P, typename T>
static hipError_t launch_p(const ConvecArgs<T> &a, hipStream_t s)
{
    if (a.plan.nb <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_convec<P, T>), dim3((unsigned)a.plan.nb), dim3(Shape<P>::NT), 0, s, a);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_convec(int porder, const ConvecArgs<T> &a, hipStream_t s)
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

template hipError_t launch_convec<double>(int, const ConvecArgs<double> &, hipStream_t);
template hipError_t launch_convec<float>(int, const ConvecArgs<float> &, hipStream_t);

}  // namespace msod

// Convective residual kernel (counterpart of full_convec_ijk,
// src/kernels/sources/elem_convec.f90:198-431, with the energy fluxes of
// full_convec_ijk_split, :674-832).
//
// One workgroup per batch of CShape<P>::BATCH elements, processed in passes of
// CShape<P>::EPB elements with one thread per element node.  Per pass:
//  1. gather: each thread reads its global node (conn, element-major) and
//     writes 10 nodal values to LDS:
//       rho, p, H = E + k, rhoH = rho E + rho k, u(3), q(3)     (k = |u|^2/2)
//  2. sweeps: for each reference direction d the thread runs the 1-D
//     derivative D[I_d][n] along its grid line.  Rather than differentiating
//     the 39 flux components the reference forms, the He column of the
//     TARGET node is folded into the fluxes at the source nodes:
//     U_d(s) = sum_i He(i,d)u_i(s).  Since He is taken at the target node,
//       div(q (x) u)_i = sum_d D_d[ q_i U_d ],  div u = sum_d He(:,d).D_d[u], ...
//     Per line the thread forms the plain 1-D derivatives of r, p, H, rH, u, q and the
//     derivatives of the products with U_d; the node's own r, u, q, H and He
//     column then contract them (eleven accumulators over the three
//     directions; the sweep code is shared by the directions).  LDS holds the
//     ten values of a node contiguously, so a source costs one address and
//     five 16-byte reads (nine values, rH formed in the sweep, where that
//     fits another workgroup per CU: convec_drop_rh).
//     The result equals the reference's algebra exactly in real arithmetic;
//     rounding differs at the 1e-16 level.
//  3. residual x gpvol, added into the batch's per-node sums in LDS in a
//     fixed order (add_pass).
// After the passes the batch's distinct nodes are written once each
// (write_batch_nodes: plain store / read-add-store per colour, or FP64 atomics).
//
// Storage types TG (geometry) and TF (nodal fields), FP64 arithmetic
// throughout (msod_kernels.h).  Without a pressure array the pressure is
// recovered as p = (gamma-1)(E - rho|u|^2/2) from the stored state.
#include "msod_kernels.h"

namespace msod {

// Inputs of one element node, loaded one pass ahead of their use so that
// the gathers of the next pass are in flight while this pass computes.
template <class A>
struct ConvecNode {
    typename A::F r, E, q[3];
    typename A::U u[3];
    typename A::Pt p;
    uint32_t li;
};

template <int NN, class A>
__device__ __forceinline__ void convec_fetch(const A &a, int64_t e, int l, uint32_t gnode, ConvecNode<A> &x)
{
    const int64_t g = gnode;
    const int64_t np = a.npoin;
    x.r = a.rho[g];
    x.E = a.E[g];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        x.u[d] = a.u[g + d * np];
        x.q[d] = a.q[g + d * np];
    }
    x.p = a.pr ? a.pr[g] : (typename A::Pt)0;
    x.li = __builtin_nontemporal_load(&a.plan.lidx[e * NN + l]);
}

// LDS bytes of a workgroup with NF values per node record.
template <int P, typename TA>
constexpr size_t convec_lds(int nf)
{
    using S = CShape<P>;
    return sizeof(TA) * (S::Q * S::Q + (size_t)S::EPB * S::NN * nf) + sizeof(double) * 5 * S::UMAX +
           3 * sizeof(uint32_t) * kMaxKpb;
}
// rH = rho E + rho k is left out of the records (formed as r H in the sweep)
// where that fits one more workgroup per CU, for FP32 element arrays with
// FP64 arithmetic at porder 6 and 7 (p=7: 82.6 -> 78.5 KB, two 512-thread
// workgroups instead of one, at most 128 VGPRs: 44 B of spills per lane;
// 48^3 p=7: 3.66 -> 3.35 ms for the FP32 split, 3.51 -> 3.21 ms with every
// input in FP32).  With FP64 element arrays the 128-VGPR bound spills 108 B
// per lane and ran slower (3.98 -> 4.18 ms): one workgroup there.
template <int P, typename TA, typename TG>
constexpr bool convec_drop_rh()
{
    return sizeof(TG) < sizeof(TA) && 163840 / convec_lds<P, TA>(9) > 163840 / convec_lds<P, TA>(10);
}
// Launch bound in waves per SIMD: 3 (up to 168 VGPRs, room for the loads in
// flight; 4 would force spills at p=4); 4 for 512-thread workgroups whose
// LDS lets two share a CU.
template <int P, typename TA, typename TG>
constexpr int convec_waves()
{
    return CShape<P>::NT == 512 && convec_drop_rh<P, TA, TG>() ? 4 : 3;
}

// TA: arithmetic type of the gathers, fluxes and sweeps (FP64, or FP32 for
// the FP32-arithmetic form); the node sums and the assembly stay FP64.
template <int P, class A, typename TA = double>
__global__ void __launch_bounds__(CShape<P>::NT, (convec_waves<P, TA, typename A::G>())) k_convec(A a)
{
    using TG = typename A::G;
    using S = CShape<P>;
    using T = TA;  // arithmetic
    constexpr int Q = S::Q, NN = S::NN, EPB = S::EPB, UMAX = S::UMAX, NPASS = S::NPASS;
    // record: r, p, H, [rH,] u(3), q(3)
    constexpr bool kRH = !convec_drop_rh<P, TA, TG>();
    constexpr int NF = kRH ? 10 : 9, iU = kRH ? 4 : 3, iQ = iU + 3;
    __shared__ T sD[Q * Q];
    __shared__ __attribute__((aligned(16))) T sF[EPB][NN][NF];  // node-major: one base address per source
    __shared__ double sR[5][UMAX];
    __shared__ uint32_t sB[kMaxKpb], sU0[kMaxKpb], sU1[kMaxKpb];
    constexpr int NT = S::NT, UPT = (UMAX + NT - 1) / NT;

    const Plan &pl = a.plan;
    const int64_t first = (int64_t)xcd_slot(blockIdx.x, gridDim.x) * pl.kpb;
    const int nbat = (int)(pl.nb - first < pl.kpb ? pl.nb - first : pl.kpb);
    const int npass = nbat * NPASS;  // passes of this block, over its batches
    const int t = threadIdx.x;
    const int le = t / NN;  // >= EPB for the few idle threads
    const int l = t - le * NN;
    const int I = l % Q, J = (l / Q) % Q, K = l / (Q * Q);
    // Loads are unconditional, so that the compiler can count them exactly
    // and wait for each one only where it is used: idle threads and the
    // elements past the end of the mesh load a real element, and the result
    // is not used.
    const int lec = le < EPB ? le : EPB - 1;
    auto elem = [&](int64_t bb, int p) -> int64_t {
        const int64_t e = bb * S::BATCH + p * EPB + lec;
        return e < a.nelem ? e : a.nelem - 1;
    };
    // element of pass g of this block (g clamped to the last pass)
    auto elem_of = [&](int g) -> int64_t {
        const int gc = g < npass ? g : npass - 1;
        return elem((int64_t)sB[gc / NPASS], gc % NPASS);
    };

    for (int i = t; i < nbat; i += blockDim.x) {
        const uint32_t bi = plan_batch(pl.batches, first + i);
        sB[i] = bi;
        sU0[i] = pl.uoff[bi];
        sU1[i] = pl.uoff[bi + 1];
    }
    for (int i = t; i < Q * Q; i += blockDim.x) sD[i] = (T)a.tab.D[i];

    // prologue: nodal values of pass 0, connectivity of pass 1
    ConvecNode<A> nd[2];
    uint32_t gnext;  // global node of the next pass's element node
    {
        const int64_t b0 = plan_batch(pl.batches, first);
        const int64_t e0 = elem(b0, 0);
        convec_fetch<NN>(a, e0, l, __builtin_nontemporal_load(&pl.conn[e0 * NN + l]), nd[0]);
        const int64_t b1 = npass > 1 ? (int64_t)plan_batch(pl.batches, first + 1 / NPASS) : b0;
        const int64_t e1 = npass > 1 ? elem(b1, 1 % NPASS) : e0;
        gnext = __builtin_nontemporal_load(&pl.conn[e1 * NN + l]);
    }

    for (int kb = 0; kb < nbat; ++kb) {
        __syncthreads();  // sB, sD visible; the previous batch's write phase is done with sR
        const int64_t b = sB[kb];
        const uint32_t u0 = sU0[kb], U = sU1[kb] - u0;
        uint32_t wn[UPT];  // this thread's entries of the batch's node list, for the write phase
#pragma unroll
        for (int j = 0; j < UPT; ++j) {
            const uint32_t i = t + j * NT;
            wn[j] = pl.unode[u0 + (i < U ? i : U - 1)];
        }
        for (uint32_t i = t; i < U; i += blockDim.x)
#pragma unroll
            for (int c = 0; c < 5; ++c) sR[c][i] = 0.0;
#pragma unroll
        for (int p = 0; p < NPASS; ++p) {
            ConvecNode<A> &cur = nd[p & 1];
            ConvecNode<A> &nxt = nd[(p + 1) & 1];
            const int g = kb * NPASS + p;
            const int64_t e = b * S::BATCH + p * EPB + le;
            const bool active = (le < EPB) && (e < a.nelem);
            if (p > 0) __syncthreads();  // the previous pass is done with sF
            if (active) {
                const T r = cur.r, Ev = cur.E;
                const T ux = cur.u[0], uy = cur.u[1], uz = cur.u[2];
                const T k = ux * ux * (T)0.5 + uy * uy * (T)0.5 + uz * uz * (T)0.5;
                // p = (gamma-1)(E - rho|u|^2/2) when no pressure array is given, in
                // FP64 whatever T: the difference cancels
                const double p64 = a.gamma_m1 * ((double)Ev - (double)r * 0.5 *
                                                     ((double)ux * ux + (double)uy * uy + (double)uz * uz));
                const T pv = a.pr ? (T)cur.p : (T)p64;
                const T H = Ev + k;
                const T rH = r * Ev + r * k;
                sF[le][l][0] = r;
                sF[le][l][1] = pv;
                sF[le][l][2] = H;
                if constexpr (kRH) sF[le][l][3] = rH;
                sF[le][l][iU] = ux;
                sF[le][l][iU + 1] = uy;
                sF[le][l][iU + 2] = uz;
                sF[le][l][iQ] = cur.q[0];
                sF[le][l][iQ + 1] = cur.q[1];
                sF[le][l][iQ + 2] = cur.q[2];
            }
            const uint32_t li = cur.li;
            // issued now, used later: He and gpvol of this pass (after the
            // barrier), the next pass's nodal values (next pass), the
            // connectivity of the pass after it.  The element arrays are read
            // once: non-temporal loads keep them from evicting the gathered
            // nodal lines, which neighbouring elements read again, from L2.
            const int64_t ec = elem(b, p);
            // kept in the storage type until used: a conversion here would
            // wait for the load before the barrier
            const GeoNode<TG> gn = load_geo<NN, true>(a.geo, ec, l);  // He(i,d) = gn.he(i + 3d)
            convec_fetch<NN>(a, elem_of(g + 1), l, gnext, nxt);
            gnext = __builtin_nontemporal_load(&pl.conn[elem_of(g + 2) * NN + l]);
            __syncthreads();
            double v[5] = {0, 0, 0, 0, 0};
            if (active) {
                // Accumulators, each a sum over the three directions and the
                // q sources of one line (r, H, rH, u, q: the node's own values):
                //   aU = div u, aQ = div q                       (from the line derivatives)
                //   aE = sum D[(U.)(rH + r H)],  aM_i = sum D[(U.)(q_i + r u_i)]
                //   sRho = sum U_d D_d[r],  sE = sum Q_d D_d[H] + U_d D_d[rH]
                //   mM_i = sum Q_d D_d[u_i] + U_d D_d[q_i] + 4 He(i,d) D_d[p]
                const T r = (T)cur.r;
                T aQ = 0, aU = 0, aE = 0, aM[3] = {0, 0, 0};
                T sRho = 0, sE = 0, mM[3] = {0, 0, 0};
#pragma unroll 1
                for (int d = 0; d < 3; ++d) {
                    // one copy of the sweep code for the three directions
                    const T h0 = (T)(d == 0 ? gn.he(0) : (d == 1 ? gn.he(3) : gn.he(6)));
                    const T h1 = (T)(d == 0 ? gn.he(1) : (d == 1 ? gn.he(4) : gn.he(7)));
                    const T h2 = (T)(d == 0 ? gn.he(2) : (d == 1 ? gn.he(5) : gn.he(8)));
                    const int Id = d == 0 ? I : (d == 1 ? J : K);
                    const int stride = d == 0 ? 1 : (d == 1 ? Q : Q * Q);
                    const int base = l - Id * stride;
                    const T Ua = h0 * sF[le][l][iU] + h1 * sF[le][l][iU + 1] + h2 * sF[le][l][iU + 2];
                    const T Qa = h0 * sF[le][l][iQ] + h1 * sF[le][l][iQ + 1] + h2 * sF[le][l][iQ + 2];
                    // 1-D derivatives along this line of the stored values
                    T dp = 0, du[3] = {0, 0, 0}, dq[3] = {0, 0, 0}, dr = 0, dH = 0, drH = 0;
#pragma unroll 1
                    for (int n = 0; n < Q; ++n) {
                        const int s = base + n * stride;
                        const T Dv = sD[Id * Q + n];
                        const T rs = sF[le][s][0], ps = sF[le][s][1], Hs = sF[le][s][2];
                        const T rHs = kRH ? sF[le][s][kRH ? 3 : 0] : rs * Hs;
                        const T us0 = sF[le][s][iU], us1 = sF[le][s][iU + 1], us2 = sF[le][s][iU + 2];
                        const T qs0 = sF[le][s][iQ], qs1 = sF[le][s][iQ + 1], qs2 = sF[le][s][iQ + 2];
                        dr += Dv * rs;
                        dH += Dv * Hs;
                        drH += Dv * rHs;
                        dp += Dv * ps;
                        du[0] += Dv * us0; du[1] += Dv * us1; du[2] += Dv * us2;
                        dq[0] += Dv * qs0; dq[1] += Dv * qs1; dq[2] += Dv * qs2;
                        // products with the target's contravariant velocity
                        const T DU = Dv * (h0 * us0 + h1 * us1 + h2 * us2);
                        aE += DU * (rHs + r * Hs);
                        aM[0] += DU * (qs0 + r * us0);
                        aM[1] += DU * (qs1 + r * us1);
                        aM[2] += DU * (qs2 + r * us2);
                    }
                    sRho += Ua * dr;
                    sE += Qa * dH + Ua * drH;
                    const T dp4 = (T)4 * dp;
                    mM[0] += Qa * du[0] + Ua * dq[0] + h0 * dp4;
                    mM[1] += Qa * du[1] + Ua * dq[1] + h1 * dp4;
                    mM[2] += Qa * du[2] + Ua * dq[2] + h2 * dp4;
                    aU += h0 * du[0] + h1 * du[1] + h2 * du[2];
                    aQ += h0 * dq[0] + h1 * dq[1] + h2 * dq[2];
                }
                const T H = sF[le][l][2], rH = kRH ? sF[le][l][kRH ? 3 : 0] : r * H;
                const T uu[3] = {sF[le][l][iU], sF[le][l][iU + 1], sF[le][l][iU + 2]};
                const T qq[3] = {sF[le][l][iQ], sF[le][l][iQ + 1], sF[le][l][iQ + 2]};
                const T Rm = (T)0.5 * (aQ + r * aU) + (T)0.5 * sRho;
                const T Re = (T)0.25 * (aE + H * aQ + rH * aU) + (T)0.25 * (sE + H * sRho);
                T Rmo[3];
#pragma unroll
                for (int i = 0; i < 3; ++i)
                    Rmo[i] = (T)0.25 * (aM[i] + uu[i] * aQ + qq[i] * aU + uu[i] * sRho + mM[i]);
                const T gv = (T)gn.gvol();
                v[0] = (double)(gv * Rm);
                v[1] = (double)(gv * Rmo[0]);
                v[2] = (double)(gv * Rmo[1]);
                v[3] = (double)(gv * Rmo[2]);
                v[4] = (double)(gv * Re);
            }
            add_pass<UMAX, EPB>(sR, pl.mode, le, active, li, v);
        }
        if (pl.mode == kAtomic) __syncthreads();  // every LDS atomic of the batch done
        write_batch_nodes<UMAX, NT>(pl.mode, wn, U, sR, a.Rmass, a.Rmom, a.Rener, a.npoin);
        if (NPASS & 1) nd[0] = nd[1];  // odd pass count: the next batch starts from slot 0
    }
}

template <int P, class A, typename TA = double>
static hipError_t launch_p(const A &a, hipStream_t s)
{
    if (a.plan.nb <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_convec<P, A, TA>), dim3(plan_blocks(a.plan)), dim3(CShape<P>::NT), 0, s, a);
    return hipGetLastError();
}

template <class A>
hipError_t launch_convec(int porder, const A &a, hipStream_t s)
{
    switch (porder) {
    case 1: return launch_p<1, A>(a, s);
    case 2: return launch_p<2, A>(a, s);
    case 3: return launch_p<3, A>(a, s);
    case 4: return launch_p<4, A>(a, s);
    case 5: return launch_p<5, A>(a, s);
    case 6: return launch_p<6, A>(a, s);
    case 7: return launch_p<7, A>(a, s);
    case 8: return launch_p<8, A>(a, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_convec_fp32(int porder, const ConvecArgs<float, float> &a, hipStream_t s)
{
    switch (porder) {
    case 1: return launch_p<1, ConvecArgs<float, float>, float>(a, s);
    case 2: return launch_p<2, ConvecArgs<float, float>, float>(a, s);
    case 3: return launch_p<3, ConvecArgs<float, float>, float>(a, s);
    case 4: return launch_p<4, ConvecArgs<float, float>, float>(a, s);
    case 5: return launch_p<5, ConvecArgs<float, float>, float>(a, s);
    case 6: return launch_p<6, ConvecArgs<float, float>, float>(a, s);
    case 7: return launch_p<7, ConvecArgs<float, float>, float>(a, s);
    case 8: return launch_p<8, ConvecArgs<float, float>, float>(a, s);
    default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_convec(int, const ConvecArgs<double, double> &, hipStream_t);
template hipError_t launch_convec(int, const ConvecArgs<float, float> &, hipStream_t);
template hipError_t launch_convec(int, const ConvecArgs<float, double> &, hipStream_t);
template hipError_t launch_convec(int, const ConvecArgs<float, float, float> &, hipStream_t);
template hipError_t launch_convec(int, const ConvecSplitArgs &, hipStream_t);

}  // namespace msod

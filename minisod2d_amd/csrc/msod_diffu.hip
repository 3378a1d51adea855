// Diffusive residual kernel (counterpart of full_diffusion_ijk,
// src/kernels/sources/elem_diffu.f90:10-197).
//
// One workgroup per batch of Shape<P>::BATCH elements, passes of
// Shape<P>::EPB elements with one thread per element node.  Per pass:
//  pass 1 (elem_diffu.f90:67-139): sweeps of rho, T, u along the node's three
//    grid lines; He of the node gives grad rho, grad T and G = grad u; then
//    the viscous stress tau and tau.u, kept in registers.
//  pass 2 (elem_diffu.f90:141-193): the reference re-reads He, gpvol and the
//    six flux vectors Phi (grad rho, tau.u, grad T, tau rows) at every source
//    node of the target's three lines.  Here each thread publishes, one
//    reference direction d at a time, the contravariant components
//        F_d = gpvol * sum_i He(i,d) Phi_i        (6 values per node)
//    and the target accumulates the transposed sweep
//        W[Phi](a) += sum_n D[n][I_d(a)] F_d(s_d(n))
//    (LDS holds 6 values per node instead of 18, node-major, in the buffer
//    pass 1 used for rho, T, u).  Then
//    Rmass += nu_e W[grad rho], Rener += W[tau.u] + kappa_e W[grad T],
//    Rmom_r += W[tau_r], with nu_e and kappa_e of the target node (the
//    reference multiplies kappa_e inside the sum; same value, other rounding).
// Each element node's five values are added into the batch's per-node sums
// in LDS in a fixed order (add_pass); the batch's distinct nodes are then
// written once (write_batch_nodes).
//
// Storage types TG (geometry, mu_e, mu_sgs) and TF (nodal fields), FP64
// arithmetic throughout (msod_kernels.h).  CLOSURE (FP64 storage only):
// mu_sgs is not read but predicted per
// element node by the learned closure (msod_mlp.h) from the pass-1 velocity
// gradient, evaluated with bf16 MFMA inside the kernel before the stress is
// formed; its LDS scratch overlays the field buffers between the passes.
#include <type_traits>

#include "msod_kernels.h"
#include "msod_mlp.h"
#include "msod_closure.h"

namespace msod {

// Inputs of one element node, loaded one pass ahead of their use so that
// the gathers of the next pass are in flight while this pass computes.
template <class A>
struct DiffuNode {
    typename A::F r, T_, muf;
    typename A::U u[3];
};

template <class A>
__device__ __forceinline__ void diffu_fetch(const A &a, uint32_t gnode, DiffuNode<A> &x)
{
    const int64_t g = gnode;
    const int64_t np = a.npoin;
    x.r = a.rho[g];
    x.T_ = a.Tem[g];
    x.muf = a.mu_fluid[g];
#pragma unroll
    for (int d = 0; d < 3; ++d) x.u[d] = a.u[g + d * np];
}

// mu_e and mu_sgs of ME consecutive elements from e0 on into LDS.  Their
// reference layout (nelem,ngaus) puts the values of one element node
// 8*nelem bytes apart, so a wave reading one element's nodes would touch 64
// lines (and pages) per load; here ME neighbouring lanes read ME
// consecutive elements of one node, and the batches that follow in mesh
// order reuse the lines.
template <int NN, int ME, int NT, bool CLOSURE, class A, typename TG = typename A::G>
__device__ __forceinline__ void diffu_stage_mu(const A &a, int64_t e0, TG (*sMu)[ME][NN])
{
    // every load unconditional and issued before any LDS write (entries past
    // the table or the mesh re-read a valid entry and are not written or not
    // used), so that the loads of one thread are in flight together
    constexpr int IT = (ME * NN + NT - 1) / NT;
    int32_t gl[IT];
    TG v0[IT], v1[IT];
#pragma unroll
    for (int k = 0; k < IT; ++k) {
        const int i = (int)threadIdx.x + k * NT < ME * NN ? (int)threadIdx.x + k * NT : ME * NN - 1;
        gl[k] = a.tab.l2g[i / ME];
    }
#pragma unroll
    for (int k = 0; k < IT; ++k) {
        const int i = (int)threadIdx.x + k * NT < ME * NN ? (int)threadIdx.x + k * NT : ME * NN - 1;
        const int64_t e = e0 + i % ME < a.nelem ? e0 + i % ME : a.nelem - 1;
        const int64_t ge = e + a.nelem * (int64_t)gl[k];  // reference layout (nelem,ngaus)
        v0[k] = a.mu_e[ge];
        v1[k] = (!CLOSURE && a.mu_sgs) ? a.mu_sgs[ge] : (TG)0;  // NULL: closure fallback
    }
#pragma unroll
    for (int k = 0; k < IT; ++k) {
        const int i = (int)threadIdx.x + k * NT;
        if (i < ME * NN) {
            sMu[0][i % ME][i / ME] = v0[k];
            sMu[1][i % ME][i / ME] = v1[k];
        }
    }
}

// Launch bound in waves per SIMD.  The closure form is bounded to 3
// workgroups per CU: the compiler then keeps the MFMA accumulators in VGPRs
// (about 148 in all) instead of 188 VGPRs and 32 AGPRs, 3 waves per SIMD
// instead of 2 (4.7 against 6.2 ms at 96^3 p=3).  512-thread workgroups with
// FP32 element arrays and FP64 arithmetic: 6, three workgroups per CU (their
// LDS fits three), at 80 VGPRs with 20 B of spills per lane (48^3 p=7: FP32
// split 3.96 -> 3.45 ms, every input in FP32 4.21 -> 3.41 ms); the FP64 form
// spilled 28 B there and ran slower (3.61 -> 3.90 ms), so it keeps two.
template <int P, typename TG, typename TA, bool CLOSURE>
constexpr int diffu_waves()
{
    return CLOSURE ? (Shape<P>::NT == 256 ? 3 : 1)
                   : (Shape<P>::NT == 512 && sizeof(TG) < sizeof(TA) ? 6 : 1);
}

template <int P, class A, bool CLOSURE, typename TA = double>
__global__ void __launch_bounds__(Shape<P>::NT, (diffu_waves<P, typename A::G, TA, CLOSURE>())) k_diffu(A a, const float *mlp)
{
    using TG = typename A::G;
    using S = Shape<P>;
    using T = TA;  // arithmetic: FP64, or FP32 for the FP32-arithmetic form (node sums stay FP64)
    constexpr int Q = S::Q, NN = S::NN, EPB = S::EPB, UMAX = S::UMAX;
    // one buffer of node records (sN) serves pass 1 (rho, T, u) and pass 2
    // (fluxes of one direction); the MLP scratch of the closure overlays it
    // between the two
    constexpr size_t kRec = sizeof(T) * EPB * 6 * NN;
    constexpr size_t kScr = CLOSURE ? (size_t)(S::NT / 64) * 64 * kMlpH * sizeof(short) : 0;
    constexpr size_t kBuf = kRec > kScr ? kRec : kScr;
    __shared__ T sD[Q * Q];
    __shared__ __attribute__((aligned(16))) unsigned char sbuf[kBuf];
    __shared__ double sR[5][UMAX];
    // mu_e, mu_sgs of the batch, staged once; of the pass, staged per pass,
    // where the batch's would cost a workgroup per CU (four fit 160 KB of LDS)
    constexpr size_t kLdsRest = sizeof(T) * Q * Q + kBuf + sizeof(double) * 5 * UMAX +
                                (CLOSURE ? sizeof(MlpShared) : 0);
    constexpr size_t kLdsCu = 160 * 1024 / 4;
    constexpr bool kMuPerPass = kLdsRest + 2 * sizeof(TG) * S::BATCH * NN > kLdsCu &&
                                kLdsRest + 2 * sizeof(TG) * EPB * NN <= kLdsCu;
    constexpr int ME = kMuPerPass ? EPB : S::BATCH;
    __shared__ TG sMu[2][ME][NN];
    __shared__ typename std::conditional<CLOSURE, MlpShared, char>::type smlp;
    // node-major records of 6 values (pass 1: rho, T, u0, u1, u2, unused;
    // pass 2: the contravariant fluxes of one direction), read in pairs: in
    // FP64 three ds_read_b128 per source instead of ds_read2_b64 pairs
    // (diffusion 13.71 -> 13.51 ms at 128^3 p=4, 3.81 -> 3.67 ms at 48^3 p=7)
    struct alignas(2 * sizeof(T)) V2 { T x, y; };
    auto sN = reinterpret_cast<V2 (*)[NN][3]>(sbuf);

    const Plan &pl = a.plan;
    constexpr int NT = S::NT, UPT = (UMAX + NT - 1) / NT;
    const int64_t b = plan_batch(pl.batches, xcd_slot(blockIdx.x, gridDim.x));
    const uint32_t u0 = pl.uoff[b], U = pl.uoff[b + 1] - u0;
    const int t = threadIdx.x;
    uint32_t wn[UPT];  // this thread's entries of the batch's node list, for the write phase
#pragma unroll
    for (int j = 0; j < UPT; ++j) {
        const uint32_t i = t + j * NT;
        wn[j] = pl.unode[u0 + (i < U ? i : U - 1)];
    }
    const int le = t / NN;
    const int l = t - le * NN;
    const int I = l % Q, J = (l / Q) % Q, K = l / (Q * Q);
    const T Cp = (T)a.Cp, Pr = (T)a.Pr;

    DiffuNode<A> cur, nxt;
    // the connectivity is read a pass ahead of the gathers that use it, so
    // that the next pass's gathers issue without a trip to memory (idle
    // threads and elements past the mesh read a real entry)
    uint32_t gnext = 0;
    auto conn_of = [&](int p) -> uint32_t {
        const int64_t ep = b * S::BATCH + p * EPB + (le < EPB ? le : 0);
        return __builtin_nontemporal_load(&pl.conn[(ep < a.nelem ? ep : a.nelem - 1) * NN + l]);
    };
    static_assert(Q * Q <= S::NT, "one derivative-table entry per thread");
    {
        // prologue: the tables and the connectivity first, then the values
        // that depend on them, then the LDS writes; every load unconditional
        // (idle threads read a real entry), so that each group is in flight
        // together
        const T dv = (T)a.tab.D[t < Q * Q ? t : Q * Q - 1];
        constexpr int IT = (ME * NN + NT - 1) / NT;
        int32_t gl[IT];
        TG v0[IT], v1[IT];
        auto mu_i = [&](int k) {
            return (int)t + k * NT < ME * NN ? (int)t + k * NT : ME * NN - 1;
        };
        if (!kMuPerPass) {
#pragma unroll
            for (int k = 0; k < IT; ++k) gl[k] = a.tab.l2g[mu_i(k) / ME];
        }
        const int64_t e = b * S::BATCH + le;
        const int64_t ec = e < a.nelem ? e : a.nelem - 1;
        const uint32_t gcur = __builtin_nontemporal_load(&pl.conn[ec * NN + l]);
        if (S::NPASS > 1) gnext = conn_of(1);
        diffu_fetch(a, gcur, cur);
        if (!kMuPerPass) {
#pragma unroll
            for (int k = 0; k < IT; ++k) {
                const int i = mu_i(k);
                const int64_t em = b * S::BATCH + i % ME < a.nelem ? b * S::BATCH + i % ME : a.nelem - 1;
                const int64_t ge = em + a.nelem * (int64_t)gl[k];
                v0[k] = a.mu_e[ge];
                v1[k] = (!CLOSURE && a.mu_sgs) ? a.mu_sgs[ge] : (TG)0;
            }
        }
        for (uint32_t i = t; i < U; i += blockDim.x)
#pragma unroll
            for (int c = 0; c < 5; ++c) sR[c][i] = 0.0;
        if (t < Q * Q) sD[t] = dv;
        if (!kMuPerPass) {
#pragma unroll
            for (int k = 0; k < IT; ++k) {
                const int i = (int)t + k * NT;
                if (i < ME * NN) {
                    sMu[0][i % ME][i / ME] = v0[k];
                    sMu[1][i % ME][i / ME] = v1[k];
                }
            }
        }
    }
    if constexpr (CLOSURE) mlp_stage(smlp, mlp);

#pragma unroll
    for (int pass = 0; pass < S::NPASS; ++pass) {
        const int64_t e = b * S::BATCH + pass * EPB + le;
        const bool active = (le < EPB) && (e < a.nelem);
        T r = 0, muf = 0, mue = 0, mus = 0, uu[3] = {0, 0, 0};
        // the previous pass read its sMu entries before its pass-2 barriers
        if (kMuPerPass) diffu_stage_mu<NN, ME, NT, CLOSURE>(a, b * S::BATCH + pass * EPB, sMu);
        __syncthreads();  // previous pass done with sN (and sR zeroed, sMu staged)
        if (active) {
            r = cur.r;
            muf = cur.muf;
            mue = (T)sMu[0][kMuPerPass ? le : pass * EPB + le][l];
            mus = (T)sMu[1][kMuPerPass ? le : pass * EPB + le][l];
            for (int d = 0; d < 3; ++d) uu[d] = cur.u[d];
            sN[le][l][0] = V2{r, (T)cur.T_};
            sN[le][l][1] = V2{uu[0], uu[1]};
            sN[le][l][2] = V2{uu[2], (T)0};
        }
        if (pass + 1 < S::NPASS) {  // next pass's inputs: in flight during this pass
            const int64_t en = e + EPB;
            if (le < EPB && en < a.nelem) diffu_fetch(a, gnext, nxt);
            if (pass + 2 < S::NPASS) gnext = conn_of(pass + 2);
        }
        // this pass's element arrays, issued before the barrier;
        // unconditional, idle threads read a real element; kept in the
        // storage type until used.  Plain loads: with the non-temporal hint
        // the 128^3 p=4 diffusion ran 4% slower in FP64 and 5-7% with FP32
        // element arrays (the hint still pays in the convection)
        GeoNode<TG> gn;  // He(i,d) = gn.he(i + 3d), gpvol = gn.gvol()
        uint32_t li;
        {
            const int64_t ec = e < a.nelem ? (le < EPB ? e : b * S::BATCH + pass * EPB) : a.nelem - 1;
            gn = load_geo<NN, false>(a.geo, ec, l);
            li = pl.lidx[ec * NN + l];
        }
        __syncthreads();

        // ---- pass 1: gradients, stress ------------------------------------
        // He(i,d) = gn.he(i + 3d), converted where used: a conversion placed
        // ahead of the sweeps would wait for every load there
        T gR[3] = {0, 0, 0}, gT[3] = {0, 0, 0}, tau[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        T tauU[3] = {0, 0, 0}, gv = 0, G[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        if (active) {
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
                    const V2 p0 = sN[le][s][0], p1 = sN[le][s][1], p2 = sN[le][s][2];
                    aR += Dv * p0.x;
                    aT += Dv * p0.y;
                    aU0 += Dv * p1.x;
                    aU1 += Dv * p1.y;
                    aU2 += Dv * p2.x;
                }
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    const T hid = (T)gn.he(i + 3 * d);
                    gR[i] += hid * aR;
                    gT[i] += hid * aT;
                    // G[i][j] = d u_i / d x_j = sum_d He(j,d) D_d[u_i]
                    G[0][i] += hid * aU0;
                    G[1][i] += hid * aU1;
                    G[2][i] += hid * aU2;
                }
            }
        }
        if constexpr (CLOSURE) {
            __syncthreads();  // every pass-1 sweep over sN done: the MLP scratch overlays it
            float f[kMlpIn];
            if (active) {
                T hinv[3][3];
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int d = 0; d < 3; ++d) hinv[i][d] = (T)gn.he(i + 3 * d);
                mlp_features<T>(G, r, uu, hinv, f);
            } else
#pragma unroll
                for (int k = 0; k < kMlpIn; ++k) f[k] = smlp.mean[k];
            const float z3 = mlp_forward_z3(smlp, f, reinterpret_cast<short *>(sbuf) + (t >> 6) * 64 * kMlpH);
            mus = active ? (T)(smlp.scale * softplusf_(z3)) : (T)0;
        }
        if (active) {
            const T divU = G[0][0] + G[1][1] + G[2][2];
            const T mu_f = muf + r * mus;
            const T mu_t = mu_f + mue;
            const T lam = mu_f * (T)(2.0 / 3.0) * divU;
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
            gv = (T)gn.gvol();
        }

        // ---- pass 2: transposed sweeps, one reference direction at a time -
        __syncthreads();  // every pass-1 sweep over sN done: the fluxes overwrite it
        T W[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            if (active) {
                const T c0 = gv * (T)gn.he(3 * d), c1 = gv * (T)gn.he(1 + 3 * d), c2 = gv * (T)gn.he(2 + 3 * d);
                sN[le][l][0] = V2{c0 * gR[0] + c1 * gR[1] + c2 * gR[2], c0 * tauU[0] + c1 * tauU[1] + c2 * tauU[2]};
                sN[le][l][1] = V2{c0 * gT[0] + c1 * gT[1] + c2 * gT[2], c0 * tau[0][0] + c1 * tau[0][1] + c2 * tau[0][2]};
                sN[le][l][2] = V2{c0 * tau[1][0] + c1 * tau[1][1] + c2 * tau[1][2], c0 * tau[2][0] + c1 * tau[2][1] + c2 * tau[2][2]};
            }
            __syncthreads();
            if (active) {
                const int Id = d == 0 ? I : (d == 1 ? J : K);
                const int stride = d == 0 ? 1 : (d == 1 ? Q : Q * Q);
                const int base = l - Id * stride;
#pragma unroll 1
                for (int n = 0; n < Q; ++n) {
                    const int s = base + n * stride;
                    const T Dv = sD[n * Q + Id];  // transposed: l_{I_d(a)}'(xi_n)
#pragma unroll
                    for (int f = 0; f < 3; ++f) {
                        const V2 pf = sN[le][s][f];
                        W[2 * f] += Dv * pf.x;
                        W[2 * f + 1] += Dv * pf.y;
                    }
                }
            }
            __syncthreads();  // sweep done before sN is overwritten
        }
        double v[5] = {0, 0, 0, 0, 0};
        if (active) {
            const T nu_e = (T)a.c_rho * mue / r;
            const T kappa = muf * Cp / Pr + (T)a.c_ener * mue * Cp / Pr + Cp * r * mus / (T)0.9;
            v[0] = (double)(nu_e * W[0]);
            v[1] = (double)W[3];
            v[2] = (double)W[4];
            v[3] = (double)W[5];
            v[4] = (double)(W[1] + kappa * W[2]);
        }
        add_pass<UMAX, EPB>(sR, pl.mode, le, active, li, v);
        if (pass + 1 < S::NPASS) cur = nxt;
    }
    __syncthreads();  // every node sum of the batch done
    write_batch_nodes<UMAX, NT>(pl.mode, wn, U, sR, a.Rmass, a.Rmom, a.Rener, a.npoin);
}

template <int P, class A, bool CLOSURE = false, typename TA = double>
static hipError_t launch_p(const A &a, hipStream_t s, const float *mlp = nullptr)
{
    if (a.plan.nb <= 0) return hipSuccess;
    hipLaunchKernelGGL((k_diffu<P, A, CLOSURE, TA>), dim3((unsigned)a.plan.nb), dim3(Shape<P>::NT), 0, s,
                       a, mlp);
    return hipGetLastError();
}

template <class A>
hipError_t launch_diffu(int porder, const A &a, hipStream_t s)
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

hipError_t launch_diffu_fp32(int porder, const DiffuArgs<float, float> &a, hipStream_t s)
{
    switch (porder) {
    case 1: return launch_p<1, DiffuArgs<float, float>, false, float>(a, s);
    case 2: return launch_p<2, DiffuArgs<float, float>, false, float>(a, s);
    case 3: return launch_p<3, DiffuArgs<float, float>, false, float>(a, s);
    case 4: return launch_p<4, DiffuArgs<float, float>, false, float>(a, s);
    case 5: return launch_p<5, DiffuArgs<float, float>, false, float>(a, s);
    case 6: return launch_p<6, DiffuArgs<float, float>, false, float>(a, s);
    case 7: return launch_p<7, DiffuArgs<float, float>, false, float>(a, s);
    case 8: return launch_p<8, DiffuArgs<float, float>, false, float>(a, s);
    default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_diffu(int, const DiffuArgs<double, double> &, hipStream_t);
template hipError_t launch_diffu(int, const DiffuArgs<float, float> &, hipStream_t);
template hipError_t launch_diffu(int, const DiffuArgs<float, double> &, hipStream_t);
template hipError_t launch_diffu(int, const DiffuArgs<float, float, float> &, hipStream_t);
template hipError_t launch_diffu(int, const DiffuSplitArgs &, hipStream_t);

hipError_t launch_diffu_closure(int porder, const DiffuArgs<double, double> &a, const float *mlp,
                                        hipStream_t s)
{
    switch (porder) {
    case 1: return launch_p<1, DiffuArgs<double, double>, true>(a, s, mlp);
    case 2: return launch_p<2, DiffuArgs<double, double>, true>(a, s, mlp);
    case 3: return launch_p<3, DiffuArgs<double, double>, true>(a, s, mlp);
    case 4: return launch_p<4, DiffuArgs<double, double>, true>(a, s, mlp);
    case 5: return launch_p<5, DiffuArgs<double, double>, true>(a, s, mlp);
    case 6: return launch_p<6, DiffuArgs<double, double>, true>(a, s, mlp);
    case 7: return launch_p<7, DiffuArgs<double, double>, true>(a, s, mlp);
    case 8: return launch_p<8, DiffuArgs<double, double>, true>(a, s, mlp);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace msod

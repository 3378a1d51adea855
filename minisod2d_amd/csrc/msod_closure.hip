// Learned subgrid-viscosity closure: the GPU side of training and inference.
//
// The diffusive residual is affine in mu_sgs (elem_diffu.f90:70,146): at an
// element node a, mu_sgs enters the stress through mu_f' = mu_f + rho*mu_sgs
// and the heat conductivity through kappa += Cp*rho*mu_sgs/0.9.  Hence, for a
// nodal error field r (momentum x3, energy), the gradient of <r, R(mu_sgs)>
// with respect to mu_sgs(e,a) is
//   g(e,a) = rho gpvol sum_d sum_i He(i,d) [ sum_k D_d[r_mom_k](a) S_ki
//                                             + D_d[r_ener](a) (S u)_i ]
//            + r_ener(a) * Cp rho / 0.9 * W[grad T](a)
// with S = G + G^T - 2/3 div(u) I (G = grad u), D_d the 1-D derivative along
// direction d inside the element, and W[grad T] the element's transposed
// sweep of the heat flux (the same quantities the forward kernel forms).
// k_adjoint computes g for every element node; k_mlp_grad back-propagates it
// through the MLP with bf16 MFMA products and FP32 accumulation.
#include <cmath>

#include "msod_kernels.h"
#include "msod_mlp.h"
#include "msod_closure.h"

namespace msod {

// ---------------------------------------------------------------------------
// element kernels: features (MODE 0) and mu_sgs gradients (MODE 1)
// ---------------------------------------------------------------------------
template <int P, int MODE>
__global__ void __launch_bounds__(Shape<P>::NT) k_closure_elem(ClosureElemArgs a)
{
    using S = Shape<P>;
    constexpr int Q = S::Q, NN = S::NN, EPB = S::EPB;
    __shared__ double sD[Q * Q];
    __shared__ double sA[EPB][5][NN];                  // rho, T, u0, u1, u2
    __shared__ double sE[EPB][MODE ? 4 : 1][NN];       // error field (MODE 1)
    __shared__ double sF[EPB][MODE ? 3 : 1][NN];       // heat flux, contravariant

    const int64_t groups = (a.nelem + EPB - 1) / EPB;
    const int64_t grp = blockIdx.x;
    if (grp >= groups) return;  // whole block
    const int t = threadIdx.x;
    for (int i = t; i < Q * Q; i += blockDim.x) sD[i] = a.tab.D[i];
    const int le = t / NN;
    const int l = t - le * NN;
    const int64_t e = grp * EPB + le;
    const bool active = (le < EPB) && (e < a.nelem);
    const int I = l % Q, J = (l / Q) % Q, K = l / (Q * Q);
    double r = 0, uu[3] = {0, 0, 0}, rerr_ener = 0;
    if (active) {
        const int64_t g = a.conn[e * NN + l];
        const int64_t np = a.npoin;
        r = a.rho[g];
        sA[le][0][l] = r;
        sA[le][1][l] = MODE ? a.Tem[g] : 0.0;
        for (int d = 0; d < 3; ++d) {
            uu[d] = a.u[g + d * np];
            sA[le][2 + d][l] = uu[d];
        }
        if (MODE) {
            for (int c = 0; c < 4; ++c) sE[le][c % (MODE ? 4 : 1)][l] = a.rerr[g + c * np];
            rerr_ener = a.rerr[g + 3 * np];
        }
    }
    __syncthreads();

    double h[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, G[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    double gT[3] = {0, 0, 0}, V[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, gv = 0;
    if (active) {
        const GeoNode<double> gn = load_geo<NN, false>(a.geo, e, l);
#pragma unroll
        for (int d = 0; d < 3; ++d)
#pragma unroll
            for (int i = 0; i < 3; ++i) h[i][d] = gn.he(i + 3 * d);
        gv = gn.gvol();
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int Id = d == 0 ? I : (d == 1 ? J : K);
            const int stride = d == 0 ? 1 : (d == 1 ? Q : Q * Q);
            const int base = l - Id * stride;
            double aT = 0, aU0 = 0, aU1 = 0, aU2 = 0, aE[4] = {0, 0, 0, 0};
#pragma unroll 1
            for (int n = 0; n < Q; ++n) {
                const int s = base + n * stride;
                const double Dv = sD[Id * Q + n];
                aT += Dv * sA[le][1][s];
                aU0 += Dv * sA[le][2][s];
                aU1 += Dv * sA[le][3][s];
                aU2 += Dv * sA[le][4][s];
                if (MODE)
#pragma unroll
                    for (int c = 0; c < 4; ++c) aE[c] += Dv * sE[le][c % (MODE ? 4 : 1)][s];
            }
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                gT[i] += h[i][d] * aT;
                G[0][i] += h[i][d] * aU0;
                G[1][i] += h[i][d] * aU1;
                G[2][i] += h[i][d] * aU2;
            }
            if (MODE)
#pragma unroll
                for (int c = 0; c < 4; ++c) V[c][d] = aE[c];
        }
    }
    if (MODE == 0) {
        if (active) {
            float f[kMlpIn];
            mlp_features<double>(G, r, uu, h, f);
#pragma unroll
            for (int k = 0; k < kMlpIn; ++k) a.feat[(e * NN + l) * kMlpIn + k] = f[k];
        }
        return;
    }
    // heat-flux transposed sweep W[grad T] at this node
    if (active)
#pragma unroll
        for (int d = 0; d < 3; ++d)
            sF[le][d % (MODE ? 3 : 1)][l] = gv * (h[0][d] * gT[0] + h[1][d] * gT[1] + h[2][d] * gT[2]);
    __syncthreads();
    if (!active) return;
    double WT = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const int Id = d == 0 ? I : (d == 1 ? J : K);
        const int stride = d == 0 ? 1 : (d == 1 ? Q : Q * Q);
        const int base = l - Id * stride;
#pragma unroll 1
        for (int n = 0; n < Q; ++n) WT += sD[n * Q + Id] * sF[le][d % (MODE ? 3 : 1)][base + n * stride];
    }
    const double divU = G[0][0] + G[1][1] + G[2][2];
    double Sm[3][3], Su[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double acc = 0;
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            Sm[i][j] = G[i][j] + G[j][i];
            acc += Sm[i][j] * uu[j];
        }
        Su[i] = acc - (2.0 / 3.0) * divU * uu[i];
        Sm[i][i] -= (2.0 / 3.0) * divU;
    }
    double acc = 0;
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int i = 0; i < 3; ++i)
            acc += h[i][d] * (V[0][d] * Sm[0][i] + V[1][d] * Sm[1][i] + V[2][d] * Sm[2][i] + V[3][d] * Su[i]);
    const double gnu = r * gv * acc + rerr_ener * a.Cp * r / 0.9 * WT;
    a.gnu[e * NN + l] = (float)gnu;
}

// ---------------------------------------------------------------------------
// MLP: inference into mu_sgs (reference layout), and parameter gradients
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_mlp_predict(const float *P, const float *feat, int64_t nelem,
                                                     int nn, const int32_t *l2g, double *mu_sgs)
{
    __shared__ MlpShared s;
    __shared__ short scratch[4][64 * kMlpH];
    mlp_stage(s, P);
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = i < nelem * nn;
    float f[kMlpIn];
#pragma unroll
    for (int k = 0; k < kMlpIn; ++k) f[k] = valid ? feat[i * kMlpIn + k] : s.mean[k];
    const float z3 = mlp_forward_z3(s, f, scratch[threadIdx.x >> 6]);
    if (valid) {
        const int64_t e = i / nn;
        const int l = (int)(i - e * nn);
        mu_sgs[e + nelem * (int64_t)l2g[l]] = (double)(s.scale * softplusf_(z3));
    }
}

// One wave per block; each block accumulates the gradient of its chunks of
// 64 samples and writes it to partial[blockIdx.x][kMlpTrain].
__global__ void __launch_bounds__(64) k_mlp_grad(const float *P, const float *feat, const float *gnu,
                                                 int64_t nsamp, float *partial)
{
    __shared__ MlpShared s;
    __shared__ short X[64][16];
    __shared__ float Z1[64][kMlpH + 1];   // pre-activations of layer 1, then dZ1
    __shared__ short H1[64][kMlpH];
    __shared__ float Z2[64][kMlpH + 1];   // pre-activations of layer 2
    __shared__ short dZ2[64][kMlpH];
    __shared__ short dZ1[64][kMlpH];
    __shared__ float dz3s[64];
    __shared__ float gacc[kMlpTrain];
    mlp_stage(s, P);
    for (int j = threadIdx.x; j < kMlpTrain; j += 64) gacc[j] = 0.f;
    __syncthreads();
    const int lane = threadIdx.x, r = lane & 15, g = lane >> 4;
    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
    for (int64_t chunk = blockIdx.x; chunk * 64 < nsamp; chunk += gridDim.x) {
        const int64_t i = chunk * 64 + lane;
        const bool valid = i < nsamp;
#pragma unroll
        for (int k = 0; k < 16; ++k)
            X[lane][k] = (valid && k < kMlpIn) ? f2bf((feat[i * kMlpIn + k] - s.mean[k]) * s.istd[k]) : (short)0;
        const float gv = valid ? gnu[i] : 0.f;
        __syncthreads();
        // layer 1
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            s16x4 av;
#pragma unroll
            for (int q = 0; q < 4; ++q) av[q] = X[mt * 16 + r][4 * g + q];
#pragma unroll
            for (int nt = 0; nt < 2; ++nt) {
                s16x4 bv;
#pragma unroll
                for (int q = 0; q < 4; ++q) bv[q] = s.W1[4 * g + q][nt * 16 + r];
                const f32x4 z = mfma16(av, bv, zero);
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int m = mt * 16 + 4 * g + q, n = nt * 16 + r;
                    const float zz = z[q] + s.b1[n];
                    Z1[m][n] = zz;
                    H1[m][n] = f2bf(siluf_(zz));
                }
            }
        }
        __syncthreads();
        // layer 2
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            s16x4 a0, a1;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                a0[q] = H1[mt * 16 + r][4 * g + q];
                a1[q] = H1[mt * 16 + r][16 + 4 * g + q];
            }
#pragma unroll
            for (int nt = 0; nt < 2; ++nt) {
                s16x4 b0, b1;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    b0[q] = s.W2[4 * g + q][nt * 16 + r];
                    b1[q] = s.W2[16 + 4 * g + q][nt * 16 + r];
                }
                f32x4 z = mfma16(a0, b0, zero);
                z = mfma16(a1, b1, z);
#pragma unroll
                for (int q = 0; q < 4; ++q) Z2[mt * 16 + 4 * g + q][nt * 16 + r] = z[q] + s.b2[nt * 16 + r];
            }
        }
        __syncthreads();
        // output layer and its gradient, per node (lane = node)
        float z3 = s.b3;
#pragma unroll 4
        for (int c = 0; c < kMlpH; ++c) z3 += siluf_(Z2[lane][c]) * s.w3[c];
        const float dz3 = valid ? gv * s.scale * sigmoidf_(z3) : 0.f;
        dz3s[lane] = dz3;
#pragma unroll 4
        for (int c = 0; c < kMlpH; ++c) dZ2[lane][c] = f2bf(dz3 * s.w3[c] * dsiluf_(Z2[lane][c]));
        __syncthreads();
        if (lane < kMlpH) {
            float sw = 0.f, sb = 0.f;
            for (int m = 0; m < 64; ++m) {
                const float d = dz3s[m], z = Z2[m][lane];
                sw += d * siluf_(z);
                sb += d * s.w3[lane] * dsiluf_(z);
            }
            gacc[kOw3 + lane] += sw;
            gacc[kOb2 + lane] += sb;
        } else if (lane == kMlpH) {
            float sum = 0.f;
            for (int m = 0; m < 64; ++m) sum += dz3s[m];
            gacc[kOb3] += sum;
        }
        // dW2 = H1^T dZ2   (M = 32 inputs, N = 32 outputs, K = 64 samples)
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
#pragma unroll
            for (int nt = 0; nt < 2; ++nt) {
                f32x4 acc = zero;
#pragma unroll
                for (int kt = 0; kt < 4; ++kt) {
                    s16x4 av, bv;
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        av[q] = H1[kt * 16 + 4 * g + q][mt * 16 + r];
                        bv[q] = dZ2[kt * 16 + 4 * g + q][nt * 16 + r];
                    }
                    acc = mfma16(av, bv, acc);
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) gacc[kOW2 + (mt * 16 + 4 * g + q) * kMlpH + nt * 16 + r] += acc[q];
            }
        // dH1 = dZ2 W2^T (M = 64 samples, N = 32, K = 32), then dZ1 = dH1 * silu'(Z1)
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int nt = 0; nt < 2; ++nt) {
                f32x4 acc = zero;
#pragma unroll
                for (int kt = 0; kt < 2; ++kt) {
                    s16x4 av, bv;
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        av[q] = dZ2[mt * 16 + r][kt * 16 + 4 * g + q];
                        bv[q] = s.W2[nt * 16 + r][kt * 16 + 4 * g + q];
                    }
                    acc = mfma16(av, bv, acc);
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int m = mt * 16 + 4 * g + q, k = nt * 16 + r;
                    const float d = acc[q] * dsiluf_(Z1[m][k]);
                    Z1[m][k] = d;
                    dZ1[m][k] = f2bf(d);
                }
            }
        __syncthreads();
        if (lane < kMlpH) {
            float sum = 0.f;
            for (int m = 0; m < 64; ++m) sum += Z1[m][lane];
            gacc[kOb1 + lane] += sum;
        }
        // dW1 = X^T dZ1   (M = 16 inputs, N = 32, K = 64 samples)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            f32x4 acc = zero;
#pragma unroll
            for (int kt = 0; kt < 4; ++kt) {
                s16x4 av, bv;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    av[q] = X[kt * 16 + 4 * g + q][r];
                    bv[q] = dZ1[kt * 16 + 4 * g + q][nt * 16 + r];
                }
                acc = mfma16(av, bv, acc);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int k = 4 * g + q;
                if (k < kMlpIn) gacc[kOW1 + k * kMlpH + nt * 16 + r] += acc[q];
            }
        }
        __syncthreads();
    }
    for (int j = threadIdx.x; j < kMlpTrain; j += 64) partial[(int64_t)blockIdx.x * kMlpTrain + j] = gacc[j];
}

__global__ void k_grad_reduce(const float *partial, int nblocks, float *grad)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= kMlpTrain) return;
    double s = 0.0;
    for (int b = 0; b < nblocks; ++b) s += partial[(int64_t)b * kMlpTrain + j];
    grad[j] = (float)s;
}

// Adam (Kingma & Ba 2015, Algorithm 1).  ob1 = 1 - b1, ob2 = 1 - b2 and the
// bias corrections c1, c2 come from the host in double.
__global__ void k_adam(float *P, const float *grad, float *m, float *v, float lr, float b1, float b2,
                       float ob1, float ob2, float eps, float c1, float c2)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= kMlpTrain) return;
    const float gj = grad[j];
    const float mj = b1 * m[j] + ob1 * gj;
    const float vj = b2 * v[j] + ob2 * gj * gj;
    m[j] = mj;
    v[j] = vj;
    P[j] -= lr * (mj / c1) / (sqrtf(vj / c2) + eps);
}

// r = 2 w_c (R_c - target_c) for the momentum and energy residuals; loss
// partial sums per block (w_c (R_c - target_c)^2).
__global__ void k_residual_error(const double *Rmom, const double *Rener, const double *target,
                                 const double *w, int64_t npoin, double *rerr, double *lpart)
{
    __shared__ double red[256];
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < npoin;
         i += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const double R = c < 3 ? Rmom[i + c * npoin] : Rener[i];
            const double d = R - target[i + c * npoin];
            rerr[i + c * npoin] = 2.0 * w[c] * d;
            acc += w[c] * d * d;
        }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) lpart[blockIdx.x] = red[0];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <int P, int MODE>
static hipError_t launch_elem_p(const ClosureElemArgs &a, hipStream_t s)
{
    using S = Shape<P>;
    const int64_t groups = (a.nelem + S::EPB - 1) / S::EPB;
    hipLaunchKernelGGL((k_closure_elem<P, MODE>), dim3((unsigned)groups), dim3(S::NT), 0, s, a);
    return hipGetLastError();
}

template <int MODE>
static hipError_t launch_elem(int porder, const ClosureElemArgs &a, hipStream_t s)
{
    switch (porder) {
    case 1: return launch_elem_p<1, MODE>(a, s);
    case 2: return launch_elem_p<2, MODE>(a, s);
    case 3: return launch_elem_p<3, MODE>(a, s);
    case 4: return launch_elem_p<4, MODE>(a, s);
    case 5: return launch_elem_p<5, MODE>(a, s);
    case 6: return launch_elem_p<6, MODE>(a, s);
    case 7: return launch_elem_p<7, MODE>(a, s);
    case 8: return launch_elem_p<8, MODE>(a, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_closure_features(int porder, const ClosureElemArgs &a, hipStream_t s)
{
    return launch_elem<0>(porder, a, s);
}

hipError_t launch_closure_adjoint(int porder, const ClosureElemArgs &a, hipStream_t s)
{
    return launch_elem<1>(porder, a, s);
}

hipError_t launch_mlp_predict(const float *P, const float *feat, int64_t nelem, int nn,
                              const int32_t *l2g, double *mu_sgs, hipStream_t s)
{
    const int64_t n = nelem * nn;
    hipLaunchKernelGGL(k_mlp_predict, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, P, feat, nelem, nn,
                       l2g, mu_sgs);
    return hipGetLastError();
}

hipError_t launch_mlp_grad(const float *P, const float *feat, const float *gnu, int64_t nsamp,
                           float *partial, int nblocks, float *grad, hipStream_t s)
{
    hipLaunchKernelGGL(k_mlp_grad, dim3(nblocks), dim3(64), 0, s, P, feat, gnu, nsamp, partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_grad_reduce, dim3((kMlpTrain + 255) / 256), dim3(256), 0, s, partial, nblocks, grad);
    return hipGetLastError();
}

hipError_t launch_adam(float *P, const float *grad, float *m, float *v, int step, double lr, double b1,
                       double b2, double eps, hipStream_t s)
{
    const double c1 = 1.0 - std::pow(b1, step), c2 = 1.0 - std::pow(b2, step);
    hipLaunchKernelGGL(k_adam, dim3((kMlpTrain + 255) / 256), dim3(256), 0, s, P, grad, m, v, (float)lr,
                       (float)b1, (float)b2, (float)(1.0 - b1), (float)(1.0 - b2), (float)eps, (float)c1,
                       (float)c2);
    return hipGetLastError();
}

hipError_t launch_residual_error(const double *Rmom, const double *Rener, const double *target,
                                 const double *w, int64_t npoin, double *rerr, double *lpart, int nblocks,
                                 hipStream_t s)
{
    hipLaunchKernelGGL(k_residual_error, dim3(nblocks), dim3(256), 0, s, Rmom, Rener, target, w, npoin, rerr,
                       lpart);
    return hipGetLastError();
}

}  // namespace msod

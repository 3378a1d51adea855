// Learned subgrid-viscosity closure: a per-node MLP
//     features (12) -> 32 -> 32 -> 1,  SiLU hidden layers, softplus output
//     mu_sgs = scale * softplus(z3)
// evaluated with bf16 MFMA (FP32 accumulation) by whole wavefronts: lane L of
// a wave owns node L, the 64 nodes of the wave form the M dimension of the
// layer GEMMs (v_mfma_f32_16x16x16_bf16: four 16-row tiles).
//
// Features of an element node (all from the diffusion kernel's pass 1):
//   G(i,j) = du_i/dx_j (9, row-major), rho, |u|, h = 2 det(J)^(1/3)
// standardised with a mean and inverse deviation stored with the weights.
//
// Parameter block (FP32, kMlpParams floats):
//   W1[12][32] b1[32] W2[32][32] b2[32] w3[32] b3 | mean[12] istd[12] scale
// the first kMlpTrain are trained; W[in][out] row-major.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace msod {

constexpr int kMlpIn = 12;
constexpr int kMlpH = 32;
enum MlpOffsets : int {
    kOW1 = 0,
    kOb1 = kOW1 + kMlpIn * kMlpH,
    kOW2 = kOb1 + kMlpH,
    kOb2 = kOW2 + kMlpH * kMlpH,
    kOw3 = kOb2 + kMlpH,
    kOb3 = kOw3 + kMlpH,
    kMlpTrain = kOb3 + 1,
    kOMean = kMlpTrain,
    kOIstd = kOMean + kMlpIn,
    kOScale = kOIstd + kMlpIn,
    kMlpParams = kOScale + 1
};

typedef short s16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bf16 bits, round to nearest even (finite inputs)
__device__ __forceinline__ short f2bf(float f)
{
    const uint32_t u = __float_as_uint(f);
    return (short)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float bf2f(short b)
{
    return __uint_as_float(((uint32_t)(uint16_t)b) << 16);
}
// The hardware reciprocal (1 ulp) instead of a correctly rounded division:
// the network runs in bf16, whose rounding is 2^-9.
__device__ __forceinline__ float sigmoidf_(float z) { return __builtin_amdgcn_rcpf(1.f + __expf(-z)); }
__device__ __forceinline__ float siluf_(float z) { return z * sigmoidf_(z); }
__device__ __forceinline__ float dsiluf_(float z)
{
    const float s = sigmoidf_(z);
    return s * (1.f + z * (1.f - s));
}
__device__ __forceinline__ float softplusf_(float z) { return z > 20.f ? z : log1pf(__expf(z)); }

// 16x16x16 bf16 MFMA tile product on one wavefront: lane (r = lane%16,
// g = lane/16) supplies A[r][4g..4g+3] and B[4g..4g+3][r] and receives
// C[4g..4g+3][r].
__device__ __forceinline__ f32x4 mfma16(s16x4 a, s16x4 b, f32x4 c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// Block-shared copy of the weights: bf16 for the MFMA operands.
struct MlpShared {
    short W1[16][kMlpH];   // rows >= kMlpIn are zero
    short W2[kMlpH][kMlpH];
    float b1[kMlpH], b2[kMlpH], w3[kMlpH];
    float mean[kMlpIn], istd[kMlpIn];
    float b3, scale;
};

// All threads of the block; the caller synchronises afterwards.
__device__ inline void mlp_stage(MlpShared &s, const float *P)
{
    for (int i = threadIdx.x; i < 16 * kMlpH; i += blockDim.x) {
        const int k = i / kMlpH, n = i % kMlpH;
        s.W1[k][n] = k < kMlpIn ? f2bf(P[kOW1 + k * kMlpH + n]) : (short)0;
    }
    for (int i = threadIdx.x; i < kMlpH * kMlpH; i += blockDim.x)
        s.W2[i / kMlpH][i % kMlpH] = f2bf(P[kOW2 + i]);
    for (int i = threadIdx.x; i < kMlpH; i += blockDim.x) {
        s.b1[i] = P[kOb1 + i];
        s.b2[i] = P[kOb2 + i];
        s.w3[i] = P[kOw3 + i];
    }
    for (int i = threadIdx.x; i < kMlpIn; i += blockDim.x) {
        s.mean[i] = P[kOMean + i];
        s.istd[i] = P[kOIstd + i];
    }
    if (threadIdx.x == 0) {
        s.b3 = P[kOb3];
        s.scale = P[kOScale];
    }
}

// Features of one element node from its pass-1 quantities.
// Gij[i][j] = du_i/dx_j, hinv[i][d] = He(i,d).
template <typename T>
__device__ __forceinline__ void mlp_features(const T (&Gij)[3][3], T rho, const T (&u)[3],
                                             const T (&hinv)[3][3], float f[kMlpIn])
{
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) f[3 * i + j] = (float)Gij[i][j];
    f[9] = (float)rho;
    f[10] = sqrtf((float)(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
    const double det = (double)hinv[0][0] * ((double)hinv[1][1] * hinv[2][2] - (double)hinv[1][2] * hinv[2][1]) -
                       (double)hinv[0][1] * ((double)hinv[1][0] * hinv[2][2] - (double)hinv[1][2] * hinv[2][0]) +
                       (double)hinv[0][2] * ((double)hinv[1][0] * hinv[2][1] - (double)hinv[1][1] * hinv[2][0]);
    f[11] = 2.f * __builtin_amdgcn_rcpf(cbrtf(fabsf((float)det)));  // det J = 1/det He
}

// Orders this wave's LDS accesses (a wave's LDS operations complete in
// issue order; the fence keeps the compiler from moving them across).
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// z3 of the MLP for this lane's node (mu_sgs = scale*softplus(z3)).
// scratch: this wave's LDS area of 64*32 shorts (4 KiB), used by this wave
// only (no block barrier inside).  All 64 lanes of the wave call it
// (inactive lanes pass any finite features).
__device__ inline float mlp_forward_z3(const MlpShared &s, const float feat[kMlpIn], short *scratch)
{
    const int lane = threadIdx.x & 63, r = lane & 15, g = lane >> 4;
    short *X = scratch;  // [64][16]
#pragma unroll
    for (int k = 0; k < 16; ++k)
        X[lane * 16 + k] = k < kMlpIn ? f2bf((feat[k] - s.mean[k]) * s.istd[k]) : (short)0;
    wave_lds_sync();
    f32x4 z1[4][2];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        s16x4 a;
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = X[(mt * 16 + r) * 16 + 4 * g + i];
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            s16x4 b;
#pragma unroll
            for (int i = 0; i < 4; ++i) b[i] = s.W1[4 * g + i][nt * 16 + r];
            z1[mt][nt] = mfma16(a, b, f32x4{0.f, 0.f, 0.f, 0.f});
        }
    }
    wave_lds_sync();
    short *H = scratch;  // [64][32], over X
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int n = nt * 16 + r;
                H[(mt * 16 + 4 * g + i) * kMlpH + n] = f2bf(siluf_(z1[mt][nt][i] + s.b1[n]));
            }
    wave_lds_sync();
    float part[4][4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
        s16x4 a0, a1;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            a0[i] = H[(mt * 16 + r) * kMlpH + 4 * g + i];
            a1[i] = H[(mt * 16 + r) * kMlpH + 16 + 4 * g + i];
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) part[mt][i] = 0.f;
#pragma unroll
        for (int nt = 0; nt < 2; ++nt) {
            s16x4 b0, b1;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                b0[i] = s.W2[4 * g + i][nt * 16 + r];
                b1[i] = s.W2[16 + 4 * g + i][nt * 16 + r];
            }
            f32x4 z = mfma16(a0, b0, f32x4{0.f, 0.f, 0.f, 0.f});
            z = mfma16(a1, b1, z);
            const int n = nt * 16 + r;
#pragma unroll
            for (int i = 0; i < 4; ++i) part[mt][i] += siluf_(z[i] + s.b2[n]) * s.w3[n];
        }
    }
    // sum over the 16 hidden columns held by the lanes of each group
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int off = 1; off < 16; off <<= 1) part[mt][i] += __shfl_xor(part[mt][i], off);
    // row m = 16 mt + 4 g' + i lives in the lanes of group g'; lane m takes it
    float mine = 0.f;
    const int src = ((lane & 15) >> 2) * 16;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float v = __shfl(part[mt][i], src);
            if ((lane >> 4) == mt && (lane & 3) == i) mine = v;
        }
    wave_lds_sync();
    return mine + s.b3;
}

}  // namespace msod

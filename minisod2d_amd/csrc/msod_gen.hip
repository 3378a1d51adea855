// Synthetic inputs on the device: the conforming box mesh and the flow and
// viscosity fields of the benchmark configurations, the same numbers as the
// host generator (minisod2d_amd/fortran/msod_mesh.f90, msod_fields.f90,
// msod_rng.c), element by element and node by node.
//
// Draw i of stream `seed` is the i-th output of a SplitMix64 generator whose
// state starts at `seed`: a pure function of (seed, i), so every thread draws
// its own numbers and the result does not depend on the launch shape.
#include "msod_gen.h"

namespace msod {

__device__ __forceinline__ uint64_t splitmix64_at(uint64_t seed, uint64_t i)
{
    uint64_t z = seed + (i + 1u) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// uniform in [0,1) with 53 random bits
__device__ __forceinline__ double rng_uniform(uint64_t seed, uint64_t i)
{
    return (double)(splitmix64_at(seed, i) >> 11) * 0x1.0p-53;
}

// standard normal by Box-Muller from draws 2i and 2i+1
__device__ __forceinline__ double rng_normal(uint64_t seed, uint64_t i)
{
    const double two_pi = 6.283185307179586476925286766559;
    const double u1 = 1.0 - rng_uniform(seed, 2u * i);  // (0,1]
    const double u2 = rng_uniform(seed, 2u * i + 1u);
    return sqrt(-2.0 * log(u1)) * cos(two_pi * u2);
}

// connec(e,a) = 1 + lattice index of node a of element e (elements x fastest)
__global__ void k_box_connec(BoxArgs b)
{
    const int64_t n = b.nelem * b.nn;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        // i = e + nelem*a: consecutive threads write consecutive elements of one node
        const int64_t e = i % b.nelem;
        const int a = (int)(i / b.nelem);
        const int64_t ex = e % b.nx, ey = (e / b.nx) % b.ny, ez = e / ((int64_t)b.nx * b.ny);
        const int64_t gi = ex * b.p + b.pos[3 * a], gj = ey * b.p + b.pos[3 * a + 1],
                      gk = ez * b.p + b.pos[3 * a + 2];
        b.connec[i] = (int32_t)(1 + gi + b.n1x * (gj + b.n1y * gk));
    }
}

// coord(ip,d) on the GLL lattice; element-interior nodes jittered
__global__ void k_box_coord(BoxArgs b)
{
    for (int64_t ip = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; ip < b.npoin;
         ip += (int64_t)gridDim.x * blockDim.x) {
        const int64_t g[3] = {ip % b.n1x, (ip / b.n1x) % b.n1y, ip / (b.n1x * b.n1y)};
        int pos[3];
        double x[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            pos[d] = (int)(g[d] % b.p);
            x[d] = b.h[d] * ((double)(g[d] / b.p) + 0.5 * (b.xi[pos[d]] + 1.0));
        }
        if (b.jitter > 0.0 && pos[0] != 0 && pos[1] != 0 && pos[2] != 0) {
#pragma unroll
            for (int d = 0; d < 3; ++d)
                x[d] = x[d] + b.jitter * 0.5 * b.h[d] * b.hloc[pos[d]] *
                                  (2.0 * rng_uniform(b.seed, 3u * (uint64_t)ip + d) - 1.0);
        }
#pragma unroll
        for (int d = 0; d < 3; ++d) b.coord[ip + b.npoin * d] = x[d];
    }
}

__global__ void k_flow_fields(FlowArgs f)
{
    const double twopi = 6.283185307179586476925286766559;
    const int64_t np = f.npoin;
    for (int64_t ip = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; ip < np;
         ip += (int64_t)gridDim.x * blockDim.x) {
        const double x = f.coord[ip], y = f.coord[ip + np], z = f.coord[ip + 2 * np];
        double u[3];
        u[0] = sin(twopi * x) * cos(twopi * y) * cos(twopi * z);
        u[1] = -cos(twopi * x) * sin(twopi * y) * cos(twopi * z);
        u[2] = 0.0;
        double rho = 1.0 + 0.05 * rng_normal(f.seed_rho, (uint64_t)ip);
        if (f.rho_jump != 0.0 && x > 0.5) rho = rho * (1.0 + f.rho_jump);
        const double pr = 1.0 / f.gamma + 0.1 * cos(2.0 * twopi * x);
        double fmag = 0.0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double qd = rho * u[d];
            f.u[ip + np * d] = u[d];
            f.q[ip + np * d] = qd;
            fmag = fmag + qd * u[d];
        }
        const double eint = pr / ((f.gamma - 1.0) * rho);
        f.rho[ip] = rho;
        f.pr[ip] = pr;
        f.E[ip] = rho * eint + 0.5 * fmag;
        f.Tem[ip] = pr / (rho * f.Rgas);
        f.mu_fluid[ip] = 1e-3;
    }
}

// mu(e,g) in the reference layout (nelem,ngaus); draw index e*ngaus + g
__global__ void k_viscosity_fields(ViscArgs v)
{
    const int64_t n = v.nelem * v.ngaus;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = i % v.nelem, g = i / v.nelem;
        const uint64_t k = (uint64_t)(e * v.ngaus + g);
        v.mu_e[i] = v.mue_max * rng_uniform(v.seed_mue, k);
        v.mu_sgs[i] = v.musgs_max * rng_uniform(v.seed_musgs, k);
    }
}

static unsigned blocks_for(int64_t n)
{
    const int64_t b = (n + 255) / 256;
    return (unsigned)(b < 1 ? 1 : (b > 65536 ? 65536 : b));
}

hipError_t launch_box_mesh(const BoxArgs &b, hipStream_t s)
{
    hipLaunchKernelGGL(k_box_connec, dim3(blocks_for(b.nelem * b.nn)), dim3(256), 0, s, b);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_box_coord, dim3(blocks_for(b.npoin)), dim3(256), 0, s, b);
    return hipGetLastError();
}

hipError_t launch_flow_fields(const FlowArgs &f, hipStream_t s)
{
    hipLaunchKernelGGL(k_flow_fields, dim3(blocks_for(f.npoin)), dim3(256), 0, s, f);
    return hipGetLastError();
}

hipError_t launch_viscosity_fields(const ViscArgs &v, hipStream_t s)
{
    hipLaunchKernelGGL(k_viscosity_fields, dim3(blocks_for(v.nelem * v.ngaus)), dim3(256), 0, s, v);
    return hipGetLastError();
}

}  // namespace msod

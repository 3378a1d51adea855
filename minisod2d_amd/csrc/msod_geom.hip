// Geometry kernel: He and gpvol of every element node from the node
// coordinates (counterpart of elem_jacobian, src/fem/sources/jacobian_oper.f90:7-136,
// which the reference driver runs on the device once per mesh,
// src/app/sources/main.f90:292-295).
//
//   Je(i,j) = sum_a dNgp(i,a,g) * coord(connec(e,a), j)
//   det     = Je11 Je22 Je33 + Je12 Je23 Je31 + Je13 Je21 Je32
//           - Je31 Je22 Je13 - Je32 Je23 Je11 - Je33 Je21 Je12
//   He      = cofactor(Je)^T / det,   gpvol = wgp(g) * det
//
// The sum over a runs over the nonzero entries of dNgp(i,:,g) only, in
// increasing a (the zero terms of the reference's dense dot_product add
// nothing); for a tensor-product hexahedron these are the q nodes on the
// grid line of g in direction i.  One workgroup takes EC consecutive
// elements:
//   1. their connec into LDS, neighbouring lanes reading neighbouring
//      elements of one node: the layout (nelem,nnode) puts one element's
//      nodes 4*nelem bytes apart, so EC (32 at p=4) elements fill a line;
//   then, EB elements at a time (one thread per element node):
//   2. the nodes' coordinates into LDS;
//   3. Je, det, He, gpvol into LDS;
//   4. He and gpvol written back as the contiguous block they form in the
//      reference layouts (He(3,3,ngaus,nelem), gpvol(1,ngaus,nelem)).
#include "msod_kernels.h"
#include "msod_geom.h"

namespace msod {

template <int P>
struct GeomShape {
    static constexpr int Q = P + 1;
    static constexpr int NN = Q * Q * Q;
    static constexpr int NT = 256;
    static constexpr int EB = NN <= NT / 2 ? NT / NN : 1;  // elements per step
    // elements per workgroup: a whole number of steps, connec within 16 KB
    static constexpr int EC_RAW = 4096 / NN < 32 ? 4096 / NN : 32;
    static constexpr int EC = EC_RAW < EB ? EB : (EC_RAW / EB) * EB;
};

template <int P>
__global__ void __launch_bounds__(256) k_elem_jacobian(GeomArgs a)
{
    using S = GeomShape<P>;
    constexpr int NN = S::NN, EB = S::EB, EC = S::EC, NT = S::NT;
    __shared__ int32_t sC[EC][NN];
    __shared__ double sX[EB][NN][3];
    __shared__ double sH[EB * NN * 9];
    __shared__ double sV[EB * NN];
    const int t = threadIdx.x;
    const int64_t ec0 = (int64_t)blockIdx.x * EC;
    const int nec = (int)(a.nelem - ec0 < EC ? a.nelem - ec0 : EC);

    for (int i = t; i < EC * NN; i += NT) {
        const int eb = i % EC, n = i / EC;
        sC[eb][n] = eb < nec ? a.connec[(ec0 + eb) + a.nelem * (int64_t)n] : 1;
    }
    unsigned bad_conn = 0, bad_det = 0;
    for (int c0 = 0; c0 < nec; c0 += EB) {
    const int64_t e0 = ec0 + c0;
    const int ne = nec - c0 < EB ? nec - c0 : EB;
    __syncthreads();  // sC staged; the previous step's write-back is done with sH, sV
    for (int i = t; i < EB * NN; i += NT) {
        const int eb = i / NN, n = i - eb * NN;
        int64_t ip = (int64_t)sC[c0 + (eb < ne ? eb : 0)][n] - 1;
        if (ip < 0 || ip >= a.npoin) {  // reported, read as node 1
            bad_conn += eb < ne;
            ip = 0;
        }
#pragma unroll
        for (int j = 0; j < 3; ++j) sX[eb][n][j] = a.coord[ip + a.npoin * j];
    }
    __syncthreads();
    for (int i = t; i < EB * NN; i += NT) {
        const int eb = i / NN, g = i - eb * NN;
        double Je[3][3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double s0 = 0.0, s1 = 0.0, s2 = 0.0;
            const int base = (g * 3 + d) * a.mz;
            const int cnt = a.nzn[g * 3 + d];
            for (int k = 0; k < cnt; ++k) {
                const int n = a.nza[base + k];
                const double w = a.nzw[base + k];
                s0 += w * sX[eb][n][0];
                s1 += w * sX[eb][n][1];
                s2 += w * sX[eb][n][2];
            }
            Je[d][0] = s0;
            Je[d][1] = s1;
            Je[d][2] = s2;
        }
        const double det = Je[0][0] * Je[1][1] * Je[2][2] + Je[0][1] * Je[1][2] * Je[2][0] +
                           Je[0][2] * Je[1][0] * Je[2][1] - Je[2][0] * Je[1][1] * Je[0][2] -
                           Je[2][1] * Je[1][2] * Je[0][0] - Je[2][2] * Je[1][0] * Je[0][1];
        // cofactors c(i,j) = (-1)^(i+j) minor(i,j); He(i,j) = c(j,i) / det
        double c[3][3];
        c[0][0] = Je[1][1] * Je[2][2] - Je[2][1] * Je[1][2];
        c[0][1] = -(Je[1][0] * Je[2][2] - Je[2][0] * Je[1][2]);
        c[0][2] = Je[1][0] * Je[2][1] - Je[2][0] * Je[1][1];
        c[1][0] = -(Je[0][1] * Je[2][2] - Je[2][1] * Je[0][2]);
        c[1][1] = Je[0][0] * Je[2][2] - Je[2][0] * Je[0][2];
        c[1][2] = -(Je[0][0] * Je[2][1] - Je[2][0] * Je[0][1]);
        c[2][0] = Je[0][1] * Je[1][2] - Je[1][1] * Je[0][2];
        c[2][1] = -(Je[0][0] * Je[1][2] - Je[1][0] * Je[0][2]);
        c[2][2] = Je[0][0] * Je[1][1] - Je[1][0] * Je[0][1];
        const double rdet = 1.0 / det;
        double *H = sH + (size_t)(eb * NN + g) * 9;  // He(:,:,g,e) column-major
#pragma unroll
        for (int ii = 0; ii < 3; ++ii)
#pragma unroll
            for (int jj = 0; jj < 3; ++jj) H[ii + 3 * jj] = rdet * c[jj][ii];
        sV[eb * NN + g] = a.wgp[g] * det;
        bad_det += (eb < ne) && !(det > 0.0);
    }
    __syncthreads();
    double *He = a.He + e0 * NN * 9;
    for (int i = t; i < ne * NN * 9; i += NT) He[i] = sH[i];
    double *gv = a.gpvol + e0 * NN;
    for (int i = t; i < ne * NN; i += NT) gv[i] = sV[i];
    }
    if (bad_conn) atomicAdd(&a.counts[0], (unsigned long long)bad_conn);
    if (bad_det) atomicAdd(&a.counts[1], (unsigned long long)bad_det);
}

template <int P>
static hipError_t launch_p(const GeomArgs &a, hipStream_t s)
{
    using S = GeomShape<P>;
    const int64_t blocks = (a.nelem + S::EC - 1) / S::EC;
    hipLaunchKernelGGL(k_elem_jacobian<P>, dim3((unsigned)blocks), dim3(S::NT), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_elem_jacobian(int porder, const GeomArgs &a, hipStream_t s)
{
    if (a.nelem <= 0) return hipSuccess;
    switch (porder) {
    case 1: return launch_p<1>(a, s);
    case 2: return launch_p<2>(a, s);
    case 3: return launch_p<3>(a, s);
    case 4: return launch_p<4>(a, s);
    case 5: return launch_p<5>(a, s);
    case 6: return launch_p<6>(a, s);
    case 7: return launch_p<7>(a, s);
    case 8: return launch_p<8>(a, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace msod

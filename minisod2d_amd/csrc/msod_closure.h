// Internal interface of the learned-closure kernels (msod_closure.hip).
#pragma once

#include "msod_kernels.h"

namespace msod {

// Element kernels of the closure: features (12 per element node, element-
// major, l lexicographic) and the gradient of <rerr, R_diffu> with respect
// to mu_sgs at every element node.
struct ClosureElemArgs {
    int64_t nelem, npoin;
    const uint32_t *conn;      // [nelem][nn] 0-based nodes
    const double *geo;  // element-major geometry (GeoNode pairs)
    const double *rho, *u, *Tem;
    double Cp;
    const double *rerr;        // [4][npoin]: momentum x, y, z, energy
    float *feat;               // [nelem*nn][12]
    float *gnu;                // [nelem*nn]
    Tables tab;
};

hipError_t launch_closure_features(int porder, const ClosureElemArgs &a, hipStream_t s);
hipError_t launch_closure_adjoint(int porder, const ClosureElemArgs &a, hipStream_t s);
hipError_t launch_mlp_predict(const float *P, const float *feat, int64_t nelem, int nn,
                              const int32_t *l2g, double *mu_sgs, hipStream_t s);
hipError_t launch_mlp_grad(const float *P, const float *feat, const float *gnu, int64_t nsamp,
                           float *partial, int nblocks, float *grad, hipStream_t s);
hipError_t launch_adam(float *P, const float *grad, float *m, float *v, int step, double lr, double b1,
                       double b2, double eps, hipStream_t s);
hipError_t launch_residual_error(const double *Rmom, const double *Rener, const double *target,
                                 const double *w, int64_t npoin, double *rerr, double *lpart,
                                 int nblocks, hipStream_t s);

// Diffusion with mu_sgs from the MLP, evaluated inside the kernel (FP64).
hipError_t launch_diffu_closure(int porder, const DiffuArgs<double, double> &a, const float *mlp,
                                hipStream_t s);

}  // namespace msod

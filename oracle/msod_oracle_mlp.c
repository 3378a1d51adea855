/*
 * CPU oracle of the learned subgrid-viscosity closure.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the closure kernels; the product
 * never links or calls it.  The reference has no closure (mu_sgs is an input
 * the driver sets to 1, src/app/sources/main.f90:334), so these functions
 * restate the network's own definition:
 *
 *   x  = (f - mean) * istd                         (12 features)
 *   z1 = x W1 + b1,  h1 = silu(z1)                 (32)
 *   z2 = h1 W2 + b2, h2 = silu(z2)                 (32)
 *   z3 = h2 . w3 + b3
 *   mu_sgs = scale * softplus(z3)
 *
 * with the parameter block W1[12][32] b1[32] W2[32][32] b2[32] w3[32] b3 |
 * mean[12] istd[12] scale (1530 floats, the first 1505 trained).
 *
 * quant = 1 rounds to bfloat16 (nearest even) exactly where the GPU kernels
 * feed the matrix cores: the standardised features, W1, h1, W2 (forward);
 * the backward of quant = 1 is that of the rounded forward with the rounding
 * taken as the identity.  quant = 0 is the plain double-precision network.
 * Everything else is double.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "msod_oracle.h"

enum { NIN = 12, NH = 32 };
enum {
    OW1 = 0,
    OB1 = OW1 + NIN * NH,
    OW2 = OB1 + NH,
    OB2 = OW2 + NH * NH,
    OW3 = OB2 + NH,
    OB3 = OW3 + NH,
    NTRAIN = OB3 + 1,
    OMEAN = NTRAIN,
    OISTD = OMEAN + NIN,
    OSCALE = OISTD + NIN,
    NPARAM = OSCALE + 1
};

/* float -> nearest bfloat16 (ties to even), returned as a float */
static float bf16(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    u = (u + 0x7fffu + ((u >> 16) & 1u)) & 0xffff0000u;
    memcpy(&f, &u, 4);
    return f;
}

static double sigm(double z) { return 1.0 / (1.0 + exp(-z)); }
static double silu(double z) { return z * sigm(z); }
static double dsilu(double z)
{
    const double s = sigm(z);
    return s * (1.0 + z * (1.0 - s));
}
static double softplus(double z) { return z > 30.0 ? z : log1p(exp(z)); }

typedef struct {
    double x[NIN], z1[NH], h1[NH], z2[NH], h2[NH], z3;
} Trace;

static void forward(const float *P, const float *f, int quant, Trace *t)
{
    for (int k = 0; k < NIN; ++k) {
        const float xs = (f[k] - P[OMEAN + k]) * P[OISTD + k];
        t->x[k] = quant ? bf16(xs) : (double)xs;
    }
    for (int n = 0; n < NH; ++n) {
        double z = 0.0;
        for (int k = 0; k < NIN; ++k) {
            const double w = quant ? bf16(P[OW1 + k * NH + n]) : P[OW1 + k * NH + n];
            z += t->x[k] * w;
        }
        t->z1[n] = z + P[OB1 + n];
        const double h = silu(t->z1[n]);
        t->h1[n] = quant ? bf16((float)h) : h;
    }
    double z3 = P[OB3];
    for (int n = 0; n < NH; ++n) {
        double z = 0.0;
        for (int k = 0; k < NH; ++k) {
            const double w = quant ? bf16(P[OW2 + k * NH + n]) : P[OW2 + k * NH + n];
            z += t->h1[k] * w;
        }
        t->z2[n] = z + P[OB2 + n];
        t->h2[n] = silu(t->z2[n]);
        z3 += t->h2[n] * P[OW3 + n];
    }
    t->z3 = z3;
}

void msod_oracle_mlp_forward(const float *P, const float *features, int64_t n, int quant,
                             double *z3, double *mu_sgs)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        Trace t;
        forward(P, features + i * NIN, quant, &t);
        if (z3) z3[i] = t.z3;
        if (mu_sgs) mu_sgs[i] = P[OSCALE] * softplus(t.z3);
    }
}

/* grad[j] = sum_i gnu[i] d mu_sgs_i / d P[j], j < 1505 */
void msod_oracle_mlp_gradient(const float *P, const float *features, const double *gnu, int64_t n,
                              int quant, double *grad)
{
    memset(grad, 0, sizeof(double) * NTRAIN);
#pragma omp parallel
    {
        double g[NTRAIN];
        memset(g, 0, sizeof g);
#pragma omp for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            Trace t;
            forward(P, features + i * NIN, quant, &t);
            const double dz3 = gnu[i] * P[OSCALE] * sigm(t.z3);
            double dz2[NH], dh1[NH];
            g[OB3] += dz3;
            for (int c = 0; c < NH; ++c) {
                g[OW3 + c] += dz3 * t.h2[c];
                dz2[c] = dz3 * P[OW3 + c] * dsilu(t.z2[c]);
                g[OB2 + c] += dz2[c];
            }
            for (int k = 0; k < NH; ++k) {
                double s = 0.0;
                for (int c = 0; c < NH; ++c) {
                    g[OW2 + k * NH + c] += t.h1[k] * dz2[c];
                    const double w = quant ? bf16(P[OW2 + k * NH + c]) : P[OW2 + k * NH + c];
                    s += dz2[c] * w;
                }
                dh1[k] = s;
            }
            for (int k = 0; k < NH; ++k) {
                const double dz1 = dh1[k] * dsilu(t.z1[k]);
                g[OB1 + k] += dz1;
                for (int m = 0; m < NIN; ++m) g[OW1 + m * NH + k] += t.x[m] * dz1;
            }
        }
#pragma omp critical
        for (int j = 0; j < NTRAIN; ++j) grad[j] += g[j];
    }
}

/* one Adam step (Kingma & Ba 2015, Algorithm 1) on the trained parameters */
void msod_oracle_adam(float *P, const float *grad, float *m, float *v, int step, double lr,
                      double beta1, double beta2, double eps)
{
    const double c1 = 1.0 - pow(beta1, step), c2 = 1.0 - pow(beta2, step);
    for (int j = 0; j < NTRAIN; ++j) {
        const double mj = beta1 * m[j] + (1.0 - beta1) * grad[j];
        const double vj = beta2 * v[j] + (1.0 - beta2) * (double)grad[j] * grad[j];
        m[j] = (float)mj;
        v[j] = (float)vj;
        P[j] = (float)(P[j] - lr * (mj / c1) / (sqrt(vj / c2) + eps));
    }
}

/*
 * CPU oracle for the miniSOD2D spectral-element residual kernels.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of the reference algorithm
 * used as the checker by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py.  The product never links or calls it.
 *
 * All arrays use the reference's Fortran layouts (column-major, 1-based
 * index VALUES inside connec / invAtoIJK / AtoI/J/K):
 *   connec(nelem,nnode)        int32   connec[e + nelem*a]
 *   He(3,3,ngaus,nelem)        double  He[i + 3*k + 9*(g + ngaus*e)]
 *   gpvol(1,ngaus,nelem)       double  gpvol[g + ngaus*e]
 *   dlxigp_ip(ngaus,3,p+1)     double  dl[g + ngaus*(d + 3*n)]
 *   invAtoIJK(p+1,p+1,p+1)     int32   inv[i + q*(j + q*k)]
 *   u,q(npoin,3)               double  u[ip + npoin*d]
 *   mu_e, mu_sgs(nelem,ngaus)  double  mu[e + nelem*g]
 * with q = p+1 and nnode = ngaus = q^3.
 */
#ifndef MSOD_ORACLE_H
#define MSOD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* GLL roots and weights on [-1,1], ascending, rounded through binary32 as
 * the reference does (mod_maths.f90:18-152).  xi, w: length p+1. */
void msod_oracle_gll(int p, double *xi, double *w);

/* Exact (un-rounded) double GLL roots/weights of the same Newton iteration. */
void msod_oracle_gll_double(int p, double *xi, double *w);

/* Element tables of the reference driver (main.f90:110-255). */
void msod_oracle_tables(int p, int32_t *gmshIJK, int32_t *invAtoIJK, int32_t *AtoIJK,
                        int32_t *AtoI, int32_t *AtoJ, int32_t *AtoK, double *xgp, double *wgp,
                        double *Ngp, double *dNgp, double *dlxigp_ip);

/* He and gpvol (jacobian_oper.f90:7-136).  coord(npoin,3). */
void msod_oracle_elem_jacobian(int p, int nelem, int npoin, const int32_t *connec,
                               const double *coord, const double *dNgp, const double *wgp,
                               double *gpvol, double *He);

/* Convective residual: Rmass and Rmom as full_convec_ijk
 * (elem_convec.f90:198-431), Rener as the energy kernel of
 * full_convec_ijk_split (elem_convec.f90:674-832), the only well-defined
 * energy residual of the reference.  nthreads<=1: sequential. */
void msod_oracle_convec(int p, int nelem, int npoin, const int32_t *connec, const double *He,
                        const double *gpvol, const double *dlxigp_ip, const int32_t *invAtoIJK,
                        const int32_t *AtoI, const int32_t *AtoJ, const int32_t *AtoK,
                        const double *u, const double *q, const double *rho, const double *pr,
                        const double *E, double *Rmass, double *Rmom, double *Rener, int nthreads);

/* Diffusive residual (elem_diffu.f90:10-197). */
void msod_oracle_diffu(int p, int nelem, int npoin, const int32_t *connec, const double *He,
                       const double *gpvol, const double *dlxigp_ip, const int32_t *invAtoIJK,
                       const int32_t *AtoI, const int32_t *AtoJ, const int32_t *AtoK, double Cp,
                       double Pr, double c_rho, double c_ener, const double *rho, const double *u,
                       const double *Tem, const double *mu_fluid, const double *mu_e,
                       const double *mu_sgs, double *Rmass, double *Rmom, double *Rener,
                       int nthreads);

/* Linear tetrahedra (TET04).  Ngp(4,4) (point, node) Ngp[g + 4*n];
 * dNgp(3,4,4) dNgp[d + 3*(n + 4*g)]; He(3,3,4,nelem); gpvol(1,4,nelem);
 * connec(nelem,4); mu_e, mu_sgs(nelem,4).  Outputs are overwritten. */
/* tet_4points (quadrature_rules.f90:84-99) and tet_04 (elem_tet.f90:9-36) */
void msod_oracle_tet_tables(double *xgp, double *wgp, double *Ngp, double *dNgp);
/* tet_jacobian (jacobian_oper.f90:138-267) */
void msod_oracle_tet_jacobian(int nelem, int npoin, const int32_t *connec, const double *coord,
                              const double *dNgp, const double *wgp, double *gpvol, double *He);
/* fem_convec (elem_convec.f90:839-948) */
void msod_oracle_fem_convec(int nelem, int npoin, const int32_t *connec, const double *Ngp,
                            const double *dNgp, const double *He, const double *gpvol,
                            const double *u, const double *q, const double *rho, const double *pr,
                            const double *E, double *Rmass, double *Rmom, double *Rener);
/* fem_diffu (elem_diffu.f90:199-317) */
void msod_oracle_fem_diffu(int nelem, int npoin, const int32_t *connec, const double *Ngp,
                           const double *dNgp, const double *He, const double *gpvol, double Cp,
                           double Pr, double c_rho, double c_ener, const double *rho,
                           const double *u, const double *Tem, const double *mu_fluid,
                           const double *mu_e, const double *mu_sgs, double *Rmass, double *Rmom,
                           double *Rener);

/* Learned closure (msod_oracle_mlp.c): the MLP mu_sgs = scale*softplus(z3)
 * of a 1530-float parameter block, per sample of 12 features; quant = 1
 * rounds to bfloat16 where the GPU feeds its matrix cores. */
void msod_oracle_mlp_forward(const float *P, const float *features, int64_t n, int quant,
                             double *z3, double *mu_sgs);
/* grad[0..1504] = sum_i gnu[i] d mu_sgs_i / d P[j] */
void msod_oracle_mlp_gradient(const float *P, const float *features, const double *gnu, int64_t n,
                              int quant, double *grad);
/* one Adam step on P[0..1504] */
void msod_oracle_adam(float *P, const float *grad, float *m, float *v, int step, double lr,
                      double beta1, double beta2, double eps);

#ifdef __cplusplus
}
#endif

#endif

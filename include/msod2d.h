/*
 * msod2d.h -- C ABI of the MI355X (gfx950) spectral-element residual kernels.
 *
 * Drop-in boundary for the two timed kernels of miniSOD2D:
 *   full_convec_ijk     src/kernels/sources/elem_convec.f90:198-211
 *   full_diffusion_ijk  src/kernels/sources/elem_diffu.f90:10-23
 * The reference passes every array by reference from Fortran with OpenACC
 * device copies made once by `!$acc enter data` in the driver
 * (src/app/sources/main.f90:197-425); the kernels then run on those copies
 * (`present`).  This ABI keeps that model:
 *
 *   1. msod_data_* mirror host arrays on the GPU, keyed by host address
 *      (the `enter data` / `update` / `exit data` counterparts);
 *   2. msod_full_convec_ijk / msod_full_diffusion_ijk take the reference's
 *      argument list (porder first, since it is a compile-time constant in
 *      the reference, src/utils/sources/mod_constants.f90:16) and resolve
 *      every array pointer to its device mirror.  A pointer that is itself
 *      device memory (hipMalloc) is used as is (OpenACC `deviceptr`).
 *
 * Arrays keep the reference's Fortran layouts and 1-based index values:
 *   connec(nelem,nnode) int32; He(3,3,ngaus,nelem); gpvol(1,ngaus,nelem);
 *   dlxigp_ip(ngaus,3,porder+1); invAtoIJK(q,q,q), AtoI/J/K(nnode) int32;
 *   u,q(npoin,3); rho,pr,E,Tem,mu_fluid(npoin); mu_e,mu_sgs(nelem,ngaus);
 *   Rmass(npoin), Rmom(npoin,3), Rener(npoin)  (intent(out): zeroed inside)
 * with q = porder+1 and nnode = ngaus = q^3.  Ngp, dNgp and xgp are accepted
 * and ignored: the reference kernels never read them.
 *
 * Every function returns a status (0 = MSOD_OK); msod_last_error() gives the
 * message of the last failure.  Calls are synchronous by default (like the
 * reference's compute regions); msod_set_async(1) lets kernel calls return
 * once queued on the library's stream.
 */
#ifndef MSOD2D_H
#define MSOD2D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum msod_status {
    MSOD_OK = 0,
    MSOD_ERR_ARG = 1,         /* invalid scalar argument (sizes, porder range, NULL) */
    MSOD_ERR_NOT_PRESENT = 2, /* array neither in the data region nor device memory */
    MSOD_ERR_PORDER = 3,      /* porder outside the compiled range 1..8 */
    MSOD_ERR_TABLE = 4,       /* invAtoIJK / AtoI,J,K / dlxigp_ip inconsistent */
    MSOD_ERR_HIP = 5,         /* HIP runtime failure */
    MSOD_ERR_STATE = 6,       /* library not initialised */
    MSOD_ERR_MESH = 7,        /* connec entry outside 1..npoin or array too small */
    MSOD_ERR_MODEL = 8        /* learned-closure weights missing or not finite */
};

/* ---- library state ---------------------------------------------------- */
int msod_init(int device);
int msod_finalize(void);
const char *msod_status_string(int status);
const char *msod_last_error(void);

/* ---- data region: replaces `!$acc enter data / update / exit data` of
 *      src/app/sources/main.f90:197-425,539-542 --------------------------- */
int msod_data_copyin(const void *host, size_t bytes);
int msod_data_create(const void *host, size_t bytes);
int msod_data_update_device(const void *host, size_t bytes);
int msod_data_update_host(void *host, size_t bytes);
int msod_data_delete(const void *host);
int msod_data_present(const void *host);         /* 1 if present, else 0 */
void *msod_data_deviceptr(const void *host);     /* NULL if not present */
/* Layouts derived from an array (the assembly plan from connec, the
 * element-major He and gpvol, the TET04 mesh analysis) are cached per array
 * version.  Arrays of the data region change version with
 * msod_data_update_device, msod_data_invalidate and the library's own device
 * routines that write them; the device address of a present array
 * (msod_data_deviceptr) shares its version.  Device memory outside the data
 * region has no version: its derived layouts are rebuilt on every call (the
 * reference re-reads every array per call, elem_convec.f90:228).
 *
 * msod_data_register enters a device array the caller owns (hipMalloc, a
 * torch tensor) into the data region, keyed by its device address, so that
 * its derived layouts are cached; the caller then calls msod_data_invalidate
 * after changing its contents, and msod_data_delete to remove it (the
 * library never frees it).  No counterpart in the reference. */
int msod_data_register(const void *dev, size_t bytes);
/* Forget layouts derived from an array changed outside msod_data_* */
int msod_data_invalidate(const void *ptr);

/* ---- execution control -------------------------------------------------- */
/* How the element contributions meet at shared nodes (no counterpart in the
 * reference, which uses `!$acc atomic update`, elem_convec.f90:414-425).
 * Both modes first sum the contributions of a batch of elements per node in
 * LDS, in a fixed order.
 *   MSOD_ASSEMBLY_ATOMIC (default): one launch over all batches in mesh
 *     order, FP64 atomics of the batch sums into zeroed outputs; the fastest
 *     (neighbouring batches meet in the same L2).
 *   MSOD_ASSEMBLY_COLORED: batches coloured so that one launch per colour
 *     writes each node without atomics; the summation order is fixed and the
 *     results are bitwise reproducible from run to run. */
enum { MSOD_ASSEMBLY_COLORED = 0, MSOD_ASSEMBLY_ATOMIC = 1 };
int msod_set_assembly(int mode);
int msod_set_async(int async);
int msod_synchronize(void);
/* HIP-event timing of every kernel launch on the library stream */
int msod_timing_enable(int enable);
/* kernel: 0 zero-fill, 1 convec, 2 diffu, 3 convec mixed, 4 diffu mixed,
 *         5 diffu learned closure, 6 closure training kernels (features,
 *         predict, sensitivity, gradient, adam, loss), 7 convec geo32,
 *         8 diffu geo32, 9 geometry (msod_elem_jacobian), 10 input
 *         generator (msod_box_mesh, msod_flow_fields,
 *         msod_viscosity_fields), 11 TET convec (msod_fem_convec),
 *         12 TET diffu (msod_fem_diffu), 13 TET geometry
 *         (msod_tet_jacobian), 14 convec fp32, 15 diffu fp32, 16 convec
 *         rp=4 drop-in, 17 diffu rp=4 drop-in, 18 convec f32split, 19 diffu
 *         f32split.  Returns accumulated ms and launch count. */
int msod_timing_get(int kernel, double *total_ms, int64_t *launches);
int msod_timing_reset(void);

/* ---- geometry ------------------------------------------------------------
 * elem_jacobian (src/fem/sources/jacobian_oper.f90:7-136), which the
 * reference driver runs on the device once per mesh (main.f90:292-295):
 * He and gpvol of every element node from the node coordinates,
 *   Je(i,j) = sum_a dNgp(i,a,g) coord(connec(e,a),j),
 *   He = cofactor(Je)^T / det(Je),  gpvol = wgp(g) det(Je),
 * with the reference's operation order for det and the cofactors.
 * coord(npoin,3), dNgp(3,nnode,ngaus), wgp(ngaus); He(3,3,ngaus,nelem) and
 * gpvol(1,ngaus,nelem) are written in place (host-present or device).
 * Synchronous.  *nbad (if not NULL) receives the number of element nodes
 * with det(Je) <= 0, which the reference does not check; connec entries
 * outside 1..npoin return MSOD_ERR_MESH. */
int msod_elem_jacobian(int porder, int nelem, int npoin, const int32_t *connec, const double *coord,
                       const double *dNgp, const double *wgp, double *gpvol, double *He,
                       int64_t *nbad);

/* ---- synthetic inputs on the device ---------------------------------------
 * No counterpart in the reference, whose driver builds disjoint copies of one
 * element (geometry/sources/mod_mesh.f90:11-65) with fixed initial
 * conditions (main.f90:317-361).  These are the device forms of the
 * generator of the benchmark configurations (the host library's box mesh,
 * flow and viscosity fields): the same numbers, drawn from counter-based
 * SplitMix64 streams, so they do not depend on who generates them.
 *   msod_box_mesh: conforming nx*ny*nz box of the unit cube, order porder,
 *     npoin = (nx p+1)(ny p+1)(nz p+1), nodes on the GLL points taken from
 *     xgp(nnode,3) and gmshAtoI/J/K; element-interior nodes moved by up to
 *     `jitter` of the local 1-D spacing; connec(nelem,nnode) 1-based,
 *     elements and nodes numbered x fastest; coord(npoin,3).
 *   msod_flow_fields: u, q(npoin,3), rho, pr, E, Tem, mu_fluid(npoin) from
 *     coord (u a Taylor-Green field, rho = 1 + 0.05 N(0,1) times
 *     1 + rho_jump for x > 0.5, pr = 1/gamma + 0.1 cos 4 pi x, q = rho u,
 *     E = pr/(gamma-1) + rho|u|^2/2, Tem = pr/(rho Rgas), mu_fluid = 1e-3).
 *   msod_viscosity_fields: mu_e ~ U[0,mue_max), mu_sgs ~ U[0,musgs_max) in
 *     the reference layout (nelem,ngaus).
 * Outputs are host-present (create) or device arrays; no host traffic. */
int msod_box_mesh(int nx, int ny, int nz, int porder, double jitter, int64_t seed, const double *xgp,
                  const int32_t *gmshAtoI, const int32_t *gmshAtoJ, const int32_t *gmshAtoK,
                  int32_t *connec, double *coord);
int msod_flow_fields(int64_t npoin, const double *coord, double gamma, double Rgas, int64_t seed_rho,
                     double rho_jump, double *u, double *q, double *rho, double *pr, double *E,
                     double *Tem, double *mu_fluid);
int msod_viscosity_fields(int64_t nelem, int ngaus, int64_t seed_mue, int64_t seed_musgs,
                          double mue_max, double musgs_max, double *mu_e, double *mu_sgs);

/* ---- residual kernels --------------------------------------------------- */

/* Convective residual, FP64.  Rmass/Rmom as full_convec_ijk
 * (elem_convec.f90:198-431); Rener with the kinetic-energy flux terms
 * fuel = E*u, fukl = k*u of full_convec_ijk_split (elem_convec.f90:694-697),
 * which the fused reference kernel leaves uninitialised (:218,:248-250). */
int msod_full_convec_ijk(int porder, int nelem, int npoin, const int32_t *connec,
                         const double *Ngp, const double *dNgp, const double *He,
                         const double *gpvol, const double *dlxigp_ip, const double *xgp,
                         const int32_t *invAtoIJK, const int32_t *gmshAtoI,
                         const int32_t *gmshAtoJ, const int32_t *gmshAtoK, const double *u,
                         const double *q, const double *rho, const double *pr,
                         const double *E, double *Rmass, double *Rmom, double *Rener);

/* Diffusive residual, FP64 (elem_diffu.f90:10-197).  c_rho and c_ener are
 * module globals of the reference (mod_numerical_params.f90:26-27); set them
 * with msod_set_diffusion_coefficients (default 1, 1). */
int msod_full_diffusion_ijk(int porder, int nelem, int npoin, const int32_t *connec,
                            const double *Ngp, const double *dNgp, const double *He,
                            const double *gpvol, const double *dlxigp_ip, const double *xgp,
                            const int32_t *invAtoIJK, const int32_t *gmshAtoI,
                            const int32_t *gmshAtoJ, const int32_t *gmshAtoK, double Cp,
                            double Pr, const double *rho, const double *u, const double *Tem,
                            const double *mu_fluid, const double *mu_e, const double *mu_sgs,
                            double *Rmass, double *Rmom, double *Rener);
int msod_set_diffusion_coefficients(double c_rho, double c_ener);

/* Precision splits (no counterpart in the reference, whose single real kind
 * rp is global, mod_constants.f90:5).  Inputs stored in FP32 halve their HBM
 * bytes; the kernels compute in FP64 (on gfx950 an FP64 FMA issues at the
 * rate of an unpacked FP32 one) and assemble in FP64, so the only error is
 * the FP32 rounding of the inputs.
 *   _mixed: every input in FP32.  The convective form takes no pressure
 *     array: it recovers p = (gamma-1)(E - rho|u|^2/2) in FP64 from the FP32
 *     conservative state.
 *   _geo32: the element-sized arrays (He, gpvol; mu_e, mu_sgs) in FP32, the
 *     nodal fields in FP64.  The element arrays enter the residuals as factors
 *     and are never differenced, so their rounding is not amplified by the
 *     derivative sweeps (relative error about 3e-7 where the FP32 nodal fields
 *     of _mixed give about 1e-5 at 128^3, growing like 1/h). */
int msod_full_convec_ijk_mixed(int porder, int nelem, int npoin, const int32_t *connec,
                               const float *He, const float *gpvol, const double *dlxigp_ip,
                               const int32_t *invAtoIJK, const int32_t *gmshAtoI,
                               const int32_t *gmshAtoJ, const int32_t *gmshAtoK,
                               const float *u, const float *q, const float *rho, const float *E,
                               double gamma_gas, double *Rmass, double *Rmom, double *Rener);
int msod_full_diffusion_ijk_mixed(int porder, int nelem, int npoin, const int32_t *connec,
                                  const float *He, const float *gpvol, const double *dlxigp_ip,
                                  const int32_t *invAtoIJK, const int32_t *gmshAtoI,
                                  const int32_t *gmshAtoJ, const int32_t *gmshAtoK, double Cp,
                                  double Pr, const float *rho, const float *u, const float *Tem,
                                  const float *mu_fluid, const float *mu_e, const float *mu_sgs,
                                  double *Rmass, double *Rmom, double *Rener);

int msod_full_convec_ijk_geo32(int porder, int nelem, int npoin, const int32_t *connec,
                               const float *He, const float *gpvol, const double *dlxigp_ip,
                               const int32_t *invAtoIJK, const int32_t *gmshAtoI,
                               const int32_t *gmshAtoJ, const int32_t *gmshAtoK, const double *u,
                               const double *q, const double *rho, const double *pr,
                               const double *E, double *Rmass, double *Rmom, double *Rener);
int msod_full_diffusion_ijk_geo32(int porder, int nelem, int npoin, const int32_t *connec,
                                  const float *He, const float *gpvol, const double *dlxigp_ip,
                                  const int32_t *invAtoIJK, const int32_t *gmshAtoI,
                                  const int32_t *gmshAtoJ, const int32_t *gmshAtoK, double Cp,
                                  double Pr, const double *rho, const double *u, const double *Tem,
                                  const double *mu_fluid, const float *mu_e, const float *mu_sgs,
                                  double *Rmass, double *Rmom, double *Rener);
/*   _fp32: the inputs of _mixed, with the gathers, fluxes and derivative
 *     sweeps in FP32 arithmetic as well; the node sums, the assembly and the
 *     pressure recovery stay FP64.  The error adds FP32 rounding inside the
 *     sweeps (with their cancellation) to the rounding of the inputs; the
 *     values are reported beside the timings. */
int msod_full_convec_ijk_fp32(int porder, int nelem, int npoin, const int32_t *connec,
                              const float *He, const float *gpvol, const double *dlxigp_ip,
                              const int32_t *invAtoIJK, const int32_t *gmshAtoI,
                              const int32_t *gmshAtoJ, const int32_t *gmshAtoK, const float *u,
                              const float *q, const float *rho, const float *E, double gamma_gas,
                              double *Rmass, double *Rmom, double *Rener);
int msod_full_diffusion_ijk_fp32(int porder, int nelem, int npoin, const int32_t *connec,
                                 const float *He, const float *gpvol, const double *dlxigp_ip,
                                 const int32_t *invAtoIJK, const int32_t *gmshAtoI,
                                 const int32_t *gmshAtoJ, const int32_t *gmshAtoK, double Cp,
                                 double Pr, const float *rho, const float *u, const float *Tem,
                                 const float *mu_fluid, const float *mu_e, const float *mu_sgs,
                                 double *Rmass, double *Rmom, double *Rener);

/*   _f32split: the FP32 split at the accuracy of the FP64 form's inputs:
 *     every input in FP32 except the velocity u and the pressure pr, which
 *     stay FP64; FP64 arithmetic, node sums and outputs.  The choice is
 *     measured field by field: rounding u to FP32 alone puts the diffusive
 *     momentum residual at 2e-5 (u is differentiated twice, the error grows
 *     like 1/h), rounding pr alone the convective momentum residual at 1e-6;
 *     each other field costs at most 6e-7 of relative L2 on any residual.
 *     Canonical bytes per pair: 1.57x fewer than FP64. */
int msod_full_convec_ijk_f32split(int porder, int nelem, int npoin, const int32_t *connec, const float *He,
                                  const float *gpvol, const double *dlxigp_ip, const int32_t *invAtoIJK,
                                  const int32_t *gmshAtoI, const int32_t *gmshAtoJ, const int32_t *gmshAtoK,
                                  const double *u, const float *q, const float *rho, const double *pr,
                                  const float *E, double *Rmass, double *Rmom, double *Rener);
int msod_full_diffusion_ijk_f32split(int porder, int nelem, int npoin, const int32_t *connec, const float *He,
                                     const float *gpvol, const double *dlxigp_ip, const int32_t *invAtoIJK,
                                     const int32_t *gmshAtoI, const int32_t *gmshAtoJ, const int32_t *gmshAtoK,
                                     double Cp, double Pr, const float *rho, const double *u, const float *Tem,
                                     const float *mu_fluid, const float *mu_e, const float *mu_sgs,
                                     double *Rmass, double *Rmom, double *Rener);

/* The drop-in for a caller built at the reference's default real kind
 * (rp = 4, src/utils/sources/mod_constants.f90:5): full_convec_ijk and
 * full_diffusion_ijk with the reference's full argument lists
 * (elem_convec.f90:198-211, elem_diffu.f90:10-23) in float, Cp and Pr
 * included, and float outputs.  The kernels compute in FP64 from the float
 * inputs and sum each batch's contributions to a node in FP64; each batch's
 * sum is rounded once to float and added with a float atomic (a node shared
 * by k batches takes k roundings), so the errors are the float storage of
 * inputs and outputs and those few additions (the reference at rp = 4
 * rounds inside every product and sum). */
int msod_full_convec_ijk_r4(int porder, int nelem, int npoin, const int32_t *connec, const float *Ngp,
                            const float *dNgp, const float *He, const float *gpvol,
                            const float *dlxigp_ip, const float *xgp, const int32_t *invAtoIJK,
                            const int32_t *gmshAtoI, const int32_t *gmshAtoJ, const int32_t *gmshAtoK,
                            const float *u, const float *q, const float *rho, const float *pr,
                            const float *E, float *Rmass, float *Rmom, float *Rener);
int msod_full_diffusion_ijk_r4(int porder, int nelem, int npoin, const int32_t *connec, const float *Ngp,
                               const float *dNgp, const float *He, const float *gpvol,
                               const float *dlxigp_ip, const float *xgp, const int32_t *invAtoIJK,
                               const int32_t *gmshAtoI, const int32_t *gmshAtoJ, const int32_t *gmshAtoK,
                               float Cp, float Pr, const float *rho, const float *u, const float *Tem,
                               const float *mu_fluid, const float *mu_e, const float *mu_sgs,
                               float *Rmass, float *Rmom, float *Rener);

/* ---- learned subgrid-viscosity closure ----------------------------------
 * No counterpart in the reference: its driver sets mu_sgs = 1
 * (main.f90:334) and its LES flags are unused (mod_numerical_params.f90:13-17).
 * A per-node MLP (12 features -> 32 -> 32 -> 1, SiLU, softplus) predicts
 * mu_sgs from the velocity gradient, rho, |u| and the element size; weights
 * are a block of MSOD_CLOSURE_NPARAM floats (layout in
 * minisod2d_amd/csrc/msod_mlp.h), host or device memory. */
#define MSOD_CLOSURE_NPARAM 1530
#define MSOD_CLOSURE_NTRAIN 1505
#define MSOD_CLOSURE_NFEAT 12

/* full_diffusion_ijk with mu_sgs predicted inside the kernel (bf16 MFMA).
 * Missing (NULL) or non-finite weights fall back to mu_sgs = 0; *used_model
 * (if not NULL) tells which was used. */
int msod_full_diffusion_ijk_closure(int porder, int nelem, int npoin, const int32_t *connec,
                                    const double *Ngp, const double *dNgp, const double *He,
                                    const double *gpvol, const double *dlxigp_ip,
                                    const double *xgp, const int32_t *invAtoIJK,
                                    const int32_t *gmshAtoI, const int32_t *gmshAtoJ,
                                    const int32_t *gmshAtoK, double Cp, double Pr,
                                    const double *rho, const double *u, const double *Tem,
                                    const double *mu_fluid, const double *mu_e,
                                    const float *mlp_params, int *used_model, double *Rmass,
                                    double *Rmom, double *Rener);

/* Training building blocks (all arrays host-present or device pointers).
 * features: [nelem*nnode][12], element-major, nodes in lexicographic SOD
 * order l = (I-1) + q((J-1) + q(K-1)); gnu has the same layout. */
int msod_closure_features(int porder, int nelem, int npoin, const int32_t *connec,
                          const double *He, const double *gpvol, const double *dlxigp_ip,
                          const int32_t *invAtoIJK, const int32_t *gmshAtoI,
                          const int32_t *gmshAtoJ, const int32_t *gmshAtoK, const double *rho,
                          const double *u, float *features);
/* mu_sgs(nelem,ngaus) (reference layout) from the MLP */
int msod_closure_predict(int porder, int nelem, const int32_t *invAtoIJK, const float *mlp_params,
                         const float *features, double *mu_sgs);
/* gnu(e,a) = d <rerr, (Rmom, Rener)> / d mu_sgs(e,a) of full_diffusion_ijk;
 * rerr: [4][npoin] (momentum x, y, z, energy) */
int msod_closure_sensitivity(int porder, int nelem, int npoin, const int32_t *connec,
                             const double *He, const double *gpvol, const double *dlxigp_ip,
                             const int32_t *invAtoIJK, const int32_t *gmshAtoI,
                             const int32_t *gmshAtoJ, const int32_t *gmshAtoK, double Cp,
                             const double *rho, const double *u, const double *Tem,
                             const double *rerr, float *gnu);
/* grad[MSOD_CLOSURE_NTRAIN] = sum_samples gnu * d mu_sgs / d params */
int msod_closure_gradient(int64_t nsamples, const float *mlp_params, const float *features,
                          const float *gnu, float *grad);
/* msod_closure_gradient that allocates no device memory.  No counterpart in
 * the reference; it is for callers that cannot allocate (code under HIP graph
 * capture, a tight training loop, harnesses that forbid hipMalloc).  The
 * per-workgroup partial sums go to a fixed library array of 1024 x 1505
 * floats (6.2 MB per GPU, held until the process ends) instead of a library
 * buffer grown on demand; the workgroups stride over the samples, so any
 * nsamples fits.  mlp_params must be present or device memory (a host block
 * would need a device copy).  Same kernels and results as
 * msod_closure_gradient.  Every call on one GPU shares that array: calls that
 * overlap on one GPU (two library contexts on one device, each with its own
 * stream) race. */
int msod_closure_gradient_noalloc(int64_t nsamples, const float *mlp_params, const float *features,
                                  const float *gnu, float *grad);
/* one Adam step on the trained part of mlp_params (m, v: moment buffers) */
int msod_closure_adam(float *mlp_params, const float *grad, float *m, float *v, int step,
                      double lr, double beta1, double beta2, double eps);
/* rerr = 2 w_c (R_c - target_c), *loss = sum w_c (R_c - target_c)^2 over the
 * momentum and energy residuals; target, rerr: [4][npoin]; weights[4] host */
int msod_closure_loss(int npoin, const double *Rmom, const double *Rener, const double *target,
                      const double *weights, double *rerr, double *loss);

/* ---- linear tetrahedra (TET04) ---------------------------------------------
 * The reference driver times two P1-tetrahedron kernels beside the hex ones
 * (main.f90:464-491), on a disjoint mesh of npoin/4 tetrahedra built over the
 * same nodal fields (main.f90:367-407).  Layouts: connec(nelem,4) int32,
 * 1-based; Ngp(4,4) (point, node); dNgp(3,4,4); wgp(4); coord(npoin,3);
 * He(3,3,4,nelem); gpvol(1,4,nelem); mu_e, mu_sgs(nelem,4).  The element
 * contributions are added with FP64 atomics into zeroed outputs, which is
 * correct on any mesh (the reference's fem_diffu adds without atomics,
 * elem_diffu.f90:306-312, relying on the mesh being disjoint).  A connec entry
 * outside 1..npoin returns MSOD_ERR_MESH (checked on the device once per
 * array version).
 *
 * tet_jacobian (src/fem/sources/jacobian_oper.f90:138-267): He and gpvol from
 * coord, in the reference's operation order; *nbad = number of points with
 * det(Je) <= 0. */
int msod_tet_jacobian(int nelem, int npoin, const int32_t *connec, const double *coord,
                      const double *dNgp, const double *wgp, double *gpvol, double *He,
                      int64_t *nbad);
/* fem_convec (src/kernels/sources/elem_convec.f90:839-948) */
int msod_fem_convec(int nelem, int npoin, const int32_t *connec, const double *Ngp,
                    const double *dNgp, const double *He, const double *gpvol, const double *u,
                    const double *q, const double *rho, const double *pr, const double *E,
                    double *Rmass, double *Rmom, double *Rener);
/* fem_diffu (src/kernels/sources/elem_diffu.f90:199-317); c_rho and c_ener as
 * set by msod_set_diffusion_coefficients.  Like the reference, the fluid
 * properties at Gauss point g weight node n by Ngp(n,g). */
int msod_fem_diffu(int nelem, int npoin, const int32_t *connec, const double *Ngp,
                   const double *dNgp, const double *He, const double *gpvol, double Cp, double Pr,
                   const double *rho, const double *u, const double *Tem, const double *mu_fluid,
                   const double *mu_e, const double *mu_sgs, double *Rmass, double *Rmom,
                   double *Rener);

#ifdef __cplusplus
}
#endif

#endif /* MSOD2D_H */

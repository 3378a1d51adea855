// Linear tetrahedra (TET04): arguments of the kernels of msod_tet.hip, the
// counterparts of fem_convec (src/kernels/sources/elem_convec.f90:839-948),
// fem_diffu (src/kernels/sources/elem_diffu.f90:199-317) and tet_jacobian
// (src/fem/sources/jacobian_oper.f90:138-267).  All pointers are device
// pointers to arrays in the reference's Fortran layouts:
//   connec(nelem,4) int32, 1-based;  He(3,3,4,nelem);  gpvol(1,4,nelem)
//   u,q(npoin,3); rho,pr,E,Tem,mu_fluid(npoin); mu_e,mu_sgs(nelem,4)
//   Rmass(npoin), Rmom(npoin,3), Rener(npoin)
// The element tables travel in the kernel arguments.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace msod {

struct TetTables {
    double Ngp[16];   // Ngp(4,4): Ngp[i + 4j] = Ngp(i+1, j+1)
    double dNgp[48];  // dNgp(3,4,4): dNgp[d + 3(a + 4g)]
    double wgp[4];
};

struct TetConvecArgs {
    int64_t nelem, npoin;
    const int32_t *connec;
    const double *He, *gpvol;
    const double *u, *q, *rho, *pr, *E;
    double *Rmass, *Rmom, *Rener;
    int store;  // 1: every node has at most one element node (plain stores)
    TetTables t;
};

struct TetDiffuArgs {
    int64_t nelem, npoin;
    const int32_t *connec;
    const double *He, *gpvol;
    const double *rho, *u, *Tem, *mu_fluid, *mu_e, *mu_sgs;
    double Cp, Pr, c_rho, c_ener;
    double *Rmass, *Rmom, *Rener;
    int store;
    TetTables t;
};

struct TetJacobianArgs {
    int64_t nelem, npoin;
    const int32_t *connec;
    const double *coord;  // (npoin,3)
    double *He, *gpvol;
    unsigned long long *counts;  // [0] connec entries outside 1..npoin, [1] det(Je) <= 0
    TetTables t;
};

// Element residuals reach the nodes in one of two ways (store):
//   0: FP64 atomics into zero-filled outputs (any mesh);
//   1: plain stores, for a mesh where no node belongs to two element nodes
//      (the reference's disjoint TET04 mesh): the value stored is the one an
//      atomic add onto zero would leave.  Nodes no element references must
//      be zeroed by the caller.
hipError_t launch_fem_convec(const TetConvecArgs &a, hipStream_t s);
hipError_t launch_fem_diffu(const TetDiffuArgs &a, hipStream_t s);
hipError_t launch_tet_jacobian(const TetJacobianArgs &a, hipStream_t s);

// Mesh analysis, once per connectivity: refs[npoin] (zeroed by the caller)
// counts the element nodes of every node; stats[0] += entries outside
// 1..npoin (not counted), stats[1] += nodes with refs > 1, stats[2] += nodes
// with refs == 0; list_unreferenced writes those nodes (0-based, any order)
// to list[0 .. stats[2]) using *cursor (zeroed by the caller).
hipError_t launch_count_refs(const int32_t *connec, int64_t n, int64_t npoin, uint32_t *refs,
                             unsigned long long *stats, hipStream_t s);
hipError_t launch_ref_summary(const uint32_t *refs, int64_t npoin, unsigned long long *stats,
                              hipStream_t s);
hipError_t launch_list_unreferenced(const uint32_t *refs, int64_t npoin, uint32_t *list,
                                    unsigned long long *cursor, hipStream_t s);

}  // namespace msod

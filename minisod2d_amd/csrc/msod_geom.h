// Device geometry: arguments of the elem_jacobian kernel (msod_geom.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace msod {

struct GeomArgs {
    const int32_t *connec;   // (nelem,nnode) reference layout, 1-based
    const double *coord;     // (npoin,3)
    const double *wgp;       // [ngaus]
    // nonzero entries of dNgp(d,:,g): nza/nzw[(g*3 + d)*mz + k], k < nzn[g*3 + d]
    const uint16_t *nza;     // local node a (0-based)
    const double *nzw;       // dNgp(d,a,g)
    const uint16_t *nzn;     // count
    int64_t nelem, npoin;
    int mz;
    double *He;              // He(3,3,ngaus,nelem)
    double *gpvol;           // gpvol(1,ngaus,nelem)
    unsigned long long *counts;  // [0] connec entries outside 1..npoin, [1] det(Je) <= 0
};

hipError_t launch_elem_jacobian(int porder, const GeomArgs &a, hipStream_t s);

}  // namespace msod

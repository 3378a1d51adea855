// Device generator of the synthetic inputs: kernel arguments (msod_gen.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace msod {

constexpr int kGenMaxP = 8;

struct BoxArgs {
    int nx, ny, nz, p, nn;
    int64_t nelem, npoin, n1x, n1y;
    double h[3];
    double xi[kGenMaxP + 1];    // ascending 1-D GLL points (binary32-rounded)
    double hloc[kGenMaxP + 1];  // smallest neighbour gap of each interior position
    const int8_t *pos;          // [nnode][3] lattice position (0..p) of node a, per axis
    double jitter;
    uint64_t seed;
    int32_t *connec;  // (nelem,nnode), 1-based
    double *coord;    // (npoin,3)
};

struct FlowArgs {
    int64_t npoin;
    const double *coord;
    double gamma, Rgas, rho_jump;
    uint64_t seed_rho;
    double *u, *q, *rho, *pr, *E, *Tem, *mu_fluid;
};

struct ViscArgs {
    int64_t nelem;
    int ngaus;
    uint64_t seed_mue, seed_musgs;
    double mue_max, musgs_max;
    double *mu_e, *mu_sgs;  // (nelem,ngaus)
};

hipError_t launch_box_mesh(const BoxArgs &b, hipStream_t s);
hipError_t launch_flow_fields(const FlowArgs &f, hipStream_t s);
hipError_t launch_viscosity_fields(const ViscArgs &v, hipStream_t s);

}  // namespace msod

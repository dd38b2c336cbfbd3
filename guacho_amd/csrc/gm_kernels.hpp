// gm_kernels.hpp -- launch interface of the step kernels (gm_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "gm_physics.hpp"

namespace gm {

// Grid geometry of one subdomain; bc[] are the six boundary kinds
// (left, right, bottom, top, out, in).
struct Geom {
  int nx, ny, nz;
  int NX, NY, NZ;     // extents with two ghost layers: n + 4
  long long ncell;    // NX * NY * NZ
  int bc[6];
};

void launch_prim(hipStream_t s, const Geom& g, const Phys& ph, const double* U, double* P, int lo);
void launch_flux(hipStream_t s, const Geom& g, const Phys& ph, int solver, int order, int axis,
                 const double* P, double* F, int* err);
void launch_efield(hipStream_t s, const Geom& g, const double* Fx, const double* Fy, const double* Fz,
                   double* E);
void launch_update(hipStream_t s, const Geom& g, const double* Uin, double* Uout, const double* Fx,
                   const double* Fy, const double* Fz, const double* E, double dt, double dx, double dy,
                   double dz, int flux_cd);
void launch_cfl(hipStream_t s, const Geom& g, const Phys& ph, const double* U, double dx, double dy, double dz,
                unsigned long long* dtmin, int* err);
void launch_ghost1(hipStream_t s, const Geom& g, double* U, int axis, int side, int kind);
void launch_aos_to_soa(hipStream_t s, long long ncell, const double* A, double* S);
void launch_soa_to_aos(hipStream_t s, long long ncell, int nvar, const double* S, double* A);
void launch_prim_all(hipStream_t s, long long ncell, const Phys& ph, const double* U, double* Paos, double* T);
void launch_divb(hipStream_t s, const Geom& g, const double* U, double dx, double dy, double dz, double* D);

}  // namespace gm

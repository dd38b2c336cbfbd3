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

// Launchers, instantiated for S = double (the FP64 drop-in) and S = float
// (the mixed form's storage).  Kernels compute primitives, the CFL
// candidates and the conserved update in FP64 for either storage type.
template <typename S>
void launch_prim(hipStream_t s, const Geom& g, const Phys& ph, const S* U, S* P, int lo);
template <typename S>
void launch_flux(hipStream_t s, const Geom& g, const Phys& ph, int solver, int order, int axis, const S* P,
                 S* F, int* err);
template <typename S>
void launch_efield(hipStream_t s, const Geom& g, const S* Fx, const S* Fy, const S* Fz, S* E);
template <typename S>
void launch_update(hipStream_t s, const Geom& g, const S* Uin, S* Uout, const S* Fx, const S* Fy, const S* Fz,
                   const S* E, double dt, double dx, double dy, double dz, int flux_cd);
template <typename S>
void launch_cfl(hipStream_t s, const Geom& g, const Phys& ph, const S* U, double dx, double dy, double dz,
                unsigned long long* dtmin, int* err);
template <typename S>
void launch_ghost1(hipStream_t s, const Geom& g, S* U, int axis, int side, int kind);
template <typename S>
void launch_aos_to_soa(hipStream_t s, long long ncell, const double* A, S* D);
template <typename S>
void launch_soa_to_aos(hipStream_t s, long long ncell, const S* D, double* A);
template <typename S>
void launch_prim_all(hipStream_t s, long long ncell, const Phys& ph, const S* U, double* Paos, double* T);
template <typename S>
void launch_divb(hipStream_t s, const Geom& g, const S* U, double dx, double dy, double dz, double* D);

}  // namespace gm

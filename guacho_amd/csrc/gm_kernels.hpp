// gm_kernels.hpp -- launch interface of the step kernels (gm_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "gm_physics.hpp"
#include "guacho_amd.h"

namespace gm {

// Grid geometry of one subdomain; bc[] are the six boundary kinds
// (left, right, bottom, top, out, in).
struct Geom {
  int nx, ny, nz;
  int NX, NY, NZ;     // extents with two ghost layers: n + 4
  long long ncell;    // NX * NY * NZ
  int bc[6];
};

// One box (subdomain) as the kernels see it: geometry, physics and its
// device buffers.  F holds the three axes' fluxes back to back (axis a at
// F + a*8*ncell).
template <typename S>
struct Box {
  Geom g;
  Phys ph;
  S* U;   // U^n, updated in place by the corrector
  S* Us;  // U* (predictor result)
  S* P;   // primitives
  S* F;   // fluxes, 3 x 8 components
  S* E;   // cell-centred E, 3 components
  double dx, dy, dz;
  int slot;  // index of the box's clock in a batch's clock table
};

// Per-box clock of a batch: the reference's time-loop state (time,
// currentIteration) plus the CFL minimum of the current step and the stop
// condition (t_end reached by the dump clamp, or max_steps taken).
struct BoxClock {
  unsigned long long dtmin;  // bit pattern of the CFL minimum (positive double)
  double dt;                 // dt of the current step (0 when done)
  double time;
  double t_end;
  double cfl;
  int it;                    // currentIteration
  int n_iter;                // CFL ramp length (src/main.f90:97: 10)
  int steps;                 // steps taken
  int max_steps;             // 0: no step limit
  int done;
  int pad;
};

// Launch shape: threads per tile (64, 128 or 256) and z-pencil length of
// the column-marching kernels.
struct Shape {
  int block;
  int kchunk;
};

// Validates a subdomain configuration; GM_OK or a GM_ERR_* code with msg.
int check_params(const gm_params* p, std::string& msg);

// Kernel classes of one step, for per-class tile tables.
enum KClass { KC_PRIM1 = 0, KC_PRIM2, KC_FLUXX, KC_FLUXY, KC_FLUXZ, KC_EFIELD, KC_UPDATE, KC_CFLRED, KC_COUNT };

// Tiles (workgroups) one box needs for a kernel class under a shape.
unsigned box_tiles(const Geom& g, int kclass, const Shape& sh);

// Launchers, instantiated for S = double (the FP64 drop-in) and S = float
// (the mixed form's storage).  Kernels compute primitives, the CFL
// candidates and the conserved update in FP64 for either storage type.
template <typename S>
void launch_prim(hipStream_t s, const Box<S>& b, int stage, const Shape& sh);
template <typename S>
void launch_flux(hipStream_t s, const Box<S>& b, int solver, int order, int axis, int* err, const Shape& sh);
template <typename S>
void launch_efield(hipStream_t s, const Box<S>& b, const Shape& sh);
template <typename S>
void launch_update(hipStream_t s, const Box<S>& b, double dt, int stage, int flux_cd, const Shape& sh);
template <typename S>
void launch_cfl(hipStream_t s, const Box<S>& b, unsigned long long* dtmin, int* err, const Shape& sh);

// Batched launchers: `tiles` is the total tile count of the class over all
// boxes and `prefix` (device, nbox+1 entries) its prefix sums.
template <typename S>
void launch_prim_batch(hipStream_t s, const Box<S>* boxes, const unsigned* prefix, int nbox, unsigned tiles,
                       int stage, const Shape& sh);
template <typename S>
void launch_flux_batch(hipStream_t s, const Box<S>* boxes, const unsigned* prefix, int nbox, unsigned tiles,
                       int solver, int order, int axis, int* err, const Shape& sh);
template <typename S>
void launch_efield_batch(hipStream_t s, const Box<S>* boxes, const unsigned* prefix, int nbox, unsigned tiles,
                         const Shape& sh);
template <typename S>
void launch_update_batch(hipStream_t s, const Box<S>* boxes, const unsigned* prefix, int nbox, unsigned tiles,
                         const BoxClock* clk, int stage, int flux_cd, const Shape& sh);
template <typename S>
void launch_cfl_batch(hipStream_t s, const Box<S>* boxes, const unsigned* prefix, int nbox, unsigned tiles,
                      BoxClock* clk, int* err, const Shape& sh);
void launch_clock_begin(hipStream_t s, BoxClock* clk, int nbox);
void launch_clock_dt(hipStream_t s, BoxClock* clk, int nbox, int tuning);

// Layout conversion, ghost fill and diagnostics of one box's state.
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

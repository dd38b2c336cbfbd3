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
  int gmem;           // 1: ghosts are materialised in memory (user boundary
                      // hooks); 0: resolved by boundary mapping
  int npas;           // passive scalars after the 8 MHD variables (0..2):
                      // variable q of a cell at q*ncell, q < 8 + npas
};

// One box (subdomain) as the kernels see it: geometry, physics and its
// device buffers.  W carries a stage's partial results from the xy-face
// kernel to the z-face kernel, in FP64 for either storage type: the
// conserved update after its x and y terms (components 0..4 with flux-CD,
// 0..7 without), with flux-CD the x/y parts of E1 (component 5) and E2
// (component 6), and the passive scalars' partials (components 8..).
template <typename S>
struct Box {
  Geom g;
  Phys ph;
  S* U;       // U^n, updated in place by the corrector
  S* Us;      // U* (predictor result)
  double* W;  // stage partials, (8 + npas) x ncell doubles
  S* E;       // cell-centred E, 3 components
  double dx, dy, dz;
  int slot;  // index of the box's clock in a batch's clock table
  int src;   // source terms: SRC_EIGHT_WAVE | SRC_USER (W holds the user's s)
  int srcin;  // 1: the stage kernels add the source terms as they store each
              // cell (no passive scalars); else the source pass does
  const double* usrc;  // with srcin: the user's s of the stage (neq x ncell), or null
  // The pressure of each interior cell as the reference's primit holds it
  // after table cooling: coolingdmc / coolingchi set primit(5) of every
  // interior cell (cooled where T > 1e4 K) and tstep then recomputes only
  // the ghost primitives (hydro_solver.f90:218-223), so the next
  // get_timestep and predictor read that pressure rather than one recovered
  // from the cooled E.  When set, the first stage's primitive recovery and
  // the CFL minimum take an interior cell's pressure from here (index as
  // U's); the table-cooling kernel writes it.  Null otherwise.
  double* pc = nullptr;
  // artificial viscosity of viscous_copy (src/hydro_solver.f90:47-65) for
  // batched boxes (launch_viscous_batch); single boxes pass it to
  // launch_viscous
  double eta = 0.0;
};

enum : int { SRC_EIGHT_WAVE = 1, SRC_USER = 2 };

// Table cooling after the step (coolingdmc / coolingchi): the reference's
// cooltab(2, 41) (T_i, Lambda_i pairs), the sqrt-law coefficient above
// 1e8 K and the parameters module's Tempsc, Psc, tsc.
struct Cooling {
  int kind;            // 1 (H-rate), 3 (DMC) or 4 (CHIANTI)
  double tempsc, psc, tsc, hicoef;
  double tab[82];
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

// Launch shape of the column-marching kernels (z faces, field update, CFL):
// threads per tile (64, 128 or 256) and z-pencil length.  The xy-face
// kernel has a fixed 32 x 16 plane tile.
struct Shape {
  int block;
  int kchunk;
  int slab;        // planes per slab of the 2.5-D stage kernel; 0: the two-kernel stage
  int slab_both;   // batches: 1 when the corrector runs as the slab kernel too
};

// Validates a subdomain configuration; GM_OK or a GM_ERR_* code with msg.
int check_params(const gm_params* p, std::string& msg);

// Tile classes of one step, for per-class tile tables: 32 x 16 plane tiles
// of the xy-face kernel, column pencils of the z-face, field-update and CFL
// kernels, and 32 x 8 column tiles of the slab kernel (Shape::slab planes).
enum KClass { KC_XY = 0, KC_COL, KC_SLAB, KC_COUNT };

// Tiles (workgroups) one box needs for a kernel class under a shape.
unsigned box_tiles(const Geom& g, int kclass, const Shape& sh);

// Launchers, instantiated for S = double (the FP64 drop-in) and S = float
// (the mixed form's storage).  Kernels compute primitives, the CFL
// candidates and the conserved update in FP64 for either storage type.
// A stage is one call of each of launch_xy, launch_z and (with flux-CD)
// launch_bfield; stage 1 is the predictor (first order, dt/2 passed in
// `dt`), stage 2 the corrector.
template <typename S>
void launch_xy(hipStream_t s, const Box<S>& b, double dt, int stage, int solver, int flux_cd, int* err);
template <typename S>
void launch_z(hipStream_t s, const Box<S>& b, double dt, int stage, int solver, int flux_cd, int* err,
              const Shape& sh);
// 2.5-D form of one stage: launch_xy and launch_z in one kernel that marches
// 32 x 8 column tiles up slabs of `kchunk` planes (no W round trip); same
// results bit for bit.  Not for boxes whose stage kernels add source terms
// (Box::srcin).
template <typename S>
void launch_slab(hipStream_t s, const Box<S>& b, double dt, int stage, int solver, int flux_cd, int* err,
                 int kchunk);
// default slab length for a grid, and the workgroups of a slab launch
int slab_kchunk(const Geom& g, int resident);
unsigned slab_tiles(const Geom& g, int kchunk);
template <typename S>
void launch_bfield(hipStream_t s, const Box<S>& b, double dt, int stage, const Shape& sh);
template <typename S>
void launch_cfl(hipStream_t s, const Box<S>& b, unsigned long long* dtmin, int* err, const Shape& sh);
// the corrector's field update (stage 2 of launch_bfield) plus launch_cfl's
// reduction over the updated state; *dtmin must hold 1e30 beforehand
template <typename S>
void launch_bfield_cfl(hipStream_t s, const Box<S>& b, double dt, unsigned long long* dtmin, int* err,
                       const Shape& sh);
// cooling pass on U^{n+1} (src/hydro_solver.f90:206-210) with the step's dt
// viscous_copy with eta != 0 (W as scratch), after the corrector
template <typename S>
void launch_viscous(hipStream_t s, const Box<S>& b, double eta);
template <typename S>
void launch_cooling(hipStream_t s, const Box<S>& b, const Cooling& cl, double dt);

// Thermal conduction (thermal_cond.f90): kind 1 isotropic, 2 anisotropic;
// the parameters module's scalings.  Ghosts live in memory while it is on.
struct Conduction {
  int kind, sat;
  double tempsc, rsc, rhosc, mu, psc, tsc, vsc2, bsc;
};
// min over interior cells of rho / Ksp(T), as the bit pattern of a positive
// double atomically reduced into *out (get_dt_cond's loop, :78-110)
template <typename S>
void launch_tc_dtmin(hipStream_t s, const Box<S>& b, const Conduction& tc, unsigned long long* out);
// one sub-step: heat fluxes of cells 0..n (into W), the E update of the
// interior with dts, thermal_bounds on E
template <typename S>
void launch_tc_substep(hipStream_t s, const Box<S>& b, const Conduction& tc, double dts);
// source terms of step() (src/hydro_solver.f90:115-121): up += dt * s after
// the stage's update, s from the stage's input state
template <typename S>
void launch_sources(hipStream_t s, const Box<S>& b, double dt, int stage, const Shape& sh);

// Batched launchers: `tiles` is the total tile count of the class over all
// boxes and `prefix` (device, nbox+1 entries) its prefix sums.  Each box's
// dt comes from its clock.  `npas`: the passive scalars of every box of the
// launch (0..2).
template <typename S>
void launch_xy_batch(hipStream_t s, const Box<S>* boxes, const unsigned* prefix, int nbox, unsigned tiles,
                     const BoxClock* clk, int stage, int solver, int flux_cd, int* err, int npas);
template <typename S>
void launch_z_batch(hipStream_t s, const Box<S>* boxes, const unsigned* prefix, int nbox, unsigned tiles,
                    const BoxClock* clk, int stage, int solver, int flux_cd, int* err, const Shape& sh, int npas);
// the 2.5-D stage kernel over a batch (tiles of class KC_SLAB, sh.slab planes)
template <typename S>
void launch_slab_batch(hipStream_t s, const Box<S>* boxes, const unsigned* prefix, int nbox, unsigned tiles,
                       const BoxClock* clk, int stage, int solver, int flux_cd, int* err, const Shape& sh, int npas);
template <typename S>
void launch_bfield_batch(hipStream_t s, const Box<S>* boxes, const unsigned* prefix, int nbox, unsigned tiles,
                         const BoxClock* clk, int stage, const Shape& sh);
template <typename S>
void launch_cfl_batch(hipStream_t s, const Box<S>* boxes, const unsigned* prefix, int nbox, unsigned tiles,
                      BoxClock* clk, int* err, const Shape& sh, int npas);
// the corrector's field update with each box's next CFL minimum folded into
// its clock (valid when no source pass follows the field update)
template <typename S>
void launch_bfield_cfl_batch(hipStream_t s, const Box<S>* boxes, const unsigned* prefix, int nbox, unsigned tiles,
                             BoxClock* clk, int* err, const Shape& sh, int npas);
template <typename S>
void launch_sources_batch(hipStream_t s, const Box<S>* boxes, const unsigned* prefix, int nbox, unsigned tiles,
                          const BoxClock* clk, int stage, const Shape& sh);
// viscous_copy after the corrector for the boxes with Box::eta != 0 whose
// step has dt > 0 (tiles of class KC_COL)
template <typename S>
void launch_viscous_batch(hipStream_t s, const Box<S>* boxes, const unsigned* prefix, int nbox, unsigned tiles,
                          const BoxClock* clk, const Shape& sh);
// clock kernels over the boxes of a launch: slots[0..nbox) (device) are
// their clock slots
void launch_clock_begin(hipStream_t s, BoxClock* clk, const int* slots, int nbox);
void launch_clock_dt(hipStream_t s, BoxClock* clk, const int* slots, int nbox, int tuning);

// Layout conversion, ghost fill and diagnostics of one box's state.
// layers 1: one face of boundaryI (src/boundaries.f90:44-247); layers 2:
// one face of boundaryII (:256-512), two mirrored layers, full transverse range
template <typename S>
void launch_ghost(hipStream_t s, const Geom& g, S* U, int axis, int side, int kind, int layers);
template <typename S>
void launch_aos_to_soa(hipStream_t s, long long ncell, const double* A, S* D, int nq = 8);
template <typename S>
void launch_soa_to_aos(hipStream_t s, long long ncell, const S* D, double* A, int nq = 8);
// pc (Box::pc): interior cells take that pressure; null: every cell's from U
template <typename S>
void launch_prim_all(hipStream_t s, long long ncell, const Phys& ph, const S* U, double* Paos, double* T, int nq,
                     const Geom& g, const double* pc);
template <typename S>
void launch_divb(hipStream_t s, const Geom& g, const S* U, double dx, double dy, double dz, double* D);

}  // namespace gm

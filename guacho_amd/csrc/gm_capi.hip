// gm_capi.hip -- the C ABI of include/guacho_amd.h over the step kernels.
//
// A context owns one subdomain's device state and one HIP stream.  The
// step (gm_tstep) follows tstep (src/hydro_solver.f90:134-229):
//   predictor: first-order fluxes of U^n, U* = U^n - dt/2 div F
//   corrector: limited fluxes of U*,      U^{n+1} = U^n - dt div F
// each stage as three kernels (xy faces, z faces, flux-CD field update)
// with ghosts resolved by boundary mapping inside the kernels, so neither
// boundaryII nor boundaryI runs on the step path; gm_download_u and
// gm_boundary_u materialise boundaryI's ghost layer when a caller needs it.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "gm_kernels.hpp"
#include "guacho_amd.h"
#include "guacho_amd_user.h"

namespace {

const char* kClassNames[GM_NKCLASS] = {"xy_predictor", "xy_corrector", "z_predictor", "z_corrector",
                                       "bfield", "cfl_dt", "ghost", "layout", "sources", "cooling",
                                       "conduction", "viscous", "bfield_cfl", "slab_predictor",
                                       "slab_corrector"};
enum { KC_XY1, KC_XY2, KC_Z1, KC_Z2, KC_BFIELD, KC_CFL, KC_GHOST, KC_LAYOUT, KC_SOURCES, KC_COOLING, KC_CONDUCTION, KC_VISCOUS,
       KC_BFIELD_CFL, KC_SLAB1, KC_SLAB2 };

thread_local std::string g_create_error;

// After table cooling the reference refreshes only the ghost primitives
// (src/hydro_solver.f90:218-223), so its conduction starts from the
// temperature of the half-step state in the interior; that combination is
// not built.  H-rate cooling is followed by a full calcprim and combines.
const char* const kTableCoolCond =
    "table cooling (GM_COOL_DMC / GM_COOL_CHI) with thermal conduction is not built: the reference's conduction "
    "then starts from the half-step temperature of the interior cells";

struct Scalars {
  unsigned long long dtmin;
  unsigned long long tcmin;  // conduction: min rho / Ksp(T) (bits of a positive double)
  int err;
  int pad;
};

__global__ void k_init_dtmin(Scalars* s) {
  if (threadIdx.x == 0 && blockIdx.x == 0) s->dtmin = (unsigned long long)__double_as_longlong(1e30);
}

__global__ void k_init_tcmin(Scalars* s) {
  if (threadIdx.x == 0 && blockIdx.x == 0) s->tcmin = (unsigned long long)__double_as_longlong(1.79769313486231570815e+308);
}

__global__ void k_clear_err(Scalars* s) {
  if (threadIdx.x == 0 && blockIdx.x == 0) s->err = 0;
}

}  // namespace

struct gm_ctx {
  gm_params p;
  gm::Geom g;
  gm::Phys ph;
  hipStream_t stream = nullptr;
  size_t esize = 8;      // bytes per stored value: 8 (FP64) or 4 (mixed, FP32 storage)
  void* U = nullptr;     // state U^n (SoA, neq vars, ghosts as uploaded / filled)
  void* Us = nullptr;    // U* after the predictor
  void* W = nullptr;     // stage partials (neq doubles per cell); also the staging
                         // buffer for layout conversion
  void* E = nullptr;     // cell-centred E (3 vars)
  void* Ssrc = nullptr;  // the user source plugin's s of a stage (neq doubles per cell)
  void* Dl = nullptr;    // staging of the Temp / divB downloads (one double per cell), on first use
  Scalars* dscal = nullptr;  // device scalars
  Scalars* hscal = nullptr;  // pinned host mirror
  bool ghosts_dirty = true;
  // the last gm_tstep reduced the CFL minimum of its result into hscal and
  // nothing has changed U's interior or the EOS since
  bool cfl_ready = false;
  bool cfl_pending = false;  // this gm_tstep queued the reduction (ready once its launches succeeded)
  // Table cooling: the pressure primit carries for the interior cells
  // (gm::Box::pc), allocated with the first table-cooling setting; live
  // from a step that ended with table cooling until the next step or upload
  double* Pc = nullptr;
  bool pc_live = false;
  double time = 0.0;                    // simulation time the user hooks see
  hipModule_t usrc_mod = nullptr;       // user source plugin
  hipFunction_t usrc_fn = nullptr;
  hipModule_t ubc_mod = nullptr;        // user boundary plugin (ghosts then live in memory)
  hipFunction_t ubc_fn = nullptr;
  gm::Cooling cool{};                   // kind 0: no cooling
  gm::Conduction cond{};                // kind 0: no thermal conduction
  int tc_substeps = 0;                  // sub-steps of the last step's conduction
  gm::Shape shape = {256, 16};  // tile width, z-pencil length
  // planes per slab of the 2.5-D stage kernel, per stage (predictor,
  // corrector); 0: the two-kernel stage (xy plane tiles, then z pencils
  // through W)
  int slab_kchunk[2] = {0, 0};
  int timing = 0;
  struct Rec { int cls; hipEvent_t a, b; };
  std::vector<Rec> recs;
  std::vector<hipEvent_t> pool;
  double ms[GM_NKCLASS] = {0};
  int64_t launches[GM_NKCLASS] = {0};
  int64_t bytes = 0;
  std::string msg;
};

namespace {

// The stage kernels of a step, per stage: planes per slab of the 2.5-D
// kernel, or 0 for the two-kernel stage.  On a grid that fills the GPU the
// predictor runs as a slab kernel and the corrector as the two-kernel
// stage, in both precisions:
// at 256^3, same box, the first-order slab predictor takes 1.43-1.47 ms
// against 1.55 for k_xy + k_z (FP64; mixed 0.77 against 0.80), while the
// limited slab corrector, held to two waves per SIMD by its window, takes
// 2.17 against 1.92 (mixed 1.52 against 1.05); per step FP64 4.13-4.14 ms
// against 4.32-4.33 (slab both stages) and 4.23-4.24 (two-kernel both),
// mixed 2.29 against 2.76 and 2.34-2.35.  With passive scalars
// both stages run as the two-kernel stage: the slab kernel's registers
// (238-256 VGPRs without them) do not hold the extra components, and it
// spills 12-132 bytes a lane.  GUACHO_AMD_SLAB_KCHUNK (planes per slab;
// 0: two-kernel) and GUACHO_AMD_SLAB_STAGES ("1", "2", "12": the stages
// that use the slab kernel) override the choice.
void stage_policy(int precision, const gm::Geom& g, int kslab[2]) {
  // slab workgroups resident at once: two per CU in FP64 (256 VGPRs), three
  // in the mixed form (the FP32 slab kernels hold 154 and 190)
  const int k = gm::slab_kchunk(g, precision == GM_PRECISION_MIXED ? 768 : 512);
  kslab[0] = g.npas == 0 ? k : 0;
  // a grid whose corrector tiles (32 x 16 x 1) do not reach two per CU runs
  // the corrector as a slab kernel too: one kernel instead of two, and the
  // two-kernel stage's occupancy advantage needs a full GPU.  Per step, FP64
  // (mixed): 128x128x4 0.077 ms against 0.108 before (0.064), 1024x4x4
  // 0.066 against 0.096, 32^3 0.092 against 0.193, 64^3 0.137 against
  // 0.209; 128^3 and up unchanged.
  const long long xy_tiles = (long long)((g.nx + 31) / 32) * ((g.ny + 15) / 16) * g.nz;
  kslab[1] = (g.npas == 0 && xy_tiles <= 512) ? k : 0;
  int len = k;
  if (const char* e = std::getenv("GUACHO_AMD_SLAB_KCHUNK")) {
    const long v = std::strtol(e, nullptr, 10);
    if (v >= 0 && v <= (1 << 20)) {
      len = (int)v;
      kslab[0] = len;
      kslab[1] = kslab[1] ? len : 0;
    }
  }
  if (const char* e = std::getenv("GUACHO_AMD_SLAB_STAGES")) {
    const std::string st(e);
    kslab[0] = st.find('1') != std::string::npos ? len : 0;
    kslab[1] = st.find('2') != std::string::npos ? len : 0;
  }
}

int fail(gm_ctx* c, int code, const std::string& m) {
  if (c) c->msg = m;
  else g_create_error = m;
  return code;
}

int hip_check(gm_ctx* c, hipError_t e, const char* what) {
  if (e == hipSuccess) return GM_OK;
  return fail(c, GM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define GM_HIP(c, call)                                   \
  do {                                                    \
    int _rc = hip_check((c), (call), #call);              \
    if (_rc != GM_OK) return _rc;                         \
  } while (0)

hipEvent_t take_event(gm_ctx* c) {
  if (!c->pool.empty()) {
    hipEvent_t e = c->pool.back();
    c->pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// The records into the per-class totals, their events back to the pool;
// the stream has been synchronised.
void fold_recs(gm_ctx* c) {
  for (auto& r : c->recs) {
    float t = 0.f;
    if (hipEventElapsedTime(&t, r.a, r.b) == hipSuccess) {
      c->ms[r.cls] += t;
      c->launches[r.cls] += 1;
    }
    c->pool.push_back(r.a);
    c->pool.push_back(r.b);
  }
  c->recs.clear();
}

template <class Fn>
void timed(gm_ctx* c, int cls, Fn&& fn) {
  if (!c->timing) {
    fn();
    return;
  }
  if (c->recs.size() >= 4096) {  // bounded while nobody reads the times
    (void)hipStreamSynchronize(c->stream);
    fold_recs(c);
  }
  hipEvent_t a = take_event(c), b = take_event(c);
  if (!a || !b) {
    fn();
    return;
  }
  (void)hipEventRecord(a, c->stream);
  fn();
  (void)hipEventRecord(b, c->stream);
  c->recs.push_back({cls, a, b});
}

int validate(const gm_params* p) {
  std::string msg;
  const int rc = gm::check_params(p, msg);
  if (rc != GM_OK) return fail(nullptr, rc, msg);
  return GM_OK;
}

}  // namespace

namespace gm {

// Configuration check shared by gm_create and gm_batch_add.
int check_params(const gm_params* p, std::string& msg) {
  auto fail = [&msg](gm_ctx*, int code, const std::string& m) {
    msg = m;
    return code;
  };
  if (!p) return fail(nullptr, GM_ERR_ARG, "params is NULL");
  if (p->neq < 8 || p->neq > 10)
    return fail(nullptr, GM_ERR_CONFIG, "neq must be 8 (BFIELD) plus 0, 1 or 2 passive scalars");
  if (p->nghost != 2) return fail(nullptr, GM_ERR_CONFIG, "nghost must be 2");
  if (p->nx < 2 || p->ny < 2 || p->nz < 2)
    return fail(nullptr, GM_ERR_CONFIG, "nx, ny, nz must each be at least 2 (two-layer ghost mirror)");
  if ((long long)(p->nx + 4) * (p->ny + 4) * (p->nz + 4) > (1LL << 36))
    return fail(nullptr, GM_ERR_CONFIG, "grid too large");
  if (p->riemann_solver < GM_SOLVER_HLL || p->riemann_solver > GM_SOLVER_HLLD)
    return fail(nullptr, GM_ERR_CONFIG, "riemann_solver must be HLL, HLLC, HLLE or HLLD (split solvers not built)");
  if (!p->mhd && !p->pmhd)
    return fail(nullptr, GM_ERR_CONFIG, "pure hydro is a neq=5 build in the reference; set mhd or pmhd");
  if (p->mhd && p->pmhd) return fail(nullptr, GM_ERR_CONFIG, "mhd and pmhd are exclusive");
  if (p->riemann_solver == GM_SOLVER_HLLC && p->mhd)
    return fail(nullptr, GM_ERR_CONFIG,
                "HLLC with full MHD leaves the star-state field undefined in the reference (hllc.f90:89-93); use pmhd");
  if (p->slope_limiter < GM_LIMITER_NO_AVERAGE || p->slope_limiter > GM_LIMITER_SUPERBEE)
    return fail(nullptr, GM_ERR_CONFIG, "slope_limiter out of range");
  for (int b = 0; b < 6; b++) {
    if (p->bc[b] < GM_BC_OUTFLOW || p->bc[b] > GM_BC_OTHER)
      return fail(nullptr, GM_ERR_CONFIG, "boundary kind out of range");
  }
  for (int a = 0; a < 3; a++)
    if ((p->bc[2 * a] == GM_BC_PERIODIC) != (p->bc[2 * a + 1] == GM_BC_PERIODIC))
      return fail(nullptr, GM_ERR_CONFIG, "a periodic boundary must pair with a periodic boundary on the same axis");
  if (!std::isfinite(p->eta)) return fail(nullptr, GM_ERR_CONFIG, "eta (artificial viscosity) must be finite");
  if (p->precision != GM_PRECISION_FP64 && p->precision != GM_PRECISION_MIXED)
    return fail(nullptr, GM_ERR_CONFIG, "precision must be GM_PRECISION_FP64 or GM_PRECISION_MIXED");
  if (!(p->dx > 0) || !(p->dy > 0) || !(p->dz > 0) || !std::isfinite(p->dx) || !std::isfinite(p->dy) ||
      !std::isfinite(p->dz))
    return fail(nullptr, GM_ERR_CONFIG, "dx, dy, dz must be positive and finite");
  if (!(p->cv > 0) || !(p->gamma > 0) || !std::isfinite(p->cv) || !std::isfinite(p->gamma))
    return fail(nullptr, GM_ERR_CONFIG, "cv and gamma must be positive and finite");
  if (p->reserved != 0) return fail(nullptr, GM_ERR_CONFIG, "reserved must be 0");
  return GM_OK;
}

}  // namespace gm

namespace {

void release(gm_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (auto& r : c->recs) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (auto e : c->pool) (void)hipEventDestroy(e);
  void* bufs[] = {c->U, c->Us, c->W, c->E, c->Ssrc, c->Dl, c->Pc};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  if (c->usrc_mod) (void)hipModuleUnload(c->usrc_mod);
  if (c->ubc_mod) (void)hipModuleUnload(c->ubc_mod);
  if (c->dscal) (void)hipFree(c->dscal);
  if (c->hscal) (void)hipHostFree(c->hscal);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

int set_device(gm_ctx* c) { return hip_check(c, hipSetDevice(c->p.device), "hipSetDevice"); }

// Staging of the one-value-per-cell downloads.  Not E or U*: E's ghosts at
// BC_OTHER faces and U*'s ghosts outside a boundary hook's reach are read
// as stored, so they carry state from step to step.
int download_staging(gm_ctx* c, double** out) {
  if (!c->Dl) {
    const size_t n = (size_t)c->g.ncell * sizeof(double);
    if (hipMalloc(&c->Dl, n) != hipSuccess) {
      (void)hipGetLastError();
      c->Dl = nullptr;
      return fail(c, GM_ERR_OOM, "device allocation for the download staging failed");
    }
    c->bytes += (int64_t)n;
  }
  *out = (double*)c->Dl;
  return GM_OK;
}

// Arguments of a user plugin launch.
gm_user_args user_args(const gm_ctx* c, void* u, double* sv, double dt, int stage) {
  gm_user_args a{};
  a.nx = c->g.nx; a.ny = c->g.ny; a.nz = c->g.nz;
  a.NX = c->g.NX; a.NY = c->g.NY; a.NZ = c->g.NZ;
  a.ncell = c->g.ncell;
  a.dx = c->p.dx; a.dy = c->p.dy; a.dz = c->p.dz;
  a.cv = c->p.cv; a.gamma = c->p.gamma;
  a.time = c->time;
  a.dt = dt;
  a.stage = stage;
  a.precision = c->esize == 8 ? GM_PRECISION_FP64 : GM_PRECISION_MIXED;
  a.mhd = c->p.mhd; a.pmhd = c->p.pmhd;
  for (int f = 0; f < 6; f++) a.bc[f] = c->p.bc[f];
  a.u = u;
  a.s = sv;
  a.eos = c->ph.eos;
  a.npas = c->g.npas;
  a.tempsc = c->ph.tempsc;
  return a;
}

void launch_plugin(gm_ctx* c, hipFunction_t fn, gm_user_args a, long long items) {
  void* params[] = {&a};
  (void)hipModuleLaunchKernel(fn, (unsigned)((items + 255) / 256), 1, 1, 256, 1, 1, 0, c->stream, params, nullptr);
}

// The user's impose_user_bc(u, order) over every stored cell of u.
void user_bc(gm_ctx* c, void* u, int order) {
  if (c->ubc_fn) launch_plugin(c, c->ubc_fn, user_args(c, u, nullptr, 0.0, order), c->g.ncell);
}

// boundaryI ghost layer on U, in the reference's serial order: periodic
// x, y, z; then closed left..in; then outflow left..in; then the user hook
// (src/boundaries.f90:245).
template <typename S>
void fill_ghost_t(gm_ctx* c, S* U, int layers) {
  const gm::Geom& g = c->g;
  timed(c, KC_GHOST, [&] {
    for (int a = 0; a < 3; a++)
      if (g.bc[2 * a] == GM_BC_PERIODIC) {
        gm::launch_ghost(c->stream, g, U, a, 0, GM_BC_PERIODIC, layers);
        gm::launch_ghost(c->stream, g, U, a, 1, GM_BC_PERIODIC, layers);
      }
    for (int kind : {GM_BC_CLOSED, GM_BC_OUTFLOW})
      for (int f = 0; f < 6; f++)
        if (g.bc[f] == kind) gm::launch_ghost(c->stream, g, U, f / 2, f % 2, kind, layers);
    user_bc(c, U, layers);
  });
}

void fill_ghost1(gm_ctx* c) {
  if (c->ubc_fn) c->cfl_ready = false;  // the hook may write inside the box
  if (c->esize == 8) fill_ghost_t<double>(c, (double*)c->U, 1);
  else fill_ghost_t<float>(c, (float*)c->U, 1);
  c->ghosts_dirty = false;
}

// this context's box, as the kernels see it
template <typename S>
gm::Box<S> box_of(const gm_ctx* c) {
  gm::Box<S> b;
  b.g = c->g;
  b.ph = c->ph;
  b.U = (S*)c->U;
  b.Us = (S*)c->Us;
  b.W = (double*)c->W;
  b.E = (S*)c->E;
  b.dx = c->p.dx;
  b.dy = c->p.dy;
  b.dz = c->p.dz;
  b.slot = 0;
  b.src = (c->p.eight_wave ? gm::SRC_EIGHT_WAVE : 0) | (c->usrc_fn ? gm::SRC_USER : 0);
  b.srcin = (b.src != 0 && c->g.npas == 0) ? 1 : 0;
  b.usrc = (b.srcin && c->usrc_fn) ? (const double*)c->Ssrc : nullptr;
  b.g.gmem = (c->ubc_fn || c->cond.kind) ? 1 : 0;
  b.pc = c->pc_live ? c->Pc : nullptr;
  return b;
}

// The user's source kernel over the interior cells of one stage, on a
// zeroed s (the W block, free once the stage's z kernel has run).
template <typename S>
void user_sources_t(gm_ctx* c, const gm::Box<S>& b, double dt, int stage) {
  (void)hipMemsetAsync(b.W, 0, (size_t)c->g.ncell * c->p.neq * sizeof(double), c->stream);
  void* u = stage == 1 ? (void*)b.U : (void*)b.Us;
  launch_plugin(c, c->usrc_fn, user_args(c, u, b.W, dt, stage), (long long)c->g.nx * c->g.ny * c->g.nz);
}

// With the source terms added by the stage kernels (Box::srcin): the user's
// s of the stage, from its input state, before the stage's kernels run.
template <typename S>
void stage_user_s_t(gm_ctx* c, const gm::Box<S>& b, double dt, int stage) {
  if (!b.srcin || !(b.src & gm::SRC_USER)) return;
  timed(c, KC_SOURCES, [&] {
    (void)hipMemsetAsync(c->Ssrc, 0, (size_t)c->g.ncell * c->p.neq * sizeof(double), c->stream);
    void* u = stage == 1 ? (void*)b.U : (void*)b.Us;
    launch_plugin(c, c->usrc_fn, user_args(c, u, (double*)c->Ssrc, dt, stage), (long long)c->g.nx * c->g.ny * c->g.nz);
  });
}

template <typename S>
void stage_sources_t(gm_ctx* c, const gm::Box<S>& b, double dt, int stage) {
  if (!b.src || b.srcin) return;  // b.srcin: the stage kernels added the terms
  timed(c, KC_SOURCES, [&] {
    if (b.src & gm::SRC_USER) user_sources_t(c, b, dt, stage);
    gm::launch_sources(c->stream, b, dt, stage, c->shape);
  });
}

// superstep / substep / ST_steps (src/thermal_cond.f90:625-681), in the
// reference's arithmetic: real**integer as the compiler's right-to-left
// binary exponentiation, cos from libm
double powi_rl(double x, int n) {
  double r = 1.0, bb = x;
  while (n) {
    if (n & 1) r = r * bb;
    bb = bb * bb;
    n >>= 1;
  }
  return r;
}
double tc_superstep(int n, double snu) {
  const double a = powi_rl(1 + snu, 2 * n), bq = powi_rl(1 - snu, 2 * n);
  return (((double)n / (2. * snu)) * (a - bq)) / (a + bq);
}
double tc_substep(int j, int n, double nu) {
  const double pi = std::acos(-1.);
  return 1. / (((nu - 1.) * std::cos((pi * (double)(2 * j - 1)) / (2. * (double)n)) + nu) + 1.);
}

// thermal_conduction (src/thermal_cond.f90:690-768): dt_cond from the
// device minimum, the super-step count on the host, then the sub-steps
template <typename S>
void conduction_t(gm_ctx* c, const gm::Box<S>& b, double dt) {
  const gm::Conduction& tc = c->cond;
  timed(c, KC_CONDUCTION, [&] {
    hipLaunchKernelGGL(k_init_tcmin, dim3(1), dim3(1), 0, c->stream, c->dscal);
    gm::launch_tc_dtmin(c->stream, b, tc, &c->dscal->tcmin);
  });
  (void)hipMemcpyAsync(&c->hscal->tcmin, &c->dscal->tcmin, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                       c->stream);
  (void)hipStreamSynchronize(c->stream);
  double dtp;
  std::memcpy(&dtp, &c->hscal->tcmin, sizeof dtp);
  double ddx = c->p.dx < c->p.dy ? c->p.dx : c->p.dy;
  ddx = ddx < c->p.dz ? ddx : c->p.dz;
  const double x = ddx * tc.rsc;
  const double dt_cond = ((((((0.25 * 0.5) * (x * x)) * c->p.cv) * 8.3145e7) * dtp) * tc.rhosc) / tc.mu;
  const double dt_hydro = dt * tc.tsc;
  const double nu = 0.01, snu = std::sqrt(0.01);
  bool super = true;
  int nsteps;
  double fstep;
  if (dt_cond < dt_hydro) {
    int j;
    for (j = 1; j <= 199; j++)
      if (tc_superstep(j, snu) > dt_hydro / dt_cond) break;
    nsteps = j;
    fstep = (dt_hydro / dt_cond) / tc_superstep(nsteps, snu);
  } else {
    super = false;
    fstep = dt_hydro / dt_cond;
    nsteps = 1;
  }
  c->tc_substeps = nsteps;
  timed(c, KC_CONDUCTION, [&] {
    for (int n = 1; n <= nsteps; n++) {
      const double dts = super ? (((dt_cond * fstep) * tc_substep(n, nsteps, nu)) / tc.psc) / tc.rsc
                               : ((dt_hydro / (double)nsteps) / tc.psc) / tc.rsc;
      gm::launch_tc_substep(c->stream, b, tc, dts);
    }
  });
}

// the step of one precision: tstep (src/hydro_solver.f90:134-229)
template <typename S>
void tstep_t(gm_ctx* c, double dt) {
  const gm_params& p = c->p;
  hipStream_t s = c->stream;
  int* err = &c->dscal->err;
  const gm::Box<S> b = box_of<S>(c);
  const gm::Shape& sh = c->shape;
  const int solver = p.riemann_solver, fcd = p.enable_flux_cd ? 1 : 0;
  const double dtm = dt / 2.0;  // hydro_solver.f90:152
  // predictor: first-order fluxes of U^n, U* = U^n - dt/2 div F (+ dt/2 s)
  // the 2.5-D stage kernel unless the stage kernels add source terms
  const int kslab1 = b.srcin ? 0 : c->slab_kchunk[0], kslab2 = b.srcin ? 0 : c->slab_kchunk[1];
  stage_user_s_t(c, b, dtm, 1);
  if (kslab1) {
    timed(c, KC_SLAB1, [&] { gm::launch_slab(s, b, dtm, 1, solver, fcd, err, kslab1); });
  } else {
    timed(c, KC_XY1, [&] { gm::launch_xy(s, b, dtm, 1, solver, fcd, err); });
    timed(c, KC_Z1, [&] { gm::launch_z(s, b, dtm, 1, solver, fcd, err, sh); });
  }
  if (fcd) timed(c, KC_BFIELD, [&] { gm::launch_bfield(s, b, dtm, 1, sh); });
  stage_sources_t(c, b, dtm, 1);
  // corrector: limited fluxes of U*, U^{n+1} = U^n - dt div F (+ dt s), in
  // place
  // boundaryII on U* (src/hydro_solver.f90:169) when ghosts live in memory
  if (b.g.gmem) fill_ghost_t<S>(c, b.Us, 2);
  stage_user_s_t(c, b, dt, 2);
  if (kslab2) {
    timed(c, KC_SLAB2, [&] { gm::launch_slab(s, b, dt, 2, solver, fcd, err, kslab2); });
  } else {
    timed(c, KC_XY2, [&] { gm::launch_xy(s, b, dt, 2, solver, fcd, err); });
    timed(c, KC_Z2, [&] { gm::launch_z(s, b, dt, 2, solver, fcd, err, sh); });
  }
  // when nothing after the field update touches U's interior (no sources,
  // viscosity, cooling, conduction, or boundary hook that may write inside
  // the box), the next get_timestep's minimum is reduced in the same pass
  const bool fuse_cfl = fcd && !b.src && p.eta == 0.0 && !c->cool.kind && !c->cond.kind && !c->ubc_fn;
  if (fuse_cfl) {
    hipLaunchKernelGGL(k_init_dtmin, dim3(1), dim3(1), 0, s, c->dscal);
    timed(c, KC_BFIELD_CFL, [&] { gm::launch_bfield_cfl(s, b, dt, &c->dscal->dtmin, err, sh); });
    (void)hipMemcpyAsync(c->hscal, c->dscal, sizeof(Scalars), hipMemcpyDeviceToHost, s);
    c->cfl_pending = true;
  } else if (fcd) {
    timed(c, KC_BFIELD, [&] { gm::launch_bfield(s, b, dt, 2, sh); });
  }
  stage_sources_t(c, b, dt, 2);
  // viscous_copy (src/hydro_solver.f90:188); with eta = 0 the identity on
  // values, so no pass
  if (p.eta != 0.0) timed(c, KC_VISCOUS, [&] { gm::launch_viscous(s, b, p.eta); });
  // cooling (src/hydro_solver.f90:206-210); table cooling also leaves the
  // interior cells' primit pressure for the next step (Box::pc)
  if (c->cool.kind) {
    gm::Box<S> bc = b;
    bc.pc = c->cool.kind == GM_COOL_H ? nullptr : c->Pc;
    timed(c, KC_COOLING, [&] { gm::launch_cooling(s, bc, c->cool, dt); });
  }
  // boundaryI on U^{n+1} (src/hydro_solver.f90:216) when ghosts live in memory
  if (b.g.gmem) fill_ghost_t<S>(c, b.U, 1);
  // thermal conduction (src/hydro_solver.f90:227)
  if (c->cond.kind) conduction_t<S>(c, b, dt);
}

template <typename S>
void cfl_t(gm_ctx* c) {
  const gm::Box<S> b = box_of<S>(c);
  timed(c, KC_CFL, [&] { gm::launch_cfl(c->stream, b, &c->dscal->dtmin, &c->dscal->err, c->shape); });
}

int read_errors(gm_ctx* c) {
  GM_HIP(c, hipMemcpyAsync(&c->hscal->err, &c->dscal->err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  GM_HIP(c, hipStreamSynchronize(c->stream));
  const int e = c->hscal->err;
  if (e) {
    hipLaunchKernelGGL(k_clear_err, dim3(1), dim3(1), 0, c->stream, c->dscal);
    GM_HIP(c, hipGetLastError());
    return fail(c, GM_ERR_NONFINITE,
                (e & 1) ? "Riemann solver fell through on a non-finite state (hlld.f90:316-317)"
                        : "non-finite value in the conserved state");
  }
  return GM_OK;
}

// ---------------------------------------------------------------------------
// gm_tstep_noalloc: the step on a caller-held device state with every
// buffer beyond that state -- U*, the stage partials W, E and the device
// scalars -- carved from one file-scope device array, so the call allocates
// nothing.  Each device that runs it holds its own copy of the array (the
// code object's data, until the process ends); host code reaches it only
// through the address hipGetSymbolAddress returns on that device.  Calls
// that overlap on one device share the array and race.
// ---------------------------------------------------------------------------
constexpr size_t kScratchBytes = (size_t)256 << 20;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
void* g_scratch_ptr[kMaxDevices] = {nullptr};
std::mutex g_scratch_mu;

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// device bytes gm_tstep_noalloc needs for params (U*, W, E, scalars)
size_t noalloc_bytes(const gm_params* p) {
  const size_t ncell = (size_t)(p->nx + 4) * (p->ny + 4) * (p->nz + 4);
  const size_t es = p->precision == GM_PRECISION_MIXED ? 4 : 8, nq = (size_t)p->neq;
  return align256(ncell * nq * es) + align256(ncell * nq * sizeof(double)) + align256(ncell * 3 * es) +
         align256(sizeof(Scalars));
}

// address of this device's scratch array (the device is current)
int scratch_of(int dev, unsigned char** out) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* a = nullptr;
    const hipError_t e = hipGetSymbolAddress(&a, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !a) {
      (void)hipGetLastError();
      g_create_error = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return GM_ERR_HIP;
    }
    g_scratch_ptr[dev] = a;
  }
  *out = (unsigned char*)g_scratch_ptr[dev];
  return GM_OK;
}

__global__ void k_publish(const Scalars* sc, double* dtp, int32_t* err) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    if (dtp) *dtp = __longlong_as_double((long long)sc->dtmin);
    if (err) *err = sc->err;
  }
}

// tstep (src/hydro_solver.f90:134-229) with the add-ons a context adds
// (user plugins, cooling, conduction) off: the kernels and order of
// tstep_t, then get_timestep's minimum over the new state, as gm_cfl_dtp
// computes it.  Scalars are set by kernels, not memsets, so a caller may
// capture the launches in a graph.
template <typename S>
void noalloc_tstep_t(hipStream_t s, const gm_params& p, S* u, unsigned char* scr, double dt) {
  gm::Box<S> b;
  b.g.nx = p.nx; b.g.ny = p.ny; b.g.nz = p.nz;
  b.g.NX = p.nx + 4; b.g.NY = p.ny + 4; b.g.NZ = p.nz + 4;
  b.g.ncell = (long long)b.g.NX * b.g.NY * b.g.NZ;
  for (int f = 0; f < 6; f++) b.g.bc[f] = p.bc[f];
  b.g.gmem = 0;
  b.g.npas = p.neq - 8;
  b.ph.cv = p.cv;
  b.ph.gamma = p.gamma;
  b.ph.mhd = p.mhd ? 1 : 0;
  b.ph.pmhd = p.pmhd ? 1 : 0;
  b.ph.limiter = p.slope_limiter;
  b.ph.eos = GM_EOS_ADIABATIC;
  b.ph.tempsc = 1.0;
  const size_t ncell = (size_t)b.g.ncell, nq = (size_t)p.neq;
  size_t off = 0;
  b.U = u;
  b.Us = (S*)(scr + off);
  off += align256(ncell * nq * sizeof(S));
  b.W = (double*)(scr + off);
  off += align256(ncell * nq * sizeof(double));
  b.E = (S*)(scr + off);
  off += align256(ncell * 3 * sizeof(S));
  Scalars* sc = (Scalars*)(scr + off);
  b.dx = p.dx; b.dy = p.dy; b.dz = p.dz;
  b.slot = 0;
  b.src = p.eight_wave ? gm::SRC_EIGHT_WAVE : 0;
  b.srcin = (b.src != 0 && b.g.npas == 0) ? 1 : 0;
  b.usrc = nullptr;
  const gm::Shape sh = {256, 16};
  const int solver = p.riemann_solver, fcd = p.enable_flux_cd ? 1 : 0;
  int* err = &sc->err;
  const double dtm = dt / 2.0;  // hydro_solver.f90:152
  int kslab[2];
  stage_policy(p.precision, b.g, kslab);  // as gm_create chooses
  auto stage = [&](double h, int st) {
    if (!b.srcin && kslab[st - 1]) {
      gm::launch_slab(s, b, h, st, solver, fcd, err, kslab[st - 1]);
    } else {
      gm::launch_xy(s, b, h, st, solver, fcd, err);
      gm::launch_z(s, b, h, st, solver, fcd, err, sh);
    }
  };
  hipLaunchKernelGGL(k_clear_err, dim3(1), dim3(1), 0, s, sc);
  stage(dtm, 1);
  if (fcd) gm::launch_bfield(s, b, dtm, 1, sh);
  if (b.src && !b.srcin) gm::launch_sources(s, b, dtm, 1, sh);
  stage(dt, 2);
  hipLaunchKernelGGL(k_init_dtmin, dim3(1), dim3(1), 0, s, sc);
  const bool fuse_cfl = fcd && !b.src && p.eta == 0.0;
  if (fuse_cfl) {
    gm::launch_bfield_cfl(s, b, dt, &sc->dtmin, err, sh);
  } else {
    if (fcd) gm::launch_bfield(s, b, dt, 2, sh);
    if (b.src && !b.srcin) gm::launch_sources(s, b, dt, 2, sh);
    if (p.eta != 0.0) gm::launch_viscous(s, b, p.eta);
    gm::launch_cfl(s, b, &sc->dtmin, err, sh);
  }
}

}  // namespace

extern "C" {

int gm_create(const gm_params* params, gm_ctx** out) {
  if (!out) return fail(nullptr, GM_ERR_ARG, "ctx out-pointer is NULL");
  *out = nullptr;
  int rc = validate(params);
  if (rc != GM_OK) return rc;
  gm_ctx* c = new gm_ctx();
  c->p = *params;
  gm::Geom& g = c->g;
  g.nx = params->nx; g.ny = params->ny; g.nz = params->nz;
  g.NX = g.nx + 4; g.NY = g.ny + 4; g.NZ = g.nz + 4;
  g.ncell = (long long)g.NX * g.NY * g.NZ;
  for (int b = 0; b < 6; b++) g.bc[b] = params->bc[b];
  g.gmem = 0;
  g.npas = params->neq - 8;
  c->ph.cv = params->cv;
  c->ph.gamma = params->gamma;
  c->ph.mhd = params->mhd ? 1 : 0;
  c->ph.pmhd = params->pmhd ? 1 : 0;
  c->ph.limiter = params->slope_limiter;
  c->ph.eos = GM_EOS_ADIABATIC;
  c->ph.tempsc = 1.0;
  stage_policy(params->precision, g, c->slab_kchunk);
  // GUACHO_AMD_COLUMN_SHAPE="block,kchunk" overrides the column kernels'
  // tile width (64, 128 or 256 threads) and z-pencil length
  if (const char* e = std::getenv("GUACHO_AMD_COLUMN_SHAPE")) {
    int bl = 0, kc = 0;
    if (std::sscanf(e, "%d,%d", &bl, &kc) == 2 && (bl == 64 || bl == 128 || bl == 256) && kc >= 1 && kc <= 4096)
      c->shape = {bl, kc, 0};
  }

  auto bail = [&](int code) {
    g_create_error = c->msg;
    release(c);
    return code;
  };
  if ((rc = set_device(c)) != GM_OK) return bail(rc);
  if ((rc = hip_check(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate")) != GM_OK)
    return bail(rc);
  c->esize = params->precision == GM_PRECISION_MIXED ? 4 : 8;
  const size_t nq = (size_t)params->neq;
  const size_t n8 = (size_t)g.ncell * nq * c->esize;
  const size_t n3 = (size_t)g.ncell * 3 * c->esize;
  const size_t nw = (size_t)g.ncell * nq * sizeof(double);
  struct { void** p; size_t n; } allocs[] = {{&c->U, n8}, {&c->Us, n8}, {&c->W, nw}, {&c->E, n3}};
  for (auto& a : allocs) {
    if (hipMalloc(a.p, a.n) != hipSuccess) {
      (void)hipGetLastError();
      c->msg = "device allocation failed";
      return bail(GM_ERR_OOM);
    }
    c->bytes += (int64_t)a.n;
  }
  if ((rc = hip_check(c, hipMalloc((void**)&c->dscal, sizeof(Scalars)), "hipMalloc")) != GM_OK) return bail(rc);
  if ((rc = hip_check(c, hipHostMalloc((void**)&c->hscal, sizeof(Scalars), hipHostMallocDefault),
                      "hipHostMalloc")) != GM_OK)
    return bail(rc);
  // zero everything once so no kernel ever reads uninitialised memory
  for (auto& a : allocs) (void)hipMemsetAsync(*a.p, 0, a.n, c->stream);
  (void)hipMemsetAsync(c->dscal, 0, sizeof(Scalars), c->stream);
  if ((rc = hip_check(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize")) != GM_OK) return bail(rc);
  *out = c;
  return GM_OK;
}

int gm_destroy(gm_ctx* c) {
  if (!c) return GM_ERR_ARG;
  (void)hipSetDevice(c->p.device);
  release(c);
  return GM_OK;
}

const char* gm_last_error(const gm_ctx* c) { return c ? c->msg.c_str() : g_create_error.c_str(); }

int gm_upload_u(gm_ctx* c, const double* u) {
  if (!c) return GM_ERR_ARG;
  if (!u) return fail(c, GM_ERR_ARG, "u is NULL");
  int rc = set_device(c);
  if (rc) return rc;
  c->cfl_ready = false;
  c->pc_live = false;  // the host's calcprim of the new state
  const int nq = c->p.neq;
  const size_t n8 = (size_t)c->g.ncell * nq * sizeof(double);
  GM_HIP(c, hipMemcpyAsync(c->W, u, n8, hipMemcpyHostToDevice, c->stream));
  timed(c, KC_LAYOUT, [&] {
    if (c->esize == 8) gm::launch_aos_to_soa(c->stream, c->g.ncell, (const double*)c->W, (double*)c->U, nq);
    else gm::launch_aos_to_soa(c->stream, c->g.ncell, (const double*)c->W, (float*)c->U, nq);
  });
  GM_HIP(c, hipGetLastError());
  GM_HIP(c, hipStreamSynchronize(c->stream));
  c->ghosts_dirty = false;  // the caller's ghosts are the reference's ghosts
  return GM_OK;
}

int gm_boundary_u(gm_ctx* c) {
  if (!c) return GM_ERR_ARG;
  int rc = set_device(c);
  if (rc) return rc;
  c->cfl_ready = false;  // a boundary hook may write inside the box
  fill_ghost1(c);
  GM_HIP(c, hipGetLastError());
  return GM_OK;
}

int gm_download_u(gm_ctx* c, double* u) {
  if (!c) return GM_ERR_ARG;
  if (!u) return fail(c, GM_ERR_ARG, "u is NULL");
  int rc = set_device(c);
  if (rc) return rc;
  if (c->ghosts_dirty) fill_ghost1(c);
  const int nq = c->p.neq;
  timed(c, KC_LAYOUT, [&] {
    if (c->esize == 8) gm::launch_soa_to_aos(c->stream, c->g.ncell, (const double*)c->U, (double*)c->W, nq);
    else gm::launch_soa_to_aos(c->stream, c->g.ncell, (const float*)c->U, (double*)c->W, nq);
  });
  GM_HIP(c, hipGetLastError());
  GM_HIP(c, hipMemcpyAsync(u, c->W, (size_t)c->g.ncell * nq * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  GM_HIP(c, hipStreamSynchronize(c->stream));
  return GM_OK;
}

int gm_download_primit(gm_ctx* c, double* primit, double* temp) {
  if (!c) return GM_ERR_ARG;
  if (!primit) return fail(c, GM_ERR_ARG, "primit is NULL");
  int rc = set_device(c);
  if (rc) return rc;
  if (c->ghosts_dirty) fill_ghost1(c);
  double* T = nullptr;
  if (temp && (rc = download_staging(c, &T)) != GM_OK) return rc;
  const int nq = c->p.neq;
  timed(c, KC_LAYOUT, [&] {
    const double* pc = c->pc_live ? c->Pc : nullptr;
    if (c->esize == 8)
      gm::launch_prim_all(c->stream, c->g.ncell, c->ph, (const double*)c->U, (double*)c->W, T, nq, c->g, pc);
    else gm::launch_prim_all(c->stream, c->g.ncell, c->ph, (const float*)c->U, (double*)c->W, T, nq, c->g, pc);
  });
  GM_HIP(c, hipGetLastError());
  GM_HIP(c, hipMemcpyAsync(primit, c->W, (size_t)c->g.ncell * nq * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  if (temp)
    GM_HIP(c, hipMemcpyAsync(temp, T, (size_t)c->g.ncell * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  GM_HIP(c, hipStreamSynchronize(c->stream));
  return GM_OK;
}

int gm_download_divb(gm_ctx* c, double* divb) {
  if (!c) return GM_ERR_ARG;
  if (!divb) return fail(c, GM_ERR_ARG, "divb is NULL");
  int rc = set_device(c);
  if (rc) return rc;
  double* D = nullptr;
  if ((rc = download_staging(c, &D)) != GM_OK) return rc;
  if (c->esize == 8) gm::launch_divb(c->stream, c->g, (const double*)c->U, c->p.dx, c->p.dy, c->p.dz, D);
  else gm::launch_divb(c->stream, c->g, (const float*)c->U, c->p.dx, c->p.dy, c->p.dz, D);
  GM_HIP(c, hipGetLastError());
  const size_t n = (size_t)c->g.nx * c->g.ny * c->g.nz;
  GM_HIP(c, hipMemcpyAsync(divb, D, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  GM_HIP(c, hipStreamSynchronize(c->stream));
  return GM_OK;
}

int gm_cfl_dtp(gm_ctx* c, double* dtp) {
  if (!c) return GM_ERR_ARG;
  if (!dtp) return fail(c, GM_ERR_ARG, "dtp is NULL");
  int rc = set_device(c);
  if (rc) return rc;
  if (!c->cfl_ready) {
    hipLaunchKernelGGL(k_init_dtmin, dim3(1), dim3(1), 0, c->stream, c->dscal);
    if (c->esize == 8) cfl_t<double>(c);
    else cfl_t<float>(c);
    GM_HIP(c, hipGetLastError());
    GM_HIP(c, hipMemcpyAsync(c->hscal, c->dscal, sizeof(Scalars), hipMemcpyDeviceToHost, c->stream));
  }
  c->cfl_ready = false;
  GM_HIP(c, hipStreamSynchronize(c->stream));
  const unsigned long long bits = c->hscal->dtmin;
  std::memcpy(dtp, &bits, sizeof(double));
  if (c->hscal->err) return read_errors(c);
  return GM_OK;
}

int gm_tstep(gm_ctx* c, double dt) {
  if (!c) return GM_ERR_ARG;
  if (!std::isfinite(dt)) return fail(c, GM_ERR_ARG, "dt is not finite");
  int rc = set_device(c);
  if (rc) return rc;
  if (!c->ubc_fn)
    for (int f = 0; f < 6; f++)
      if (c->p.bc[f] == GM_BC_OTHER)
        return fail(c, GM_ERR_CONFIG, "BC_OTHER faces need a user boundary plugin (gm_set_user_bc)");
  c->cfl_ready = c->cfl_pending = false;
  if (c->esize == 8) tstep_t<double>(c, dt);
  else tstep_t<float>(c, dt);
  // after table cooling tstep recomputes only the ghost primitives
  // (hydro_solver.f90:218-223); otherwise all of them
  c->pc_live = c->cool.kind && c->cool.kind != GM_COOL_H;
  c->ghosts_dirty = c->ubc_fn == nullptr && !c->cond.kind;  // else the step ends with boundaryI
  GM_HIP(c, hipGetLastError());
  c->cfl_ready = c->cfl_pending;
  return GM_OK;
}

int gm_sync(gm_ctx* c) {
  if (!c) return GM_ERR_ARG;
  int rc = set_device(c);
  if (rc) return rc;
  GM_HIP(c, hipStreamSynchronize(c->stream));
  return read_errors(c);
}

int gm_set_timing(gm_ctx* c, int enable) {
  if (!c) return GM_ERR_ARG;
  c->timing = enable ? 1 : 0;
  return GM_OK;
}

int gm_kernel_times(gm_ctx* c, double* ms, int64_t* launches, const char** names) {
  if (!c) return GM_ERR_ARG;
  int rc = set_device(c);
  if (rc) return rc;
  GM_HIP(c, hipStreamSynchronize(c->stream));
  fold_recs(c);
  for (int k = 0; k < GM_NKCLASS; k++) {
    if (ms) ms[k] = c->ms[k];
    if (launches) launches[k] = c->launches[k];
    if (names) names[k] = kClassNames[k];
    c->ms[k] = 0;
    c->launches[k] = 0;
  }
  return GM_OK;
}

int gm_set_time(gm_ctx* c, double time) {
  if (!c) return GM_ERR_ARG;
  if (!std::isfinite(time)) return fail(c, GM_ERR_ARG, "time is not finite");
  c->time = time;
  return GM_OK;
}

int gm_set_user_sources(gm_ctx* c, const char* code_object, const char* kernel) {
  if (!c) return GM_ERR_ARG;
  int rc = set_device(c);
  if (rc) return rc;
  GM_HIP(c, hipStreamSynchronize(c->stream));
  if (c->usrc_mod) (void)hipModuleUnload(c->usrc_mod);
  c->usrc_mod = nullptr;
  c->usrc_fn = nullptr;
  if (!code_object) return GM_OK;
  if (!kernel) return fail(c, GM_ERR_ARG, "kernel name is NULL");
  hipModule_t m = nullptr;
  if (hipModuleLoad(&m, code_object) != hipSuccess) {
    (void)hipGetLastError();
    return fail(c, GM_ERR_CONFIG, std::string("cannot load code object ") + code_object);
  }
  hipFunction_t f = nullptr;
  if (hipModuleGetFunction(&f, m, kernel) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipModuleUnload(m);
    return fail(c, GM_ERR_CONFIG, std::string("no kernel ") + kernel + " in " + code_object);
  }
  if (!c->Ssrc) {
    const size_t n = (size_t)c->g.ncell * c->p.neq * sizeof(double);
    if (hipMalloc(&c->Ssrc, n) != hipSuccess) {
      (void)hipGetLastError();
      c->Ssrc = nullptr;
      (void)hipModuleUnload(m);
      return fail(c, GM_ERR_OOM, "device allocation for the user source terms failed");
    }
    c->bytes += (int64_t)n;
  }
  c->usrc_mod = m;
  c->usrc_fn = f;
  return GM_OK;
}

int gm_set_user_bc(gm_ctx* c, const char* code_object, const char* kernel) {
  if (!c) return GM_ERR_ARG;
  int rc = set_device(c);
  if (rc) return rc;
  GM_HIP(c, hipStreamSynchronize(c->stream));
  if (c->ubc_mod) (void)hipModuleUnload(c->ubc_mod);
  c->ubc_mod = nullptr;
  c->ubc_fn = nullptr;
  if (!code_object) return GM_OK;
  if (!kernel) return fail(c, GM_ERR_ARG, "kernel name is NULL");
  hipModule_t m = nullptr;
  if (hipModuleLoad(&m, code_object) != hipSuccess) {
    (void)hipGetLastError();
    return fail(c, GM_ERR_CONFIG, std::string("cannot load code object ") + code_object);
  }
  hipFunction_t f = nullptr;
  if (hipModuleGetFunction(&f, m, kernel) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipModuleUnload(m);
    return fail(c, GM_ERR_CONFIG, std::string("no kernel ") + kernel + " in " + code_object);
  }
  c->ubc_mod = m;
  c->ubc_fn = f;
  // from here on the kernels read U's ghosts as stored: materialise them
  if (c->ghosts_dirty) fill_ghost1(c);
  GM_HIP(c, hipGetLastError());
  return GM_OK;
}

int gm_set_cooling(gm_ctx* c, int32_t kind, const double* cooltab, double tempsc, double psc, double tsc) {
  if (!c) return GM_ERR_ARG;
  if (kind == GM_COOL_NONE) {
    c->cool.kind = 0;
    return GM_OK;
  }
  if (kind != GM_COOL_DMC && kind != GM_COOL_CHI && kind != GM_COOL_H)
    return fail(c, GM_ERR_CONFIG, "cooling kind must be GM_COOL_NONE, GM_COOL_H, GM_COOL_DMC or GM_COOL_CHI");
  if (kind == GM_COOL_H && c->g.npas < 1)
    return fail(c, GM_ERR_CONFIG, "GM_COOL_H needs the neutral density as a passive scalar (neq >= 9)");
  if (!cooltab && kind != GM_COOL_H) return fail(c, GM_ERR_ARG, "cooltab is NULL");
  if (kind != GM_COOL_H && c->cond.kind) return fail(c, GM_ERR_CONFIG, kTableCoolCond);
  if (!(tempsc > 0) || !(psc > 0) || !(tsc > 0) || !std::isfinite(tempsc) || !std::isfinite(psc) ||
      !std::isfinite(tsc))
    return fail(c, GM_ERR_CONFIG, "tempsc, psc and tsc must be positive and finite");
  if (cooltab)
    for (int i = 0; i < 82; i++)
      if (!std::isfinite(cooltab[i])) return fail(c, GM_ERR_CONFIG, "cooltab holds a non-finite value");
  gm::Cooling cl{};
  cl.kind = kind;
  cl.tempsc = tempsc;
  cl.psc = psc;
  cl.tsc = tsc;
  cl.hicoef = kind == GM_COOL_DMC ? 0.27e-26 : 0.21e-26;  // cooling_dmc.f90:88, cooling_chi.f90
  if (cooltab) std::memcpy(cl.tab, cooltab, sizeof cl.tab);
  if (kind != GM_COOL_H && !c->Pc) {
    int rc = set_device(c);
    if (rc) return rc;
    if (hipMalloc((void**)&c->Pc, (size_t)c->g.ncell * sizeof(double)) != hipSuccess) {
      (void)hipGetLastError();
      c->Pc = nullptr;
      return fail(c, GM_ERR_OOM, "device allocation failed (table cooling's primitive pressure)");
    }
    c->bytes += (int64_t)c->g.ncell * (int64_t)sizeof(double);
  }
  c->cool = cl;
  return GM_OK;
}

int gm_set_conduction(gm_ctx* c, const gm_conduction_params* tp) {
  if (!c) return GM_ERR_ARG;
  if (!tp) return fail(c, GM_ERR_ARG, "conduction params is NULL");
  if (tp->kind == GM_TC_OFF) {
    c->cond.kind = 0;
    return GM_OK;
  }
  if (tp->kind != GM_TC_ISOTROPIC && tp->kind != GM_TC_ANISOTROPIC)
    return fail(c, GM_ERR_CONFIG, "conduction kind must be GM_TC_OFF, GM_TC_ISOTROPIC or GM_TC_ANISOTROPIC");
  if (c->cool.kind && c->cool.kind != GM_COOL_H) return fail(c, GM_ERR_CONFIG, kTableCoolCond);
  const double v[] = {tp->tempsc, tp->rsc, tp->rhosc, tp->mu, tp->psc, tp->tsc, tp->vsc2, tp->bsc};
  for (double x : v)
    if (!(x > 0) || !std::isfinite(x)) return fail(c, GM_ERR_CONFIG, "conduction scalings must be positive and finite");
  int rc = set_device(c);
  if (rc) return rc;
  gm::Conduction t{};
  t.kind = tp->kind;
  t.sat = tp->saturation ? 1 : 0;
  t.tempsc = tp->tempsc; t.rsc = tp->rsc; t.rhosc = tp->rhosc; t.mu = tp->mu;
  t.psc = tp->psc; t.tsc = tp->tsc; t.vsc2 = tp->vsc2; t.bsc = tp->bsc;
  c->cond = t;
  // from here on the kernels read U's ghosts as stored: materialise them
  if (c->ghosts_dirty) fill_ghost1(c);
  GM_HIP(c, hipGetLastError());
  return GM_OK;
}

int gm_set_eos(gm_ctx* c, int32_t eos, double tempsc) {
  if (!c) return GM_ERR_ARG;
  if (eos != GM_EOS_ADIABATIC && eos != GM_EOS_SINGLE_SPECIE && eos != GM_EOS_H_RATE)
    return fail(c, GM_ERR_CONFIG, "eq_of_state must be GM_EOS_ADIABATIC, GM_EOS_SINGLE_SPECIE or GM_EOS_H_RATE "
                                  "(EOS_CHEM needs the chemistry network)");
  if (eos == GM_EOS_H_RATE && c->g.npas < 1)
    return fail(c, GM_ERR_CONFIG, "GM_EOS_H_RATE needs the neutral density as a passive scalar (neq >= 9)");
  if (!(tempsc > 0) || !std::isfinite(tempsc)) return fail(c, GM_ERR_CONFIG, "Tempsc must be positive and finite");
  c->ph.eos = eos;
  c->ph.tempsc = tempsc;
  c->cfl_ready = false;
  return GM_OK;
}

int gm_conduction_substeps(const gm_ctx* c, int32_t* n) {
  if (!c || !n) return GM_ERR_ARG;
  *n = c->tc_substeps;
  return GM_OK;
}

int gm_device_bytes(const gm_ctx* c, int64_t* bytes) {
  if (!c || !bytes) return GM_ERR_ARG;
  *bytes = c->bytes;
  return GM_OK;
}

int gm_tstep_noalloc(const gm_params* params, void* u, double dt, void* stream, double* dtp_next, int32_t* err) {
  int rc = validate(params);
  if (rc != GM_OK) return rc;
  for (int f = 0; f < 6; f++)
    if (params->bc[f] == GM_BC_OTHER)
      return fail(nullptr, GM_ERR_CONFIG, "BC_OTHER faces need a context with a user boundary plugin");
  const size_t need = noalloc_bytes(params);
  if (need > kScratchBytes)
    return fail(nullptr, GM_ERR_CONFIG,
                "grid too large for gm_tstep_noalloc: needs " + std::to_string(need) + " scratch bytes, the limit is " +
                    std::to_string(kScratchBytes) + " (256 MiB)");
  if (!u) return fail(nullptr, GM_ERR_ARG, "u is NULL");
  if (!std::isfinite(dt)) return fail(nullptr, GM_ERR_ARG, "dt is not finite");
  if (params->device < 0 || params->device >= kMaxDevices) return fail(nullptr, GM_ERR_ARG, "device out of range");
  if ((rc = hip_check(nullptr, hipSetDevice(params->device), "hipSetDevice")) != GM_OK) return rc;
  unsigned char* scr = nullptr;
  if ((rc = scratch_of(params->device, &scr)) != GM_OK) return rc;
  hipStream_t s = (hipStream_t)stream;
  if (params->precision == GM_PRECISION_MIXED) noalloc_tstep_t<float>(s, *params, (float*)u, scr, dt);
  else noalloc_tstep_t<double>(s, *params, (double*)u, scr, dt);
  const size_t off = noalloc_bytes(params) - align256(sizeof(Scalars));
  if (dtp_next || err) hipLaunchKernelGGL(k_publish, dim3(1), dim3(1), 0, s, (const Scalars*)(scr + off), dtp_next, err);
  return hip_check(nullptr, hipGetLastError(), "kernel launch");
}

}  // extern "C"

// gm_batch.hip -- batches of independent boxes behind the gm_batch_* C ABI.
//
// Each box owns its device buffers (admitted against a memory budget).  A
// step of the batch is the same kernel sequence as one box's tstep, but
// every launch covers the tiles of all active boxes: a device table of box
// views plus per-kernel-class prefix sums of their tile counts lets each
// workgroup find its box.  Per-box clocks on the device hold time,
// currentIteration and the CFL minimum, so the ramp and dump clamp of
// get_timestep (src/hydro_core.f90:677-694) run there and a step needs no
// host round trip; steps_per_launch steps are captured once in a HIP graph
// (kernel nodes only) and replayed.  The tables and the graph are rebuilt
// when boxes join or leave.
//
// The autotuner times each candidate launch shape (tile width x z-pencil
// length) on the admitted boxes with dt = 0 steps: every flux is computed
// in full while the state keeps its values.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "gm_kernels.hpp"
#include "guacho_amd.h"

namespace {

unsigned long long gm_bits_of(double x) {
  unsigned long long u;
  std::memcpy(&u, &x, sizeof u);
  return u;
}

struct Slot {
  bool used = false;
  gm_params p{};
  void* U = nullptr;
  void* Us = nullptr;
  void* W = nullptr;  // stage partials, 8 doubles per cell; also upload/download staging
  void* E = nullptr;
  int64_t bytes = 0;
  gm::BoxClock clock{};  // host copy as of the last sync
};

constexpr int kMaxBoxes = 4096;

}  // namespace

struct gm_batch {
  gm_batch_params bp{};
  size_t esize = 8;
  hipStream_t stream = nullptr;  // the steps (graph launches)
  hipStream_t xfer = nullptr;    // admissions and downloads, beside running steps
  std::vector<Slot> slots;
  int64_t budget = 0;
  int64_t used = 0;
  void* dboxes = nullptr;          // Box<S>[max_boxes], active boxes packed
  unsigned* dprefix = nullptr;     // KC_COUNT x (max_boxes + 1)
  int* dslots = nullptr;           // clock slots of the packed table
  gm::BoxClock* dclk = nullptr;    // one clock per slot
  gm::BoxClock* hclk = nullptr;    // pinned mirror
  int* derr = nullptr;
  int* herr = nullptr;
  std::vector<int> order;          // slots of the packed table: admitted boxes not yet done
  // the launches in flight: their boxes' slots, the event after their clock
  // read-back, and device buffers whose release waits for them
  bool in_flight = false;
  std::vector<int> launched;
  hipEvent_t ev_launch = nullptr;
  std::vector<void*> zombies;
  bool any_sources = false;        // some active box has 8-wave source terms
  bool any_viscous = false;        // some active box has artificial viscosity
  int neq = 8;                     // variables of every box: 8 + its passive scalars (the first box sets it)
  unsigned tiles[gm::KC_COUNT] = {0};
  bool tables_dirty = true;
  bool graph_dirty = true;
  gm::Shape shape = {256, 16};
  bool tuned = false;
  double tuned_ms = 0.0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // launch geometry the graph was captured with: its nodes read the box and
  // prefix tables through stable device pointers, so new table contents with
  // the same geometry reuse it
  int cap_nbox = -1;
  unsigned cap_tiles[gm::KC_COUNT] = {0};
  gm::Shape cap_shape = {0, 0};
  bool cap_sources = false;
  bool cap_viscous = false;
  int cap_neq = 0;
  std::string msg;
};

namespace {

thread_local std::string g_batch_create_error;

int bfail(gm_batch* b, int code, const std::string& m) {
  if (b) b->msg = m;
  else g_batch_create_error = m;
  return code;
}

int bhip(gm_batch* b, hipError_t e, const char* what) {
  if (e == hipSuccess) return GM_OK;
  return bfail(b, GM_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define BHIP(b, call)                                 \
  do {                                                \
    int _rc = bhip((b), (call), #call);               \
    if (_rc != GM_OK) return _rc;                     \
  } while (0)

int64_t box_bytes_of(const gm_params& p, size_t esize) {
  const int64_t ncell = (int64_t)(p.nx + 4) * (p.ny + 4) * (p.nz + 4), nq = p.neq;
  return ncell * ((nq + nq + 3) * (int64_t)esize + nq * (int64_t)sizeof(double));
}

gm::Geom geom_of(const gm_params& p) {
  gm::Geom g;
  g.nx = p.nx; g.ny = p.ny; g.nz = p.nz;
  g.NX = p.nx + 4; g.NY = p.ny + 4; g.NZ = p.nz + 4;
  g.ncell = (long long)g.NX * g.NY * g.NZ;
  for (int k = 0; k < 6; k++) g.bc[k] = p.bc[k];
  g.gmem = 0;
  g.npas = p.neq - 8;
  return g;
}

template <typename S>
gm::Box<S> box_view(const Slot& s, int slot) {
  gm::Box<S> b;
  b.g = geom_of(s.p);
  b.ph.cv = s.p.cv;
  b.ph.gamma = s.p.gamma;
  b.ph.mhd = s.p.mhd ? 1 : 0;
  b.ph.pmhd = s.p.pmhd ? 1 : 0;
  b.ph.limiter = s.p.slope_limiter;
  b.ph.eos = GM_EOS_ADIABATIC;
  b.ph.tempsc = 1.0;
  b.U = (S*)s.U;
  b.Us = (S*)s.Us;
  b.W = (double*)s.W;
  b.E = (S*)s.E;
  b.dx = s.p.dx;
  b.dy = s.p.dy;
  b.dz = s.p.dz;
  b.slot = slot;
  b.src = s.p.eight_wave ? gm::SRC_EIGHT_WAVE : 0;
  b.srcin = 0;
  b.usrc = nullptr;
  b.eta = s.p.eta;
  return b;
}

void free_slot(Slot& s) {
  void* bufs[] = {s.U, s.Us, s.W, s.E};
  for (void* p : bufs)
    if (p) (void)hipFree(p);
  s = Slot();
}

void drop_graph(gm_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->graph_dirty = true;
}

// Pack the active boxes and their per-class tile prefix tables for `sh`.
template <typename S>
int rebuild_tables_t(gm_batch* b, const gm::Shape& sh) {
  const int cap = b->bp.max_boxes;
  b->order.clear();
  for (int i = 0; i < cap; i++)
    if (b->slots[i].used && !b->slots[i].clock.done) b->order.push_back(i);
  const int n = (int)b->order.size();
  std::vector<gm::Box<S>> boxes(n);
  std::vector<unsigned> prefix((size_t)gm::KC_COUNT * (cap + 1), 0u);
  for (int k = 0; k < gm::KC_COUNT; k++) {
    unsigned acc = 0;
    for (int i = 0; i < n; i++) {
      prefix[(size_t)k * (cap + 1) + i] = acc;
      const gm::Geom g = geom_of(b->slots[b->order[i]].p);
      const unsigned t = gm::box_tiles(g, k, sh);
      if ((unsigned long long)acc + t > 0xFFFFFFFFull / 2) return bfail(b, GM_ERR_BUSY, "batch exceeds the tile index range");
      acc += t;
    }
    prefix[(size_t)k * (cap + 1) + n] = acc;
    b->tiles[k] = acc;
  }
  b->any_sources = false;
  b->any_viscous = false;
  for (int i = 0; i < n; i++) {
    boxes[i] = box_view<S>(b->slots[b->order[i]], b->order[i]);
    b->any_sources = b->any_sources || boxes[i].src != 0;
    b->any_viscous = b->any_viscous || boxes[i].eta != 0.0;
  }
  BHIP(b, hipStreamSynchronize(b->stream));
  if (n) BHIP(b, hipMemcpy(b->dboxes, boxes.data(), sizeof(gm::Box<S>) * n, hipMemcpyHostToDevice));
  if (n) BHIP(b, hipMemcpy(b->dslots, b->order.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  BHIP(b, hipMemcpy(b->dprefix, prefix.data(), sizeof(unsigned) * prefix.size(), hipMemcpyHostToDevice));
  b->tables_dirty = false;
  return GM_OK;
}

int rebuild_tables(gm_batch* b, const gm::Shape& sh) {
  return b->esize == 8 ? rebuild_tables_t<double>(b, sh) : rebuild_tables_t<float>(b, sh);
}

// Queue `nsteps` batched steps (tuning: dt = 0, clocks untouched).
template <typename S>
void enqueue_steps_t(gm_batch* b, int nsteps, int tuning) {
  const int n = (int)b->order.size();
  if (n == 0) return;
  const int cap = b->bp.max_boxes;
  const gm::Box<S>* boxes = (const gm::Box<S>*)b->dboxes;
  auto pre = [&](int k) { return b->dprefix + (size_t)k * (cap + 1); };
  hipStream_t s = b->stream;
  const gm::Shape& sh = b->shape;
  const int solver = b->bp.riemann_solver, fcd = b->bp.enable_flux_cd;
  const unsigned txy = b->tiles[gm::KC_XY], tcol = b->tiles[gm::KC_COL], tslab = b->tiles[gm::KC_SLAB];
  const int npas = b->neq - 8;
  // With flux-CD and no source or viscous pass, each box's CFL minimum for
  // the next step is folded into the corrector's field update (its clock
  // holds it; gm_batch_add reduces a new box's first one); otherwise a CFL
  // pass starts every step.
  const bool fold = fcd && !b->any_sources && !b->any_viscous;
  for (int st = 0; st < nsteps; st++) {
    if (!fold) {
      gm::launch_clock_begin(s, b->dclk, b->dslots, n);
      gm::launch_cfl_batch(s, boxes, pre(gm::KC_COL), n, tcol, b->dclk, b->derr, sh, npas);
    }
    gm::launch_clock_dt(s, b->dclk, b->dslots, n, tuning);
    for (int stage = 1; stage <= 2; stage++) {
      if (sh.slab && (stage == 1 || sh.slab_both)) {  // the slab kernel, as single boxes run it
        gm::launch_slab_batch(s, boxes, pre(gm::KC_SLAB), n, tslab, b->dclk, stage, solver, fcd, b->derr, sh, npas);
      } else {
        gm::launch_xy_batch(s, boxes, pre(gm::KC_XY), n, txy, b->dclk, stage, solver, fcd, b->derr, npas);
        gm::launch_z_batch(s, boxes, pre(gm::KC_COL), n, tcol, b->dclk, stage, solver, fcd, b->derr, sh, npas);
      }
      if (fold && stage == 2) gm::launch_bfield_cfl_batch(s, boxes, pre(gm::KC_COL), n, tcol, b->dclk, b->derr, sh, npas);
      else if (fcd) gm::launch_bfield_batch(s, boxes, pre(gm::KC_COL), n, tcol, b->dclk, stage, sh);
      if (b->any_sources) gm::launch_sources_batch(s, boxes, pre(gm::KC_COL), n, tcol, b->dclk, stage, sh);
      // viscous_copy (src/hydro_solver.f90:188) after the corrector
      if (stage == 2 && b->any_viscous)
        gm::launch_viscous_batch(s, boxes, pre(gm::KC_COL), n, tcol, b->dclk, sh);
    }
  }
}

void enqueue_steps(gm_batch* b, int nsteps, int tuning) {
  if (b->esize == 8) enqueue_steps_t<double>(b, nsteps, tuning);
  else enqueue_steps_t<float>(b, nsteps, tuning);
}

int capture_graph(gm_batch* b) {
  drop_graph(b);
  BHIP(b, hipStreamSynchronize(b->stream));
  BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  enqueue_steps(b, b->bp.steps_per_launch, 0);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (e != hipSuccess) return bhip(b, e, "hipStreamEndCapture");
  b->graph = g;
  BHIP(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
  b->graph_dirty = false;
  b->cap_nbox = (int)b->order.size();
  std::memcpy(b->cap_tiles, b->tiles, sizeof b->tiles);
  b->cap_shape = b->shape;
  b->cap_sources = b->any_sources;
  b->cap_viscous = b->any_viscous;
  b->cap_neq = b->neq;
  return GM_OK;
}

bool graph_matches(const gm_batch* b) {
  return b->exec && b->cap_nbox == (int)b->order.size() && b->cap_sources == b->any_sources &&
         b->cap_viscous == b->any_viscous && b->cap_neq == b->neq &&
         b->cap_shape.block == b->shape.block && b->cap_shape.slab == b->shape.slab &&
         b->cap_shape.slab_both == b->shape.slab_both &&
         b->cap_shape.kchunk == b->shape.kchunk && std::memcmp(b->cap_tiles, b->tiles, sizeof b->tiles) == 0;
}

// Time candidate shapes on the admitted boxes; keep the fastest: the
// predictor's stage kernel (the two-kernel stage, or the 2.5-D slab kernel
// with 4, 8, 16 or 32 planes per slab) x the column kernels' tile width x
// z-pencil length; then, with the best tile width and pencil length, both
// stages as slab kernels of 1 to 32 planes (what a batch too small to fill
// the GPU runs fastest, as a single small box does).
int autotune(gm_batch* b) {
  static const int slabs[] = {0, 4, 8, 16, 32};
  static const int both_slabs[] = {1, 2, 4, 8, 16, 32};
  static const int blocks[] = {64, 128, 256};
  static const int chunks[] = {8, 16, 32};
  hipEvent_t e0, e1;
  BHIP(b, hipEventCreate(&e0));
  BHIP(b, hipEventCreate(&e1));
  double best = 1e300;
  gm::Shape best_sh = b->shape;
  auto time_shape = [&](const gm::Shape& sh) -> int {
    b->shape = sh;
    int rc = rebuild_tables(b, sh);
    if (rc != GM_OK) return rc;
    enqueue_steps(b, 1, 1);  // warm-up
    (void)hipEventRecord(e0, b->stream);
    enqueue_steps(b, 2, 1);
    (void)hipEventRecord(e1, b->stream);
    BHIP(b, hipEventSynchronize(e1));
    float ms = 0.f;
    BHIP(b, hipEventElapsedTime(&ms, e0, e1));
    if (ms / 2.0 < best) {
      best = ms / 2.0;
      best_sh = sh;
    }
    return GM_OK;
  };
  int rc = GM_OK;
  for (int sl : slabs)
    for (int blk : blocks)
      for (int kc : chunks)
        if (rc == GM_OK) rc = time_shape(gm::Shape{blk, kc, sl, 0});
  const gm::Shape cols = best_sh;
  for (int sl : both_slabs)
    if (rc == GM_OK) rc = time_shape(gm::Shape{cols.block, cols.kchunk, sl, 1});
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != GM_OK) return rc;
  b->shape = best_sh;
  b->tuned = true;
  b->tuned_ms = best;
  b->tables_dirty = true;
  b->graph_dirty = true;
  return GM_OK;
}

void free_zombies(gm_batch* b) {
  for (void* p : b->zombies) (void)hipFree(p);
  b->zombies.clear();
}

// The launches in flight have finished: their boxes' clocks from the
// read-back, boxes that completed leave the next launch's table, buffers
// released meanwhile are freed; a deferred device error is reported.
int finish_launch(gm_batch* b) {
  if (b->in_flight) {
    b->in_flight = false;
    for (int i : b->launched) {
      if (!b->slots[i].used) continue;
      b->slots[i].clock = b->hclk[i];
      if (b->hclk[i].done) b->tables_dirty = true;
    }
    b->launched.clear();
  }
  free_zombies(b);
  if (*b->herr) {
    const int e = *b->herr;
    *b->herr = 0;
    BHIP(b, hipMemset(b->derr, 0, sizeof(int)));
    return bfail(b, GM_ERR_NONFINITE,
                 (e & 1) ? "Riemann solver fell through on a non-finite state (hlld.f90:316-317)"
                         : "non-finite value in a box's conserved state");
  }
  return GM_OK;
}

// Wait for the launches in flight (only if some are).
int wait_launch(gm_batch* b) {
  if (!b->in_flight) return GM_OK;
  BHIP(b, hipStreamSynchronize(b->stream));
  return finish_launch(b);
}

bool in_launch(const gm_batch* b, int slot) {
  if (!b->in_flight) return false;
  for (int i : b->launched)
    if (i == slot) return true;
  return false;
}

bool valid_id(const gm_batch* b, int32_t id) {
  return id >= 0 && id < b->bp.max_boxes && b->slots[id].used;
}

template <typename S>
void ghost_fill_t(hipStream_t st, const Slot& s) {
  const gm::Geom g = geom_of(s.p);
  S* U = (S*)s.U;
  for (int a = 0; a < 3; a++)
    if (g.bc[2 * a] == GM_BC_PERIODIC) {
      gm::launch_ghost(st, g, U, a, 0, GM_BC_PERIODIC, 1);
      gm::launch_ghost(st, g, U, a, 1, GM_BC_PERIODIC, 1);
    }
  for (int kind : {GM_BC_CLOSED, GM_BC_OUTFLOW})
    for (int f = 0; f < 6; f++)
      if (g.bc[f] == kind) gm::launch_ghost(st, g, U, f / 2, f % 2, kind, 1);
}

}  // namespace

extern "C" {

int gm_batch_create(const gm_batch_params* bp, gm_batch** out) {
  if (!out) return bfail(nullptr, GM_ERR_ARG, "batch out-pointer is NULL");
  *out = nullptr;
  if (!bp) return bfail(nullptr, GM_ERR_ARG, "params is NULL");
  if (bp->precision != GM_PRECISION_FP64 && bp->precision != GM_PRECISION_MIXED)
    return bfail(nullptr, GM_ERR_CONFIG, "precision must be GM_PRECISION_FP64 or GM_PRECISION_MIXED");
  if (bp->riemann_solver < GM_SOLVER_HLL || bp->riemann_solver > GM_SOLVER_HLLD)
    return bfail(nullptr, GM_ERR_CONFIG, "riemann_solver out of range");
  if (bp->max_boxes < 1 || bp->max_boxes > kMaxBoxes)
    return bfail(nullptr, GM_ERR_CONFIG, "max_boxes must be in 1..4096");
  if (bp->steps_per_launch < 1 || bp->steps_per_launch > 1000)
    return bfail(nullptr, GM_ERR_CONFIG, "steps_per_launch must be in 1..1000");
  if (bp->reserved != 0 || bp->memory_budget < 0) return bfail(nullptr, GM_ERR_CONFIG, "bad reserved / budget field");
  gm_batch* b = new gm_batch();
  b->bp = *bp;
  // stage kernels before (or without) autotuning: the two-kernel stage, or
  // GUACHO_AMD_BATCH_SLAB planes per slab of the 2.5-D kernel for the
  // predictor (negative: for both stages)
  if (const char* e = std::getenv("GUACHO_AMD_BATCH_SLAB")) {
    const long v = std::strtol(e, nullptr, 10);
    if (v >= -(1 << 20) && v <= (1 << 20)) {
      b->shape.slab = (int)(v < 0 ? -v : v);
      b->shape.slab_both = v < 0 ? 1 : 0;
    }
  }
  b->esize = bp->precision == GM_PRECISION_MIXED ? 4 : 8;
  b->slots.resize(bp->max_boxes);
  auto bail = [&](int code) {
    g_batch_create_error = b->msg;
    gm_batch_destroy(b);
    return code;
  };
  int rc;
  if ((rc = bhip(b, hipSetDevice(bp->device), "hipSetDevice")) != GM_OK) return bail(rc);
  if ((rc = bhip(b, hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate")) != GM_OK)
    return bail(rc);
  if ((rc = bhip(b, hipStreamCreateWithFlags(&b->xfer, hipStreamNonBlocking), "hipStreamCreate")) != GM_OK)
    return bail(rc);
  if ((rc = bhip(b, hipEventCreateWithFlags(&b->ev_launch, hipEventDisableTiming), "hipEventCreate")) != GM_OK)
    return bail(rc);
  size_t freeb = 0, totalb = 0;
  if ((rc = bhip(b, hipMemGetInfo(&freeb, &totalb), "hipMemGetInfo")) != GM_OK) return bail(rc);
  b->budget = bp->memory_budget > 0 ? bp->memory_budget : (int64_t)(0.8 * (double)freeb);
  const int cap = bp->max_boxes;
  if ((rc = bhip(b, hipMalloc(&b->dboxes, sizeof(gm::Box<double>) * cap), "hipMalloc")) != GM_OK) return bail(rc);
  if ((rc = bhip(b, hipMalloc((void**)&b->dprefix, sizeof(unsigned) * gm::KC_COUNT * (cap + 1)), "hipMalloc")) !=
      GM_OK)
    return bail(rc);
  if ((rc = bhip(b, hipMalloc((void**)&b->dclk, sizeof(gm::BoxClock) * cap), "hipMalloc")) != GM_OK) return bail(rc);
  if ((rc = bhip(b, hipMalloc((void**)&b->dslots, sizeof(int) * cap), "hipMalloc")) != GM_OK) return bail(rc);
  if ((rc = bhip(b, hipMalloc((void**)&b->derr, sizeof(int)), "hipMalloc")) != GM_OK) return bail(rc);
  if ((rc = bhip(b, hipHostMalloc((void**)&b->hclk, sizeof(gm::BoxClock) * cap, hipHostMallocDefault),
                 "hipHostMalloc")) != GM_OK)
    return bail(rc);
  if ((rc = bhip(b, hipHostMalloc((void**)&b->herr, sizeof(int), hipHostMallocDefault), "hipHostMalloc")) != GM_OK)
    return bail(rc);
  // every slot starts as a completed, empty clock
  for (int i = 0; i < cap; i++) {
    b->hclk[i] = gm::BoxClock();
    b->hclk[i].done = 1;
  }
  if ((rc = bhip(b, hipMemcpy(b->dclk, b->hclk, sizeof(gm::BoxClock) * cap, hipMemcpyHostToDevice), "hipMemcpy")) !=
      GM_OK)
    return bail(rc);
  if ((rc = bhip(b, hipMemset(b->derr, 0, sizeof(int)), "hipMemset")) != GM_OK) return bail(rc);
  *b->herr = 0;
  *out = b;
  return GM_OK;
}

int gm_batch_destroy(gm_batch* b) {
  if (!b) return GM_ERR_ARG;
  (void)hipSetDevice(b->bp.device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->xfer) (void)hipStreamSynchronize(b->xfer);
  drop_graph(b);
  free_zombies(b);
  for (auto& s : b->slots)
    if (s.used) free_slot(s);
  if (b->dboxes) (void)hipFree(b->dboxes);
  if (b->dslots) (void)hipFree(b->dslots);
  if (b->ev_launch) (void)hipEventDestroy(b->ev_launch);
  if (b->xfer) (void)hipStreamDestroy(b->xfer);
  if (b->dprefix) (void)hipFree(b->dprefix);
  if (b->dclk) (void)hipFree(b->dclk);
  if (b->derr) (void)hipFree(b->derr);
  if (b->hclk) (void)hipHostFree(b->hclk);
  if (b->herr) (void)hipHostFree(b->herr);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return GM_OK;
}

const char* gm_batch_last_error(const gm_batch* b) { return b ? b->msg.c_str() : g_batch_create_error.c_str(); }

int gm_batch_box_bytes(const gm_batch* b, const gm_params* p, int64_t* bytes) {
  if (!b || !p || !bytes) return GM_ERR_ARG;
  *bytes = box_bytes_of(*p, b->esize);
  return GM_OK;
}

int gm_batch_free_bytes(const gm_batch* b, int64_t* bytes) {
  if (!b || !bytes) return GM_ERR_ARG;
  *bytes = b->budget - b->used;
  return GM_OK;
}

int gm_batch_add(gm_batch* b, const gm_params* p, const double* u, double cfl, double t_end, int64_t max_steps,
                 int32_t* id) {
  if (!b) return GM_ERR_ARG;
  if (!p || !u || !id) return bfail(b, GM_ERR_ARG, "params, u or id is NULL");
  std::string m;
  int rc = gm::check_params(p, m);
  if (rc != GM_OK) return bfail(b, rc, m);
  for (int f = 0; f < 6; f++)
    if (p->bc[f] == GM_BC_OTHER)
      return bfail(b, GM_ERR_CONFIG, "BC_OTHER (user boundary hooks) is for single-box contexts (gm_set_user_bc)");
  int active = 0;
  for (const Slot& q : b->slots) active += q.used ? 1 : 0;
  if (active && p->neq != b->neq)
    return bfail(b, GM_ERR_CONFIG, "every box of a batch has the same passive scalars (neq)");
  if (p->precision != b->bp.precision || p->riemann_solver != b->bp.riemann_solver ||
      (p->enable_flux_cd != 0) != (b->bp.enable_flux_cd != 0))
    return bfail(b, GM_ERR_CONFIG, "box precision, solver and flux-CD must match the batch");
  if (!(cfl > 0) || !std::isfinite(cfl)) return bfail(b, GM_ERR_ARG, "cfl must be positive");
  if (!(t_end > 0)) return bfail(b, GM_ERR_ARG, "t_end must be positive (use +inf for no time limit)");
  if (max_steps < 0 || max_steps > INT32_MAX)
    return bfail(b, GM_ERR_ARG, "max_steps must be in 0..2147483647");
  if (std::isinf(t_end) && max_steps == 0) return bfail(b, GM_ERR_ARG, "a box needs t_end or max_steps");
  int slot = -1;
  for (int i = 0; i < b->bp.max_boxes; i++)
    if (!b->slots[i].used) {
      slot = i;
      break;
    }
  if (slot < 0) return bfail(b, GM_ERR_BUSY, "box table full");
  const int64_t need = box_bytes_of(*p, b->esize);
  if (b->used + need > b->budget) return bfail(b, GM_ERR_BUSY, "memory budget exhausted");
  BHIP(b, hipSetDevice(b->bp.device));
  // Admission runs on the transfer stream, beside any launch in flight (the
  // new slot is in none of their tables): allocation, upload, clock and the
  // first CFL minimum; it has completed when this call returns.
  hipStream_t x = b->xfer;
  Slot s;
  s.p = *p;
  s.p.device = b->bp.device;
  const gm::Geom g = geom_of(*p);
  const size_t nq = (size_t)p->neq;
  const size_t nu = (size_t)g.ncell * nq * b->esize, n3 = (size_t)g.ncell * 3 * b->esize;
  const size_t nw = (size_t)g.ncell * nq * sizeof(double);
  struct { void** ptr; size_t n; } al[] = {{&s.U, nu}, {&s.Us, nu}, {&s.W, nw}, {&s.E, n3}};
  for (auto& a : al) {
    if (hipMalloc(a.ptr, a.n) != hipSuccess) {
      (void)hipGetLastError();
      free_slot(s);
      return bfail(b, GM_ERR_OOM, "device allocation failed");
    }
    (void)hipMemsetAsync(*a.ptr, 0, a.n, x);
  }
  s.bytes = need;
  // upload: host AoS doubles -> staging (partials block) -> SoA state
  BHIP(b, hipMemcpyAsync(s.W, u, (size_t)g.ncell * nq * sizeof(double), hipMemcpyHostToDevice, x));
  if (b->esize == 8) gm::launch_aos_to_soa(x, g.ncell, (const double*)s.W, (double*)s.U, (int)nq);
  else gm::launch_aos_to_soa(x, g.ncell, (const double*)s.W, (float*)s.U, (int)nq);
  BHIP(b, hipGetLastError());
  gm::BoxClock c{};
  c.dtmin = (unsigned long long)gm_bits_of(1e30);
  c.dt = 0.0;
  c.time = 0.0;
  c.t_end = t_end;
  c.cfl = cfl;
  c.it = 1;
  c.n_iter = 10;
  c.steps = 0;
  c.max_steps = (int)max_steps;
  c.done = 0;
  BHIP(b, hipMemcpyAsync(b->dclk + slot, &c, sizeof(gm::BoxClock), hipMemcpyHostToDevice, x));
  // the first step's CFL minimum (get_timestep over the uploaded state),
  // which the batch's field update folds in for every later step
  if (b->esize == 8) gm::launch_cfl(x, box_view<double>(s, slot), &b->dclk[slot].dtmin, b->derr, gm::Shape{256, 16});
  else gm::launch_cfl(x, box_view<float>(s, slot), &b->dclk[slot].dtmin, b->derr, gm::Shape{256, 16});
  BHIP(b, hipGetLastError());
  BHIP(b, hipStreamSynchronize(x));
  s.used = true;
  s.clock = c;
  b->slots[slot] = s;
  b->neq = p->neq;
  b->used += need;
  b->tables_dirty = true;
  b->graph_dirty = true;
  *id = slot;
  return GM_OK;
}

int gm_batch_remove(gm_batch* b, int32_t id) {
  if (!b) return GM_ERR_ARG;
  if (!valid_id(b, id)) return bfail(b, GM_ERR_ARG, "no such box");
  BHIP(b, hipSetDevice(b->bp.device));
  int rc;
  // a box that a launch in flight steps is released once it finishes
  if (in_launch(b, id) && (rc = wait_launch(b)) != GM_OK) return rc;
  gm::BoxClock c{};
  c.done = 1;
  BHIP(b, hipMemcpyAsync(b->dclk + id, &c, sizeof(gm::BoxClock), hipMemcpyHostToDevice, b->xfer));
  BHIP(b, hipStreamSynchronize(b->xfer));
  b->used -= b->slots[id].bytes;
  Slot& sl = b->slots[id];
  if (b->in_flight) {  // freeing could wait for the launches in flight
    for (void* q : {sl.U, sl.Us, sl.W, sl.E})
      if (q) b->zombies.push_back(q);
    sl = Slot();
  } else {
    free_slot(sl);
  }
  b->tables_dirty = true;
  b->graph_dirty = true;
  return GM_OK;
}

int gm_batch_advance(gm_batch* b, int32_t nlaunch) {
  if (!b) return GM_ERR_ARG;
  if (nlaunch < 1) return bfail(b, GM_ERR_ARG, "nlaunch must be >= 1");
  BHIP(b, hipSetDevice(b->bp.device));
  int rc;
  // new tables or a new graph need the launches in flight to have finished
  // (their nodes read the tables)
  if ((b->tables_dirty || b->graph_dirty || (b->bp.autotune && !b->tuned)) && (rc = wait_launch(b)) != GM_OK)
    return rc;
  int active = 0;
  for (auto& s : b->slots) active += (s.used && !s.clock.done) ? 1 : 0;
  if (active == 0) return GM_OK;
  if (b->bp.autotune && !b->tuned) {
    if ((rc = autotune(b)) != GM_OK) return rc;
  }
  if (b->tables_dirty) {
    if ((rc = rebuild_tables(b, b->shape)) != GM_OK) return rc;
    b->graph_dirty = !graph_matches(b);
  }
  if (b->graph_dirty) {
    if ((rc = capture_graph(b)) != GM_OK) return rc;
  }
  for (int i = 0; i < nlaunch; i++) BHIP(b, hipGraphLaunch(b->exec, b->stream));
  // the clocks and the error flag come back behind the launches
  BHIP(b, hipMemcpyAsync(b->hclk, b->dclk, sizeof(gm::BoxClock) * b->bp.max_boxes, hipMemcpyDeviceToHost,
                         b->stream));
  BHIP(b, hipMemcpyAsync(b->herr, b->derr, sizeof(int), hipMemcpyDeviceToHost, b->stream));
  BHIP(b, hipEventRecord(b->ev_launch, b->stream));
  b->in_flight = true;
  b->launched = b->order;
  return GM_OK;
}

int gm_batch_sync(gm_batch* b) {
  if (!b) return GM_ERR_ARG;
  BHIP(b, hipSetDevice(b->bp.device));
  BHIP(b, hipStreamSynchronize(b->stream));
  BHIP(b, hipStreamSynchronize(b->xfer));
  return finish_launch(b);
}

int gm_batch_query(gm_batch* b, int32_t* idle) {
  if (!b || !idle) return GM_ERR_ARG;
  BHIP(b, hipSetDevice(b->bp.device));
  *idle = 1;
  if (!b->in_flight) return GM_OK;
  const hipError_t e = hipEventQuery(b->ev_launch);
  if (e == hipErrorNotReady) {
    *idle = 0;
    return GM_OK;
  }
  BHIP(b, e);
  return finish_launch(b);
}

int gm_batch_clock(gm_batch* b, int32_t id, double* time, int64_t* steps, int32_t* done) {
  if (!b) return GM_ERR_ARG;
  if (!valid_id(b, id)) return bfail(b, GM_ERR_ARG, "no such box");
  const gm::BoxClock& c = b->slots[id].clock;
  if (time) *time = c.time;
  if (steps) *steps = c.steps;
  if (done) *done = c.done;
  return GM_OK;
}

int gm_batch_download_u(gm_batch* b, int32_t id, double* u) {
  if (!b) return GM_ERR_ARG;
  if (!valid_id(b, id)) return bfail(b, GM_ERR_ARG, "no such box");
  if (!u) return bfail(b, GM_ERR_ARG, "u is NULL");
  BHIP(b, hipSetDevice(b->bp.device));
  int rc;
  // a box that a launch in flight steps is read once it finishes; any other
  // box is read on the transfer stream, beside the launches
  if (in_launch(b, id) && (rc = wait_launch(b)) != GM_OK) return rc;
  const Slot& s = b->slots[id];
  const gm::Geom g = geom_of(s.p);
  hipStream_t x = b->xfer;
  if (b->esize == 8) {
    ghost_fill_t<double>(x, s);
    gm::launch_soa_to_aos(x, g.ncell, (const double*)s.U, (double*)s.W, s.p.neq);
  } else {
    ghost_fill_t<float>(x, s);
    gm::launch_soa_to_aos(x, g.ncell, (const float*)s.U, (double*)s.W, s.p.neq);
  }
  BHIP(b, hipGetLastError());
  BHIP(b, hipMemcpyAsync(u, s.W, (size_t)g.ncell * s.p.neq * sizeof(double), hipMemcpyDeviceToHost, x));
  BHIP(b, hipStreamSynchronize(x));
  return GM_OK;
}

int gm_batch_tuning(const gm_batch* b, int32_t* block, int32_t* kchunk, int32_t* slab, double* ms_per_step) {
  if (!b) return GM_ERR_ARG;
  if (block) *block = b->shape.block;
  if (kchunk) *kchunk = b->shape.kchunk;
  if (slab) *slab = b->shape.slab_both ? -b->shape.slab : b->shape.slab;
  if (ms_per_step) *ms_per_step = b->tuned ? b->tuned_ms : 0.0;
  return GM_OK;
}

int gm_batch_active(const gm_batch* b, int32_t* n) {
  if (!b || !n) return GM_ERR_ARG;
  int a = 0;
  for (auto& s : b->slots) a += s.used ? 1 : 0;
  *n = a;
  return GM_OK;
}

}  // extern "C"

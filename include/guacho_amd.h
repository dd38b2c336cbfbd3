/*
 * guacho_amd.h -- C ABI of the MI355X Godunov MHD step library
 * (guacho_amd/libguacho_amd.so).
 *
 * The reference has no FFI: its hot path is Fortran called from the time
 * loop src/main.f90:94-125.  Each entry point below replaces one reference
 * routine on that path; a Fortran host binds them with ISO_C_BINDING
 * (host/guacho_capi.f90) and Python with ctypes (guacho_amd/_capi.py).
 *
 * Conventions
 *  - Host arrays are caller-allocated in the reference's globals layout,
 *    Fortran column-major u(neq, nxmin:nxmax, nymin:nymax, nzmin:nzmax)
 *    with neq (8 plus the passive scalars) fastest and nghost = 2 (src/globals.f90:33-38,
 *    src/init.f90:145-151).  The library stores structure-of-arrays in
 *    device memory and transposes on upload/download.
 *  - Every call returns GM_OK (0) or a positive GM_ERR_* code; the message
 *    is available from gm_last_error().  Nothing aborts the process.
 *  - One host thread per context.  Calls are ordered on the context's HIP
 *    stream; downloads and gm_cfl_dtp synchronize with it.
 */
#ifndef GUACHO_AMD_H
#define GUACHO_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GM_OK 0
#define GM_ERR_CONFIG 1    /* unsupported or invalid configuration */
#define GM_ERR_HIP 2       /* HIP runtime error */
#define GM_ERR_NONFINITE 3 /* non-finite state; mirrors the HLLD/HLLC
                              fall-through stop at hlld.f90:316-317,
                              hllc.f90:135-138 */
#define GM_ERR_OOM 4       /* device allocation failed */
#define GM_ERR_ARG 5       /* bad argument (null pointer, size) */

/* Named constants: same values as src/constants.f90:56-93. */
#define GM_SOLVER_HLL 1
#define GM_SOLVER_HLLC 2
#define GM_SOLVER_HLLE 3
#define GM_SOLVER_HLLD 4
#define GM_BC_OUTFLOW 1
#define GM_BC_CLOSED 2
#define GM_BC_PERIODIC 3
#define GM_BC_OTHER 4
#define GM_LIMITER_NO_AVERAGE (-1)
#define GM_LIMITER_NO_LIMIT 0
#define GM_LIMITER_MINMOD 1
#define GM_LIMITER_VAN_LEER 2
#define GM_LIMITER_VAN_ALBADA 3
#define GM_LIMITER_UMIST 4
#define GM_LIMITER_WOODWARD 5
#define GM_LIMITER_SUPERBEE 6

#define GM_PRECISION_FP64 0  /* the drop-in: bitwise the reference's arithmetic */
#define GM_PRECISION_MIXED 1 /* FP32 states/fluxes, FP64 pressure recovery,
                                update accumulation and reductions */

/* Run configuration: the subset of the reference's parameters module
 * (OT/parameters.f90:48-177) that the step reads, plus the grid spacing
 * computed by initmain (src/init.f90:119-122). */
typedef struct gm_params {
  int32_t nx, ny, nz;     /* interior cells of the subdomain */
  int32_t neq;            /* 8 (BFIELD) plus the passive scalars (PASSIVES
                             builds: npas = 1 or 2), so 8, 9 or 10; the
                             boxes of one batch share it */
  int32_t nghost;         /* must be 2 */
  int32_t riemann_solver; /* GM_SOLVER_* */
  int32_t slope_limiter;  /* GM_LIMITER_* */
  int32_t bc[6];          /* bc_left, bc_right, bc_bottom, bc_top, bc_out, bc_in
                             (GM_BC_OTHER: the user boundary hook's face) */
  int32_t mhd;            /* full MHD */
  int32_t pmhd;           /* passive B (HLLC) */
  int32_t enable_flux_cd; /* flux-CD div B control */
  int32_t eight_wave;     /* Powell 8-wave div B source terms (src/sources.f90:124-175) */
  int32_t precision;      /* GM_PRECISION_* */
  int32_t device;         /* HIP device ordinal */
  int32_t reserved;       /* must be 0 */
  double dx, dy, dz;
  double cv;              /* specific heat at constant volume */
  double gamma;           /* (cv+1)/cv, passed as the caller computed it */
  double eta;             /* artificial viscosity of viscous_copy
                             (src/hydro_solver.f90:47-65); finite; per box
                             in batches */
} gm_params;

typedef struct gm_ctx gm_ctx;

/* Allocate device state for one subdomain.  Replaces the allocations of
 * initmain (src/init.f90:144-159) and validates the configuration the way
 * its report does (src/init.f90:209-404). */
int gm_create(const gm_params *params, gm_ctx **ctx);

/* Free the context and its device memory. */
int gm_destroy(gm_ctx *ctx);

/* Message for the last error on ctx (or of the last failed gm_create when
 * ctx is NULL).  Valid until the next call on the same context. */
const char *gm_last_error(const gm_ctx *ctx);

/* Copy the full conserved array, ghosts included, host -> device.
 * Replaces writes to globals::u by initflow (src/init.f90:432-477). */
int gm_upload_u(gm_ctx *ctx, const double *u);

/* Copy the full conserved array, ghosts included, device -> host, after
 * the one-layer ghost fill of boundaryI.  Replaces reads of globals::u by
 * write_BIN (src/Out_BIN_Module.f90:162).  Synchronizes. */
int gm_download_u(gm_ctx *ctx, double *u);

/* Primitives (and temperature when temp != NULL) of every cell, ghosts
 * included: calcprim(u, primit) (src/hydro_core.f90:245-320).  temp has
 * the layout Temp(nxmin:nxmax, nymin:nymax, nzmin:nzmax).  Synchronizes. */
int gm_download_primit(gm_ctx *ctx, double *primit, double *temp);

/* One-layer ghost fill of the conserved state: boundaryI
 * (src/boundaries.f90:44-247).  gm_tstep applies boundary conditions
 * itself; this is for callers that change u and need the ghosts. */
int gm_boundary_u(gm_ctx *ctx);

/* The minimum over interior cells of dx/(|vx|+cfx), dy/(|vy|+cfy),
 * dz/(|vz|+cfz) (src/hydro_core.f90:642-675; csound for pmhd).  The host
 * applies the CFL ramp, the MIN allreduce and the dump clamp of
 * hydro_core.f90:677-694.  Returns GM_ERR_NONFINITE if any interior cell
 * holds a non-finite value.  Synchronizes.  Right after a gm_tstep with
 * flux-CD and no sources, viscosity, cooling, conduction or boundary hook,
 * the minimum was already reduced by that step's field update and this
 * call only waits for it; any upload, ghost fill with a hook or EOS change
 * in between makes it reduce over the stored state again. */
int gm_cfl_dtp(gm_ctx *ctx, double *dtp);

/* Advance the state by dt: tstep (src/hydro_solver.f90:134-229), with the
 * source terms, viscosity, table or H-rate cooling, thermal conduction and
 * user plugins the context has switched on; chemistry and diffuse radiation
 * are not part of the library.  Asynchronous; a non-finite flux sets an
 * error reported by the next gm_cfl_dtp or gm_sync. */
int gm_tstep(gm_ctx *ctx, double dt);

/* Wait for queued work; report any deferred device error. */
int gm_sync(gm_ctx *ctx);

/* No counterpart in the reference.  For callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc): one gm_tstep plus the next
 * gm_cfl_dtp's minimum, on a caller-held device state, with no context and
 * no allocation.  U*, the stage partials, E and the device scalars live in
 * one file-scope device array of 256 MiB that each GPU running this call
 * holds until the process ends; a grid needing more (FP64: about 117^3
 * cells, mixed: 131^3) is refused with GM_ERR_CONFIG before any device
 * work.  Calls that overlap on one GPU share that array and race.
 *   params    as for gm_create; BC_OTHER faces are refused (they need a
 *             context's user boundary plugin); EOS adiabatic
 *   u         device pointer to the state in the library's device layout:
 *             variable q of cell (i,j,k) (1-based, two ghost layers) at
 *             u[q*ncell + ((k+1)*(ny+4) + (j+1))*(nx+4) + (i+1)],
 *             ncell = (nx+4)(ny+4)(nz+4); doubles for GM_PRECISION_FP64,
 *             floats for GM_PRECISION_MIXED.  Ghost cells are neither read
 *             nor written (boundary mapping, as in gm_tstep).
 *   stream    hipStream_t to launch on, NULL for the null stream; the call
 *             only enqueues kernels (no memset, copy or synchronisation)
 *   dtp_next  device double or NULL: get_timestep's minimum
 *             (src/hydro_core.f90:642-675) over the new state
 *   err       device int32 or NULL: 0, or the kernels' error bits (1: a
 *             Riemann solver fell through on a non-finite state)
 * Errors before any launch: GM_ERR_CONFIG / GM_ERR_ARG (message in
 * gm_last_error(NULL)). */
int gm_tstep_noalloc(const gm_params *params, void *u, double dt, void *stream, double *dtp_next,
                     int32_t *err);

/* Divergence of B with the dump formula of write_BIN
 * (src/Out_BIN_Module.f90:214-222), divb(nx, ny, nz) column-major. */
int gm_download_divb(gm_ctx *ctx, double *divb);

/* ---- measurement ---------------------------------------------------- */
/* Kernel timing: while enabled, gm_tstep records HIP events around each
 * kernel on the context's stream.  gm_kernel_times returns, per kernel
 * class, the accumulated milliseconds and launch count since the last
 * reset, and the kernel names (static strings). */
#define GM_NKCLASS 15
int gm_set_timing(gm_ctx *ctx, int enable);
int gm_kernel_times(gm_ctx *ctx, double *ms, int64_t *launches, const char **names);

/* Device memory held by the context, in bytes. */
int gm_device_bytes(const gm_ctx *ctx, int64_t *bytes);

/* ---- user plugins (guacho_amd_user.h) -------------------------------- */
/* The simulation time the user hooks see (the reference's globals::time,
 * which its driver advances after each tstep, src/main.f90:94-125). */
int gm_set_time(gm_ctx *ctx, double time);

/* User source terms: replaces get_user_source_terms (called from source(),
 * src/sources.f90:206, with user_source_terms = .true.).  Loads the gfx950
 * code object at `code_object` and its kernel `kernel` (signature in
 * guacho_amd_user.h); each stage of gm_tstep zeroes s, launches the kernel
 * over the interior cells, adds the 8-wave terms when eight_wave is set and
 * applies up = up + dt*s (src/hydro_solver.f90:115-121).  code_object NULL
 * removes the plugin.  GM_ERR_CONFIG if the file or kernel cannot be
 * loaded. */
int gm_set_user_sources(gm_ctx *ctx, const char *code_object, const char *kernel);

/* User boundary hook: replaces impose_user_bc(u, order) (bc_user = .true.;
 * called at the end of boundaryI with order 1 and of boundaryII with order
 * 2, src/boundaries.f90:245, 508).  Loads `kernel` from the gfx950 code
 * object (signature in guacho_amd_user.h; launched over every stored cell,
 * ghosts included).  With a hook installed the ghost layers live in memory,
 * as in the reference: gm_boundary_u runs boundaryI then the hook (order 1);
 * each gm_tstep runs boundaryII then the hook (order 2) on U* between its
 * stages and boundaryI then the hook (order 1) on U at its end
 * (src/hydro_solver.f90:169, 216).  Faces of kind GM_BC_OTHER are left to
 * the hook: no standard fill touches them, and their E ghosts are read as
 * stored (boundaryI_ef has no branch for them).  A step with a GM_BC_OTHER
 * face and no hook fails with GM_ERR_CONFIG.  code_object NULL removes the
 * hook.  Hooks are not available in batches. */
int gm_set_user_bc(gm_ctx *ctx, const char *code_object, const char *kernel);

/* ---- post-step physics ------------------------------------------------ */
#define GM_COOL_NONE 0   /* constants.f90:72-77 */
#define GM_COOL_H    1   /* H-rate cooling: needs neq >= 9 (the neutral density) */
#define GM_COOL_DMC  3
#define GM_COOL_CHI  4

/* Cooling at the end of every gm_tstep, after the corrector and before the
 * boundary fill: coolingdmc (src/cooling_dmc.f90:115-156) for GM_COOL_DMC,
 * coolingchi (src/cooling_chi.f90) for GM_COOL_CHI, coolingh
 * (src/cooling_h.f90:41-67: the hydrogen rate equation on the neutral
 * density, the first passive scalar, and the parametrised loss ALOSS) for
 * GM_COOL_H, which takes no table (tempsc and tsc as below; psc unused) and
 * needs neq >= 9.  cooltab
 * holds the reference's cooltab(2, 41) as its read_table leaves it (column
 * major: T_1, Lambda_1, T_2, Lambda_2, ...; for DMC already 10**a,
 * 10**(-b)); tempsc, psc, tsc are the parameters module's Tempsc, Psc and
 * tsc.  GM_COOL_NONE turns cooling off (cooltab may then be NULL).  As
 * the reference keeps the cooled pressure in primit for the next step
 * (cooling_dmc.f90:145-148; tstep then recomputes only the ghost
 * primitives), the library keeps each interior cell's pressure after
 * GM_COOL_DMC / GM_COOL_CHI for the next step's CFL minimum, predictor and
 * gm_download_primit (one more double per cell, allocated here on the first
 * table setting; GM_ERR_OOM if that fails).  The device's exp/log10 are
 * within 1 ulp of the host's, so cooled runs match the reference to a
 * tolerance (1e-14 of each field's scale over 5 steps), not bit for bit. */
int gm_set_cooling(gm_ctx *ctx, int32_t kind, const double *cooltab, double tempsc, double psc, double tsc);

#define GM_TC_OFF         0   /* constants.f90:96-98 */
#define GM_TC_ISOTROPIC   1
#define GM_TC_ANISOTROPIC 2

/* Thermal conduction: th_cond and tc_saturation of the parameters module,
 * and its scalings (Tempsc, rsc, rhosc, mu, Psc, tsc, vsc2, bsc). */
typedef struct gm_conduction_params {
  int32_t kind;        /* GM_TC_* */
  int32_t saturation;  /* tc_saturation */
  double tempsc, rsc, rhosc, mu, psc, tsc, vsc2, bsc;
} gm_conduction_params;

/* Thermal conduction at the end of every gm_tstep, after the boundary fill
 * (thermal_conduction, src/thermal_cond.f90:690-768): dt_cond from the
 * Spitzer coefficient (:78-110), super-time-stepping when it is shorter than
 * the step (:625-681), per sub-step the isotropic (:189-267) or
 * field-aligned (:277-487) heat fluxes, optionally saturated, the update of
 * E, and thermal_bounds' E ghosts (:496-616).  With conduction on, ghosts
 * live in memory (as with a user boundary hook) and gm_tstep waits for the
 * dt_cond reduction.  pow and sqrt are the device's: conducting runs match
 * the reference to a tolerance.  kind GM_TC_OFF turns it off. */
int gm_set_conduction(gm_ctx *ctx, const gm_conduction_params *params);

/* Sub-steps the last gm_tstep's conduction took. */
int gm_conduction_substeps(const gm_ctx *ctx, int32_t *nsteps);

#define GM_EOS_ADIABATIC     1   /* constants.f90 EOS_ADIABATIC */
#define GM_EOS_SINGLE_SPECIE 2   /* constants.f90 EOS_SINGLE_SPECIE */
#define GM_EOS_H_RATE        3   /* constants.f90 EOS_H_RATE (needs npas >= 1) */

/* Equation of state of every primitive recovery, u2prim
 * (src/hydro_core.f90:46-129) with eq_of_state and Tempsc of the parameters
 * module.  GM_EOS_ADIABATIC (the default): T = (p / rho) Tempsc (:90-92).
 * GM_EOS_SINGLE_SPECIE (:94-99): T = max(1, (p / rho) Tempsc) and the
 * pressure becomes rho T / Tempsc -- a temperature floor of 1 K that the
 * step, the CFL reduction, the source terms, cooling, conduction,
 * gm_download_primit's temperature and plugins (gm_user_u2prim) all see.
 * The other EOS need passive scalars and are not built (GM_ERR_CONFIG).
 * Before the first call Tempsc is 1. */
int gm_set_eos(gm_ctx *ctx, int32_t eq_of_state, double tempsc);

/* ---- batches of independent boxes ----------------------------------- */
/* No counterpart in the reference (one domain per run there).  A batch
 * advances many independent subdomains ("boxes") of uneven size together:
 * every kernel of a step covers the tiles of all boxes in one launch, each
 * box keeps its own CFL dt (ramp and clamp of hydro_core.f90:677-694 on the
 * device), and the steps are captured in a HIP graph replayed per advance.
 * A box's arithmetic does not depend on the other boxes: results are the
 * same alone and in any mix. */
#define GM_ERR_BUSY 6      /* no room: memory budget or box table full */

typedef struct gm_batch_params {
  int32_t precision;        /* GM_PRECISION_*, for every box */
  int32_t riemann_solver;   /* one solver per batch */
  int32_t enable_flux_cd;   /* one setting per batch */
  int32_t device;
  int32_t max_boxes;        /* capacity of the box table (<= 4096) */
  int32_t steps_per_launch; /* steps captured in one graph launch (>= 1) */
  int32_t autotune;         /* 1: time stage-kernel x tile width x z-pencil candidates on the first advance */
  int32_t reserved;         /* must be 0 */
  int64_t memory_budget;    /* device bytes the boxes may hold; 0: 80% of free memory */
} gm_batch_params;

typedef struct gm_batch gm_batch;

int gm_batch_create(const gm_batch_params *params, gm_batch **batch);
int gm_batch_destroy(gm_batch *batch);
const char *gm_batch_last_error(const gm_batch *batch);
/* device bytes one box of this configuration holds */
int gm_batch_box_bytes(const gm_batch *batch, const gm_params *box, int64_t *bytes);
/* budget left for new boxes */
int gm_batch_free_bytes(const gm_batch *batch, int64_t *bytes);
/* Admit a box: uploads u (the reference layout, ghosts included) and
 * starts its clock at time 0, currentIteration 1.  It completes when its
 * dt clamps to t_end (the dump clamp of hydro_core.f90:691-694) or after
 * max_steps steps (0: no limit; at most 2^31 - 1, else GM_ERR_ARG).
 * GM_ERR_BUSY when it does not fit; GM_ERR_CONFIG when its neq differs
 * from that of the boxes the batch holds.  Runs beside launches in flight (on a
 * stream of its own) and has finished when it returns; the box joins the
 * next gm_batch_advance. */
int gm_batch_add(gm_batch *batch, const gm_params *box, const double *u, double cfl, double t_end,
                 int64_t max_steps, int32_t *box_id);
/* Release a box and its memory (waits only if a launch in flight steps
 * the box; the memory of a box no launch in flight touches is released
 * once those launches finish). */
int gm_batch_remove(gm_batch *batch, int32_t box_id);
/* Queue nlaunch graph launches of steps_per_launch steps for every box that
 * has not completed (as of the last sync or idle query; a box completing
 * within the launches steps with dt = 0 after that).  Asynchronous; waits
 * for launches in flight only when the set of boxes changed. */
int gm_batch_advance(gm_batch *batch, int32_t nlaunch);
/* Wait; report deferred device errors (GM_ERR_NONFINITE). */
int gm_batch_sync(gm_batch *batch);
/* Without waiting: *idle = 1 when the queued launches have finished (the
 * clocks are then those of the last launch, as after gm_batch_sync, and a
 * deferred device error is reported), 0 while they run. */
int gm_batch_query(gm_batch *batch, int32_t *idle);
/* Clock of a box after the last sync or idle query: time, steps taken,
 * completion. */
int gm_batch_clock(gm_batch *batch, int32_t box_id, double *time, int64_t *steps, int32_t *done);
/* The box's conserved array after boundaryI's ghost fill.  Waits only if a
 * launch in flight steps the box; otherwise runs beside the launches. */
int gm_batch_download_u(gm_batch *batch, int32_t box_id, double *u);
/* Launch shape in use -- the column kernels' tile width and z-pencil length,
 * the slab kernel's planes per slab (0: the two-kernel stage for both
 * stages; positive: the slab kernel for the predictor; negative: for both
 * stages, -slab planes per slab) -- and the autotuner's measured step time
 * (0 if not tuned). */
int gm_batch_tuning(const gm_batch *batch, int32_t *block, int32_t *kchunk, int32_t *slab, double *ms_per_step);
/* Number of boxes in the batch. */
int gm_batch_active(const gm_batch *batch, int32_t *n);

#ifdef __cplusplus
}
#endif
#endif /* GUACHO_AMD_H */

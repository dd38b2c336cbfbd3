/* guacho_amd_user.h -- device-side interface of user plugins.
 *
 * The reference's user extension points are Fortran routines in the
 * problem's user_mod (OT/user_mod.f90:77-122, EXO/user_mod.f90:131-206):
 *   get_user_source_terms(pp, s, i, j, k)  -- called per interior cell from
 *       source() (src/sources.f90:206) inside step(), s accumulated on zero;
 *   impose_user_bc(u, order)               -- called at the end of boundaryI
 *       (order 1, on u) and boundaryII (order 2, on up)
 *       (src/boundaries.f90:245, 508).
 * Here each is a HIP kernel in a code object the user builds for gfx950
 * (hipcc --genco --offload-arch=gfx950 -ffp-contract=off -I include
 * plugin.hip -o plugin.hsaco) and registers with gm_set_user_sources /
 * gm_set_user_bc.  The library launches it on its own stream with one
 * argument, a gm_user_args by value, in workgroups of 256 threads: over
 * the interior cells (ceil(nx*ny*nz / 256) workgroups, gm_user_cell) for
 * sources, over every stored cell (ceil(ncell / 256), gm_user_any_cell) for
 * boundary hooks.  A hook runs as a parallel loop: a thread may read any
 * cell, but a cell another thread writes in the same call must not feed
 * the value it writes.  A source kernel, for example:
 *
 *   extern "C" __global__ void my_sources(gm_user_args a) {
 *     int i, j, k;
 *     if (!gm_user_cell(a, blockIdx.x * 256 + threadIdx.x, i, j, k)) return;
 *     const int64_t c = gm_user_index(a, i, j, k);
 *     double u[GM_USER_MAXVAR], pp[GM_USER_MAXVAR];
 *     gm_user_load(a, c, u);
 *     gm_user_u2prim(a, u, pp);              // primit(:, i, j, k)
 *     a.s[2 * a.ncell + c] -= pp[0] * g;      // s(3) = s(3) - rho g
 *   }
 *
 * Indices are the reference's: interior cells i = 1..nx (likewise j, k),
 * two ghost layers (-1..n+2).  Variables are 0-based here: q = 0..7 holds
 * the reference's 1..8 (rho, rho u, rho v, rho w, E, Bx, By, Bz), q = 8..
 * the passive scalars (a.npas of them, the reference's neqdyn+1..neq).
 */
#ifndef GUACHO_AMD_USER_H
#define GUACHO_AMD_USER_H

#include <hip/hip_runtime.h>
#include <stdint.h>

typedef struct gm_user_args {
  int32_t nx, ny, nz;      /* interior cells */
  int32_t NX, NY, NZ;      /* padded extents, n + 4 */
  int64_t ncell;           /* NX * NY * NZ */
  double dx, dy, dz;
  double cv, gamma;
  double time;             /* simulation time of the step (gm_set_time) */
  double dt;               /* sources: dt of the stage (predictor dt/2); hooks: 0 */
  int32_t stage;           /* sources: 1 predictor, 2 corrector; hooks: order 1 or 2 */
  int32_t precision;       /* 0 (FP64): u is double*; 1 (mixed): u is float* */
  int32_t mhd, pmhd;
  int32_t bc[6];
  void *u;                 /* sources: the stage's input state (read only);
                              hooks: the state whose ghosts were just filled */
  double *s;               /* sources: 8 x ncell, zero on entry (interior cells
                              are read back); hooks: NULL */
  int32_t eos;             /* eq_of_state (gm_set_eos); 0 reads as adiabatic */
  int32_t npas;            /* passive scalars after the 8 MHD variables */
  double tempsc;           /* Tempsc of the equation of state */
  int64_t reserved[2];
} gm_user_args;

#ifdef __HIPCC__

/* Interior cell of linear index t (x fastest); false past the last cell. */
__device__ __forceinline__ bool gm_user_cell(const gm_user_args &a, int64_t t, int &i, int &j, int &k) {
  const int64_t n = (int64_t)a.nx * a.ny * a.nz;
  if (t < 0 || t >= n) return false;
  i = (int)(t % a.nx) + 1;
  const int64_t r = t / a.nx;
  j = (int)(r % a.ny) + 1;
  k = (int)(r / a.ny) + 1;
  return true;
}

/* Any stored cell of linear index t (x fastest), ghosts included:
 * i in -1..nx+2 (likewise j, k); false past the last cell. */
__device__ __forceinline__ bool gm_user_any_cell(const gm_user_args &a, int64_t t, int &i, int &j, int &k) {
  if (t < 0 || t >= a.ncell) return false;
  i = (int)(t % a.NX) - 1;
  const int64_t r = t / a.NX;
  j = (int)(r % a.NY) - 1;
  k = (int)(r / a.NY) - 1;
  return true;
}

/* Storage index of cell (i, j, k), ghosts included (i in -1..nx+2). */
__device__ __forceinline__ int64_t gm_user_index(const gm_user_args &a, int i, int j, int k) {
  return ((int64_t)(k + 1) * a.NY + (j + 1)) * a.NX + (i + 1);
}

/* Variable q of cell c as stored (structure of arrays: u + q * ncell). */
__device__ __forceinline__ double gm_user_get(const gm_user_args &a, int64_t c, int q) {
  return a.precision ? (double)((const float *)a.u)[q * a.ncell + c] : ((const double *)a.u)[q * a.ncell + c];
}

__device__ __forceinline__ void gm_user_set(const gm_user_args &a, int64_t c, int q, double v) {
  if (a.precision) ((float *)a.u)[q * a.ncell + c] = (float)v;
  else ((double *)a.u)[q * a.ncell + c] = v;
}

/* The 8 + a.npas variables of cell c (at most GM_USER_MAXVAR). */
#define GM_USER_MAXVAR 10
__device__ __forceinline__ void gm_user_load(const gm_user_args &a, int64_t c, double *u) {
  for (int q = 0; q < 8 + a.npas; q++) u[q] = gm_user_get(a, c, q);
}

/* a / b, correctly rounded: the gfx9 FP64 division sequence (reciprocal,
 * two Newton-Raphson refinements, quotient, one residual correction,
 * v_div_fixup) without its operand scaling.  v_div_scale only rescales when
 * |a| < 2^-969, b or a/b is subnormal, |b| > 2^1021 or exp(a) - exp(b) >=
 * 768; for all other finite nonzero operands these are the full sequence's
 * steps, so the quotient is the IEEE one.  v_div_fixup gives zero, infinite
 * and NaN operands their IEEE results and the quotient its sign.  Divisions
 * by one denominator share its refined reciprocal. */
__device__ __forceinline__ double gm_div(double a, double b) {
  const double r0 = __builtin_amdgcn_rcp(b);
  const double e0 = __builtin_fma(-b, r0, 1.0);
  const double r1 = __builtin_fma(r0, e0, r0);
  const double e1 = __builtin_fma(-b, r1, 1.0);
  const double r2 = __builtin_fma(r1, e1, r1);
  const double q0 = a * r2;
  const double rem = __builtin_fma(-b, q0, a);
  const double q1 = __builtin_fma(rem, r2, q0);
  return __builtin_amdgcn_div_fixup(q1, b, a);
}

/* sqrt(a), correctly rounded: the gfx9 FP64 sequence (reciprocal square
 * root, Goldschmidt refinement, two residual corrections) without the 2^256
 * prescaling it applies below 2^-767; zero and +inf return themselves and
 * negative operands NaN, as IEEE sqrt. */
__device__ __forceinline__ double gm_sqrt(double a) {
  const double y = __builtin_amdgcn_rsq(a);
  const double s0 = a * y;
  const double h0 = y * 0.5;
  const double r0 = __builtin_fma(-h0, s0, 0.5);
  const double h1 = __builtin_fma(h0, r0, h0);
  const double s1 = __builtin_fma(s0, r0, s0);
  const double d0 = __builtin_fma(-s1, s1, a);
  const double s2 = __builtin_fma(d0, h1, s1);
  const double d1 = __builtin_fma(-s2, s2, a);
  const double s3 = __builtin_fma(d1, h1, s2);
  return (a == 0.0 || a == __builtin_inf()) ? a : s3;
}

/* u2prim of the reference (src/hydro_core.f90:46-129, EOS_ADIABATIC):
 * density floor 1e-15, pressure floor 1e-16; the library's own primitive
 * recovery is this function. */
__device__ __forceinline__ void gm_u2prim(double cv, int mhd, const double u[8], double p[8]) {
  const double r = (u[0] > 1e-15) ? u[0] : 1e-15;
  p[0] = r;
  p[1] = gm_div(u[1], r);
  p[2] = gm_div(u[2], r);
  p[3] = gm_div(u[3], r);
  double pr;
  if (mhd)
    pr = gm_div((u[4] - (0.5 * r) * ((p[1] * p[1] + p[2] * p[2]) + p[3] * p[3]))
                - 0.5 * ((u[5] * u[5] + u[6] * u[6]) + u[7] * u[7]), cv);
  else
    pr = gm_div(u[4] - (0.5 * r) * ((p[1] * p[1] + p[2] * p[2]) + p[3] * p[3]), cv);
  p[4] = (pr > 1e-16) ? pr : 1e-16;
  p[5] = u[5];
  p[6] = u[6];
  p[7] = u[7];
}

/* The equation of state's temperature and pressure (src/hydro_core.f90:
 * 90-112): eos 2 (EOS_SINGLE_SPECIE) floors T at 1 and sets p = rho T /
 * Tempsc; eos 3 (EOS_H_RATE, p[8] the neutral density, the first passive
 * scalar) does the same with the particle density 2 rho - p[8] (floored at
 * 1e-15) in place of rho; any other eos is adiabatic, T = (p / rho) Tempsc.
 * Returns T. */
__device__ __forceinline__ double gm_eos(int eos, double tempsc, double *p) {
  if (eos == 2 || eos == 3) {
    double den = p[0];
    if (eos == 3) {
      den = 2.0 * p[0] - p[8];
      den = (den > 1e-15) ? den : 1e-15;
    }
    const double x = gm_div(p[4], den) * tempsc;
    const double t = (1.0 > x) ? 1.0 : x;
    p[4] = gm_div(den * t, tempsc);
    return t;
  }
  return gm_div(p[4], p[0]) * tempsc;
}

/* u (8 + a.npas values, as gm_user_load gives them) -> primitives p (as
 * many), passive scalars copied, the equation of state applied. */
__device__ __forceinline__ void gm_user_u2prim(const gm_user_args &a, const double *u, double *p) {
  gm_u2prim(a.cv, a.mhd, u, p);
  for (int q = 8; q < 8 + a.npas; q++) p[q] = u[q];
  if (a.eos == 2 || (a.eos == 3 && a.npas > 0)) (void)gm_eos(a.eos, a.tempsc, p);
}

#endif /* __HIPCC__ */
#endif /* GUACHO_AMD_USER_H */

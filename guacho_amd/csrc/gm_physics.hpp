// gm_physics.hpp -- per-cell and per-interface MHD physics for gfx950.
//
// Each function evaluates the same floating-point expressions, in the same
// order, as the reference routine cited beside it (Fortran keeps
// parentheses and evaluates left to right; x**2 is x*x).  Compiled with
// -ffp-contract=off, FP64 division and sqrt correctly rounded, so the FP64
// path reproduces the reference's results bit for bit.
//
// States are 8-vectors in the reference's slot order
//   conserved: rho, rho*u, rho*v, rho*w, E, Bx, By, Bz
//   primitive: rho, u, v, w, p, Bx, By, Bz
// Interface solvers work in the frame of the face normal; callers permute
// y/z faces into that frame with gm::axis_perm (the reference's
// swapy/swapz, src/hydro_core.f90:485-534).
#pragma once
#include <hip/hip_runtime.h>

namespace gm {

enum : int { SOLVER_HLL = 1, SOLVER_HLLC = 2, SOLVER_HLLE = 3, SOLVER_HLLD = 4 };
enum : int { BC_OUTFLOW = 1, BC_CLOSED = 2, BC_PERIODIC = 3, BC_OTHER = 4 };
enum : int {
  LIM_NO_AVERAGE = -1, LIM_NO_LIMIT = 0, LIM_MINMOD = 1, LIM_VAN_LEER = 2,
  LIM_VAN_ALBADA = 3, LIM_UMIST = 4, LIM_WOODWARD = 5, LIM_SUPERBEE = 6
};

// Physical constants of a run, uniform across a launch.
struct Phys {
  double cv, gamma;
  int mhd, pmhd;   // full MHD / passive B
  int limiter;
};

__device__ __forceinline__ double dmax(double a, double b) { return (a > b) ? a : b; }
__device__ __forceinline__ double dmin(double a, double b) { return (a < b) ? a : b; }
__device__ __forceinline__ double sign1(double x) { return __builtin_copysign(1.0, x); }

// y faces exchange slots (1,2) and (5,6); z faces (1,3) and (5,7).  The
// axis is a compile-time constant so the permutation is pure register
// renaming (a run-time axis makes the compiler index the array, which
// moves it to scratch memory).
template <int AXIS>
__device__ __forceinline__ void axis_perm(double v[8]) {
  if constexpr (AXIS == 1) {
    double a = v[1]; v[1] = v[2]; v[2] = a;
    a = v[5]; v[5] = v[6]; v[6] = a;
  } else if constexpr (AXIS == 2) {
    double a = v[1]; v[1] = v[3]; v[3] = a;
    a = v[5]; v[5] = v[7]; v[7] = a;
  }
}

// u2prim, src/hydro_core.f90:46-129 (EOS_ADIABATIC).  Returns prims in p.
__device__ __forceinline__ void u2prim(const Phys& ph, const double u[8], double p[8]) {
  const double r = dmax(u[0], 1e-15);
  p[0] = r;
  p[1] = u[1] / r;
  p[2] = u[2] / r;
  p[3] = u[3] / r;
  double pr;
  if (ph.mhd)
    pr = ((u[4] - (0.5 * r) * ((p[1] * p[1] + p[2] * p[2]) + p[3] * p[3]))
          - 0.5 * ((u[5] * u[5] + u[6] * u[6]) + u[7] * u[7])) / ph.cv;
  else
    pr = (u[4] - (0.5 * r) * ((p[1] * p[1] + p[2] * p[2]) + p[3] * p[3])) / ph.cv;
  p[4] = dmax(pr, 1e-16);
  p[5] = u[5];
  p[6] = u[6];
  p[7] = u[7];
}

// prim2u, src/hydro_core.f90:331-383
__device__ __forceinline__ void prim2u(const Phys& ph, const double p[8], double u[8]) {
  u[0] = p[0];
  u[1] = p[0] * p[1];
  u[2] = p[0] * p[2];
  u[3] = p[0] * p[3];
  double e = (0.5 * p[0]) * ((p[1] * p[1] + p[2] * p[2]) + p[3] * p[3]) + ph.cv * p[4];
  if (ph.mhd) e = e + 0.5 * ((p[5] * p[5] + p[6] * p[6]) + p[7] * p[7]);
  u[4] = e;
  u[5] = p[5];
  u[6] = p[6];
  u[7] = p[7];
}

// prim2f, src/hydro_core.f90:395-476 (x-direction flux)
__device__ __forceinline__ void prim2f(const Phys& ph, const double p[8], double f[8]) {
  const double rv = p[0] * p[1];
  if (ph.mhd) {
    const double b2 = (p[5] * p[5] + p[6] * p[6]) + p[7] * p[7];
    const double etot = 0.5 * (((p[0] * ((p[1] * p[1] + p[2] * p[2]) + p[3] * p[3])
                                 + p[5] * p[5]) + p[6] * p[6]) + p[7] * p[7]) + ph.cv * p[4];
    f[0] = rv;
    f[1] = (rv * p[1] + p[4]) + 0.5 * ((p[6] * p[6] + p[7] * p[7]) - p[5] * p[5]);
    f[2] = rv * p[2] - p[5] * p[6];
    f[3] = rv * p[3] - p[5] * p[7];
    f[4] = p[1] * ((etot + p[4]) + 0.5 * b2)
           - p[5] * ((p[1] * p[5] + p[2] * p[6]) + p[3] * p[7]);
  } else {
    const double etot = (0.5 * p[0]) * ((p[1] * p[1] + p[2] * p[2]) + p[3] * p[3]) + ph.cv * p[4];
    f[0] = rv;
    f[1] = rv * p[1] + p[4];
    f[2] = rv * p[2];
    f[3] = rv * p[3];
    f[4] = p[1] * (etot + p[4]);
  }
  f[5] = 0.0;
  f[6] = p[1] * p[6] - p[5] * p[2];
  f[7] = p[1] * p[7] - p[5] * p[3];
}

// csound, src/hydro_core.f90:544-553
__device__ __forceinline__ double csound(const Phys& ph, double pr, double d) {
  return __builtin_sqrt((ph.gamma * pr) / d);
}

// one component of cfast, src/hydro_core.f90:576-579
__device__ __forceinline__ double cfast1(double g, double pr, double d, double b2, double bn) {
  const double t = g * pr + b2;
  return __builtin_sqrt((0.5 * (t + __builtin_sqrt(t * t - (((4.0 * g) * pr) * bn) * bn))) / d);
}

// cfastX, src/hydro_core.f90:591-604
__device__ __forceinline__ double cfastX(const Phys& ph, const double p[8]) {
  const double g = ph.gamma;
  const double b2 = (p[5] * p[5] + p[6] * p[6]) + p[7] * p[7];
  const double cs2va2 = (g * p[4] + b2) / p[0];
  return __builtin_sqrt(0.5 * (cs2va2 + __builtin_sqrt(cs2va2 * cs2va2
                                                       - ((((4.0 * g) * p[4]) * (p[5] * p[5])) / p[0]) / p[0])));
}

// average() of the slope limiters, src/hydro_core.f90:735-796
__device__ __forceinline__ double lim_average(int lim, double a, double b) {
  double s, c, d;
  switch (lim) {
    case LIM_NO_LIMIT:
      return 0.5 * (a + b);
    case LIM_MINMOD:
      s = sign1(a);
      return s * dmax(0.0, dmin(__builtin_fabs(a), s * b));
    case LIM_VAN_LEER:
      if (a * b <= 0.0) return 0.0;
      return ((a * b) * (a + b)) / (a * a + b * b);
    case LIM_VAN_ALBADA: {
      const double delta = 1e-7;
      return (a * (b * b + delta) + b * (a * a + delta)) / ((a * a + b * b) + delta);
    }
    case LIM_UMIST:
      s = sign1(a);
      c = 0.25 * a + 0.75 * b;
      d = 0.75 * a + 0.25 * b;
      return s * dmax(0.0, dmin(dmin(dmin(2.0 * __builtin_fabs(a), (2.0 * s) * b), s * c), s * d));
    case LIM_WOODWARD:
      s = sign1(a);
      c = 0.5 * (a + b);
      return s * dmax(0.0, dmin(dmin(2.0 * __builtin_fabs(a), (2.0 * s) * b), s * c));
    case LIM_SUPERBEE: {
      s = sign1(b);
      const double av1 = dmin(2.0 * __builtin_fabs(b), s * a);
      const double av2 = dmin(__builtin_fabs(b), (2.0 * s) * a);
      return s * dmax(dmax(0.0, av1), av2);
    }
    default:  // LIM_NO_AVERAGE
      return 0.0;
  }
}

// limiter, src/hydro_core.f90:712-731 (in place on pl, pr)
__device__ __forceinline__ void limiter(int lim, const double pll[8], double pl[8], double pr[8],
                                        const double prr[8]) {
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const double dl = pl[q] - pll[q];
    const double dm = pr[q] - pl[q];
    const double dr = prr[q] - pr[q];
    const double al = lim_average(lim, dl, dm);
    const double ar = lim_average(lim, dm, dr);
    pl[q] = pl[q] + al * 0.5;
    pr[q] = pr[q] - ar * 0.5;
  }
}

// Star-state transverse velocity and field of one side, hlld.f90:116-133
// (L) and :164-181 (R); den == 0 takes the degenerate values.
__device__ __forceinline__ void hlld_star_side(const double s[8], double bx, double smu, double smsM,
                                               double sMmu, double& vst, double& wst, double& byst,
                                               double& bzst) {
  const double den = (s[0] * smu) * smsM - bx * bx;
  if (den == 0.0) {
    vst = s[2];
    wst = s[3];
    byst = 0.0;
    bzst = 0.0;
  } else {
    vst = s[2] - ((bx * s[6]) * sMmu) / den;
    wst = s[3] - ((bx * s[7]) * sMmu) / den;
    byst = (s[6] * (s[0] * (smu * smu) - bx * bx)) / den;
    bzst = (s[7] * (s[0] * (smu * smu) - bx * bx)) / den;
  }
}

// Total energy of one side, hlld.f90:113-114 / :161-162
__device__ __forceinline__ double hlld_energy(double cv, const double s[8], double bx) {
  return ((0.5 * s[0]) * ((s[1] * s[1] + s[2] * s[2]) + s[3] * s[3]) + cv * s[4])
         + 0.5 * ((bx * bx + s[6] * s[6]) + s[7] * s[7]);
}

// prim2fhlld, src/hlld.f90:48-319.  Returns false on the fall-through
// (non-finite input) that stops the reference at :316-317.
//
// Organised for a 64-wide wavefront: both sides' star states are built for
// every subsonic interface (the double-star regions need both anyway), the
// side that feeds the flux is chosen by value selects, and only the
// double-star averages (:241-252) sit in a divergent branch.  Each region's
// flux is the same expression, on the same operands, as the reference's
// branch for that region: UL* :111-156, UR* :159-204, UL** :255-283,
// UR** :286-314.
__device__ __forceinline__ bool hlld(const Phys& ph, const double pl[8], const double pr[8], double ff[8]) {
  const double cv = ph.cv;
  const double csl = cfastX(ph, pl), csr = cfastX(ph, pr);
  const double sr = dmax(pl[1] + csl, pr[1] + csr);
  const double sl = dmin(pl[1] - csl, pr[1] - csr);
  if (sl > 0) { prim2f(ph, pl, ff); return true; }
  if (sr < 0) { prim2f(ph, pr, ff); return true; }

  const double bx = 0.5 * (pl[5] + pr[5]);
  const double signbx = sign1(bx);
  const double ptl = pl[4] + 0.5 * ((bx * bx + pl[6] * pl[6]) + pl[7] * pl[7]);
  const double ptr = pr[4] + 0.5 * ((bx * bx + pr[6] * pr[6]) + pr[7] * pr[7]);
  const double slmul = sl - pl[1], srmur = sr - pr[1];
  const double rholul = pl[0] * pl[1], rhorur = pr[0] * pr[1];
  const double den_m = srmur * pr[0] - slmul * pl[0];
  const double sM = (((srmur * rhorur - slmul * rholul) - ptr) + ptl) / den_m;
  const double srmsM = sr - sM, slmsM = sl - sM;
  const double rhostl = (pl[0] * slmul) / slmsM;
  const double rhostr = (pr[0] * srmur) / srmsM;
  const double sqrl = __builtin_sqrt(rhostl), sqrr = __builtin_sqrt(rhostr);
  const double sstl = sM - __builtin_fabs(bx) / sqrl;
  const double sstr = sM + __builtin_fabs(bx) / sqrr;
  const double pst = (((srmur * pr[0]) * ptl - (slmul * pl[0]) * ptr)
                      + (((pl[0] * pr[0]) * srmur) * slmul) * (pr[1] - pl[1])) / den_m;

  double vstl, wstl, bystl, bzstl, vstr, wstr, bystr, bzstr;
  hlld_star_side(pl, bx, slmul, slmsM, sM - pl[1], vstl, wstl, bystl, bzstl);
  hlld_star_side(pr, bx, srmur, srmsM, sM - pr[1], vstr, wstr, bystr, bzstr);

  const bool single = (sstl >= 0) || (sstr <= 0);
  if (!single && !(sM >= 0) && !(sM <= 0)) return false;
  const bool left = single ? (sstl >= 0) : (sM >= 0);
  double s[8];
#pragma unroll
  for (int q = 0; q < 8; q++) s[q] = left ? pl[q] : pr[q];
  const double smu = left ? slmul : srmur;
  const double smsM = left ? slmsM : srmsM;
  const double rhost = left ? rhostl : rhostr;
  const double sqr = left ? sqrl : sqrr;
  const double pt = left ? ptl : ptr;
  const double vst = left ? vstl : vstr, wst = left ? wstl : wstr;
  const double byst = left ? bystl : bystr, bzst = left ? bzstl : bzstr;
  const double e = hlld_energy(cv, s, bx);
  const double vdotb = (s[1] * bx + s[2] * s[6]) + s[3] * s[7];
  const double vstdotbst = (sM * bx + vst * byst) + wst * bzst;
  const double est = (((smu * e - pt * s[1]) + pst * sM) + bx * (vdotb - vstdotbst)) / smsM;

  double vt = vst, wt = wst, byt = byst, bzt = bzst, et = est, vdb = vstdotbst;
  if (!single) {
    const double dd = sqrl + sqrr;
    const double vstst = ((sqrl * vstl + sqrr * vstr) + (bystr - bystl) * signbx) / dd;
    const double wstst = ((sqrl * wstl + sqrr * wstr) + (bzstr - bzstl) * signbx) / dd;
    const double sqlr = __builtin_sqrt(rhostl * rhostr);
    const double bystst = ((sqrl * bystr + sqrr * bystl) + (sqlr * (vstr - vstl)) * signbx) / dd;
    const double bzstst = ((sqrl * bzstr + sqrr * bzstl) + (sqlr * (wstr - wstl)) * signbx) / dd;
    const double vststdotbstst = (sM * bx + vstst * bystst) + wstst * bzstst;
    // eL** = eL* - sqrt(rhoL*)(...)signBx ; eR** = eR* + sqrt(rhoR*)(...)signBx
    const double corr = (sqr * (vstdotbst - vststdotbstst)) * signbx;
    et = left ? est - corr : est + corr;
    vt = vstst;
    wt = wstst;
    byt = bystst;
    bzt = bzstst;
    vdb = vststdotbstst;
  }
  ff[0] = rhost * sM;
  ff[1] = (rhost * (sM * sM) + pst) - bx * bx;
  ff[2] = (rhost * sM) * vt - bx * byt;
  ff[3] = (rhost * sM) * wt - bx * bzt;
  ff[4] = sM * (et + pst) - bx * vdb;
  ff[5] = 0.0;
  ff[6] = byt * sM - bx * vt;
  ff[7] = bzt * sM - bx * wt;
  return true;
}

// HLL average of prim2fhlle (src/hlle.f90:74-79) and prim2fhll (src/hll.f90:73-78)
__device__ __forceinline__ void hll_average(const Phys& ph, const double pl[8], const double pr[8],
                                            double sl, double sr, double ff[8]) {
  double fl[8], fr[8], ul[8], ur[8];
  prim2f(ph, pl, fl);
  prim2f(ph, pr, fr);
  prim2u(ph, pl, ul);
  prim2u(ph, pr, ur);
#pragma unroll
  for (int q = 0; q < 8; q++) ff[q] = ((sr * fl[q] - sl * fr[q]) + (sl * sr) * (ur[q] - ul[q])) / (sr - sl);
}

// prim2fhlle, src/hlle.f90:48-83
__device__ __forceinline__ bool hlle(const Phys& ph, const double pl[8], const double pr[8], double ff[8]) {
  const double csl = cfastX(ph, pl), csr = cfastX(ph, pr);
  const double sr = dmax(pl[1] + csl, pr[1] + csr);
  const double sl = dmin(pl[1] - csl, pr[1] - csr);
  if (sl > 0) { prim2f(ph, pl, ff); return true; }
  if (sr < 0) { prim2f(ph, pr, ff); return true; }
  hll_average(ph, pl, pr, sl, sr, ff);
  return true;
}

// prim2fhll, src/hll.f90:47-82
__device__ __forceinline__ bool hll(const Phys& ph, const double pl[8], const double pr[8], double ff[8]) {
  const double csl = csound(ph, pl[4], pl[0]), csr = csound(ph, pr[4], pr[0]);
  const double sr = dmax(pl[1] + csl, pr[1] + csr);
  const double sl = dmin(pl[1] - csl, pr[1] - csr);
  if (sl > 0) { prim2f(ph, pl, ff); return true; }
  if (sr < 0) { prim2f(ph, pr, ff); return true; }
  hll_average(ph, pl, pr, sl, sr, ff);
  return true;
}

// prim2fhllc, src/hllc.f90:44-140 (passive B)
__device__ __forceinline__ bool hllc(const Phys& ph, const double pl[8], const double pr[8], double ff[8]) {
  const double csl = csound(ph, pl[4], pl[0]), csr = csound(ph, pr[4], pr[0]);
  const double sr = dmax(pl[1] + csl, pr[1] + csr);
  const double sl = dmin(pl[1] - csl, pr[1] - csr);
  if (sl > 0) { prim2f(ph, pl, ff); return true; }
  if (sr < 0) { prim2f(ph, pr, ff); return true; }
  const double slmul = sl - pl[1], srmur = sr - pr[1];
  const double rholul = pl[0] * pl[1], rhorur = pr[0] * pr[1];
  const double sst = (((srmur * rhorur - slmul * rholul) - pr[4]) + pl[4]) / (srmur * pr[0] - slmul * pl[0]);
  if (!(sst >= 0.0) && !(sst <= 0.0)) return false;
  const bool left = (sst >= 0.0);
  double s[8];
#pragma unroll
  for (int q = 0; q < 8; q++) s[q] = left ? pl[q] : pr[q];
  const double smu = left ? slmul : srmur;
  const double sw = left ? sl : sr;
  const double rhost = (s[0] * smu) / (sw - sst);
  const double ek = (0.5 * s[0]) * ((s[1] * s[1] + s[2] * s[2]) + s[3] * s[3]) + ph.cv * s[4];
  double uuk[8], uu[8];
  uuk[0] = rhost;
  uuk[1] = rhost * sst;
  uuk[2] = rhost * s[2];
  uuk[3] = rhost * s[3];
  uuk[4] = rhost * (ek / s[0] + (sst - s[1]) * (sst + s[4] / (s[0] * smu)));
#pragma unroll
  for (int q = 5; q < 8; q++) uuk[q] = ph.pmhd ? (rhost * s[q]) / s[0] : 0.0;
  prim2f(ph, s, ff);
  prim2u(ph, s, uu);
#pragma unroll
  for (int q = 0; q < 8; q++) ff[q] = ff[q] + sw * (uuk[q] - uu[q]);
  return true;
}

template <int SOLVER>
__device__ __forceinline__ bool riemann(const Phys& ph, const double pl[8], const double pr[8], double ff[8]) {
  if constexpr (SOLVER == SOLVER_HLLD) return hlld(ph, pl, pr, ff);
  else if constexpr (SOLVER == SOLVER_HLLE) return hlle(ph, pl, pr, ff);
  else if constexpr (SOLVER == SOLVER_HLL) return hll(ph, pl, pr, ff);
  else return hllc(ph, pl, pr, ff);
}

}  // namespace gm

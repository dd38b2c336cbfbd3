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

#include "guacho_amd.h"
#include "guacho_amd_user.h"

namespace gm {

// Arithmetic of one precision.  FP64 division and sqrt are correctly rounded
// (gm_div, gm_sqrt: the IEEE results, on which the drop-in's bitwise parity
// depends).  FP32 -- the mixed form's
// reconstruction and fluxes -- uses the hardware reciprocal and square
// root (v_rcp_f32, v_sqrt_f32, about 1 ulp), since its accuracy is measured
// against the FP64 path by tolerance.
template <typename T>
struct Ar;
template <>
struct Ar<double> {
  static __device__ __forceinline__ double div(double a, double b) { return gm_div(a, b); }
  static __device__ __forceinline__ double sqrt(double a) { return gm_sqrt(a); }
};
template <>
struct Ar<float> {
  static __device__ __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
  static __device__ __forceinline__ float sqrt(float a) { return __builtin_amdgcn_sqrtf(a); }
};

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
  int eos;         // GM_EOS_ADIABATIC, GM_EOS_SINGLE_SPECIE or GM_EOS_H_RATE
  double tempsc;   // Tempsc of the equation of state
};

template <typename T> __device__ __forceinline__ T dmax(T a, T b) { return (a > b) ? a : b; }
template <typename T> __device__ __forceinline__ T dmin(T a, T b) { return (a < b) ? a : b; }
template <typename T> __device__ __forceinline__ T sign1(T x) { return __builtin_copysign(T(1), x); }
template <typename T> __device__ __forceinline__ T tabs(T x) { return __builtin_fabs(x); }
__device__ __forceinline__ float tabs(float x) { return __builtin_fabsf(x); }
__device__ __forceinline__ float sign1(float x) { return __builtin_copysignf(1.0f, x); }

// y faces exchange slots (1,2) and (5,6); z faces (1,3) and (5,7).  The
// axis is a compile-time constant so the permutation is pure register
// renaming (a run-time axis makes the compiler index the array, which
// moves it to scratch memory).
template <int AXIS, typename T>
__device__ __forceinline__ void axis_perm(T v[8]) {
  if constexpr (AXIS == 1) {
    T a = v[1]; v[1] = v[2]; v[2] = a;
    a = v[5]; v[5] = v[6]; v[6] = a;
  } else if constexpr (AXIS == 2) {
    T a = v[1]; v[1] = v[3]; v[3] = a;
    a = v[5]; v[5] = v[7]; v[7] = a;
  }
}

// u2prim, src/hydro_core.f90:46-129, with the equation of state's pressure
// floor (EOS_SINGLE_SPECIE, :94-99).  Always FP64: this is the pressure
// recovery E - rho v^2/2 - B^2/2, whose cancellation at low plasma beta the
// mixed form keeps in double.  One implementation, shared with user plugins
// (gm_u2prim, gm_eos in guacho_amd_user.h).
// NQ = 8 + the passive scalars, whose primitives are their conserved
// densities (:86-90); EOS_H_RATE reads the first one (the neutral density).
template <int NQ = 8>
__device__ __forceinline__ void u2prim(const Phys& ph, const double u[NQ], double p[NQ]) {
  gm_u2prim(ph.cv, ph.mhd, u, p);
#pragma unroll
  for (int a = 8; a < NQ; a++) p[a] = u[a];
  if (ph.eos == GM_EOS_SINGLE_SPECIE || (NQ > 8 && ph.eos == GM_EOS_H_RATE)) (void)gm_eos(ph.eos, ph.tempsc, p);
}

// ... and u2prim's temperature T (:90-112).
template <int NQ = 8>
__device__ __forceinline__ double u2prim_t(const Phys& ph, const double u[NQ], double p[NQ]) {
  gm_u2prim(ph.cv, ph.mhd, u, p);
#pragma unroll
  for (int a = 8; a < NQ; a++) p[a] = u[a];
  return gm_eos(NQ > 8 ? ph.eos : (ph.eos == GM_EOS_H_RATE ? GM_EOS_ADIABATIC : ph.eos), ph.tempsc, p);
}

// The temperature a post-step module uses: its own Tempsc scaling of p/rho
// under the adiabatic EOS, u2prim's floored T under EOS_SINGLE_SPECIE (one
// Tempsc for both: the parameters module's).
__device__ __forceinline__ double module_temp(const Phys& ph, double t_eos, const double p[8], double tempsc) {
  return (ph.eos == GM_EOS_SINGLE_SPECIE || ph.eos == GM_EOS_H_RATE) ? t_eos : gm_div(p[4], p[0]) * tempsc;
}

// prim2u, src/hydro_core.f90:331-383 (passive scalars :377-379)
template <int NQ = 8, typename T>
__device__ __forceinline__ void prim2u(const Phys& ph, const T p[NQ], T u[NQ]) {
  const T cv = (T)ph.cv;
  u[0] = p[0];
  u[1] = p[0] * p[1];
  u[2] = p[0] * p[2];
  u[3] = p[0] * p[3];
  T e = (T(0.5) * p[0]) * ((p[1] * p[1] + p[2] * p[2]) + p[3] * p[3]) + cv * p[4];
  if (ph.mhd) e = e + T(0.5) * ((p[5] * p[5] + p[6] * p[6]) + p[7] * p[7]);
  u[4] = e;
  u[5] = p[5];
  u[6] = p[6];
  u[7] = p[7];
#pragma unroll
  for (int a = 8; a < NQ; a++) u[a] = p[a];
}

// prim2f, src/hydro_core.f90:395-476 (x-direction flux; passive scalars
// :470-472)
template <int NQ = 8, typename T>
__device__ __forceinline__ void prim2f(const Phys& ph, const T p[NQ], T f[NQ]) {
  const T cv = (T)ph.cv;
  const T rv = p[0] * p[1];
  if (ph.mhd) {
    const T b2 = (p[5] * p[5] + p[6] * p[6]) + p[7] * p[7];
    const T etot = T(0.5) * (((p[0] * ((p[1] * p[1] + p[2] * p[2]) + p[3] * p[3])
                              + p[5] * p[5]) + p[6] * p[6]) + p[7] * p[7]) + cv * p[4];
    f[0] = rv;
    f[1] = (rv * p[1] + p[4]) + T(0.5) * ((p[6] * p[6] + p[7] * p[7]) - p[5] * p[5]);
    f[2] = rv * p[2] - p[5] * p[6];
    f[3] = rv * p[3] - p[5] * p[7];
    f[4] = p[1] * ((etot + p[4]) + T(0.5) * b2)
           - p[5] * ((p[1] * p[5] + p[2] * p[6]) + p[3] * p[7]);
  } else {
    const T etot = (T(0.5) * p[0]) * ((p[1] * p[1] + p[2] * p[2]) + p[3] * p[3]) + cv * p[4];
    f[0] = rv;
    f[1] = rv * p[1] + p[4];
    f[2] = rv * p[2];
    f[3] = rv * p[3];
    f[4] = p[1] * (etot + p[4]);
  }
  f[5] = T(0);
  f[6] = p[1] * p[6] - p[5] * p[2];
  f[7] = p[1] * p[7] - p[5] * p[3];
#pragma unroll
  for (int a = 8; a < NQ; a++) f[a] = p[a] * p[1];
}

// csound, src/hydro_core.f90:544-553
template <typename T>
__device__ __forceinline__ T csound(const Phys& ph, T pr, T d) {
  return Ar<T>::sqrt(Ar<T>::div((T)ph.gamma * pr, d));
}

// one component of cfast, src/hydro_core.f90:576-579 (the CFL candidates;
// FP32 in the mixed form)
template <typename T>
__device__ __forceinline__ T cfast1(T g, T pr, T d, T b2, T bn) {
  using A = Ar<T>;
  const T t = g * pr + b2;
  return A::sqrt(A::div(T(0.5) * (t + A::sqrt(t * t - (((T(4) * g) * pr) * bn) * bn)), d));
}

// cfastX, src/hydro_core.f90:591-604
template <typename T>
__device__ __forceinline__ T cfastX(const Phys& ph, const T p[8]) {
  using A = Ar<T>;
  const T g = (T)ph.gamma;
  const T b2 = (p[5] * p[5] + p[6] * p[6]) + p[7] * p[7];
  const T cs2va2 = A::div(g * p[4] + b2, p[0]);
  return A::sqrt(T(0.5) * (cs2va2 + A::sqrt(cs2va2 * cs2va2
                                             - A::div(A::div(((T(4) * g) * p[4]) * (p[5] * p[5]), p[0]), p[0]))));
}

// average() of the slope limiters, src/hydro_core.f90:735-796
template <typename T>
__device__ __forceinline__ T lim_average(int lim, T a, T b) {
  using A = Ar<T>;
  T s, c, d;
  switch (lim) {
    case LIM_NO_LIMIT:
      return T(0.5) * (a + b);
    case LIM_MINMOD:
      s = sign1(a);
      return s * dmax(T(0), dmin(tabs(a), s * b));
    case LIM_VAN_LEER:
      if (a * b <= T(0)) return T(0);
      return A::div((a * b) * (a + b), a * a + b * b);
    case LIM_VAN_ALBADA: {
      const T delta = T(1e-7);
      return A::div(a * (b * b + delta) + b * (a * a + delta), (a * a + b * b) + delta);
    }
    case LIM_UMIST:
      s = sign1(a);
      c = T(0.25) * a + T(0.75) * b;
      d = T(0.75) * a + T(0.25) * b;
      return s * dmax(T(0), dmin(dmin(dmin(T(2) * tabs(a), (T(2) * s) * b), s * c), s * d));
    case LIM_WOODWARD:
      s = sign1(a);
      c = T(0.5) * (a + b);
      return s * dmax(T(0), dmin(dmin(T(2) * tabs(a), (T(2) * s) * b), s * c));
    case LIM_SUPERBEE: {
      s = sign1(b);
      const T av1 = dmin(T(2) * tabs(b), s * a);
      const T av2 = dmin(tabs(b), (T(2) * s) * a);
      return s * dmax(dmax(T(0), av1), av2);
    }
    default:  // LIM_NO_AVERAGE
      return T(0);
  }
}

// limiter, src/hydro_core.f90:712-731 (in place on pl, pr)
template <int NQ = 8, typename T>
__device__ __forceinline__ void limiter(int lim, const T pll[NQ], T pl[NQ], T pr[NQ], const T prr[NQ]) {
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const T dl = pl[q] - pll[q];
    const T dm = pr[q] - pl[q];
    const T dr = prr[q] - pr[q];
    const T al = lim_average(lim, dl, dm);
    const T ar = lim_average(lim, dm, dr);
    pl[q] = pl[q] + al * T(0.5);
    pr[q] = pr[q] - ar * T(0.5);
  }
}

// Star-state transverse velocity and field of one side, hlld.f90:116-133
// (L) and :164-181 (R); den == 0 takes the degenerate values.
template <typename T>
__device__ __forceinline__ void hlld_star_side(const T s[8], T bx, T smu, T smsM, T sMmu, T& vst, T& wst,
                                               T& byst, T& bzst) {
  using A = Ar<T>;
  const T den = (s[0] * smu) * smsM - bx * bx;
  if (den == T(0)) {
    vst = s[2];
    wst = s[3];
    byst = T(0);
    bzst = T(0);
  } else {
    vst = s[2] - A::div((bx * s[6]) * sMmu, den);
    wst = s[3] - A::div((bx * s[7]) * sMmu, den);
    byst = A::div(s[6] * (s[0] * (smu * smu) - bx * bx), den);
    bzst = A::div(s[7] * (s[0] * (smu * smu) - bx * bx), den);
  }
}

// Total energy of one side, hlld.f90:113-114 / :161-162
template <typename T>
__device__ __forceinline__ T hlld_energy(T cv, const T s[8], T bx) {
  return ((T(0.5) * s[0]) * ((s[1] * s[1] + s[2] * s[2]) + s[3] * s[3]) + cv * s[4])
         + T(0.5) * ((bx * bx + s[6] * s[6]) + s[7] * s[7]);
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
template <int NQ = 8, typename T>
__device__ __forceinline__ bool hlld_cs(const Phys& ph, const T pl[NQ], const T pr[NQ], T csl, T csr, T ff[NQ]) {
  using A = Ar<T>;
  const T cv = (T)ph.cv;
  const T sr = dmax(pl[1] + csl, pr[1] + csr);
  const T sl = dmin(pl[1] - csl, pr[1] - csr);
  if (sl > T(0)) { prim2f<NQ>(ph, pl, ff); return true; }
  if (sr < T(0)) { prim2f<NQ>(ph, pr, ff); return true; }

  const T bx = T(0.5) * (pl[5] + pr[5]);
  const T signbx = sign1(bx);
  const T ptl = pl[4] + T(0.5) * ((bx * bx + pl[6] * pl[6]) + pl[7] * pl[7]);
  const T ptr = pr[4] + T(0.5) * ((bx * bx + pr[6] * pr[6]) + pr[7] * pr[7]);
  const T slmul = sl - pl[1], srmur = sr - pr[1];
  const T rholul = pl[0] * pl[1], rhorur = pr[0] * pr[1];
  const T den_m = srmur * pr[0] - slmul * pl[0];
  const T sM = A::div(((srmur * rhorur - slmul * rholul) - ptr) + ptl, den_m);
  const T srmsM = sr - sM, slmsM = sl - sM;
  const T rhostl = A::div(pl[0] * slmul, slmsM);
  const T rhostr = A::div(pr[0] * srmur, srmsM);
  const T sqrl = A::sqrt(rhostl), sqrr = A::sqrt(rhostr);
  const T sstl = sM - A::div(tabs(bx), sqrl);
  const T sstr = sM + A::div(tabs(bx), sqrr);
  const T pst = A::div(((srmur * pr[0]) * ptl - (slmul * pl[0]) * ptr)
                       + (((pl[0] * pr[0]) * srmur) * slmul) * (pr[1] - pl[1]), den_m);

  T vstl, wstl, bystl, bzstl, vstr, wstr, bystr, bzstr;
  hlld_star_side(pl, bx, slmul, slmsM, sM - pl[1], vstl, wstl, bystl, bzstl);
  hlld_star_side(pr, bx, srmur, srmsM, sM - pr[1], vstr, wstr, bystr, bzstr);

  const bool single = (sstl >= T(0)) || (sstr <= T(0));
  if (!single && !(sM >= T(0)) && !(sM <= T(0))) return false;
  const bool left = single ? (sstl >= T(0)) : (sM >= T(0));
  T s[NQ];
#pragma unroll
  for (int q = 0; q < NQ; q++) s[q] = left ? pl[q] : pr[q];
  const T smu = left ? slmul : srmur;
  const T smsM = left ? slmsM : srmsM;
  const T rhost = left ? rhostl : rhostr;
  const T sqr = left ? sqrl : sqrr;
  const T pt = left ? ptl : ptr;
  const T vst = left ? vstl : vstr, wst = left ? wstl : wstr;
  const T byst = left ? bystl : bystr, bzst = left ? bzstl : bzstr;
  const T e = hlld_energy(cv, s, bx);
  const T vdotb = (s[1] * bx + s[2] * s[6]) + s[3] * s[7];
  const T vstdotbst = (sM * bx + vst * byst) + wst * bzst;
  const T est = A::div(((smu * e - pt * s[1]) + pst * sM) + bx * (vdotb - vstdotbst), smsM);

  T vt = vst, wt = wst, byt = byst, bzt = bzst, et = est, vdb = vstdotbst;
  if (!single) {
    const T dd = sqrl + sqrr;
    const T vstst = A::div((sqrl * vstl + sqrr * vstr) + (bystr - bystl) * signbx, dd);
    const T wstst = A::div((sqrl * wstl + sqrr * wstr) + (bzstr - bzstl) * signbx, dd);
    const T sqlr = A::sqrt(rhostl * rhostr);
    const T bystst = A::div((sqrl * bystr + sqrr * bystl) + (sqlr * (vstr - vstl)) * signbx, dd);
    const T bzstst = A::div((sqrl * bzstr + sqrr * bzstl) + (sqlr * (wstr - wstl)) * signbx, dd);
    const T vststdotbstst = (sM * bx + vstst * bystst) + wstst * bzstst;
    // eL** = eL* - sqrt(rhoL*)(...)signBx ; eR** = eR* + sqrt(rhoR*)(...)signBx
    const T corr = (sqr * (vstdotbst - vststdotbstst)) * signbx;
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
  ff[5] = T(0);
  ff[6] = byt * sM - bx * vt;
  ff[7] = bzt * sM - bx * wt;
  // passive scalars of the flux side, one expression for its star and
  // double-star regions (hlld.f90:149-151, 197-199, 276-278, 307-309)
#pragma unroll
  for (int a = 8; a < NQ; a++) ff[a] = A::div((sM * s[a]) * smu, smsM);
  return true;
}

template <int NQ = 8, typename T>
__device__ __forceinline__ bool hlld(const Phys& ph, const T pl[NQ], const T pr[NQ], T ff[NQ]) {
  return hlld_cs<NQ>(ph, pl, pr, cfastX(ph, pl), cfastX(ph, pr), ff);
}

// HLL average of prim2fhlle (src/hlle.f90:74-79) and prim2fhll (src/hll.f90:73-78)
template <int NQ = 8, typename T>
__device__ __forceinline__ void hll_average(const Phys& ph, const T pl[NQ], const T pr[NQ], T sl, T sr, T ff[NQ]) {
  T fl[NQ], fr[NQ], ul[NQ], ur[NQ];
  prim2f<NQ>(ph, pl, fl);
  prim2f<NQ>(ph, pr, fr);
  prim2u<NQ>(ph, pl, ul);
  prim2u<NQ>(ph, pr, ur);
#pragma unroll
  for (int q = 0; q < NQ; q++) ff[q] = Ar<T>::div((sr * fl[q] - sl * fr[q]) + (sl * sr) * (ur[q] - ul[q]), sr - sl);
}

// prim2fhlle, src/hlle.f90:48-83
template <int NQ = 8, typename T>
__device__ __forceinline__ bool hlle_cs(const Phys& ph, const T pl[NQ], const T pr[NQ], T csl, T csr, T ff[NQ]) {
  const T sr = dmax(pl[1] + csl, pr[1] + csr);
  const T sl = dmin(pl[1] - csl, pr[1] - csr);
  if (sl > T(0)) { prim2f<NQ>(ph, pl, ff); return true; }
  if (sr < T(0)) { prim2f<NQ>(ph, pr, ff); return true; }
  hll_average<NQ>(ph, pl, pr, sl, sr, ff);
  return true;
}

// prim2fhll, src/hll.f90:47-82
template <int NQ = 8, typename T>
__device__ __forceinline__ bool hll_cs(const Phys& ph, const T pl[NQ], const T pr[NQ], T csl, T csr, T ff[NQ]) {
  const T sr = dmax(pl[1] + csl, pr[1] + csr);
  const T sl = dmin(pl[1] - csl, pr[1] - csr);
  if (sl > T(0)) { prim2f<NQ>(ph, pl, ff); return true; }
  if (sr < T(0)) { prim2f<NQ>(ph, pr, ff); return true; }
  hll_average<NQ>(ph, pl, pr, sl, sr, ff);
  return true;
}

// prim2fhllc, src/hllc.f90:44-140 (passive B)
template <int NQ = 8, typename T>
__device__ __forceinline__ bool hllc_cs(const Phys& ph, const T pl[NQ], const T pr[NQ], T csl, T csr, T ff[NQ]) {
  using A = Ar<T>;
  const T sr = dmax(pl[1] + csl, pr[1] + csr);
  const T sl = dmin(pl[1] - csl, pr[1] - csr);
  if (sl > T(0)) { prim2f<NQ>(ph, pl, ff); return true; }
  if (sr < T(0)) { prim2f<NQ>(ph, pr, ff); return true; }
  const T slmul = sl - pl[1], srmur = sr - pr[1];
  const T rholul = pl[0] * pl[1], rhorur = pr[0] * pr[1];
  const T sst = A::div(((srmur * rhorur - slmul * rholul) - pr[4]) + pl[4], srmur * pr[0] - slmul * pl[0]);
  if (!(sst >= T(0)) && !(sst <= T(0))) return false;
  const bool left = (sst >= T(0));
  T s[NQ];
#pragma unroll
  for (int q = 0; q < NQ; q++) s[q] = left ? pl[q] : pr[q];
  const T smu = left ? slmul : srmur;
  const T sw = left ? sl : sr;
  const T rhost = A::div(s[0] * smu, sw - sst);
  const T ek = (T(0.5) * s[0]) * ((s[1] * s[1] + s[2] * s[2]) + s[3] * s[3]) + (T)ph.cv * s[4];
  T uuk[NQ], uu[NQ];
  uuk[0] = rhost;
  uuk[1] = rhost * sst;
  uuk[2] = rhost * s[2];
  uuk[3] = rhost * s[3];
  uuk[4] = rhost * (A::div(ek, s[0]) + (sst - s[1]) * (sst + A::div(s[4], s[0] * smu)));
#pragma unroll
  for (int q = 5; q < 8; q++) uuk[q] = ph.pmhd ? A::div(rhost * s[q], s[0]) : T(0);
#pragma unroll
  for (int a = 8; a < NQ; a++) uuk[a] = A::div(rhost * s[a], s[0]);  // hllc.f90:94-97, 123-126
  prim2f<NQ>(ph, s, ff);
  prim2u<NQ>(ph, s, uu);
#pragma unroll
  for (int q = 0; q < NQ; q++) ff[q] = ff[q] + sw * (uuk[q] - uu[q]);
  return true;
}

// The wave-speed estimate of one side, from that side's primitives in the
// solver's axis order: cfastX for HLLD and HLLE (hlld.f90:64-65,
// hlle.f90:58-59), csound for HLL and HLLC (hll.f90:57-58, hllc.f90:55-56).
// It depends on that state alone, so in a first-order sweep, where a
// face's states are the cells' own primitives, a cell's value serves both
// of its faces on that axis: the same operation on the same operands, the
// same bits.
template <int SOLVER, typename T>
__device__ __forceinline__ T side_speed(const Phys& ph, const T p[8]) {
  if constexpr (SOLVER == SOLVER_HLLD || SOLVER == SOLVER_HLLE) return cfastX(ph, p);
  else return csound(ph, p[4], p[0]);
}

// The interface flux from both sides' states and their side_speed values.
template <int SOLVER, int NQ = 8, typename T>
__device__ __forceinline__ bool riemann_cs(const Phys& ph, const T pl[NQ], const T pr[NQ], T csl, T csr, T ff[NQ]) {
  if constexpr (SOLVER == SOLVER_HLLD) return hlld_cs<NQ>(ph, pl, pr, csl, csr, ff);
  else if constexpr (SOLVER == SOLVER_HLLE) return hlle_cs<NQ>(ph, pl, pr, csl, csr, ff);
  else if constexpr (SOLVER == SOLVER_HLL) return hll_cs<NQ>(ph, pl, pr, csl, csr, ff);
  else return hllc_cs<NQ>(ph, pl, pr, csl, csr, ff);
}

template <int SOLVER, int NQ = 8, typename T>
__device__ __forceinline__ bool riemann(const Phys& ph, const T pl[NQ], const T pr[NQ], T ff[NQ]) {
  return riemann_cs<SOLVER, NQ>(ph, pl, pr, side_speed<SOLVER>(ph, pl), side_speed<SOLVER>(ph, pr), ff);
}

}  // namespace gm

// gm_kernels.hip -- gfx950 kernels of the Godunov MHD step.
//
// Device layout: structure of arrays.  Variable q of the conserved state
// lives at base + q*ncell; cell (i, j, k), with the reference's 1-based
// indices and two ghost layers (i in -1..nx+2), sits at
//   c = ((k+1)*NY + (j+1))*NX + (i+1),   NX = nx+4, NY = ny+4.
// x is fastest, so a wavefront of 64 consecutive i reads 512 contiguous
// bytes per variable.
//
// The step never writes ghost cells of the state.  Where a stencil reaches
// past the domain, the kernels read the interior cell the reference's ghost
// fill would have copied there (boundaryI: src/boundaries.f90:44-247 for
// U^n; boundaryII: :256-512 for U*; boundaryI_ef:
// src/flux_cd_module.f90:45-237 for E), negating the momentum or E
// component that the closed boundary flips.  No interior update reads an
// edge or corner ghost, so a per-axis mapping is exact.
#include "gm_kernels.hpp"

namespace gm {

__device__ __forceinline__ long long cidx(const Geom& g, int i, int j, int k) {
  return ((long long)(k + 1) * g.NY + (j + 1)) * g.NX + (i + 1);
}

// Workgroups are dealt to the 8 XCDs round-robin by linear id.  Renumber
// them so each XCD runs one contiguous range of logical tiles (whole
// neighbouring rows and planes), so stencil neighbours hit its own L2.
constexpr int kXCD = 8;
__device__ __forceinline__ unsigned xcd_logical_block() {
  const unsigned b = blockIdx.x, nb = gridDim.x;
  const unsigned q = nb / kXCD, r = nb % kXCD;
  const unsigned x = b % kXCD, idx = b / kXCD;
  return x < r ? x * (q + 1) + idx : r * (q + 1) + (x - r) * q + idx;
}

// Logical tile -> (flat index within a plane of `plane_n` items, plane).
// Tiles of blockDim.x items; `bpp` tiles per plane.  False past the end.
__device__ __forceinline__ bool tile_item(unsigned lb, unsigned bpp, unsigned plane_n, int& t, int& plane) {
  plane = (int)(lb / bpp);
  t = (int)((lb - (unsigned)plane * bpp) * blockDim.x + threadIdx.x);
  return (unsigned)t < plane_n;
}

__device__ __forceinline__ unsigned tiles_of(unsigned items) { return (items + blockDim.x - 1) / blockDim.x; }

// Block of a batched launch -> (box, logical tile within the box), from the
// prefix sums of the boxes' tile counts (prefix[0] = 0).
__device__ __forceinline__ void batch_locate(const unsigned* __restrict__ prefix, int nbox, unsigned lb,
                                             int& box, unsigned& local) {
  int lo = 0, hi = nbox;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (prefix[mid] <= lb) lo = mid;
    else hi = mid;
  }
  box = lo;
  local = lb - prefix[lo];
}

// Interior source of ghost index i on an axis of n cells; flips when the
// boundary on that side is closed.  Valid for i in -1..n+2 (n >= 2).
__device__ __forceinline__ int map_ghost(int i, int n, int lo, int hi, bool& flip) {
  if (i < 1) {
    if (lo == BC_PERIODIC) return i + n;
    flip = (lo == BC_CLOSED);
    return 1 - i;
  }
  if (i > n) {
    if (hi == BC_PERIODIC) return i - n;
    flip = (hi == BC_CLOSED);
    return 2 * n + 1 - i;
  }
  return i;
}

// Conserved 8-vector of cell (i,j,k) in FP64, ghosts resolved through the
// boundary conditions.
template <typename S>
__device__ __forceinline__ void load_cons(const Geom& g, const S* __restrict__ U, int i, int j, int k,
                                          double u[8]) {
  bool fx = false, fy = false, fz = false;
  const int si = map_ghost(i, g.nx, g.bc[0], g.bc[1], fx);
  const int sj = map_ghost(j, g.ny, g.bc[2], g.bc[3], fy);
  const int sk = map_ghost(k, g.nz, g.bc[4], g.bc[5], fz);
  const long long c = cidx(g, si, sj, sk);
#pragma unroll
  for (int q = 0; q < 8; q++) u[q] = (double)U[q * g.ncell + c];
  if (fx) u[1] = -u[1];
  if (fy) u[2] = -u[2];
  if (fz) u[3] = -u[3];
}

// ---------------------------------------------------------------------------
// Kernel bodies.  Each works on one logical tile `lb` of one box; the
// __global__ wrappers below either launch one box (grid = its tiles, XCD
// order) or a batch of boxes (grid = all tiles of all boxes, each block
// finding its box in a prefix table).  Stage 1 is the predictor (reads U^n,
// writes U*), stage 2 the corrector (reads U*, updates U^n in place).
// ---------------------------------------------------------------------------

// primitives: calcprim over the cells the flux stencils read
// (src/hydro_core.f90:245-320), ghost values by boundary mapping; FP64
// pressure recovery for either storage type.
template <typename S>
__device__ __forceinline__ void prim_body(const Box<S>& b, int stage, unsigned lb) {
  const Geom& g = b.g;
  const int lo = stage == 1 ? 0 : -1;  // one ghost layer for first order, two for the limiter
  const int ex = g.nx + 2 - 2 * lo, ey = g.ny + 2 - 2 * lo;
  int t, kk;
  if (!tile_item(lb, tiles_of((unsigned)(ex * ey)), (unsigned)(ex * ey), t, kk)) return;
  const int i = t % ex + lo, j = t / ex + lo, k = kk + lo;
  // only face ghosts are read downstream; skip edges and corners
  const int out = (i < 1 || i > g.nx) + (j < 1 || j > g.ny) + (k < 1 || k > g.nz);
  if (out > 1) return;
  double u[8], p[8];
  load_cons(g, stage == 1 ? b.U : b.Us, i, j, k, u);
  u2prim(b.ph, u, p);
  const long long c = cidx(g, i, j, k);
#pragma unroll
  for (int q = 0; q < 8; q++) b.P[q * g.ncell + c] = (S)p[q];
}

// interface fluxes: the sweep of hlldfluxes / hllefluxes / hllfluxes /
// hllcfluxes (e.g. src/hlld.f90:331-432) for one axis.  Face f(:,i,j,k)
// lies between cells i and i+1 (i = 0..n along the axis); only faces with
// interior transverse indices feed the update.
template <int SOLVER, int ORDER, int axis, typename S>
__device__ __forceinline__ void flux_body(const Box<S>& b, unsigned lb, int* __restrict__ err) {
  const Geom& g = b.g;
  const int e0 = axis == 0 ? g.nx + 1 : g.nx;
  const int e1 = axis == 1 ? g.ny + 1 : g.ny;
  int t, k;
  if (!tile_item(lb, tiles_of((unsigned)(e0 * e1)), (unsigned)(e0 * e1), t, k)) return;
  int i = t % e0, j = t / e0;
  if (axis != 0) i += 1;
  if (axis != 1) j += 1;
  if (axis != 2) k += 1;
  const long long stride = axis == 0 ? 1 : axis == 1 ? (long long)g.NX : (long long)g.NX * g.NY;
  const long long c = cidx(g, i, j, k);
  const S* __restrict__ P = b.P;
  S* __restrict__ F = b.F + axis * 8 * g.ncell;
  S pl[8], pr[8], ff[8];
#pragma unroll
  for (int q = 0; q < 8; q++) {
    pl[q] = P[q * g.ncell + c];
    pr[q] = P[q * g.ncell + c + stride];
  }
  if constexpr (ORDER == 2) {
    S pll[8], prr[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
      pll[q] = P[q * g.ncell + c - stride];
      prr[q] = P[q * g.ncell + c + 2 * stride];
    }
    limiter(b.ph.limiter, pll, pl, pr, prr);
  }
  axis_perm<axis>(pl);
  axis_perm<axis>(pr);
  const bool ok = riemann<SOLVER>(b.ph, pl, pr, ff);
  if (!ok) atomicOr(err, 1);
  axis_perm<axis>(ff);
#pragma unroll
  for (int q = 0; q < 8; q++) F[q * g.ncell + c] = ff[q];
}

// cell-centred electric field, get_efield (src/flux_cd_module.f90:245-273)
template <typename S>
__device__ __forceinline__ void efield_body(const Box<S>& b, unsigned lb) {
  const Geom& g = b.g;
  int t, k;
  if (!tile_item(lb, tiles_of((unsigned)(g.nx * g.ny)), (unsigned)(g.nx * g.ny), t, k)) return;
  const int i = t % g.nx + 1, j = t / g.nx + 1;
  k += 1;
  const long long sx = 1, sy = g.NX, sz = (long long)g.NX * g.NY;
  const long long N = g.ncell;
  const long long c = cidx(g, i, j, k);
  const S* __restrict__ Fx = b.F;
  const S* __restrict__ Fy = b.F + 8 * N;
  const S* __restrict__ Fz = b.F + 16 * N;
  S* __restrict__ E = b.E;
  auto ld = [](const S* a, long long o) { return (double)a[o]; };
  E[0 * N + c] = (S)(0.25 * (((-ld(Fy, 7 * N + c - sy) - ld(Fy, 7 * N + c)) + ld(Fz, 6 * N + c - sz)) + ld(Fz, 6 * N + c)));
  E[1 * N + c] = (S)(0.25 * (((+ld(Fx, 7 * N + c - sx) + ld(Fx, 7 * N + c)) - ld(Fz, 5 * N + c - sz)) - ld(Fz, 5 * N + c)));
  E[2 * N + c] = (S)(0.25 * (((-ld(Fx, 6 * N + c - sx) - ld(Fx, 6 * N + c)) + ld(Fy, 5 * N + c - sy)) + ld(Fy, 5 * N + c)));
}

// ---------------------------------------------------------------------------
// upwind update, step(dt) (src/hydro_solver.f90:77-127) with flux_cd_update
// (src/flux_cd_module.f90:285-323), accumulated in FP64.  Each thread
// marches up a z-column segment of `kchunk` cells, carrying the z-face flux
// below the cell and E one plane below and at the cell, so every z
// neighbour is read once.
//
// E ghosts (boundaryI_ef): the update reads e2, e3 at x neighbours, e1, e3
// at y neighbours and e1, e2 at z neighbours.  A closed boundary negates e1
// on x ghosts and e2 on y ghosts only (flux_cd_module.f90:145-176), never a
// component read there, so every E ghost resolves to a fixed interior
// neighbour offset: periodic wraps, closed and outflow copy the edge cell.
// Stage 2 writes U^{n+1} over U^n in place (each cell reads only its own
// conserved state).
// ---------------------------------------------------------------------------

// offset (in cells along the axis) of the E ghost-resolved neighbour
__device__ __forceinline__ int e_nbr(int i, int n, int d, int bc) {
  const int t = i + d;
  if (t >= 1 && t <= n) return d;
  if (bc == BC_PERIODIC) return (t < 1 ? t + n : t - n) - i;
  return 0;  // closed / outflow: ghost takes the edge cell's value
}

template <int FLUXCD, typename S>
__device__ __forceinline__ void update_body(const Box<S>& b, unsigned lb, double dt, int stage, int kchunk) {
  constexpr int NH = FLUXCD ? 5 : 8;
  const Geom& g = b.g;
  int t, kc;
  if (!tile_item(lb, tiles_of((unsigned)(g.nx * g.ny)), (unsigned)(g.nx * g.ny), t, kc)) return;
  const int i = t % g.nx + 1, j = t / g.nx + 1;
  const int k0 = kc * kchunk + 1;
  const int k1 = min(k0 + kchunk - 1, g.nz);
  const long long sx = 1, sy = g.NX, sz = (long long)g.NX * g.NY;
  const long long N = g.ncell;
  const double dtdx = dt / b.dx, dtdy = dt / b.dy, dtdz = dt / b.dz;
  const S* Uin = b.U;
  S* Uout = stage == 1 ? b.Us : b.U;
  const S* __restrict__ Fx = b.F;
  const S* __restrict__ Fy = b.F + 8 * N;
  const S* __restrict__ Fz = b.F + 16 * N;
  const S* __restrict__ E = b.E;
  long long c = cidx(g, i, j, k0);
  double fzm[NH];
#pragma unroll
  for (int q = 0; q < NH; q++) fzm[q] = (double)Fz[q * N + c - sz];
  long long oxm = 0, oxp = 0, oym = 0, oyp = 0;
  double e1m = 0, e2m = 0, e1c = 0, e2c = 0;
  if constexpr (FLUXCD) {
    oxm = e_nbr(i, g.nx, -1, g.bc[0]) * sx;
    oxp = e_nbr(i, g.nx, +1, g.bc[1]) * sx;
    oym = e_nbr(j, g.ny, -1, g.bc[2]) * sy;
    oyp = e_nbr(j, g.ny, +1, g.bc[3]) * sy;
    const long long ozm = e_nbr(k0, g.nz, -1, g.bc[4]) * sz;
    e1m = (double)E[0 * N + c + ozm];
    e2m = (double)E[1 * N + c + ozm];
    e1c = (double)E[0 * N + c];
    e2c = (double)E[1 * N + c];
  }
#pragma unroll 1
  for (int k = k0; k <= k1; k++, c += sz) {
    double out[8];
#pragma unroll
    for (int q = 0; q < NH; q++) {
      const long long o = q * N + c;
      const double fz = (double)Fz[o];
      out[q] = (((double)Uin[o] - dtdx * ((double)Fx[o] - (double)Fx[o - sx]))
                - dtdy * ((double)Fy[o] - (double)Fy[o - sy]))
               - dtdz * (fz - fzm[q]);
      fzm[q] = fz;
    }
    if constexpr (FLUXCD) {
      const long long ozp = e_nbr(k, g.nz, +1, g.bc[5]) * sz;
      const double e1p = (double)E[0 * N + c + ozp];
      const double e2p = (double)E[1 * N + c + ozp];
      const S* E1 = E;
      const S* E2 = E + N;
      const S* E3 = E + 2 * N;
      auto d = [](S v) { return (double)v; };
      out[5] = (d(Uin[5 * N + c]) - (0.5 * dtdy) * (d(E3[c + oyp]) - d(E3[c + oym]))) + (0.5 * dtdz) * (e2p - e2m);
      out[6] = (d(Uin[6 * N + c]) + (0.5 * dtdx) * (d(E3[c + oxp]) - d(E3[c + oxm]))) - (0.5 * dtdz) * (e1p - e1m);
      out[7] = (d(Uin[7 * N + c]) - (0.5 * dtdx) * (d(E2[c + oxp]) - d(E2[c + oxm]))) + (0.5 * dtdy) * (d(E1[c + oyp]) - d(E1[c + oym]));
      e1m = e1c;
      e2m = e2c;
      e1c = e1p;
      e2c = e2p;
    }
#pragma unroll
    for (int q = 0; q < 8; q++) Uout[q * N + c] = (S)out[q];
  }
}

// ---------------------------------------------------------------------------
// CFL candidate reduction: the min loop of get_timestep
// (src/hydro_core.f90:642-675) on the primitives of the interior.  Exact
// (min is order-free); block minima meet in one 64-bit atomicMin on the
// bit pattern, which orders positive doubles like their values.  Any
// non-finite conserved value sets err bit 2.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = dmin(v, __shfl_xor(v, off, 64));
  return v;
}

template <typename S>
__device__ __forceinline__ void cfl_body(const Box<S>& b, unsigned lb, int kchunk,
                                         unsigned long long* __restrict__ dtmin, int* __restrict__ err) {
  __shared__ double red[4];
  const Geom& g = b.g;
  const Phys& ph = b.ph;
  double dtp = 1e30;
  bool bad = false;
  int t, kc;
  if (tile_item(lb, tiles_of((unsigned)(g.nx * g.ny)), (unsigned)(g.nx * g.ny), t, kc)) {
    const int i = t % g.nx + 1, j = t / g.nx + 1;
    const int k0 = kc * kchunk + 1;
    const int k1 = min(k0 + kchunk - 1, g.nz);
    for (int k = k0; k <= k1; k++) {
      const long long c = cidx(g, i, j, k);
      double u[8], p[8];
#pragma unroll
      for (int q = 0; q < 8; q++) {
        u[q] = (double)b.U[q * g.ncell + c];
        bad |= !__builtin_isfinite(u[q]);
      }
      u2prim(ph, u, p);
      if (ph.mhd) {
        const double b2 = (p[5] * p[5] + p[6] * p[6]) + p[7] * p[7];
        const double cx = cfast1(ph.gamma, p[4], p[0], b2, p[5]);
        const double cy = cfast1(ph.gamma, p[4], p[0], b2, p[6]);
        const double cz = cfast1(ph.gamma, p[4], p[0], b2, p[7]);
        dtp = dmin(dtp, b.dx / (__builtin_fabs(p[1]) + cx));
        dtp = dmin(dtp, b.dy / (__builtin_fabs(p[2]) + cy));
        dtp = dmin(dtp, b.dz / (__builtin_fabs(p[3]) + cz));
      } else {
        const double cs = csound(ph, p[4], p[0]);
        dtp = dmin(dtp, b.dx / (__builtin_fabs(p[1]) + cs));
        dtp = dmin(dtp, b.dy / (__builtin_fabs(p[2]) + cs));
        dtp = dmin(dtp, b.dz / (__builtin_fabs(p[3]) + cs));
      }
    }
  }
  if (__any(bad)) {
    if ((threadIdx.x & 63) == 0) atomicOr(err, 2);
  }
  dtp = wave_min(dtp);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = dtp;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = dmin(m, red[w]);
    atomicMin(dtmin, (unsigned long long)__double_as_longlong(m));
  }
}

// ---------------------------------------------------------------------------
// __global__ wrappers: one box ...
// ---------------------------------------------------------------------------
template <typename S>
__global__ void __launch_bounds__(256) k_prim(Box<S> b, int stage) { prim_body(b, stage, xcd_logical_block()); }

template <int SOLVER, int ORDER, int AXIS, typename S>
__global__ void __launch_bounds__(256) k_flux(Box<S> b, int* err) {
  flux_body<SOLVER, ORDER, AXIS>(b, xcd_logical_block(), err);
}

template <typename S>
__global__ void __launch_bounds__(256) k_efield(Box<S> b) { efield_body(b, xcd_logical_block()); }

template <int FLUXCD, typename S>
__global__ void __launch_bounds__(256) k_update(Box<S> b, double dt, int stage, int kchunk) {
  update_body<FLUXCD>(b, xcd_logical_block(), dt, stage, kchunk);
}

template <typename S>
__global__ void __launch_bounds__(256) k_cfl(Box<S> b, int kchunk, unsigned long long* dtmin, int* err) {
  cfl_body(b, xcd_logical_block(), kchunk, dtmin, err);
}

// ... and a batch of boxes
template <typename S>
__global__ void __launch_bounds__(256) kb_prim(const Box<S>* __restrict__ boxes, const unsigned* __restrict__ prefix,
                                               int nbox, int stage) {
  int bx;
  unsigned lb;
  batch_locate(prefix, nbox, xcd_logical_block(), bx, lb);
  prim_body(boxes[bx], stage, lb);
}

template <int SOLVER, int ORDER, int AXIS, typename S>
__global__ void __launch_bounds__(256) kb_flux(const Box<S>* __restrict__ boxes, const unsigned* __restrict__ prefix,
                                               int nbox, int* err) {
  int bx;
  unsigned lb;
  batch_locate(prefix, nbox, xcd_logical_block(), bx, lb);
  flux_body<SOLVER, ORDER, AXIS>(boxes[bx], lb, err);
}

template <typename S>
__global__ void __launch_bounds__(256) kb_efield(const Box<S>* __restrict__ boxes, const unsigned* __restrict__ prefix,
                                                 int nbox) {
  int bx;
  unsigned lb;
  batch_locate(prefix, nbox, xcd_logical_block(), bx, lb);
  efield_body(boxes[bx], lb);
}

// dt of each box from the device clock table; the predictor uses dt/2
// (hydro_solver.f90:152)
template <int FLUXCD, typename S>
__global__ void __launch_bounds__(256) kb_update(const Box<S>* __restrict__ boxes, const unsigned* __restrict__ prefix,
                                                 int nbox, const BoxClock* __restrict__ clk, int stage, int kchunk) {
  int bx;
  unsigned lb;
  batch_locate(prefix, nbox, xcd_logical_block(), bx, lb);
  const Box<S>& b = boxes[bx];
  const double dt = clk[b.slot].dt;
  update_body<FLUXCD>(b, lb, stage == 1 ? dt / 2.0 : dt, stage, kchunk);
}

template <typename S>
__global__ void __launch_bounds__(256) kb_cfl(const Box<S>* __restrict__ boxes, const unsigned* __restrict__ prefix,
                                              int nbox, int kchunk, BoxClock* clk, int* err) {
  int bx;
  unsigned lb;
  batch_locate(prefix, nbox, xcd_logical_block(), bx, lb);
  cfl_body(boxes[bx], lb, kchunk, &clk[boxes[bx].slot].dtmin, err);
}

// Start of a batched step: reset every box's CFL minimum.
__global__ void kb_clock_begin(BoxClock* clk, int nbox) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nbox) clk[b].dtmin = (unsigned long long)__double_as_longlong(1e30);
}

// get_timestep's ramp and clamp on the device (src/hydro_core.f90:677-694):
// dt = cfl * 2^-(n_iter+1-it) * dtp for it <= n_iter, cfl * dtp after;
// clamped to land on t_end, which completes the box.  A box that is done,
// or a timing-only launch (tuning != 0), steps with dt = 0: the fluxes are
// computed in full and the state keeps its values.
__global__ void kb_clock_dt(BoxClock* clk, int nbox, int tuning) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbox) return;
  BoxClock c = clk[b];
  if (c.done || tuning) {
    clk[b].dt = 0.0;
    return;
  }
  double dtp = __longlong_as_double((long long)c.dtmin);
  if (c.it <= c.n_iter) dtp = (c.cfl * __builtin_ldexp(1.0, -(c.n_iter + 1 - c.it))) * dtp;
  else dtp = c.cfl * dtp;
  double dt = dtp;
  bool finish = false;
  if ((c.time + dt) >= c.t_end) {
    dt = c.t_end - c.time;
    finish = true;
  }
  c.dt = dt;
  c.time = c.time + dt;
  c.it += 1;
  c.steps += 1;
  if (finish || (c.max_steps > 0 && c.steps >= c.max_steps)) c.done = 1;
  clk[b] = c;
}

// ---------------------------------------------------------------------------
// ghost fill of one face of the conserved state, boundaryI
// (src/boundaries.f90:44-247, serial branch): periodic copies over the full
// transverse range, closed/outflow over 0..n+1.
// ---------------------------------------------------------------------------
template <typename S>
__global__ void __launch_bounds__(256) k_ghost1(Geom g, S* __restrict__ U, int axis, int side, int kind) {
  const int n0 = axis == 0 ? g.ny : g.nx;  // transverse extents (interior)
  const int n1 = axis == 2 ? g.ny : g.nz;
  const int na = axis == 0 ? g.nx : axis == 1 ? g.ny : g.nz;
  const int lo = kind == BC_PERIODIC ? -1 : 0;
  const int e0 = n0 + 2 - 2 * lo, e1 = n1 + 2 - 2 * lo;
  const long long n = (long long)e0 * e1;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const int a = (int)(t % e0) + lo;
    const int b = (int)(t / e0) + lo;
    const int dst = side == 0 ? 0 : na + 1;
    int src;
    if (kind == BC_PERIODIC) src = side == 0 ? na : 1;
    else src = side == 0 ? 1 : na;
    int di, dj, dk, si, sj, sk;
    if (axis == 0) { di = dst; si = src; dj = sj = a; dk = sk = b; }
    else if (axis == 1) { dj = dst; sj = src; di = si = a; dk = sk = b; }
    else { dk = dst; sk = src; di = si = a; dj = sj = b; }
    const long long cd = cidx(g, di, dj, dk), cs = cidx(g, si, sj, sk);
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const S v = U[q * g.ncell + cs];
      U[q * g.ncell + cd] = (kind == BC_CLOSED && q == 1 + axis) ? -v : v;
    }
  }
}

// ---------------------------------------------------------------------------
// layout conversion between the reference's AoS u(neq,i,j,k) and SoA
// ---------------------------------------------------------------------------
template <typename S>
__global__ void __launch_bounds__(256) k_aos_to_soa(long long ncell, const double* __restrict__ A,
                                                    S* __restrict__ D) {
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < ncell;
       c += (long long)gridDim.x * blockDim.x) {
#pragma unroll
    for (int q = 0; q < 8; q++) D[q * ncell + c] = (S)A[c * 8 + q];
  }
}

template <typename S>
__global__ void __launch_bounds__(256) k_soa_to_aos(long long ncell, const S* __restrict__ D,
                                                    double* __restrict__ A) {
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < ncell;
       c += (long long)gridDim.x * blockDim.x) {
#pragma unroll
    for (int q = 0; q < 8; q++) A[c * 8 + q] = (double)D[q * ncell + c];
  }
}

// primitives and temperature of every cell as stored (calcprim over the
// whole array, hydro_core.f90:304-316), AoS output
template <typename S>
__global__ void __launch_bounds__(256) k_prim_all(long long ncell, Phys ph, const S* __restrict__ U,
                                                  double* __restrict__ Paos, double* __restrict__ T) {
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < ncell;
       c += (long long)gridDim.x * blockDim.x) {
    double u[8], p[8];
#pragma unroll
    for (int q = 0; q < 8; q++) u[q] = (double)U[q * ncell + c];
    u2prim(ph, u, p);
#pragma unroll
    for (int q = 0; q < 8; q++) Paos[c * 8 + q] = p[q];
    if (T) T[c] = (p[4] / p[0]) * 1.0;
  }
}

// divB of the dump (src/Out_BIN_Module.f90:214-222), divb(nx,ny,nz)
template <typename S>
__global__ void __launch_bounds__(256) k_divb(Geom g, const S* __restrict__ U, double dx, double dy,
                                              double dz, double* __restrict__ D) {
  const long long n = (long long)g.nx * g.ny * g.nz;
  const long long sx = 1, sy = g.NX, sz = (long long)g.NX * g.NY;
  const long long N = g.ncell;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(t % g.nx) + 1;
    const long long r = t / g.nx;
    const int j = (int)(r % g.ny) + 1;
    const int k = (int)(r / g.ny) + 1;
    const long long c = cidx(g, i, j, k);
    auto d = [](S v) { return (double)v; };
    D[t] = ((d(U[5 * N + c + sx]) - d(U[5 * N + c - sx])) / (2.0 * dx)
            + (d(U[6 * N + c + sy]) - d(U[6 * N + c - sy])) / (2.0 * dy))
           + (d(U[7 * N + c + sz]) - d(U[7 * N + c - sz])) / (2.0 * dz);
  }
}

// ---------------------------------------------------------------------------
// host-side launchers, instantiated for FP64 and FP32 storage
// ---------------------------------------------------------------------------
static inline unsigned grid_for(long long n, int block = 256) {
  long long b = (n + block - 1) / block;
  const long long cap = 256LL * 1024;  // grid-stride loops cover the rest
  if (b > cap) b = cap;
  if (b < 1) b = 1;
  return (unsigned)b;
}

static inline unsigned ceil_div(long long a, long long b) { return (unsigned)((a + b - 1) / b); }

unsigned box_tiles(const Geom& g, int kclass, const Shape& sh) {
  const long long B = sh.block;
  const unsigned nchunk = ceil_div(g.nz, sh.kchunk);
  switch (kclass) {
    case KC_PRIM1: return ceil_div((long long)(g.nx + 2) * (g.ny + 2), B) * (unsigned)(g.nz + 2);
    case KC_PRIM2: return ceil_div((long long)(g.nx + 4) * (g.ny + 4), B) * (unsigned)(g.nz + 4);
    case KC_FLUXX: return ceil_div((long long)(g.nx + 1) * g.ny, B) * (unsigned)g.nz;
    case KC_FLUXY: return ceil_div((long long)g.nx * (g.ny + 1), B) * (unsigned)g.nz;
    case KC_FLUXZ: return ceil_div((long long)g.nx * g.ny, B) * (unsigned)(g.nz + 1);
    case KC_EFIELD: return ceil_div((long long)g.nx * g.ny, B) * (unsigned)g.nz;
    case KC_UPDATE:
    case KC_CFLRED: return ceil_div((long long)g.nx * g.ny, B) * nchunk;
  }
  return 0;
}

template <typename S>
void launch_prim(hipStream_t s, const Box<S>& b, int stage, const Shape& sh) {
  const unsigned n = box_tiles(b.g, stage == 1 ? KC_PRIM1 : KC_PRIM2, sh);
  hipLaunchKernelGGL(k_prim<S>, dim3(n), dim3(sh.block), 0, s, b, stage);
}

template <int SOLVER, int ORDER, int AXIS, typename S>
static void launch_flux_a(hipStream_t s, const Box<S>& b, int* err, const Shape& sh) {
  const unsigned n = box_tiles(b.g, KC_FLUXX + AXIS, sh);
  hipLaunchKernelGGL((k_flux<SOLVER, ORDER, AXIS, S>), dim3(n), dim3(sh.block), 0, s, b, err);
}

template <int SOLVER, int ORDER, int AXIS, typename S>
static void launch_flux_ab(hipStream_t s, const Box<S>* boxes, const unsigned* prefix, int nbox, unsigned tiles,
                           int* err, const Shape& sh) {
  hipLaunchKernelGGL((kb_flux<SOLVER, ORDER, AXIS, S>), dim3(tiles), dim3(sh.block), 0, s, boxes, prefix, nbox, err);
}

// solver x order x axis dispatch for both launch forms
#define GM_FLUX_DISPATCH(CALL)                                                       \
  switch (solver) {                                                                  \
    case SOLVER_HLLD: GM_FLUX_OA(SOLVER_HLLD, CALL) break;                            \
    case SOLVER_HLLE: GM_FLUX_OA(SOLVER_HLLE, CALL) break;                            \
    case SOLVER_HLL: GM_FLUX_OA(SOLVER_HLL, CALL) break;                              \
    default: GM_FLUX_OA(SOLVER_HLLC, CALL) break;                                     \
  }
#define GM_FLUX_OA(SV, CALL)                                                         \
  if (order == 1) {                                                                  \
    if (axis == 0) CALL(SV, 1, 0) else if (axis == 1) CALL(SV, 1, 1) else CALL(SV, 1, 2) \
  } else {                                                                           \
    if (axis == 0) CALL(SV, 2, 0) else if (axis == 1) CALL(SV, 2, 1) else CALL(SV, 2, 2) \
  }

template <typename S>
void launch_flux(hipStream_t s, const Box<S>& b, int solver, int order, int axis, int* err, const Shape& sh) {
#define GM_CALL1(SV, O, A) { launch_flux_a<SV, O, A, S>(s, b, err, sh); }
  GM_FLUX_DISPATCH(GM_CALL1)
#undef GM_CALL1
}

template <typename S>
void launch_flux_batch(hipStream_t s, const Box<S>* boxes, const unsigned* prefix, int nbox, unsigned tiles,
                       int solver, int order, int axis, int* err, const Shape& sh) {
#define GM_CALLB(SV, O, A) { launch_flux_ab<SV, O, A, S>(s, boxes, prefix, nbox, tiles, err, sh); }
  GM_FLUX_DISPATCH(GM_CALLB)
#undef GM_CALLB
}
#undef GM_FLUX_OA
#undef GM_FLUX_DISPATCH

template <typename S>
void launch_efield(hipStream_t s, const Box<S>& b, const Shape& sh) {
  hipLaunchKernelGGL(k_efield<S>, dim3(box_tiles(b.g, KC_EFIELD, sh)), dim3(sh.block), 0, s, b);
}

template <typename S>
void launch_update(hipStream_t s, const Box<S>& b, double dt, int stage, int flux_cd, const Shape& sh) {
  const unsigned n = box_tiles(b.g, KC_UPDATE, sh);
  if (flux_cd) hipLaunchKernelGGL((k_update<1, S>), dim3(n), dim3(sh.block), 0, s, b, dt, stage, sh.kchunk);
  else hipLaunchKernelGGL((k_update<0, S>), dim3(n), dim3(sh.block), 0, s, b, dt, stage, sh.kchunk);
}

template <typename S>
void launch_cfl(hipStream_t s, const Box<S>& b, unsigned long long* dtmin, int* err, const Shape& sh) {
  hipLaunchKernelGGL(k_cfl<S>, dim3(box_tiles(b.g, KC_CFLRED, sh)), dim3(sh.block), 0, s, b, sh.kchunk, dtmin, err);
}

template <typename S>
void launch_prim_batch(hipStream_t s, const Box<S>* boxes, const unsigned* prefix, int nbox, unsigned tiles,
                       int stage, const Shape& sh) {
  hipLaunchKernelGGL(kb_prim<S>, dim3(tiles), dim3(sh.block), 0, s, boxes, prefix, nbox, stage);
}

template <typename S>
void launch_efield_batch(hipStream_t s, const Box<S>* boxes, const unsigned* prefix, int nbox, unsigned tiles,
                         const Shape& sh) {
  hipLaunchKernelGGL(kb_efield<S>, dim3(tiles), dim3(sh.block), 0, s, boxes, prefix, nbox);
}

template <typename S>
void launch_update_batch(hipStream_t s, const Box<S>* boxes, const unsigned* prefix, int nbox, unsigned tiles,
                         const BoxClock* clk, int stage, int flux_cd, const Shape& sh) {
  if (flux_cd)
    hipLaunchKernelGGL((kb_update<1, S>), dim3(tiles), dim3(sh.block), 0, s, boxes, prefix, nbox, clk, stage,
                       sh.kchunk);
  else
    hipLaunchKernelGGL((kb_update<0, S>), dim3(tiles), dim3(sh.block), 0, s, boxes, prefix, nbox, clk, stage,
                       sh.kchunk);
}

template <typename S>
void launch_cfl_batch(hipStream_t s, const Box<S>* boxes, const unsigned* prefix, int nbox, unsigned tiles,
                      BoxClock* clk, int* err, const Shape& sh) {
  hipLaunchKernelGGL(kb_cfl<S>, dim3(tiles), dim3(sh.block), 0, s, boxes, prefix, nbox, sh.kchunk, clk, err);
}

void launch_clock_begin(hipStream_t s, BoxClock* clk, int nbox) {
  hipLaunchKernelGGL(kb_clock_begin, dim3(ceil_div(nbox, 64)), dim3(64), 0, s, clk, nbox);
}

void launch_clock_dt(hipStream_t s, BoxClock* clk, int nbox, int tuning) {
  hipLaunchKernelGGL(kb_clock_dt, dim3(ceil_div(nbox, 64)), dim3(64), 0, s, clk, nbox, tuning);
}

template <typename S>
void launch_ghost1(hipStream_t s, const Geom& g, S* U, int axis, int side, int kind) {
  const int n0 = axis == 0 ? g.ny : g.nx;
  const int n1 = axis == 2 ? g.ny : g.nz;
  const long long n = (long long)(n0 + 4) * (n1 + 4);
  hipLaunchKernelGGL(k_ghost1<S>, dim3(grid_for(n)), dim3(256), 0, s, g, U, axis, side, kind);
}

template <typename S>
void launch_aos_to_soa(hipStream_t s, long long ncell, const double* A, S* D) {
  hipLaunchKernelGGL(k_aos_to_soa<S>, dim3(grid_for(ncell)), dim3(256), 0, s, ncell, A, D);
}

template <typename S>
void launch_soa_to_aos(hipStream_t s, long long ncell, const S* D, double* A) {
  hipLaunchKernelGGL(k_soa_to_aos<S>, dim3(grid_for(ncell)), dim3(256), 0, s, ncell, D, A);
}

template <typename S>
void launch_prim_all(hipStream_t s, long long ncell, const Phys& ph, const S* U, double* Paos, double* T) {
  hipLaunchKernelGGL(k_prim_all<S>, dim3(grid_for(ncell)), dim3(256), 0, s, ncell, ph, U, Paos, T);
}

template <typename S>
void launch_divb(hipStream_t s, const Geom& g, const S* U, double dx, double dy, double dz, double* D) {
  const long long n = (long long)g.nx * g.ny * g.nz;
  hipLaunchKernelGGL(k_divb<S>, dim3(grid_for(n)), dim3(256), 0, s, g, U, dx, dy, dz, D);
}

#define GM_INSTANTIATE(S)                                                                                        \
  template void launch_prim<S>(hipStream_t, const Box<S>&, int, const Shape&);                                    \
  template void launch_flux<S>(hipStream_t, const Box<S>&, int, int, int, int*, const Shape&);                    \
  template void launch_efield<S>(hipStream_t, const Box<S>&, const Shape&);                                       \
  template void launch_update<S>(hipStream_t, const Box<S>&, double, int, int, const Shape&);                     \
  template void launch_cfl<S>(hipStream_t, const Box<S>&, unsigned long long*, int*, const Shape&);               \
  template void launch_prim_batch<S>(hipStream_t, const Box<S>*, const unsigned*, int, unsigned, int,             \
                                     const Shape&);                                                               \
  template void launch_flux_batch<S>(hipStream_t, const Box<S>*, const unsigned*, int, unsigned, int, int, int,   \
                                     int*, const Shape&);                                                         \
  template void launch_efield_batch<S>(hipStream_t, const Box<S>*, const unsigned*, int, unsigned, const Shape&); \
  template void launch_update_batch<S>(hipStream_t, const Box<S>*, const unsigned*, int, unsigned,                \
                                       const BoxClock*, int, int, const Shape&);                                  \
  template void launch_cfl_batch<S>(hipStream_t, const Box<S>*, const unsigned*, int, unsigned, BoxClock*, int*,  \
                                    const Shape&);                                                                \
  template void launch_ghost1<S>(hipStream_t, const Geom&, S*, int, int, int);                                    \
  template void launch_aos_to_soa<S>(hipStream_t, long long, const double*, S*);                                  \
  template void launch_soa_to_aos<S>(hipStream_t, long long, const S*, double*);                                  \
  template void launch_prim_all<S>(hipStream_t, long long, const Phys&, const S*, double*, double*);              \
  template void launch_divb<S>(hipStream_t, const Geom&, const S*, double, double, double, double*);
GM_INSTANTIATE(double)
GM_INSTANTIATE(float)
#undef GM_INSTANTIATE

}  // namespace gm

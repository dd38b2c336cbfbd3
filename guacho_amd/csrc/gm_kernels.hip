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
// Tiles of 256 items; `bpp` tiles per plane.  Returns false past the end.
__device__ __forceinline__ bool tile_item(unsigned bpp, unsigned plane_n, int& t, int& plane) {
  const unsigned lb = xcd_logical_block();
  plane = (int)(lb / bpp);
  t = (int)((lb - (unsigned)plane * bpp) * 256u + threadIdx.x);
  return (unsigned)t < plane_n;
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
// primitives: calcprim over the cells the flux stencils read
// (src/hydro_core.f90:245-320), ghost values by boundary mapping.
// ---------------------------------------------------------------------------
template <typename S>
__global__ void __launch_bounds__(256) k_prim(Geom g, Phys ph, const S* __restrict__ U, S* __restrict__ P,
                                              int lo, unsigned bpp) {
  // cells lo..n+1-lo on each axis (lo = 0: one ghost layer, lo = -1: two)
  const int ex = g.nx + 2 - 2 * lo, ey = g.ny + 2 - 2 * lo;
  int t, kk;
  if (!tile_item(bpp, (unsigned)(ex * ey), t, kk)) return;
  const int i = t % ex + lo, j = t / ex + lo, k = kk + lo;
  // only face ghosts are read downstream; skip edges and corners
  const int out = (i < 1 || i > g.nx) + (j < 1 || j > g.ny) + (k < 1 || k > g.nz);
  if (out > 1) return;
  double u[8], p[8];
  load_cons(g, U, i, j, k, u);
  u2prim(ph, u, p);
  const long long c = cidx(g, i, j, k);
#pragma unroll
  for (int q = 0; q < 8; q++) P[q * g.ncell + c] = (S)p[q];
}

// ---------------------------------------------------------------------------
// interface fluxes: the sweep of hlldfluxes / hllefluxes / hllfluxes /
// hllcfluxes (e.g. src/hlld.f90:331-432) for one axis.  Face f(:,i,j,k)
// lies between cells i and i+1 (i = 0..n along the axis); only faces with
// interior transverse indices feed the update.
// ---------------------------------------------------------------------------
template <int SOLVER, int ORDER, int axis, typename S>
__global__ void __launch_bounds__(256) k_flux(Geom g, Phys ph, const S* __restrict__ P, S* __restrict__ F,
                                              int* __restrict__ err, unsigned bpp) {
  // face grid: along-axis index 0..n, transverse indices interior
  const int e0 = axis == 0 ? g.nx + 1 : g.nx;
  const int e1 = axis == 1 ? g.ny + 1 : g.ny;
  int t, k;
  if (!tile_item(bpp, (unsigned)(e0 * e1), t, k)) return;
  int i = t % e0, j = t / e0;
  if (axis != 0) i += 1;
  if (axis != 1) j += 1;
  if (axis != 2) k += 1;
  const long long stride = axis == 0 ? 1 : axis == 1 ? (long long)g.NX : (long long)g.NX * g.NY;
  const long long c = cidx(g, i, j, k);
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
    limiter(ph.limiter, pll, pl, pr, prr);
  }
  axis_perm<axis>(pl);
  axis_perm<axis>(pr);
  const bool ok = riemann<SOLVER>(ph, pl, pr, ff);
  if (!ok) atomicOr(err, 1);
  axis_perm<axis>(ff);
#pragma unroll
  for (int q = 0; q < 8; q++) F[q * g.ncell + c] = ff[q];
}

// ---------------------------------------------------------------------------
// cell-centred electric field, get_efield (src/flux_cd_module.f90:245-273)
// ---------------------------------------------------------------------------
template <typename S>
__global__ void __launch_bounds__(256) k_efield(Geom g, const S* __restrict__ Fx, const S* __restrict__ Fy,
                                                const S* __restrict__ Fz, S* __restrict__ E, unsigned bpp) {
  int t, k;
  if (!tile_item(bpp, (unsigned)(g.nx * g.ny), t, k)) return;
  const int i = t % g.nx + 1, j = t / g.nx + 1;
  k += 1;
  const long long sx = 1, sy = g.NX, sz = (long long)g.NX * g.NY;
  const long long N = g.ncell;
  const long long c = cidx(g, i, j, k);
  auto ld = [](const S* a, long long o) { return (double)a[o]; };
  E[0 * N + c] = (S)(0.25 * (((-ld(Fy, 7 * N + c - sy) - ld(Fy, 7 * N + c)) + ld(Fz, 6 * N + c - sz)) + ld(Fz, 6 * N + c)));
  E[1 * N + c] = (S)(0.25 * (((+ld(Fx, 7 * N + c - sx) + ld(Fx, 7 * N + c)) - ld(Fz, 5 * N + c - sz)) - ld(Fz, 5 * N + c)));
  E[2 * N + c] = (S)(0.25 * (((-ld(Fx, 6 * N + c - sx) - ld(Fx, 6 * N + c)) + ld(Fy, 5 * N + c - sy)) + ld(Fy, 5 * N + c)));
}

// ---------------------------------------------------------------------------
// upwind update, step(dt) (src/hydro_solver.f90:77-127) with flux_cd_update
// (src/flux_cd_module.f90:285-323).  Each thread marches up a z-column
// segment of KCHUNK cells, carrying the z-face flux below the cell and E
// one plane below and at the cell, so every z neighbour is read once.
//
// E ghosts (boundaryI_ef): the update reads e2, e3 at x neighbours, e1, e3
// at y neighbours and e1, e2 at z neighbours.  A closed boundary negates e1
// on x ghosts and e2 on y ghosts only (flux_cd_module.f90:145-176), never a
// component read there, so every E ghost resolves to a fixed interior
// neighbour offset: periodic wraps, closed and outflow copy the edge cell.
// Uout may alias Uin (each cell reads only its own conserved state).
// ---------------------------------------------------------------------------
constexpr int KCHUNK = 16;

// offset (in cells along the axis) of the E ghost-resolved neighbour
__device__ __forceinline__ int e_nbr(int i, int n, int d, int bc) {
  const int t = i + d;
  if (t >= 1 && t <= n) return d;
  if (bc == BC_PERIODIC) return (t < 1 ? t + n : t - n) - i;
  return 0;  // closed / outflow: ghost takes the edge cell's value
}

template <int FLUXCD, typename S>
__global__ void __launch_bounds__(256) k_update(Geom g, const S* Uin, S* Uout, const S* __restrict__ Fx,
                                                const S* __restrict__ Fy, const S* __restrict__ Fz,
                                                const S* __restrict__ E, double dt, double dx, double dy,
                                                double dz, unsigned bpp) {
  constexpr int NH = FLUXCD ? 5 : 8;
  int t, kc;
  if (!tile_item(bpp, (unsigned)(g.nx * g.ny), t, kc)) return;
  const int i = t % g.nx + 1, j = t / g.nx + 1;
  const int k0 = kc * KCHUNK + 1;
  const int k1 = min(k0 + KCHUNK - 1, g.nz);
  const long long sx = 1, sy = g.NX, sz = (long long)g.NX * g.NY;
  const long long N = g.ncell;
  const double dtdx = dt / dx, dtdy = dt / dy, dtdz = dt / dz;
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
__global__ void __launch_bounds__(256) k_cfl(Geom g, Phys ph, const S* __restrict__ U, double dx,
                                             double dy, double dz, unsigned long long* __restrict__ dtmin,
                                             int* __restrict__ err, unsigned bpp) {
  __shared__ double red[4];
  double dtp = 1e30;
  bool bad = false;
  int t, kc;
  if (tile_item(bpp, (unsigned)(g.nx * g.ny), t, kc)) {
    const int i = t % g.nx + 1, j = t / g.nx + 1;
    const int k0 = kc * KCHUNK + 1;
    const int k1 = min(k0 + KCHUNK - 1, g.nz);
    for (int k = k0; k <= k1; k++) {
      const long long c = cidx(g, i, j, k);
      double u[8], p[8];
#pragma unroll
      for (int q = 0; q < 8; q++) {
        u[q] = (double)U[q * g.ncell + c];
        bad |= !__builtin_isfinite(u[q]);
      }
      u2prim(ph, u, p);
      if (ph.mhd) {
        const double b2 = (p[5] * p[5] + p[6] * p[6]) + p[7] * p[7];
        const double cx = cfast1(ph.gamma, p[4], p[0], b2, p[5]);
        const double cy = cfast1(ph.gamma, p[4], p[0], b2, p[6]);
        const double cz = cfast1(ph.gamma, p[4], p[0], b2, p[7]);
        dtp = dmin(dtp, dx / (__builtin_fabs(p[1]) + cx));
        dtp = dmin(dtp, dy / (__builtin_fabs(p[2]) + cy));
        dtp = dmin(dtp, dz / (__builtin_fabs(p[3]) + cz));
      } else {
        const double cs = csound(ph, p[4], p[0]);
        dtp = dmin(dtp, dx / (__builtin_fabs(p[1]) + cs));
        dtp = dmin(dtp, dy / (__builtin_fabs(p[2]) + cs));
        dtp = dmin(dtp, dz / (__builtin_fabs(p[3]) + cs));
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

static inline unsigned tiles(long long items) { return (unsigned)((items + 255) / 256); }

template <typename S>
void launch_prim(hipStream_t s, const Geom& g, const Phys& ph, const S* U, S* P, int lo) {
  const int ex = g.nx + 2 - 2 * lo, ey = g.ny + 2 - 2 * lo, ez = g.nz + 2 - 2 * lo;
  const unsigned bpp = tiles((long long)ex * ey);
  hipLaunchKernelGGL(k_prim<S>, dim3(bpp * ez), dim3(256), 0, s, g, ph, U, P, lo, bpp);
}

template <int SOLVER, int ORDER, int AXIS, typename S>
static void launch_flux_a(hipStream_t s, const Geom& g, const Phys& ph, const S* P, S* F, int* err) {
  const long long e0 = AXIS == 0 ? g.nx + 1 : g.nx;
  const long long e1 = AXIS == 1 ? g.ny + 1 : g.ny;
  const long long e2 = AXIS == 2 ? g.nz + 1 : g.nz;
  const unsigned bpp = tiles(e0 * e1);
  hipLaunchKernelGGL((k_flux<SOLVER, ORDER, AXIS, S>), dim3(bpp * (unsigned)e2), dim3(256), 0, s, g, ph, P, F,
                     err, bpp);
}

template <int SOLVER, int ORDER, typename S>
static void launch_flux_t(hipStream_t s, const Geom& g, const Phys& ph, int axis, const S* P, S* F, int* err) {
  if (axis == 0) launch_flux_a<SOLVER, ORDER, 0>(s, g, ph, P, F, err);
  else if (axis == 1) launch_flux_a<SOLVER, ORDER, 1>(s, g, ph, P, F, err);
  else launch_flux_a<SOLVER, ORDER, 2>(s, g, ph, P, F, err);
}

template <typename S>
void launch_flux(hipStream_t s, const Geom& g, const Phys& ph, int solver, int order, int axis, const S* P,
                 S* F, int* err) {
#define GM_FLUX_CASE(SV)                                                     \
  case SV:                                                                   \
    if (order == 1) launch_flux_t<SV, 1>(s, g, ph, axis, P, F, err);         \
    else launch_flux_t<SV, 2>(s, g, ph, axis, P, F, err);                    \
    break;
  switch (solver) {
    GM_FLUX_CASE(SOLVER_HLLD)
    GM_FLUX_CASE(SOLVER_HLLE)
    GM_FLUX_CASE(SOLVER_HLL)
    GM_FLUX_CASE(SOLVER_HLLC)
  }
#undef GM_FLUX_CASE
}

template <typename S>
void launch_efield(hipStream_t s, const Geom& g, const S* Fx, const S* Fy, const S* Fz, S* E) {
  const unsigned bpp = tiles((long long)g.nx * g.ny);
  hipLaunchKernelGGL(k_efield<S>, dim3(bpp * g.nz), dim3(256), 0, s, g, Fx, Fy, Fz, E, bpp);
}

template <typename S>
void launch_update(hipStream_t s, const Geom& g, const S* Uin, S* Uout, const S* Fx, const S* Fy, const S* Fz,
                   const S* E, double dt, double dx, double dy, double dz, int flux_cd) {
  const unsigned bpp = tiles((long long)g.nx * g.ny);
  const unsigned nchunk = (unsigned)((g.nz + KCHUNK - 1) / KCHUNK);
  if (flux_cd)
    hipLaunchKernelGGL((k_update<1, S>), dim3(bpp * nchunk), dim3(256), 0, s, g, Uin, Uout, Fx, Fy, Fz, E, dt,
                       dx, dy, dz, bpp);
  else
    hipLaunchKernelGGL((k_update<0, S>), dim3(bpp * nchunk), dim3(256), 0, s, g, Uin, Uout, Fx, Fy, Fz, E, dt,
                       dx, dy, dz, bpp);
}

template <typename S>
void launch_cfl(hipStream_t s, const Geom& g, const Phys& ph, const S* U, double dx, double dy, double dz,
                unsigned long long* dtmin, int* err) {
  const unsigned bpp = tiles((long long)g.nx * g.ny);
  const unsigned nchunk = (unsigned)((g.nz + KCHUNK - 1) / KCHUNK);
  hipLaunchKernelGGL(k_cfl<S>, dim3(bpp * nchunk), dim3(256), 0, s, g, ph, U, dx, dy, dz, dtmin, err, bpp);
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

#define GM_INSTANTIATE(S)                                                                                       \
  template void launch_prim<S>(hipStream_t, const Geom&, const Phys&, const S*, S*, int);                        \
  template void launch_flux<S>(hipStream_t, const Geom&, const Phys&, int, int, int, const S*, S*, int*);        \
  template void launch_efield<S>(hipStream_t, const Geom&, const S*, const S*, const S*, S*);                    \
  template void launch_update<S>(hipStream_t, const Geom&, const S*, S*, const S*, const S*, const S*, const S*, \
                                 double, double, double, double, int);                                           \
  template void launch_cfl<S>(hipStream_t, const Geom&, const Phys&, const S*, double, double, double,           \
                              unsigned long long*, int*);                                                        \
  template void launch_ghost1<S>(hipStream_t, const Geom&, S*, int, int, int);                                   \
  template void launch_aos_to_soa<S>(hipStream_t, long long, const double*, S*);                                 \
  template void launch_soa_to_aos<S>(hipStream_t, long long, const S*, double*);                                 \
  template void launch_prim_all<S>(hipStream_t, long long, const Phys&, const S*, double*, double*);             \
  template void launch_divb<S>(hipStream_t, const Geom&, const S*, double, double, double, double*);
GM_INSTANTIATE(double)
GM_INSTANTIATE(float)
#undef GM_INSTANTIATE

}  // namespace gm

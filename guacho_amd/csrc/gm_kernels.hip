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

#ifndef GM_KERNELS_PART
#define GM_KERNELS_PART 0
#endif
#define GM_PART(n) (GM_KERNELS_PART == 0 || GM_KERNELS_PART == (n))

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

// Source index of ghost index i: the boundary mapping, or i itself when
// the ghosts are materialised in memory (user boundary hooks).
__device__ __forceinline__ int ghost_src(const Geom& g, int i, int n, int lo, int hi, bool& flip) {
  if (g.gmem) return i;
  return map_ghost(i, n, lo, hi, flip);
}

// Powell's 8-wave terms of cell (i,j,k) (divbcorr_8w_source,
// src/sources.f90:137-168) from the stage's input state U, added to sv as
// source() adds them (sv - divB*..., :188-220): divB by central differences
// of the neighbours' field (ghosts by boundary mapping; the field is never
// flipped), v and B of the cell as u2prim gives them.
template <typename S>
__device__ __forceinline__ void w8_terms(const Box<S>& b, const S* __restrict__ U, int i, int j, int k, double sv[8]) {
  const Geom& g = b.g;
  const long long N = g.ncell, sy = g.NX, sz = (long long)g.NX * g.NY;
  const long long c = cidx(g, i, j, k);
  bool f = false;
  const long long oxm = ghost_src(g, i - 1, g.nx, g.bc[0], g.bc[1], f) - i;
  const long long oxp = ghost_src(g, i + 1, g.nx, g.bc[0], g.bc[1], f) - i;
  const long long oym = (ghost_src(g, j - 1, g.ny, g.bc[2], g.bc[3], f) - j) * sy;
  const long long oyp = (ghost_src(g, j + 1, g.ny, g.bc[2], g.bc[3], f) - j) * sy;
  const long long ozm = (ghost_src(g, k - 1, g.nz, g.bc[4], g.bc[5], f) - k) * sz;
  const long long ozp = (ghost_src(g, k + 1, g.nz, g.bc[4], g.bc[5], f) - k) * sz;
  auto d = [&](int q, long long o) { return (double)U[q * N + c + o]; };
  const double divb = (gm_div(d(5, oxp) - d(5, oxm), 2.0 * b.dx) + gm_div(d(6, oyp) - d(6, oym), 2.0 * b.dy))
                      + gm_div(d(7, ozp) - d(7, ozm), 2.0 * b.dz);
  const double r0 = d(0, 0);
  const double r = (r0 > 1e-15) ? r0 : 1e-15;
  const double v1 = gm_div(d(1, 0), r), v2 = gm_div(d(2, 0), r), v3 = gm_div(d(3, 0), r);
  const double b1 = d(5, 0), b2 = d(6, 0), b3 = d(7, 0);
  sv[1] = sv[1] - divb * b1;
  sv[2] = sv[2] - divb * b2;
  sv[3] = sv[3] - divb * b3;
  sv[4] = sv[4] - divb * ((v1 * b1 + v2 * b2) + v3 * b3);
  sv[5] = sv[5] - divb * v1;
  sv[6] = sv[6] - divb * v2;
  sv[7] = sv[7] - divb * v3;
}

// Conserved variables (NQ: the 8 MHD ones and the passive scalars) of
// cell (i,j,k) in FP64, ghosts resolved through the boundary conditions.
template <int NQ = 8, typename S>
__device__ __forceinline__ void load_cons(const Geom& g, const S* __restrict__ U, int i, int j, int k,
                                          double u[NQ]) {
  bool fx = false, fy = false, fz = false;
  const int si = ghost_src(g, i, g.nx, g.bc[0], g.bc[1], fx);
  const int sj = ghost_src(g, j, g.ny, g.bc[2], g.bc[3], fy);
  const int sk = ghost_src(g, k, g.nz, g.bc[4], g.bc[5], fz);
  const long long c = cidx(g, si, sj, sk);
#pragma unroll
  for (int q = 0; q < NQ; q++) u[q] = (double)U[q * g.ncell + c];
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
//
// One stage of tstep (src/hydro_solver.f90:134-229) is three kernels:
//   xy:     primitives of a 32 x 16 plane tile and its halo into LDS
//           (calcprim, src/hydro_core.f90:245-320), the x and y interface
//           fluxes (e.g. hlldfluxes, src/hlld.f90:331-432), the x and y terms
//           of the update, E3 and the x/y parts of E1, E2;
//   z:      per column, a march up a z-pencil: primitives, the z interface
//           fluxes, the z term of the update and E1, E2
//           (get_efield, src/flux_cd_module.f90:245-273);
//   bfield: the flux-CD field update (flux_cd_update,
//           src/flux_cd_module.f90:285-323) from the cell-centred E.
// Neither primitives nor fluxes reach HBM.  Every sum keeps the reference's
// evaluation order (step, src/hydro_solver.f90:77-127: the x, y, z terms
// subtracted from U in that order; get_efield's four-term sums), with the
// partial sums carried between kernels in FP64, so the split is exact.
// ---------------------------------------------------------------------------

// Plane tile of the xy kernel: TX x TY interior cells of one z-plane, one
// thread per cell.  Wave w holds rows 2w (lanes 0-31) and 2w+1 (lanes
// 32-63), so a row is half a wavefront and x neighbours are adjacent lanes.
// The slab kernel (2.5-D) uses the same layout with TYS = 8 rows.
constexpr int TX = 32, TY = 16, XY_THREADS = TX * TY, XY_WAVES = XY_THREADS / 64;
static_assert(TX == 32 && XY_THREADS == 512, "two tile rows per 64-lane wavefront");
constexpr int TYS = 8, SL_THREADS = TX * TYS;

// LDS of a plane tile of TYT rows for reconstruction order ORDER (halo
// H = ORDER).  Primitives [8][RY][RX] of the tile and its x and y halos
// (corners are not loaded); after the face solves that area holds the
// y-face fluxes [8][TYT][TX].  Then the fluxes of the tile's closing faces:
// y-face TYT of each column [8][TX] and x-face TX of each row [8][TYT].
// Second order adds reconstructed face states: the upper x state of the
// halo cell (-1, r), of the tile's last cell (TX-1, r) and the lower x
// state of the halo cell (TX, r) [8][TYT] each; the lower y state of the
// halo cell (c, TYT) [8][TX]; and the upper y states of rows -1, 1, 3, ..,
// TYT-1 [8][TYT/2+1][TX] (the rows an even row, held in the other
// wavefront, reads below itself).
template <int ORDER, int NQ = 8, int TYT = TY>
struct XYLds {
  static constexpr int H = ORDER;
  static constexpr int RX = TX + 2 * H, RY = TYT + 2 * H;
  static constexpr int PS = RX * RY;  // component stride of the primitives
  static constexpr int PRIM = NQ * PS;
  static constexpr int NLOAD = 2 * H * TX + TYT * RX;  // primitives loaded
  static constexpr int TOP = PRIM;
  static constexpr int XEND = TOP + NQ * TX;
  static constexpr int XL = XEND + NQ * TYT;
  static constexpr int XU = XL + NQ * TYT;
  static constexpr int XR = XU + NQ * TYT;
  static constexpr int YD = XR + NQ * TYT;
  static constexpr int YU = YD + NQ * TX;
  static constexpr int NYU = TYT / 2 + 1;
  static constexpr int SIZE = ORDER == 1 ? XL : YU + NQ * NYU * TX;
  static_assert(NQ * TYT * TX <= PRIM, "y-face fluxes must fit the primitive area");
};

// Reconstructed states of one cell along one axis, from the primitives in
// LDS (`at` the cell, `stride` one step along the axis): up is the left
// state of the face above the cell, dn the right state of the face below.
// limiter() (src/hydro_core.f90:712-731) forms both from the one slope
// average(P - P-, P+ - P) of the cell, so each cell's slope serves both of
// its faces.
template <int NQ, typename S>
__device__ __forceinline__ void cell_states(int lim, const S* __restrict__ lp, int ps, int at, int stride, S up[NQ],
                                            S dn[NQ]) {
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const S p = lp[q * ps + at];
    const S s = lim_average(lim, p - lp[q * ps + at - stride], lp[q * ps + at + stride] - p);
    up[q] = p + s * S(0.5);
    dn[q] = p - s * S(0.5);
  }
}

// One of the two: the upper state when `up` holds, else the lower.
template <int NQ, typename S>
__device__ __forceinline__ void cell_state(int lim, const S* __restrict__ lp, int ps, int at, int stride, bool up,
                                           S v[NQ]) {
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const S p = lp[q * ps + at];
    const S s = lim_average(lim, p - lp[q * ps + at - stride], lp[q * ps + at + stride] - p);
    const S hs = s * S(0.5);
    v[q] = up ? p + hs : p - hs;
  }
}

// y faces exchange slots (1,2) and (5,6) when `ydir` holds (axis_perm<1>
// with a run-time axis, by value selects).
template <typename T>
__device__ __forceinline__ void perm_y_if(bool ydir, T* v) {
  const T a = ydir ? v[2] : v[1], b = ydir ? v[1] : v[2];
  const T e = ydir ? v[6] : v[5], f = ydir ? v[5] : v[6];
  v[1] = a;
  v[2] = b;
  v[5] = e;
  v[6] = f;
}

// Cross-lane moves within one 64-lane wavefront by DPP and v_permlane32_swap
// (no LDS permute): lane_up1 gives lane l the value of lane l-1 (DPP
// wave_shr:1), lane_down1 that of lane l+1 (wave_shl:1), lane_up32 gives
// lanes 32..63 the values of lanes 0..31.  A lane with no source lane (lane
// 0, lane 63, lanes 0..31 in that order) reads 0, 0 and its own value; the
// callers replace those.  Every lane of the wavefront must be active.
template <int CTRL>
__device__ __forceinline__ unsigned dpp_word(unsigned v) {
  return (unsigned)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xf, 0xf, true);
}
__device__ __forceinline__ unsigned half_word(unsigned v) { return __builtin_amdgcn_permlane32_swap(v, v, false, false)[0]; }
template <int CTRL, typename S>
__device__ __forceinline__ S lane_move(S v) {
  if constexpr (sizeof(S) == 8) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = CTRL ? dpp_word<CTRL>((unsigned)u) : half_word((unsigned)u);
    const unsigned hi = CTRL ? dpp_word<CTRL>((unsigned)(u >> 32)) : half_word((unsigned)(u >> 32));
    return __builtin_bit_cast(S, ((unsigned long long)hi << 32) | lo);
  } else {
    const unsigned u = __builtin_bit_cast(unsigned, v);
    return __builtin_bit_cast(S, CTRL ? dpp_word<CTRL>(u) : half_word(u));
  }
}
template <typename S>
__device__ __forceinline__ S lane_up1(S v) { return lane_move<0x138>(v); }
template <typename S>
__device__ __forceinline__ S lane_down1(S v) { return lane_move<0x130>(v); }
template <typename S>
__device__ __forceinline__ S lane_up32(S v) { return lane_move<0>(v); }

// Faces of one plane tile, from the primitives in LDS (stored, then a
// barrier) to the four fluxes of the thread's own cell: fx (low x face),
// fxr (high x face), fy (low y face), fyu (high y face).  Thread (c, r) --
// lane c + 32 (r & 1) of wave r/2 -- solves the low faces of cell (c, r):
// x-face c of row r (between cells c-1 and c) and y-face r of column c.
// Its high x face is lane c+1's, its high y face row r+1's; the closing
// faces of a full tile (x-face TX of each row, y-face TYT of each column)
// are solved together by wave 0.  In a partial tile the closing faces are
// the low faces of the halo cells just past it, solved by the threads of
// those cells.
//
// Second order: every thread forms the x and y states of its own cell
// (cell_states), the last two waves those of the 2 (TX + TYT) halo cells
// next to the tile.  The x faces are solved before the y states are formed,
// which keeps the live register set of one axis at a time.  Every thread
// must call this (it holds barriers); the outputs are meaningful for the
// tile's own cells only.  Returns false when a solver fell through.  `tid`
// is threadIdx.x.
template <int SOLVER, int ORDER, int NQ, int TYT, typename S, bool CS = false>
__device__ __forceinline__ bool xy_solve(const Box<S>& b, S* __restrict__ lds, int wx, int wy, int tid, S fx[NQ],
                                         S fy[NQ], const S* __restrict__ cfs = nullptr) {
  using L = XYLds<ORDER, NQ, TYT>;
  constexpr int H = L::H, RX = L::RX, PS = L::PS;
  constexpr int WAVES = TX * TYT / 64;
  static_assert(TX + TYT <= 64, "closing faces in one wavefront");
  static_assert(2 * (TX + TYT) <= 128, "halo states in the last two wavefronts");
  const int w = tid >> 6, l = tid & 63;
  const int c = l & (TX - 1), r = 2 * w + (l >> 5);
  S* __restrict__ lp = lds;
  S* __restrict__ ltop = lds + L::TOP;
  S* __restrict__ lend = lds + L::XEND;
  const int at = (r + H) * RX + c + H;  // own cell in the primitive area
  const bool vx = r < wy && c <= wx;    // x-face c of row r is the tile's
  const bool vy = c < wx && r <= wy;    // y-face r of column c is the tile's
  const bool cx = w == 0 && l < TYT;                   // closing x-face of row l
  const bool cy = w == 0 && l >= TYT && l < TYT + TX;   // closing y-face of column l - TYT
  const int crow = cx ? l : 0, ccol = cy ? l - TYT : 0;
  bool ok = true;
#pragma unroll
  for (int q = 0; q < NQ; q++) fx[q] = fy[q] = S(0);

  if constexpr (ORDER == 1) {
    // CS: the cells' side speeds are in LDS beside the primitives (x at
    // cfs[cell], y at cfs[PS + cell]; side_speed of the cell's primitives
    // in that axis's order)
    S pl[NQ], pr[NQ];
    if (vx) {
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        pl[q] = lp[q * PS + at - 1];
        pr[q] = lp[q * PS + at];
      }
      if constexpr (CS) ok &= riemann_cs<SOLVER, NQ>(b.ph, pl, pr, cfs[at - 1], cfs[at], fx);
      else ok &= riemann<SOLVER, NQ>(b.ph, pl, pr, fx);
    }
    if (vy) {
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        pl[q] = lp[q * PS + at - RX];
        pr[q] = lp[q * PS + at];
      }
      axis_perm<1>(pl);
      axis_perm<1>(pr);
      if constexpr (CS) ok &= riemann_cs<SOLVER, NQ>(b.ph, pl, pr, cfs[PS + at - RX], cfs[PS + at], fy);
      else ok &= riemann<SOLVER, NQ>(b.ph, pl, pr, fy);
      axis_perm<1>(fy);
    }
    if ((cx && wx == TX && crow < wy) || (cy && wy == TYT && ccol < wx)) {
      const int hi = cy ? (TYT + H) * RX + ccol + H : (crow + H) * RX + TX + H;
      const int st = cy ? RX : 1;
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        pl[q] = lp[q * PS + hi - st];
        pr[q] = lp[q * PS + hi];
      }
      S ff[NQ];
      perm_y_if(cy, pl);
      perm_y_if(cy, pr);
      if constexpr (CS) {
        const S* __restrict__ cf = cfs + (cy ? PS : 0);
        ok &= riemann_cs<SOLVER, NQ>(b.ph, pl, pr, cf[hi - st], cf[hi], ff);
      } else {
        ok &= riemann<SOLVER, NQ>(b.ph, pl, pr, ff);
      }
      perm_y_if(cy, ff);
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        if (cy) ltop[q * TX + ccol] = ff[q];
        else lend[q * TYT + crow] = ff[q];
      }
    }
  } else {
    S* __restrict__ xl = lds + L::XL;
    S* __restrict__ xu = lds + L::XU;
    S* __restrict__ xr = lds + L::XR;
    S* __restrict__ yd = lds + L::YD;
    S* __restrict__ yu = lds + L::YU;
    const int lim = b.ph.limiter;
    // halo cells, into LDS, by the last two waves (h = 0..95): the x
    // states of (-1, r) and (TX, r), the y states of (c, -1) and (c, TYT)
    if (w >= WAVES - 2) __builtin_amdgcn_s_setprio(2);
    if (w >= WAVES - 2) {
      const int h = (w - (WAVES - 2)) * 64 + l;
      if (h < 2 * TYT + 2 * TX) {
        int hat, hst, hdst, hqs;
        bool hup;
        if (h < TYT) {
          hat = (h + H) * RX + H - 1, hst = 1, hup = true, hdst = L::XL + h, hqs = TYT;
        } else if (h < 2 * TYT) {
          hat = (h - TYT + H) * RX + H + TX, hst = 1, hup = false, hdst = L::XR + h - TYT, hqs = TYT;
        } else if (h < 2 * TYT + TX) {
          hat = (H - 1) * RX + h - 2 * TYT + H, hst = RX, hup = true, hdst = L::YU + h - 2 * TYT, hqs = L::NYU * TX;
        } else {
          hat = (H + TYT) * RX + h - 2 * TYT - TX + H, hst = RX, hup = false, hdst = L::YD + h - 2 * TYT - TX, hqs = TX;
        }
        S hs[NQ];
        cell_state<NQ>(lim, lp, PS, hat, hst, hup, hs);
#pragma unroll
        for (int q = 0; q < NQ; q++) lds[hdst + q * hqs] = hs[q];
      }
    }
    if (w >= WAVES - 2) __builtin_amdgcn_s_setprio(0);
    // x: the upper state passes one lane up (lane c-1 holds cell c-1)
    S pl[NQ], pr[NQ];
    {
      S pxu[NQ];
      cell_states<NQ>(lim, lp, PS, at, 1, pxu, pr);
      if (c == TX - 1) {
#pragma unroll
        for (int q = 0; q < NQ; q++) xu[q * TYT + r] = pxu[q];
      }
#pragma unroll
      for (int q = 0; q < NQ; q++) pl[q] = lane_up1(pxu[q]);
    }
    __syncthreads();  // halo states in LDS
    if (c == 0) {
#pragma unroll
      for (int q = 0; q < NQ; q++) pl[q] = xl[q * TYT + r];
    }
    if (vx) ok &= riemann<SOLVER, NQ>(b.ph, pl, pr, fx);
    // y: odd rows hand their upper state to the even row above them (the
    // next wavefront) through LDS, even rows theirs to the odd row above
    // (lanes 32 up in the same wavefront)
    {
      S pyu[NQ];
      cell_states<NQ>(lim, lp, PS, at, RX, pyu, pr);
      if (r & 1) {
#pragma unroll
        for (int q = 0; q < NQ; q++) yu[(q * L::NYU + (r + 1) / 2) * TX + c] = pyu[q];
      }
#pragma unroll
      for (int q = 0; q < NQ; q++) pl[q] = lane_up32(pyu[q]);
    }
    __syncthreads();  // every read of the primitives is done; y states in LDS
    if (w == 0) __builtin_amdgcn_s_setprio(2);
    if (!(r & 1)) {
#pragma unroll
      for (int q = 0; q < NQ; q++) pl[q] = yu[(q * L::NYU + r / 2) * TX + c];
    }
    if (vy) {
      axis_perm<1>(pl);
      axis_perm<1>(pr);
      ok &= riemann<SOLVER, NQ>(b.ph, pl, pr, fy);
      axis_perm<1>(fy);
    }
    if ((cx && wx == TX && crow < wy) || (cy && wy == TYT && ccol < wx)) {
      S ff[NQ];
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        pl[q] = cy ? yu[(q * L::NYU + TYT / 2) * TX + ccol] : xu[q * TYT + crow];
        pr[q] = cy ? yd[q * TX + ccol] : xr[q * TYT + crow];
      }
      perm_y_if(cy, pl);
      perm_y_if(cy, pr);
      ok &= riemann<SOLVER, NQ>(b.ph, pl, pr, ff);
      perm_y_if(cy, ff);
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        if (cy) ltop[q * TX + ccol] = ff[q];
        else lend[q * TYT + crow] = ff[q];
      }
    }
  }
  if (w == 0) __builtin_amdgcn_s_setprio(0);
  return ok;
}

// ... then the exchange: each thread's high x face (lane c+1's low face, or
// the row's closing face) into fxr, the tile's y-face fluxes into LDS for
// xy_fyu.  Every thread calls this (barriers); the cell's state for the
// update is best fetched between xy_solve and this call.
template <int ORDER, int NQ, int TYT, typename S>
__device__ __forceinline__ void xy_exchange(S* __restrict__ lds, int wx, int wy, int tid, const S fx[NQ],
                                            const S fy[NQ], S fxr[NQ]) {
  using L = XYLds<ORDER, NQ, TYT>;
  const int w = tid >> 6, l = tid & 63;
  const int c = l & (TX - 1), r = 2 * w + (l >> 5);
  const bool vy = c < wx && r <= wy;
  S* __restrict__ lend = lds + L::XEND;
  if constexpr (ORDER == 1) __syncthreads();  // every read of the primitives is done
  S* __restrict__ lfy = lds;
  if (vy) {
#pragma unroll
    for (int q = 0; q < NQ; q++) lfy[(q * TYT + r) * TX + c] = fy[q];
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const S v = lane_down1(fx[q]);
    fxr[q] = c == TX - 1 ? lend[q * TYT + r] : v;
  }
}

// the high y face of the thread's cell (row r+1's low face, or the
// column's closing face), after xy_exchange
template <int ORDER, int NQ, int TYT, typename S>
__device__ __forceinline__ void xy_fyu(const S* __restrict__ lds, int tid, S fyu[NQ]) {
  using L = XYLds<ORDER, NQ, TYT>;
  const int w = tid >> 6, l = tid & 63;
  const int c = l & (TX - 1), r = 2 * w + (l >> 5);
  const S* __restrict__ ltop = lds + L::TOP;
#pragma unroll
  for (int q = 0; q < NQ; q++) fyu[q] = r == TYT - 1 ? ltop[q * TX + c] : lds[(q * TYT + r + 1) * TX + c];
}

// xy kernel: the primitives of a 32 x 16 plane tile and its halo into LDS,
// the tile's x and y faces, then the x/y terms of the update, E3 and the
// x/y parts of E1, E2 into W for the z kernel.  The faces are solved as in
// xy_solve / xy_exchange / xy_fyu (thread (c, r) solves the low faces of
// cell (c, r), wave 0 the closing faces, the last two waves the halo
// states), written out here in one piece: at 4 waves per SIMD this kernel
// sits at 127-128 VGPRs, and the same code through those three functions
// is scheduled differently and spills 21-22 of them (about 10% slower).
// The two forms are checked equal bit for bit by the slab-vs-two-kernel
// tests.
template <int SOLVER, int ORDER, int FLUXCD, typename S, int NP = 0>
__device__ __forceinline__ void xy_body(const Box<S>& b, unsigned lb, double dt, int* __restrict__ err,
                                        S* __restrict__ lds) {
  constexpr int NQ = 8 + NP;  // the MHD variables and the passive scalars
  using L = XYLds<ORDER, NQ>;
  constexpr int H = L::H, RX = L::RX, PS = L::PS;
  constexpr int NH = FLUXCD ? 5 : 8;
  const Geom& g = b.g;
  const int ntx = (g.nx + TX - 1) / TX, nty = (g.ny + TY - 1) / TY;
  const unsigned per_plane = (unsigned)(ntx * nty);
  const int k = (int)(lb / per_plane) + 1;
  const int rem = (int)(lb - (unsigned)(k - 1) * per_plane);
  const int i0 = (rem % ntx) * TX + 1, j0 = (rem / ntx) * TY + 1;
  const int wx = min(TX, g.nx - i0 + 1), wy = min(TY, g.ny - j0 + 1);
  const int tid = (int)threadIdx.x, w = tid >> 6, l = tid & 63;
  const int c = l & (TX - 1), r = 2 * w + (l >> 5);
  const S* __restrict__ Usrc = ORDER == 1 ? b.U : b.Us;
  S* __restrict__ lp = lds;
  S* __restrict__ ltop = lds + L::TOP;
  S* __restrict__ lend = lds + L::XEND;

  // primitives of the tile, its x halo on the tile's rows and its y halo
  // on the tile's columns (H rows below, TY rows of RX, H rows above);
  // ghosts by boundary mapping
  // (at most two cells a thread; both cells' loads are issued before either
  // recovery, so their latencies overlap)
  static_assert(L::NLOAD <= 2 * XY_THREADS, "two primitive loads per thread");
  // Wave priorities: a new workgroup issues its tile loads ahead of the
  // older workgroup's arithmetic (issue is oldest-first otherwise), and the
  // waves with extra work before a barrier (halo states, closing faces) run
  // ahead of the rest.  Priority changes issue order only.
  __builtin_amdgcn_s_setprio(3);
  {
    double u[2][NQ];
    int dst[2];
#pragma unroll
    for (int it = 0; it < 2; it++) {
      const int idx = tid + it * XY_THREADS;
      int rx, ry;
      if (idx < H * TX) {
        ry = idx / TX;
        rx = H + idx % TX;
      } else if (idx < H * TX + TY * RX) {
        const int t = idx - H * TX;
        ry = H + t / RX;
        rx = t % RX;
      } else {
        const int t = idx - H * TX - TY * RX;
        ry = H + TY + t / TX;
        rx = H + t % TX;
      }
      dst[it] = -1;
      if (idx < L::NLOAD && rx < wx + 2 * H && ry < wy + 2 * H) {
        dst[it] = ry * RX + rx;
        load_cons<NQ>(g, Usrc, i0 - H + rx, j0 - H + ry, k, u[it]);
      }
    }
    __builtin_amdgcn_s_setprio(0);
#pragma unroll
    for (int it = 0; it < 2; it++) {
      if (dst[it] < 0) continue;
      double p[NQ];
      u2prim<NQ>(b.ph, u[it], p);
      if (ORDER == 1 && b.pc) {  // Box::pc: interior cells only
        const int hi = i0 - H + dst[it] % RX, hj = j0 - H + dst[it] / RX;
        if (hi >= 1 && hi <= g.nx && hj >= 1 && hj <= g.ny) p[4] = b.pc[cidx(g, hi, hj, k)];
      }
#pragma unroll
      for (int q = 0; q < NQ; q++) lp[q * PS + dst[it]] = (S)p[q];
    }
  }
  __syncthreads();

  const int at = (r + H) * RX + c + H;  // own cell in the primitive area
  const bool vx = r < wy && c <= wx;    // x-face c of row r is the tile's
  const bool vy = c < wx && r <= wy;    // y-face r of column c is the tile's
  const bool cx = w == 0 && l < TY;                   // closing x-face of row l
  const bool cy = w == 0 && l >= TY && l < TY + TX;   // closing y-face of column l - TY
  const int crow = cx ? l : 0, ccol = cy ? l - TY : 0;
  bool ok = true;
  S fx[NQ], fy[NQ];
#pragma unroll
  for (int q = 0; q < NQ; q++) fx[q] = fy[q] = S(0);

  if constexpr (ORDER == 1) {
    S pl[NQ], pr[NQ];
    if (vx) {
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        pl[q] = lp[q * PS + at - 1];
        pr[q] = lp[q * PS + at];
      }
      ok &= riemann<SOLVER, NQ>(b.ph, pl, pr, fx);
    }
    if (vy) {
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        pl[q] = lp[q * PS + at - RX];
        pr[q] = lp[q * PS + at];
      }
      axis_perm<1>(pl);
      axis_perm<1>(pr);
      ok &= riemann<SOLVER, NQ>(b.ph, pl, pr, fy);
      axis_perm<1>(fy);
    }
    if ((cx && wx == TX && crow < wy) || (cy && wy == TY && ccol < wx)) {
      const int hi = cy ? (TY + H) * RX + ccol + H : (crow + H) * RX + TX + H;
      const int st = cy ? RX : 1;
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        pl[q] = lp[q * PS + hi - st];
        pr[q] = lp[q * PS + hi];
      }
      S ff[NQ];
      perm_y_if(cy, pl);
      perm_y_if(cy, pr);
      ok &= riemann<SOLVER, NQ>(b.ph, pl, pr, ff);
      perm_y_if(cy, ff);
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        if (cy) ltop[q * TX + ccol] = ff[q];
        else lend[q * TY + crow] = ff[q];
      }
    }
  } else {
    S* __restrict__ xl = lds + L::XL;
    S* __restrict__ xu = lds + L::XU;
    S* __restrict__ xr = lds + L::XR;
    S* __restrict__ yd = lds + L::YD;
    S* __restrict__ yu = lds + L::YU;
    const int lim = b.ph.limiter;
    // halo cells, into LDS, by the last two waves (h = 0..95): the x
    // states of (-1, r) and (TX, r), the y states of (c, -1) and (c, TY)
    if (w >= XY_WAVES - 2) __builtin_amdgcn_s_setprio(2);
    if (w >= XY_WAVES - 2) {
      const int h = (w - (XY_WAVES - 2)) * 64 + l;
      if (h < 2 * TY + 2 * TX) {
        int hat, hst, hdst, hqs;
        bool hup;
        if (h < TY) {
          hat = (h + H) * RX + H - 1, hst = 1, hup = true, hdst = L::XL + h, hqs = TY;
        } else if (h < 2 * TY) {
          hat = (h - TY + H) * RX + H + TX, hst = 1, hup = false, hdst = L::XR + h - TY, hqs = TY;
        } else if (h < 2 * TY + TX) {
          hat = (H - 1) * RX + h - 2 * TY + H, hst = RX, hup = true, hdst = L::YU + h - 2 * TY, hqs = L::NYU * TX;
        } else {
          hat = (H + TY) * RX + h - 2 * TY - TX + H, hst = RX, hup = false, hdst = L::YD + h - 2 * TY - TX, hqs = TX;
        }
        S hs[NQ];
        cell_state<NQ>(lim, lp, PS, hat, hst, hup, hs);
#pragma unroll
        for (int q = 0; q < NQ; q++) lds[hdst + q * hqs] = hs[q];
      }
    }
    if (w >= XY_WAVES - 2) __builtin_amdgcn_s_setprio(0);
    // x: the upper state passes one lane up (lane c-1 holds cell c-1)
    S pl[NQ], pr[NQ];
    {
      S pxu[NQ];
      cell_states<NQ>(lim, lp, PS, at, 1, pxu, pr);
      if (c == TX - 1) {
#pragma unroll
        for (int q = 0; q < NQ; q++) xu[q * TY + r] = pxu[q];
      }
#pragma unroll
      for (int q = 0; q < NQ; q++) pl[q] = lane_up1(pxu[q]);
    }
    __syncthreads();  // halo states in LDS
    if (c == 0) {
#pragma unroll
      for (int q = 0; q < NQ; q++) pl[q] = xl[q * TY + r];
    }
    if (vx) ok &= riemann<SOLVER, NQ>(b.ph, pl, pr, fx);
    // y: odd rows hand their upper state to the even row above them (the
    // next wavefront) through LDS, even rows theirs to the odd row above
    // (lanes 32 up in the same wavefront)
    {
      S pyu[NQ];
      cell_states<NQ>(lim, lp, PS, at, RX, pyu, pr);
      if (r & 1) {
#pragma unroll
        for (int q = 0; q < NQ; q++) yu[(q * L::NYU + (r + 1) / 2) * TX + c] = pyu[q];
      }
#pragma unroll
      for (int q = 0; q < NQ; q++) pl[q] = lane_up32(pyu[q]);
    }
    __syncthreads();  // every read of the primitives is done; y states in LDS
    if (w == 0) __builtin_amdgcn_s_setprio(2);
    if (!(r & 1)) {
#pragma unroll
      for (int q = 0; q < NQ; q++) pl[q] = yu[(q * L::NYU + r / 2) * TX + c];
    }
    if (vy) {
      axis_perm<1>(pl);
      axis_perm<1>(pr);
      ok &= riemann<SOLVER, NQ>(b.ph, pl, pr, fy);
      axis_perm<1>(fy);
    }
    if ((cx && wx == TX && crow < wy) || (cy && wy == TY && ccol < wx)) {
      S ff[NQ];
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        pl[q] = cy ? yu[(q * L::NYU + TY / 2) * TX + ccol] : xu[q * TY + crow];
        pr[q] = cy ? yd[q * TX + ccol] : xr[q * TY + crow];
      }
      perm_y_if(cy, pl);
      perm_y_if(cy, pr);
      ok &= riemann<SOLVER, NQ>(b.ph, pl, pr, ff);
      perm_y_if(cy, ff);
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        if (cy) ltop[q * TX + ccol] = ff[q];
        else lend[q * TY + crow] = ff[q];
      }
    }
  }
  if (w == 0) __builtin_amdgcn_s_setprio(0);
  if (!ok) atomicOr(err, 1);
  // the cell's state for the update, fetched ahead of the flux exchange
  const bool own = c < wx && r < wy;
  const long long N = g.ncell;
  const long long cc = cidx(g, i0 + c, j0 + r, k);
  // FP64 carries the reference's partial sums u - dtx*df - dty*dg to the z
  // kernel; the mixed form carries the FP32 flux increment -dtx*df - dty*dg
  // and the z kernel adds it to u in FP64
  constexpr bool INC = sizeof(S) == 4;
  double u0[NH], u0p[NP > 0 ? NP : 1];
  if (own && !INC) {
#pragma unroll
    for (int q = 0; q < NH; q++) u0[q] = (double)b.U[q * N + cc];
#pragma unroll
    for (int a = 0; a < NP; a++) u0p[a] = (double)b.U[(8 + a) * N + cc];
  }
  if constexpr (ORDER == 1) __syncthreads();  // every read of the primitives is done
  S* __restrict__ lfy = lds;
  if (vy) {
#pragma unroll
    for (int q = 0; q < NQ; q++) lfy[(q * TY + r) * TX + c] = fy[q];
  }
  __syncthreads();

  S fxr[NQ], fyu[NQ];
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    const S v = lane_down1(fx[q]);
    fxr[q] = c == TX - 1 ? lend[q * TY + r] : v;
  }
  if (!own) return;
#pragma unroll
  for (int q = 0; q < NQ; q++) fyu[q] = r == TY - 1 ? ltop[q * TX + c] : lfy[(q * TY + r + 1) * TX + c];
  const double dtdx = dt / b.dx, dtdy = dt / b.dy;
  if constexpr (INC) {
    float* __restrict__ Wf = reinterpret_cast<float*>(b.W);
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      if (q >= NH && q < 8) continue;
      Wf[q * N + cc] = (float)((-(dtdx * ((double)fxr[q] - (double)fx[q]))) - dtdy * ((double)fyu[q] - (double)fy[q]));
    }
    if constexpr (FLUXCD) {
      Wf[5 * N + cc] = (float)((-(double)fy[7]) - (double)fyu[7]);
      Wf[6 * N + cc] = (float)((+(double)fx[7]) + (double)fxr[7]);
    }
  } else {
    double* __restrict__ W = b.W;
#pragma unroll
    for (int q = 0; q < NH; q++)
      W[q * N + cc] = (u0[q] - dtdx * ((double)fxr[q] - (double)fx[q])) - dtdy * ((double)fyu[q] - (double)fy[q]);
    // passive scalars: the same partial sums (flux_cd_module.f90:303-308)
#pragma unroll
    for (int a = 0; a < NP; a++) {
      const int q = 8 + a;
      W[q * N + cc] = (u0p[a] - dtdx * ((double)fxr[q] - (double)fx[q])) - dtdy * ((double)fyu[q] - (double)fy[q]);
    }
    if constexpr (FLUXCD) {
      W[5 * N + cc] = (-(double)fy[7]) - (double)fyu[7];
      W[6 * N + cc] = (+(double)fx[7]) + (double)fxr[7];
    }
  }
  if constexpr (FLUXCD)
    b.E[2 * N + cc] = (S)(0.25 * ((((-(double)fx[6]) - (double)fxr[6]) + (double)fy[5]) + (double)fyu[5]));
}

// z kernel.  A thread marches up the pencil k0..k1 of column (i, j),
// solving z-faces k0-1 .. k1 (face k between cells k and k+1).  Primitives
// enter a sliding window one cell ahead of the faces; the raw state of the
// next cell is fetched an iteration early.  The second-order window holds
// P(k), P(k+1) and the upper-face state of cell k: limiter()'s
// reconstruction of a cell depends on the cell and its two neighbours
// only, so each cell's slope is formed once and serves both its faces
// (pl + al/2 for the face above, pr - ar/2 for the face below).
// SRCIN: the source terms (the user's s, Powell's 8-wave terms) added to
// each stored cell (Box::srcin).
template <int SOLVER, int ORDER, int FLUXCD, typename S, int NP = 0, bool SRCIN = false>
__device__ __forceinline__ void z_body(const Box<S>& b, unsigned lb, double dt, int kchunk, int* __restrict__ err) {
  constexpr int NH = FLUXCD ? 5 : 8;
  constexpr int NQ = 8 + NP;  // the MHD variables and the passive scalars
  const Geom& g = b.g;
  int t, kc;
  if (!tile_item(lb, tiles_of((unsigned)(g.nx * g.ny)), (unsigned)(g.nx * g.ny), t, kc)) return;
  const int i = t % g.nx + 1, j = t / g.nx + 1;
  const int k0 = kc * kchunk + 1;
  const int k1 = min(k0 + kchunk - 1, g.nz);
  const int klast = k1 + ORDER;  // last cell the window takes in
  const long long N = g.ncell;
  const long long col = cidx(g, i, j, -1);  // cell k of the column at col + (k+1)*sz
  const long long sz = (long long)g.NX * g.NY;
  const double dtdz = dt / b.dz;
  const S* __restrict__ Usrc = ORDER == 1 ? b.U : b.Us;
  S* __restrict__ Uout = ORDER == 1 ? b.Us : b.U;
  const double* __restrict__ W = b.W;
  const int lim = b.ph.limiter;
  constexpr bool INC = sizeof(S) == 4;  // W holds FP32 x/y flux increments
  constexpr bool w8 = SRCIN;

  S ur[NQ];
  bool uflip = false;
  int uk = 0;  // cell of the fetched state
  auto fetch = [&](int k) {
    bool f = false;
    const int sk = ghost_src(g, k, g.nz, g.bc[4], g.bc[5], f);
    const long long cc = col + (long long)(sk + 1) * sz;
#pragma unroll
    for (int q = 0; q < NQ; q++) ur[q] = Usrc[q * N + cc];
    uflip = f;
    uk = k;
  };
  auto take = [&](S p[NQ]) {
    double u[NQ], pd[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) u[q] = (double)ur[q];
    if (uflip) u[3] = -u[3];
    u2prim<NQ>(b.ph, u, pd);
    // first stage: an interior cell's pressure as primit carries it after
    // table cooling (Box::pc)
    if (ORDER == 1 && b.pc && uk >= 1 && uk <= g.nz) pd[4] = b.pc[col + (long long)(uk + 1) * sz];
#pragma unroll
    for (int q = 0; q < NQ; q++) p[q] = (S)pd[q];
  };

  S pm[NQ], pc[NQ], pu[NQ];  // P(kf), P(kf+1), upper-face state of cell kf
  if constexpr (ORDER == 2) {
    S p0[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) p0[q] = S(0);
    fetch(k0 - 2);
    take(p0);
    fetch(k0 - 1);
    take(pm);
    fetch(k0);
    take(pc);
#pragma unroll
    for (int q = 0; q < NQ; q++) pu[q] = pm[q] + lim_average(lim, pm[q] - p0[q], pc[q] - pm[q]) * S(0.5);
    fetch(k0 + 1);
  } else {
    fetch(k0 - 1);
    take(pc);
    fetch(k0);
  }
  // first order: the side speed of P(kf), kept from the face below (a
  // cell's value serves both of its z faces)
  S cz = S(0);
  if constexpr (ORDER == 1) {
    S t[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) t[q] = pc[q];
    axis_perm<2>(t);
    cz = side_speed<SOLVER>(b.ph, t);
  }
  S fl[NQ];
#pragma unroll
  for (int q = 0; q < NQ; q++) fl[q] = S(0);
#pragma unroll 1
  for (int kf = k0 - 1; kf <= k1; kf++) {
    S pn[NQ];
    take(pn);  // P(kf+ORDER)
    if (kf + ORDER + 1 <= klast) fetch(kf + ORDER + 1);
    // partials of the cell this face closes, loaded ahead of the solve
    const int kc0 = max(kf, k0);
    const long long c = col + (long long)(kc0 + 1) * sz;
    double wv[NQ];
    if constexpr (INC) {
      // U^n (the predictor's input, the corrector's output) plus the FP32
      // x/y flux increment, in FP64
      const float* __restrict__ Wf = reinterpret_cast<const float*>(W);
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        if (q >= NH && q < 8) continue;
        wv[q] = (double)b.U[q * N + c] + (double)Wf[q * N + c];
      }
      if constexpr (FLUXCD) {
        wv[5] = (double)Wf[5 * N + c];
        wv[6] = (double)Wf[6 * N + c];
      }
    } else {
#pragma unroll
      for (int q = 0; q < NH; q++) wv[q] = W[q * N + c];
      if constexpr (FLUXCD) {
        wv[5] = W[5 * N + c];
        wv[6] = W[6 * N + c];
      }
#pragma unroll
      for (int q = 8; q < NQ; q++) wv[q] = W[q * N + c];
    }
    S left[NQ], right[NQ];
    if constexpr (ORDER == 2) {
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        const S s = lim_average(lim, pc[q] - pm[q], pn[q] - pc[q]);
        left[q] = pu[q];
        right[q] = pc[q] - s * S(0.5);
        pu[q] = pc[q] + s * S(0.5);
        pm[q] = pc[q];
        pc[q] = pn[q];
      }
    } else {
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        left[q] = pc[q];
        right[q] = pn[q];
        pc[q] = pn[q];
      }
    }
    axis_perm<2>(left);
    axis_perm<2>(right);
    S ff[NQ];
    bool fok;
    if constexpr (ORDER == 1) {
      const S csr = side_speed<SOLVER>(b.ph, right);
      fok = riemann_cs<SOLVER, NQ>(b.ph, left, right, cz, csr, ff);
      cz = csr;
    } else {
      fok = riemann<SOLVER, NQ>(b.ph, left, right, ff);
    }
    if (!fok) atomicOr(err, 1);
    axis_perm<2>(ff);
    if (kf >= k0) {
      // source terms (source(), src/sources.f90:188-220): added to the
      // stored update here rather than in a pass of their own over U
      double sv[8];
#pragma unroll
      for (int q = 0; q < 8; q++) sv[q] = (w8 && b.usrc) ? b.usrc[q * N + c] : 0.0;
      if (w8 && (b.src & SRC_EIGHT_WAVE)) w8_terms(b, Usrc, i, j, kf, sv);
#pragma unroll
      for (int q = 0; q < NH; q++) {
        S r = (S)(wv[q] - dtdz * ((double)ff[q] - (double)fl[q]));
        if (w8) r = (S)((double)r + dt * sv[q]);
        Uout[q * N + c] = r;
      }
#pragma unroll
      for (int q = 8; q < NQ; q++) {
        S r = (S)(wv[q] - dtdz * ((double)ff[q] - (double)fl[q]));
        if (w8) r = (S)((double)r + dt * 0.0);
        Uout[q * N + c] = r;
      }
      if constexpr (FLUXCD) {
        b.E[0 * N + c] = (S)(0.25 * ((wv[5] + (double)fl[6]) + (double)ff[6]));
        b.E[1 * N + c] = (S)(0.25 * ((wv[6] - (double)fl[5]) - (double)ff[5]));
      }
    }
#pragma unroll
    for (int q = 0; q < NQ; q++) fl[q] = ff[q];
  }
}

// ---------------------------------------------------------------------------
// Slab kernel (2.5-D blocking): a whole stage of tstep for a 32 x 8 column
// tile over a slab of `kchunk` planes, marching up z.  Per plane m: the
// tile's primitives into LDS -- the own cells' from the column window, the
// halo's from HBM -- then the z face above each own cell (z_body's march,
// one face per plane), the tile's x and y faces (xy_solve, xy_exchange), and the cell's
// whole update: the x, y and z terms of step in the reference's order, E3
// and E1, E2 (get_efield's four-term sums), stored once.  The x/y partial
// sums that the xy kernel hands the z kernel through W stay in registers
// here, rounded exactly as W holds them (FP64; FP32 increments in the mixed
// form), so the two paths give the same bits.  A stage then reads U (and
// U^n's own cells for the corrector) and writes the updated hydro part and
// E: no W round trip and no second pass over U.
//
// The column window: the primitives of the two cells above the current
// plane sit in two LDS slots per thread (ping-pong by plane parity), the
// upper-face state of the current cell and the z flux below it in
// registers; the next cell's state is fetched at the top of each plane
// with the halo.
// ---------------------------------------------------------------------------
template <int ORDER, int NQ, typename S>
struct SlabLds {
  using X = XYLds<ORDER, NQ, TYS>;
  static constexpr int WIN = X::SIZE;             // [2][NQ][SL_THREADS] window slots
  static constexpr int CF = WIN + 2 * NQ * SL_THREADS;  // first order: [2][PS] x and y side speeds
  static constexpr int SIZE = CF + (ORDER == 1 ? 2 * X::PS : 0);
};

template <int SOLVER, int ORDER, int FLUXCD, typename S, int NP = 0>
__device__ __forceinline__ void slab_body(const Box<S>& b, unsigned lb, double dt, int kchunk,
                                          int* __restrict__ err, S* __restrict__ lds) {
  constexpr int NQ = 8 + NP;  // the MHD variables and the passive scalars
  constexpr int TYT = TYS;
  using L = XYLds<ORDER, NQ, TYT>;
  constexpr int H = L::H, RX = L::RX, PS = L::PS;
  constexpr int NH = FLUXCD ? 5 : 8;
  constexpr bool INC = sizeof(S) == 4;  // mixed: FP32 x/y flux increments
  static_assert(L::NLOAD <= 2 * SL_THREADS, "two halo loads per thread");
  const Geom& g = b.g;
  const int ntx = (g.nx + TX - 1) / TX, nty = (g.ny + TYT - 1) / TYT;
  const unsigned per_plane = (unsigned)(ntx * nty);
  const int kc = (int)(lb / per_plane);
  const int rem = (int)(lb - (unsigned)kc * per_plane);
  const int i0 = (rem % ntx) * TX + 1, j0 = (rem / ntx) * TYT + 1;
  const int wx = min(TX, g.nx - i0 + 1), wy = min(TYT, g.ny - j0 + 1);
  const int k0 = kc * kchunk + 1;
  if (k0 > g.nz) return;  // (whole workgroup: the grid has whole slabs only)
  const int k1 = min(k0 + kchunk - 1, g.nz);
  const long long N = g.ncell;
  const long long sz = (long long)g.NX * g.NY;
  const double dtdx = dt / b.dx, dtdy = dt / b.dy, dtdz = dt / b.dz;
  const S* __restrict__ Usrc = ORDER == 1 ? b.U : b.Us;
  S* __restrict__ Uout = ORDER == 1 ? b.Us : b.U;
  const int lim = b.ph.limiter;
  S* __restrict__ lp = lds;
  S* __restrict__ win = lds + SlabLds<ORDER, NQ, S>::WIN;  // win[(slot * NQ + q) * SL_THREADS + tid]
  S* __restrict__ cfs = lds + SlabLds<ORDER, NQ, S>::CF;   // first order: side speeds, x [PS] then y [PS]
  bool ok = true;
  // lane (c, r) of the tile: c = tid & 31, r = 2 (tid >> 6) + ((tid & 63) >> 5);
  // cell k of its column at col + (k+1)*sz
  auto own_of = [&](int t) {
    const int c = t & (TX - 1), r = 2 * (t >> 6) + ((t & 63) >> 5);
    return c < wx && r < wy;
  };
  auto col_of = [&](int t, bool own) {
    const int c = t & (TX - 1), r = 2 * (t >> 6) + ((t & 63) >> 5);
    return cidx(g, own ? i0 + c : i0, own ? j0 + r : j0, -1);
  };

  // primitives of cell kk of the own column (ghosts by boundary mapping)
  auto fetch = [&](long long col, int kk, S ur[NQ], bool& flip) {
    bool f = false;
    const int sk = ghost_src(g, kk, g.nz, g.bc[4], g.bc[5], f);
    const long long cc = col + (long long)(sk + 1) * sz;
#pragma unroll
    for (int q = 0; q < NQ; q++) ur[q] = Usrc[q * N + cc];
    flip = f;
  };
  // kk: the cell's plane; the first stage takes an interior cell's
  // pressure as primit carries it after table cooling (Box::pc)
  auto take = [&](const S ur[NQ], bool flip, S p[NQ], long long col, int kk) {
    double u[NQ], pd[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) u[q] = (double)ur[q];
    if (flip) u[3] = -u[3];
    u2prim<NQ>(b.ph, u, pd);
    if (ORDER == 1 && b.pc && kk >= 1 && kk <= g.nz) pd[4] = b.pc[col + (long long)(kk + 1) * sz];
#pragma unroll
    for (int q = 0; q < NQ; q++) p[q] = (S)pd[q];
  };
  auto wput = [&](int tid, int slot, const S p[NQ]) {
#pragma unroll
    for (int q = 0; q < NQ; q++) win[(slot * NQ + q) * SL_THREADS + tid] = p[q];
  };
  auto wget = [&](int tid, int slot, S p[NQ]) {
#pragma unroll
    for (int q = 0; q < NQ; q++) p[q] = win[(slot * NQ + q) * SL_THREADS + tid];
  };
  // z face kf (between cells kf and kf+1) from P(kf) (`lo`), the window
  // slot holding P(kf+1) and P(kf+ORDER) (`nx`); ORDER 2 moves pu to cell
  // kf+1's upper state
  // first order: cz is the side speed of the cell below the face, kept
  // from the face below it (a cell's value serves both its z faces)
  S pu[NQ];
  S cz = S(0);
#pragma unroll
  for (int q = 0; q < NQ; q++) pu[q] = S(0);
  auto z_face = [&](int tid, const S lo[NQ], int slot1, const S nx[NQ], S ff[NQ]) {
    S left[NQ], right[NQ];
    if constexpr (ORDER == 2) {
      S p1[NQ];
      wget(tid, slot1, p1);
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        const S sl = lim_average(lim, p1[q] - lo[q], nx[q] - p1[q]);
        left[q] = pu[q];
        right[q] = p1[q] - sl * S(0.5);
        pu[q] = p1[q] + sl * S(0.5);
      }
    } else {
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        left[q] = lo[q];
        right[q] = nx[q];
      }
    }
    axis_perm<2>(left);
    axis_perm<2>(right);
    if constexpr (ORDER == 1) {
      const S csr = side_speed<SOLVER>(b.ph, right);
      ok &= riemann_cs<SOLVER, NQ>(b.ph, left, right, cz, csr, ff);
      cz = csr;
    } else {
      ok &= riemann<SOLVER, NQ>(b.ph, left, right, ff);
    }
    axis_perm<2>(ff);
  };
  // the x and y side speeds of a cell into LDS (first order)
  auto put_speeds = [&](int cell, const S p[NQ], bool xs, bool ys) {
    if (xs) cfs[cell] = side_speed<SOLVER>(b.ph, p);
    if (ys) {
      S t[NQ];
#pragma unroll
      for (int q = 0; q < NQ; q++) t[q] = p[q];
      axis_perm<1>(t);
      cfs[L::PS + cell] = side_speed<SOLVER>(b.ph, t);
    }
  };

  // Window before the first plane: slot (k0 & 1) holds P(k0), the other
  // P(k0 + 1) (ORDER 2); the face below the slab into fl.
  S fl[NQ];
#pragma unroll
  for (int q = 0; q < NQ; q++) fl[q] = S(0);
  {
    const int tid = (int)threadIdx.x;
    const bool own = own_of(tid);
    const long long col = col_of(tid, own);
    if (own) {
      S ur[NQ], pa[NQ], pb[NQ];
      bool f;
      if constexpr (ORDER == 2) {
        S p0[NQ];
        fetch(col, k0 - 2, ur, f);
        take(ur, f, p0, col, k0 - 2);
        fetch(col, k0 - 1, ur, f);
        take(ur, f, pa, col, k0 - 1);      // P(k0-1)
        fetch(col, k0, ur, f);
        take(ur, f, pb, col, k0);          // P(k0)
#pragma unroll
        for (int q = 0; q < NQ; q++) pu[q] = pa[q] + lim_average(lim, pa[q] - p0[q], pb[q] - pa[q]) * S(0.5);
        wput(tid, k0 & 1, pb);
        fetch(col, k0 + 1, ur, f);
        take(ur, f, p0, col, k0 + 1);      // P(k0+1)
        z_face(tid, pa, k0 & 1, p0, fl);   // face k0-1; pu -> upper state of k0
        wput(tid, (k0 + 1) & 1, p0);
      } else {
        fetch(col, k0 - 1, ur, f);
        take(ur, f, pa, col, k0 - 1);      // P(k0-1)
        fetch(col, k0, ur, f);
        take(ur, f, pb, col, k0);          // P(k0)
        {
          S t[NQ];
#pragma unroll
          for (int q = 0; q < NQ; q++) t[q] = pa[q];
          axis_perm<2>(t);
          cz = side_speed<SOLVER>(b.ph, t);
        }
        z_face(tid, pa, 0, pb, fl);        // face k0-1
        wput(tid, k0 & 1, pb);
      }
    }
  }

#pragma unroll 1
  for (int m = k0; m <= k1; m++) {
    if (m > k0) __syncthreads();  // the previous plane's reads of LDS are done
    // Every per-lane index below is recomputed in each plane from an
    // opaque copy of the lane id: hoisted out of the march, the per-lane
    // addresses (a 64-bit pointer per variable and access) would hold
    // registers for the whole loop.
    int tv = (int)threadIdx.x;
    asm volatile("" : "+v"(tv));
    const int tid = tv;
    const int c = tid & (TX - 1), r = 2 * (tid >> 6) + ((tid & 63) >> 5);
    const bool own = c < wx && r < wy;
    const long long col = col_of(tid, own);
    const int at = (r + H) * RX + c + H;  // own cell in the primitive area
    // the next cell of the window (P(m+ORDER)), the own cell into the
    // primitive area, the halo (and, in a partial tile, the cells past its
    // edge) from HBM; ghosts by boundary mapping
    S un[NQ];
    bool unf = false;
    if (own) fetch(col, m + ORDER, un, unf);
    if (own) {
      S pm[NQ];
      wget(tid, m & 1, pm);
#pragma unroll
      for (int q = 0; q < NQ; q++) lp[q * PS + at] = pm[q];
      if constexpr (ORDER == 1) put_speeds(at, pm, true, true);
    }
    __builtin_amdgcn_s_setprio(3);
    {
      double u[2][NQ];
      int dst[2];
#pragma unroll
      for (int it = 0; it < 2; it++) {
        const int idx = tid + it * SL_THREADS;
        int rx, ry;
        if (idx < H * TX) {
          ry = idx / TX;
          rx = H + idx % TX;
        } else if (idx < H * TX + TYT * RX) {
          const int t = idx - H * TX;
          ry = H + t / RX;
          rx = t % RX;
        } else {
          const int t = idx - H * TX - TYT * RX;
          ry = H + TYT + t / TX;
          rx = H + t % TX;
        }
        const bool own_pos = rx >= H && rx < H + wx && ry >= H && ry < H + wy;
        dst[it] = -1;
        if (idx < L::NLOAD && rx < wx + 2 * H && ry < wy + 2 * H && !own_pos) {
          dst[it] = ry * RX + rx;
          load_cons<NQ>(g, Usrc, i0 - H + rx, j0 - H + ry, m, u[it]);
        }
      }
      __builtin_amdgcn_s_setprio(0);
#pragma unroll
      for (int it = 0; it < 2; it++) {
        if (dst[it] < 0) continue;
        double p[NQ];
        u2prim<NQ>(b.ph, u[it], p);
        if (ORDER == 1 && b.pc) {  // Box::pc: interior cells only
          const int hi = i0 - H + dst[it] % RX, hj = j0 - H + dst[it] / RX;
          if (hi >= 1 && hi <= g.nx && hj >= 1 && hj <= g.ny) p[4] = b.pc[cidx(g, hi, hj, m)];
        }
        S ps[NQ];
#pragma unroll
        for (int q = 0; q < NQ; q++) {
          ps[q] = (S)p[q];
          lp[q * PS + dst[it]] = ps[q];
        }
        // a halo cell on one of the tile's rows closes an x face, any other
        // a y face
        if constexpr (ORDER == 1) {
          const int hy = dst[it] / RX;
          const bool xs = hy >= H && hy < H + wy;
          put_speeds(dst[it], ps, xs, !xs);
        }
      }
    }
    // z face m (above the own cell): P(m) from the window slot, P(m+ORDER)
    // just fetched; the window moves up one cell
    S ff[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) ff[q] = S(0);
    if (own) {
      S pm[NQ], pn[NQ];
      wget(tid, m & 1, pm);
      take(un, unf, pn, col, m + ORDER);
      if constexpr (ORDER == 2) {
        z_face(tid, pm, (m + 1) & 1, pn, ff);
        wput(tid, m & 1, pn);  // slot of P(m) now holds P(m+2)
      } else {
        z_face(tid, pm, 0, pn, ff);
        wput(tid, (m + 1) & 1, pn);
      }
    }
    __syncthreads();

    const long long cc = col + (long long)(m + 1) * sz;
    S fx[NQ], fy[NQ], fxr[NQ], fyu[NQ];
    double u0[NQ];
    ok &= xy_solve<SOLVER, ORDER, NQ, TYT, S, ORDER == 1>(b, lds, wx, wy, tid, fx, fy, cfs);
    if (own) {
#pragma unroll
      for (int q = 0; q < NQ; q++)
        if (q < NH || q >= 8) u0[q] = (double)b.U[q * N + cc];
    }
    xy_exchange<ORDER, NQ, TYT>(lds, wx, wy, tid, fx, fy, fxr);
    if (own) {
      xy_fyu<ORDER, NQ, TYT>(lds, tid, fyu);
      // the x/y partials, rounded as the xy kernel stores them in W, then
      // the z term (step, hydro_solver.f90:105-107 / flux_cd_module.f90:299-308)
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        if (q >= NH && q < 8) continue;
        double wv;
        if constexpr (INC)
          wv = u0[q] + (double)(float)((-(dtdx * ((double)fxr[q] - (double)fx[q]))) -
                                       dtdy * ((double)fyu[q] - (double)fy[q]));
        else
          wv = (u0[q] - dtdx * ((double)fxr[q] - (double)fx[q])) - dtdy * ((double)fyu[q] - (double)fy[q]);
        Uout[q * N + cc] = (S)(wv - dtdz * ((double)ff[q] - (double)fl[q]));
      }
      if constexpr (FLUXCD) {
        const double w5 = (double)(S)((-(double)fy[7]) - (double)fyu[7]);
        const double w6 = (double)(S)((+(double)fx[7]) + (double)fxr[7]);
        b.E[2 * N + cc] = (S)(0.25 * ((((-(double)fx[6]) - (double)fxr[6]) + (double)fy[5]) + (double)fyu[5]));
        b.E[0 * N + cc] = (S)(0.25 * ((w5 + (double)fl[6]) + (double)ff[6]));
        b.E[1 * N + cc] = (S)(0.25 * ((w6 - (double)fl[5]) - (double)ff[5]));
      }
    }
#pragma unroll
    for (int q = 0; q < NQ; q++) fl[q] = ff[q];
  }
  if (!ok) atomicOr(err, 1);
}

// ---------------------------------------------------------------------------
// flux-CD field update (src/flux_cd_module.f90:285-323, with step's
// ordering of terms), accumulated in FP64.  Each thread marches up a
// z-pencil carrying E1, E2 one plane below and at the cell.
//
// E ghosts (boundaryI_ef): the update reads e2, e3 at x neighbours, e1, e3
// at y neighbours and e1, e2 at z neighbours.  A closed boundary negates e1
// on x ghosts and e2 on y ghosts only (flux_cd_module.f90:145-176), never a
// component read there, so every E ghost resolves to a fixed interior
// neighbour offset: periodic wraps, closed and outflow copy the edge cell.
// Stage 2 writes over U^n in place (each cell reads only its own field).
// ---------------------------------------------------------------------------

// offset (in cells along the axis) of the E ghost-resolved neighbour; a
// BC_OTHER face's E ghosts are never filled (boundaryI_ef has no branch for
// it) and are read as stored
__device__ __forceinline__ int e_nbr(int i, int n, int d, int bc) {
  const int t = i + d;
  if (t >= 1 && t <= n) return d;
  if (bc == BC_OTHER) return d;
  if (bc == BC_PERIODIC) return (t < 1 ? t + n : t - n) - i;
  return 0;  // closed / outflow: ghost takes the edge cell's value
}

template <int NQ, typename C>
__device__ __forceinline__ double cfl_cell(const Phys& ph, double dx, double dy, double dz, const double* u,
                                           double dtp, const double* pcell = nullptr);

// CFL = 1 (corrector only): also get_timestep's candidate of each updated
// cell into *dtp, non-finite values into *bad
template <typename S, int CFL = 0, int NQ = 8>
__device__ __forceinline__ void bfield_body(const Box<S>& b, unsigned lb, double dt, int stage, int kchunk,
                                            double* dtp = nullptr, bool* bad = nullptr) {
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
  const S* __restrict__ E1 = b.E;
  const S* __restrict__ E2 = b.E + N;
  const S* __restrict__ E3 = b.E + 2 * N;
  auto d = [](S v) { return (double)v; };
  long long c = cidx(g, i, j, k0);
  const long long oxm = e_nbr(i, g.nx, -1, g.bc[0]) * sx;
  const long long oxp = e_nbr(i, g.nx, +1, g.bc[1]) * sx;
  const long long oym = e_nbr(j, g.ny, -1, g.bc[2]) * sy;
  const long long oyp = e_nbr(j, g.ny, +1, g.bc[3]) * sy;
  const long long ozm = e_nbr(k0, g.nz, -1, g.bc[4]) * sz;
  double e1m = d(E1[c + ozm]), e2m = d(E2[c + ozm]);
  double e1c = d(E1[c]), e2c = d(E2[c]);
#pragma unroll 1
  for (int k = k0; k <= k1; k++, c += sz) {
    const long long ozp = e_nbr(k, g.nz, +1, g.bc[5]) * sz;
    const double e1p = d(E1[c + ozp]);
    const double e2p = d(E2[c + ozp]);
    const double b5 = (d(Uin[5 * N + c]) - (0.5 * dtdy) * (d(E3[c + oyp]) - d(E3[c + oym]))) + (0.5 * dtdz) * (e2p - e2m);
    const double b6 = (d(Uin[6 * N + c]) + (0.5 * dtdx) * (d(E3[c + oxp]) - d(E3[c + oxm]))) - (0.5 * dtdz) * (e1p - e1m);
    const double b7 = (d(Uin[7 * N + c]) - (0.5 * dtdx) * (d(E2[c + oxp]) - d(E2[c + oxm]))) + (0.5 * dtdy) * (d(E1[c + oyp]) - d(E1[c + oym]));
    if (!CFL && b.srcin) {  // (the CFL fold runs only without source terms)
      // the field's source terms, from the stage's input state
      double sv[8];
#pragma unroll
      for (int q = 0; q < 8; q++) sv[q] = b.usrc ? b.usrc[q * N + c] : 0.0;
      if (b.src & SRC_EIGHT_WAVE) w8_terms(b, stage == 1 ? b.U : b.Us, i, j, k, sv);
      Uout[5 * N + c] = (S)((double)(S)b5 + dt * sv[5]);
      Uout[6 * N + c] = (S)((double)(S)b6 + dt * sv[6]);
      Uout[7 * N + c] = (S)((double)(S)b7 + dt * sv[7]);
    } else {
      Uout[5 * N + c] = (S)b5;
      Uout[6 * N + c] = (S)b6;
      Uout[7 * N + c] = (S)b7;
    }
    if (CFL) {
      // this cell of U^{n+1} is final: its hydro part and passives were
      // written by the z kernel, its field just now (as stored, in S)
      double u[NQ];
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        u[q] = q == 5   ? (double)(S)b5
               : q == 6 ? (double)(S)b6
               : q == 7 ? (double)(S)b7
                        : (q < 8 + g.npas ? (double)Uout[q * N + c] : 0.0);
        *bad |= !__builtin_isfinite(u[q]);
      }
      *dtp = cfl_cell<NQ, S>(b.ph, b.dx, b.dy, b.dz, u, *dtp);
    }
    e1m = e1c;
    e2m = e2c;
    e1c = e1p;
    e2c = e2p;
  }
}

// ---------------------------------------------------------------------------
// Source terms of step() (src/hydro_solver.f90:115-121): after the stage's
// update, up = up + dt*s with s from source() (src/sources.f90:188-220):
// zero, plus the user's terms (already in W, written by the user plugin on
// a zeroed W), then Powell's 8-wave terms (w8_terms) from the stage's input
// state -- U^n for the predictor, U* for the corrector.  Without passive
// scalars (Box::srcin) k_z and the field update add the terms as they store
// each cell instead, the user's s computed before the stage.  Column pencils.
// ---------------------------------------------------------------------------
template <typename S>
__device__ __forceinline__ void sources_body(const Box<S>& b, unsigned lb, double dt, int stage, int kchunk) {
  const Geom& g = b.g;
  if (!b.src || b.srcin) return;  // b.srcin: the stage kernels added the terms
  int t, kc;
  if (!tile_item(lb, tiles_of((unsigned)(g.nx * g.ny)), (unsigned)(g.nx * g.ny), t, kc)) return;
  const int i = t % g.nx + 1, j = t / g.nx + 1;
  const int k0 = kc * kchunk + 1;
  const int k1 = min(k0 + kchunk - 1, g.nz);
  const long long N = g.ncell;
  const S* __restrict__ Uin = stage == 1 ? b.U : b.Us;
  S* __restrict__ Uout = stage == 1 ? b.Us : b.U;
  for (int k = k0; k <= k1; k++) {
    const long long c = cidx(g, i, j, k);
    double sv[10];
#pragma unroll
    for (int q = 0; q < 10; q++) sv[q] = (q < 8 + g.npas && (b.src & SRC_USER)) ? b.W[q * N + c] : 0.0;
    if (b.src & SRC_EIGHT_WAVE) w8_terms(b, Uin, i, j, k, sv);
#pragma unroll
    for (int q = 0; q < 10; q++)
      if (q < 8 + g.npas) Uout[q * N + c] = (S)((double)Uout[q * N + c] + dt * sv[q]);
  }
}

// ---------------------------------------------------------------------------
// Table cooling, coolingdmc (src/cooling_dmc.f90:115-156) / coolingchi
// (src/cooling_chi.f90): every interior cell above Tmin = 1e4 K loses
// thermal energy, p -> p exp(-Ce dt tsc), Ce = Lambda(T) rho^2 / (cv p Psc),
// Lambda from cooldmc / coolchi (sqrt law above 1e8 K, else linear
// interpolation at table index int(10 log10 T) - 39).  E changes by the
// thermal energy lost.  exp and log10 are the device's (within 1 ulp of the
// host's libm).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double cool_lambda(const Cooling& cl, double T) {
  if (T > 1e8) return cl.hicoef * __builtin_sqrt(T);
  int if1 = (int)(log10(T) * 10) - 39;
  if1 = if1 < 1 ? 1 : (if1 > 40 ? 40 : if1);  // T = 1e8 exactly would read past the 41-point table
  const double T0 = cl.tab[2 * (if1 - 1)], c0 = cl.tab[2 * (if1 - 1) + 1];
  const double T1 = cl.tab[2 * if1], c1 = cl.tab[2 * if1 + 1];
  return ((c1 - c0) * (T - T0)) / (T1 - T0) + c0;
}

template <typename S>
__global__ void __launch_bounds__(256) k_cooling(Box<S> b, Cooling cl, double dt) {
  const Geom& g = b.g;
  const long long n = (long long)g.nx * g.ny * g.nz;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = (int)(t % g.nx) + 1;
  const long long r = t / g.nx;
  const int j = (int)(r % g.ny) + 1, k = (int)(r / g.ny) + 1;
  const long long c = cidx(g, i, j, k), N = g.ncell;
  double u[10], p[10];
#pragma unroll
  for (int q = 0; q < 10; q++) u[q] = q < 8 + g.npas ? (double)b.U[q * N + c] : 0.0;
  const double T = module_temp(b.ph, u2prim_t<10>(b.ph, u, p), p, cl.tempsc);
  if (!(T > 10000.)) {
    if (b.pc) b.pc[c] = p[4];  // primit(5) = u2prim's pressure
    return;
  }
  const double eth0 = b.ph.cv * p[4];
  const double aloss = cool_lambda(cl, T);
  const double dens = p[0];
  const double ce = (aloss * (dens * dens)) / (eth0 * cl.psc);
  const double pn = p[4] * exp(-(ce * (dt * cl.tsc)));
  b.U[4 * N + c] = (S)((u[4] - eth0) + b.ph.cv * pn);
  if (b.pc) b.pc[c] = pn;  // primit(5), cooled
}

// ---------------------------------------------------------------------------
// H-rate cooling, coolingh (src/cooling_h.f90:41-67): per interior cell,
// atomic(dt tsc, u, 1, 1) without diffuse radiation (:226-374) -- the
// neutral fraction's rate equation solved over the step (recombination
// alpha :80-88, collisional ionisation colf :108-116), the radiative loss
// ALOSS (:141-218, with betah :124-137) and the relaxed temperature -- then
// the neutral density (the first passive scalar) and E.  Evaluation order
// is the reference's; the reference runs alpha, colf and betah with REAL(16)
// constants, the device in FP64, and exp/log/pow are the device's.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double hr_betah(double t) {
  const double a = 157890. / t;
  return (1.133e-24 / __builtin_sqrt(a)) * ((-0.0713 + 0.5 * log(a)) + 0.640 * pow(a, -0.33333));
}

__device__ double hr_aloss(double x1, double den, double dh0, double te0) {
  const double XION = 2.179e-11, XO = 1.e-3;
  const double C0 = 0.5732, C1 = 1.8288e-5, C2 = -1.15822e-10, C3 = 9.4288e-16;
  const double D0 = 0.5856, D1 = 1.55083e-5, D2 = -9.669e-12, D3 = 5.716e-19;
  const double ENK = 118409., EN = 1.634E-11;
  const double te = dmax(te0, 10.);
  const double dh = den;
  const double dhp = (1. - x1) * dh;
  const double de = dhp + 1.e-4 * dh;
  const double doi = XO * dh0;
  const double doii = XO * dhp;
  if (te <= 1e4) return 0.;
  double omega = 0.;
  if (te <= 55000.) omega = C0 + te * (C1 + te * (C2 + te * C3));
  if (te >= 72000.) omega = D0 + te * (D1 + te * (D2 + te * D3));
  if (te > 55000. && te < 72000.) {
    const double omegal = C0 + te * (C1 + te * (C2 + te * C3));
    const double omegah = D0 + te * (D1 + te * (D2 + te * D3));
    const double frac = (te - 55000.) / 17000.;
    omega = (1. - frac) * omegal + frac * omegah;
  }
  const double qla = ((8.6287E-6 / (2. * __builtin_sqrt(te))) * omega) * exp(-(ENK / te));
  const double ecoll = dmax(((de * dh0) * qla) * EN, 0.);
  const double cion = (5.834E-11 * __builtin_sqrt(te)) * exp(-(1.579E5 / te));
  const double eion = ((de * dh0) * cion) * XION;
  const double erec = dmax((de * dhp) * hr_betah(te), 0.);
  const double tm = 1. / te;
  const double t2 = tm * tm;
  const double eoi = dmax((de * doi) * pow(10., ((1381465. * t2) - (12328.69 * tm)) - 19.82621), 0.);
  const double eoii = dmax((de * doii) * pow(10., (-(2061075. * t2) - (14596.24 * tm)) - 19.01402), 0.);
  const double betaf = (1.3 * 1.42E-27) * pow(te, 0.5);
  const double hiicool = (de * dhp) * betaf;
  const double equil = (((1.0455E-18 / pow(te, 0.63)) * (1. - exp(-pow(te * 1.E-5, 1.63)))) * de) * den + hiicool;
  double fr = 0.;
  if (te >= 54770.) fr = 1.;
  if (te > 44770. && te < 54770.) {
    const double ex2 = exp(-((2. * (te - 49770.)) / 500.));
    fr = 0.5 * (1. + (1. - ex2) / (1. + ex2));
  }
  return ((ecoll + eion) + ((erec + 7.033 * (eoi + eoii)) * (1. - fr))) + equil * fr;
}

template <typename S>
__global__ void __launch_bounds__(256) k_cooling_h(Box<S> b, Cooling cl, double dt) {
  const Geom& g = b.g;
  const long long n = (long long)g.nx * g.ny * g.nz;
  const long long tt = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (tt >= n) return;
  const int i = (int)(tt % g.nx) + 1;
  const long long r = tt / g.nx;
  const int j = (int)(r % g.ny) + 1, k = (int)(r / g.ny) + 1;
  const long long c = cidx(g, i, j, k), N = g.ncell;
  double u[10], p[10];
#pragma unroll
  for (int q = 0; q < 10; q++) u[q] = q < 8 + g.npas ? (double)b.U[q * N + c] : 0.0;
  const double t = module_temp(b.ph, u2prim_t<10>(b.ph, u, p), p, cl.tempsc);
  const double dts = dt * cl.tsc;
  const double xi = 1.e-4, boltzm = 1.3807e-16;
  const double col = (5.83e-11 * __builtin_sqrt(t)) * exp(-(157828. / t));
  const double rec = 2.55e-13 * pow(1.e4 / t, 0.79);
  const double y0 = u[8] / u[0];
  const double dh = u[0];
  const double a = rec + col;
  const double bq = -((2. + xi) * rec + (1. + xi) * col);
  const double cq = (1. + xi) * rec;
  const double d = __builtin_sqrt(bq * bq - (4. * a) * cq);
  const double g0 = ((((2. * a) * y0) + bq) + d) / ((((2. * a) * y0) + bq) - d);
  const double e = exp(-((d * dh) * dts));
  double y1 = ((-bq) - ((d * (1. + g0 * e)) / (1. - g0 * e))) / (2. * a);
  y1 = dmin(y1, 0.9999);
  y1 = dmax(y1, 0.);
  const double al = hr_aloss(y0, dh, u[8], t) / (dh * dh);
  const double tprime = 10.;
  const double ce = ((2. * dh) * al) / ((3. * boltzm) * t);
  double t1 = tprime + (t - tprime) * exp(-(ce * dts));
  t1 = dmax(t1, 0.1 * t);
  t1 = dmin(t1, 10. * t);
  const double u9 = y1 * u[0];
  double e5 = (((b.ph.cv * ((2. * u[0]) - u9)) * t1) / cl.tempsc)
              + (0.5 * p[0]) * ((p[1] * p[1] + p[2] * p[2]) + p[3] * p[3]);
  if (b.ph.mhd) e5 = e5 + 0.5 * ((p[5] * p[5] + p[6] * p[6]) + p[7] * p[7]);
  b.U[8 * N + c] = (S)u9;
  b.U[4 * N + c] = (S)e5;
}

// ---------------------------------------------------------------------------
// Artificial viscosity, viscous_copy (src/hydro_solver.f90:47-65):
//   u = up + eta ((((((up(i+1) + up(i-1)) + up(j+1)) + up(j-1)) + up(k+1))
//                   + up(k-1)) - 6 up)
// on interior cells.  up is the corrector's result (U, updated in place);
// at that point up's ghosts still hold boundaryII's fill of the predictor
// state U* (step() writes interior cells only), so a neighbour outside the
// domain is U* through the boundary mapping (or U*'s stored ghost when
// ghosts live in memory).  The result goes to W in FP64 and a second pass
// stores it, so no cell is overwritten while a neighbour still reads it.
// ---------------------------------------------------------------------------
struct ViscNbr {
  long long c;  // storage index
  int flip;     // momentum component negated by a closed face (-1: none)
  bool star;    // read from U* (outside the domain)
};

__device__ __forceinline__ ViscNbr visc_nbr(const Geom& g, int i, int j, int k) {
  ViscNbr n;
  n.star = i < 1 || i > g.nx || j < 1 || j > g.ny || k < 1 || k > g.nz;
  n.flip = -1;
  if (!n.star) {
    n.c = cidx(g, i, j, k);
    return n;
  }
  bool fx = false, fy = false, fz = false;
  const int si = ghost_src(g, i, g.nx, g.bc[0], g.bc[1], fx);
  const int sj = ghost_src(g, j, g.ny, g.bc[2], g.bc[3], fy);
  const int sk = ghost_src(g, k, g.nz, g.bc[4], g.bc[5], fz);
  n.c = cidx(g, si, sj, sk);
  n.flip = fx ? 1 : fy ? 2 : fz ? 3 : -1;
  return n;
}

template <typename S>
__device__ __forceinline__ double visc_val(const Box<S>& b, const ViscNbr& n, int q) {
  const double v = (double)(n.star ? b.Us : b.U)[q * b.g.ncell + n.c];
  return q == n.flip ? -v : v;
}

// viscous_copy of interior cell (i, j, k): up + eta * (the six neighbours'
// sum - 6 up) into W, neighbours outside the domain from U* through the
// boundary mapping; then W back into U
template <typename S>
__device__ __forceinline__ void visc_cell(const Box<S>& b, int i, int j, int k, double eta) {
  const Geom& g = b.g;
  const long long c = cidx(g, i, j, k), N = g.ncell;
  const ViscNbr xp = visc_nbr(g, i + 1, j, k), xm = visc_nbr(g, i - 1, j, k);
  const ViscNbr yp = visc_nbr(g, i, j + 1, k), ym = visc_nbr(g, i, j - 1, k);
  const ViscNbr zp = visc_nbr(g, i, j, k + 1), zm = visc_nbr(g, i, j, k - 1);
  for (int q = 0; q < 8 + g.npas; q++) {
    const double up = (double)b.U[q * N + c];
    const double lap = (((((visc_val(b, xp, q) + visc_val(b, xm, q)) + visc_val(b, yp, q)) + visc_val(b, ym, q))
                         + visc_val(b, zp, q)) + visc_val(b, zm, q)) - 6.0 * up;
    b.W[q * N + c] = up + eta * lap;
  }
}

template <typename S>
__device__ __forceinline__ void visc_store_cell(const Box<S>& b, int i, int j, int k) {
  const long long c = cidx(b.g, i, j, k), N = b.g.ncell;
  for (int q = 0; q < 8 + b.g.npas; q++) b.U[q * N + c] = (S)b.W[q * N + c];
}

template <typename S>
__global__ void __launch_bounds__(256) k_viscous(Box<S> b, double eta) {
  const Geom& g = b.g;
  const long long n = (long long)g.nx * g.ny * g.nz;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const long long r = t / g.nx;
  visc_cell(b, (int)(t % g.nx) + 1, (int)(r % g.ny) + 1, (int)(r / g.ny) + 1, eta);
}

template <typename S>
__global__ void __launch_bounds__(256) k_visc_store(Box<S> b) {
  const Geom& g = b.g;
  const long long n = (long long)g.nx * g.ny * g.nz;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const long long r = t / g.nx;
  visc_store_cell(b, (int)(t % g.nx) + 1, (int)(r % g.ny) + 1, (int)(r / g.ny) + 1);
}

// ---------------------------------------------------------------------------
// Thermal conduction, thermal_conduction (src/thermal_cond.f90:690-768),
// sub-step kernels.  Temperatures and primitives come from the stored state
// (ghosts in memory), T = (p / rho) Tempsc as calcprim gives them.
// Spitzer coefficients Ksp = 6e-7 T^2.5 (:142-149), Ksp_parl = 9.2181e-7
// T^2.5, Ksp_perp = 0.30089e33 rho / (B^2 sqrt(T)) rho (:157-178); pow and
// sqrt are the device's.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double tc_ksp(double t) { return 6.e-7 * pow(t, 2.5); }
__device__ __forceinline__ double tc_ksp_parl(double t) { return 9.2181e-7 * pow(t, 2.5); }
__device__ __forceinline__ double tc_ksp_perp(double t, double d, double b2) {
  return ((0.30089e+33 * d) / (b2 * __builtin_sqrt(t))) * d;
}
__device__ __forceinline__ double tc_min(double a, double b) { return (a < b) ? a : b; }

struct TcCell {
  double rho, p, t, b[3];
};

template <typename S>
__device__ __forceinline__ TcCell tc_cell(const Box<S>& b, const Conduction& tc, int i, int j, int k) {
  const long long c = cidx(b.g, i, j, k), N = b.g.ncell;
  double u[10], p[10];
#pragma unroll
  for (int q = 0; q < 10; q++) u[q] = q < 8 + b.g.npas ? (double)b.U[q * N + c] : 0.0;
  const double te = u2prim_t<10>(b.ph, u, p);
  TcCell r;
  r.rho = p[0];
  r.p = p[4];
  r.t = module_temp(b.ph, te, p, tc.tempsc);
  r.b[0] = p[5];
  r.b[1] = p[6];
  r.b[2] = p[7];
  return r;
}

template <typename S>
__global__ void __launch_bounds__(256) k_tc_dtmin(Box<S> b, Conduction tc, unsigned long long* out) {
  __shared__ double red[4];
  const Geom& g = b.g;
  const long long n = (long long)g.nx * g.ny * g.nz;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double v = 1.79769313486231570815e+308;  // huge(1.)
  if (t < n) {
    const int i = (int)(t % g.nx) + 1;
    const long long r = t / g.nx;
    const TcCell cl = tc_cell(b, tc, i, (int)(r % g.ny) + 1, (int)(r / g.ny) + 1);
    v = cl.rho / tc_ksp(cl.t);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = tc_min(v, __shfl_xor(v, off, 64));
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = red[0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = tc_min(m, red[w]);
    atomicMin(out, (unsigned long long)__double_as_longlong(m));
  }
}

// heatfluxes (:189-267) / MHD_heatfluxes (:277-487): the three forward faces
// of every cell (i, j, k) in 0..n, written to W (f, g, h at components 0..2)
template <typename S>
__global__ void __launch_bounds__(256) k_tc_flux(Box<S> b, Conduction tc) {
  const Geom& g = b.g;
  const int ex = g.nx + 1, ey = g.ny + 1;
  const long long n = (long long)ex * ey * (g.nz + 1);
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = (int)(t % ex);
  const long long r = t / ex;
  const int j = (int)(r % ey), k = (int)(r / ey);
  const double hh[3] = {b.dx, b.dy, b.dz};
  const TcCell c0 = tc_cell(b, tc, i, j, k);
  TcCell nb[3];
  nb[0] = tc_cell(b, tc, i + 1, j, k);
  nb[1] = tc_cell(b, tc, i, j + 1, k);
  nb[2] = tc_cell(b, tc, i, j, k + 1);
  double out[3];
  if (tc.kind == 1) {
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const TcCell& c1 = nb[a];
      if (c0.t == c1.t) {
        out[a] = 0.;
        continue;
      }
      const double meanp = 0.5 * (c0.p + c1.p);
      const double meand = 0.5 * (c0.rho + c1.rho);
      const double meant = 0.5 * (c0.t + c1.t);
      const double dtg = (c1.t - c0.t) / (hh[a] * tc.rsc);
      double coef;
      if (tc.sat) {
        double cs = __builtin_sqrt((b.ph.gamma * meanp) / meand);
        cs = tc_min(cs * __builtin_sqrt(tc.vsc2), 3.e10);
        coef = tc_min(tc_ksp(meant), ((((5. * 0.4) * cs) * meanp) * tc.psc) / __builtin_fabs(dtg));
      } else {
        coef = tc_ksp(meant);
      }
      out[a] = -((coef * dtg) * 1.);
    }
  } else {
    const double alpha = 5.0 * 0.3;
    double bx = c0.b[0], by = c0.b[1], bz = c0.b[2];
    const double b2 = (bx * bx + by * by) + bz * bz;
    const double modb = __builtin_sqrt(b2);
    bx = bx / modb;
    by = by / modb;
    bz = bz / modb;
    const double bsc2 = tc.bsc * tc.bsc;
    double grad[3], kpar[3], kperp[3], csat[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const TcCell& c1 = nb[a];
      if (__builtin_fabs(c0.t - c1.t) < 1.0e-14) {
        grad[a] = 0.0;
        kpar[a] = 0.0;
        kperp[a] = 0.0;
        csat[a] = 0.0;
        continue;
      }
      const double meand = 0.5 * (c0.rho + c1.rho);
      const double meant = 0.5 * (c0.t + c1.t);
      if (tc.sat) {
        const double meanp = 0.5 * (c0.p + c1.p);
        double cs = __builtin_sqrt((b.ph.gamma * meanp) / meand);
        cs = tc_min(cs * __builtin_sqrt(tc.vsc2), 3.e10);
        csat[a] = (alpha * meand) * (cs * (cs * cs));  // cs**3 as the compiler's powi forms it
      } else {
        csat[a] = 0.0;
      }
      grad[a] = (c1.t - c0.t) / (hh[a] * tc.rsc);
      kpar[a] = tc_ksp_parl(meant);
      kperp[a] = tc_ksp_perp(meant, meand * tc.rhosc, b2 * bsc2);
    }
    const double bgradt = (bx * grad[0] + by * grad[1]) + bz * grad[2];
    const double gpar[3] = {bgradt * bx, bgradt * by, bgradt * bz};
    const double gperp[3] = {grad[0] - gpar[0], grad[1] - gpar[1], grad[2] - gpar[2]};
    if (!tc.sat) {
#pragma unroll
      for (int a = 0; a < 3; a++) out[a] = (-(kpar[a] * gpar[a])) - kperp[a] * gperp[a];
    } else {
      const double gperpm = __builtin_sqrt((gperp[0] * gperp[0] + gperp[1] * gperp[1]) + gperp[2] * gperp[2]);
#pragma unroll
      for (int a = 0; a < 3; a++)
        out[a] = (-((1. / (1. / (kpar[a] + 1.e-14) + bgradt / (csat[a] + 1.e-14))) * gpar[a]))
                 - (1. / (1. / (kperp[a] + 1.e-14) + gperpm / (csat[a] + 1.e-14))) * gperp[a];
    }
  }
  const long long c = cidx(g, i, j, k), N = g.ncell;
  b.W[c] = out[0];
  b.W[N + c] = out[1];
  b.W[2 * N + c] = out[2];
}

// u(5) = u(5) - dts (df/dx + dg/dy + dh/dz) on the interior (:752-758)
template <typename S>
__global__ void __launch_bounds__(256) k_tc_update(Box<S> b, double dts) {
  const Geom& g = b.g;
  const long long n = (long long)g.nx * g.ny * g.nz;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = (int)(t % g.nx) + 1;
  const long long r = t / g.nx;
  const int j = (int)(r % g.ny) + 1, k = (int)(r / g.ny) + 1;
  const long long c = cidx(g, i, j, k), N = g.ncell;
  const long long sy = g.NX, sz = (long long)g.NX * g.NY;
  const double* F = b.W;
  const double* G = b.W + N;
  const double* H = b.W + 2 * N;
  const double e = (double)b.U[4 * N + c];
  b.U[4 * N + c] = (S)(e - dts * (((F[c] - F[c - 1]) / b.dx + (G[c] - G[c - sy]) / b.dy) + (H[c] - H[c - sz]) / b.dz));
}

// thermal_bounds (:496-616) on E, one face: periodic copies over the full
// transverse range (periodic != 0), else the edge-cell copy over 0..n+1
template <typename S>
__global__ void __launch_bounds__(256) k_tc_bound(Geom g, S* __restrict__ U, int axis, int side, int periodic) {
  const int n0 = axis == 0 ? g.ny : g.nx;
  const int n1 = axis == 2 ? g.ny : g.nz;
  const int na = axis == 0 ? g.nx : axis == 1 ? g.ny : g.nz;
  const int lo = periodic ? -1 : 0;
  const int e0 = n0 + 2 - 2 * lo, e1 = n1 + 2 - 2 * lo;
  const long long n = (long long)e0 * e1;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int a = (int)(t % e0) + lo, bq = (int)(t / e0) + lo;
  const int dst = side == 0 ? 0 : na + 1;
  const int src = periodic ? (side == 0 ? na : 1) : (side == 0 ? 1 : na);
  int di, dj, dk, si, sj, sk;
  if (axis == 0) { di = dst; si = src; dj = sj = a; dk = sk = bq; }
  else if (axis == 1) { dj = dst; sj = src; di = si = a; dk = sk = bq; }
  else { dk = dst; sk = src; di = si = a; dj = sj = bq; }
  U[4 * g.ncell + cidx(g, di, dj, dk)] = U[4 * g.ncell + cidx(g, si, sj, sk)];
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

// get_timestep's candidate of one interior cell (src/hydro_core.f90:641-672):
// min over the axes of d / (|v| + c), c the fast speed along the axis (MHD)
// or the sound speed.  The primitives and the minimum are FP64; the wave
// speeds and candidates are C (FP32 in the mixed form).
template <int NQ, typename C>
__device__ __forceinline__ double cfl_cell(const Phys& ph, double dx, double dy, double dz, const double* u,
                                           double dtp, const double* pcell) {
  using A = Ar<C>;
  double pd[NQ];
  u2prim<NQ>(ph, u, pd);
  if (pcell) pd[4] = *pcell;  // the pressure primit carries after table cooling
  C p[8];
#pragma unroll
  for (int q = 0; q < 8; q++) p[q] = (C)pd[q];
  const C g = (C)ph.gamma;
  if (ph.mhd) {
    const C b2 = (p[5] * p[5] + p[6] * p[6]) + p[7] * p[7];
    const C cx = cfast1(g, p[4], p[0], b2, p[5]);
    const C cy = cfast1(g, p[4], p[0], b2, p[6]);
    const C cz = cfast1(g, p[4], p[0], b2, p[7]);
    dtp = dmin(dtp, (double)A::div((C)dx, tabs(p[1]) + cx));
    dtp = dmin(dtp, (double)A::div((C)dy, tabs(p[2]) + cy));
    dtp = dmin(dtp, (double)A::div((C)dz, tabs(p[3]) + cz));
  } else {
    const C cs = csound(ph, p[4], p[0]);
    dtp = dmin(dtp, (double)A::div((C)dx, tabs(p[1]) + cs));
    dtp = dmin(dtp, (double)A::div((C)dy, tabs(p[2]) + cs));
    dtp = dmin(dtp, (double)A::div((C)dz, tabs(p[3]) + cs));
  }
  return dtp;
}

// the workgroup's minimum into *dtmin (positive doubles order as their bit
// patterns); a non-finite conserved value sets bit 2 of *err.  Every thread
// of the workgroup (at most 256) calls it.
__device__ __forceinline__ void cfl_reduce(double dtp, bool bad, unsigned long long* __restrict__ dtmin,
                                           int* __restrict__ err) {
  __shared__ double red[4];
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

template <typename S, int NQ = 8>
__device__ __forceinline__ void cfl_body(const Box<S>& b, unsigned lb, int kchunk,
                                         unsigned long long* __restrict__ dtmin, int* __restrict__ err) {
  const Geom& g = b.g;
  double dtp = 1e30;
  bool bad = false;
  int t, kc;
  if (tile_item(lb, tiles_of((unsigned)(g.nx * g.ny)), (unsigned)(g.nx * g.ny), t, kc)) {
    const int i = t % g.nx + 1, j = t / g.nx + 1;
    const int k0 = kc * kchunk + 1;
    const int k1 = min(k0 + kchunk - 1, g.nz);
    for (int k = k0; k <= k1; k++) {
      const long long c = cidx(g, i, j, k);
      double u[NQ];
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        u[q] = q < 8 + g.npas ? (double)b.U[q * g.ncell + c] : 0.0;
        bad |= !__builtin_isfinite(u[q]);
      }
      dtp = cfl_cell<NQ, S>(b.ph, b.dx, b.dy, b.dz, u, dtp, b.pc ? b.pc + c : nullptr);
    }
  }
  cfl_reduce(dtp, bad, dtmin, err);
}

// ---------------------------------------------------------------------------
// __global__ wrappers: one box ...
// ---------------------------------------------------------------------------
// NP passive scalars: their larger LDS tile leaves one workgroup per CU, so
// the register budget is not squeezed for them.
template <int SOLVER, int ORDER, int FLUXCD, typename S, int NP>
__global__ void __launch_bounds__(XY_THREADS) __attribute__((amdgpu_waves_per_eu(NP == 0 ? 4 : 2)))
k_xy(Box<S> b, double dt, int* err) {
  __shared__ S lds[XYLds<ORDER, 8 + NP>::SIZE];
  xy_body<SOLVER, ORDER, FLUXCD, S, NP>(b, xcd_logical_block(), dt, err, lds);
}

template <int SOLVER, int ORDER, int FLUXCD, typename S, int NP, bool W8 = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ORDER == 1 && FLUXCD && NP == 0 && !W8 ? 3 : 1)))
k_z(Box<S> b, double dt, int kchunk, int* err) {
  // tiles in the reverse of the xy kernel's order, so the first pencils
  // read the partial sums k_xy wrote last, still in the memory-side cache;
  // the corrector's field update then runs forward over what this kernel
  // wrote last
  z_body<SOLVER, ORDER, FLUXCD, S, NP, W8>(b, gridDim.x - 1 - xcd_logical_block(), dt, kchunk, err);
}

// the 2.5-D stage kernel: up to 256 VGPRs (two waves per SIMD, two
// workgroups per CU) for the column window carried across the plane march;
// the FP64 corrector with passive scalars needs more LDS than two
// workgroups per CU leave, so it runs at one (the compiler says so)
template <int SOLVER, int ORDER, int FLUXCD, typename S, int NP>
__global__ void __launch_bounds__(SL_THREADS) __attribute__((amdgpu_waves_per_eu(2)))
k_slab(Box<S> b, double dt, int kchunk, int* err) {
  __shared__ S lds[SlabLds<ORDER, 8 + NP, S>::SIZE];
  slab_body<SOLVER, ORDER, FLUXCD, S, NP>(b, xcd_logical_block(), dt, kchunk, err, lds);
}

template <typename S>
__global__ void __launch_bounds__(256) k_bfield(Box<S> b, double dt, int stage, int kchunk) {
  // the predictor's field update runs its tiles in the reverse of the slab
  // kernel's order (E written last is read first, from the memory-side
  // cache); the corrector's, after the reversed z kernel, runs forward
  const unsigned lb = xcd_logical_block();
  bfield_body(b, stage == 1 ? gridDim.x - 1 - lb : lb, dt, stage, kchunk);
}

// the corrector's field update with the next step's CFL minimum folded in:
// U^{n+1} is complete cell by cell once its field is written, so when
// nothing after the field update changes U's interior the reduction needs
// no pass of its own over U
template <typename S, int NQ>
__global__ void __launch_bounds__(256) k_bfield_cfl(Box<S> b, double dt, int kchunk, unsigned long long* dtmin,
                                                    int* err) {
  double dtp = 1e30;
  bool bad = false;
  bfield_body<S, 1, NQ>(b, xcd_logical_block(), dt, 2, kchunk, &dtp, &bad);
  cfl_reduce(dtp, bad, dtmin, err);
}

template <typename S, int NQ>
__global__ void __launch_bounds__(256) k_cfl(Box<S> b, int kchunk, unsigned long long* dtmin, int* err) {
  cfl_body<S, NQ>(b, xcd_logical_block(), kchunk, dtmin, err);
}

template <typename S>
__global__ void __launch_bounds__(256) k_sources(Box<S> b, double dt, int stage, int kchunk) {
  sources_body(b, xcd_logical_block(), dt, stage, kchunk);
}

// ... and a batch of boxes.  dt of each box from the device clock table;
// the predictor uses dt/2 (hydro_solver.f90:152).
__device__ __forceinline__ double stage_dt(const BoxClock* __restrict__ clk, int slot, int stage) {
  const double dt = clk[slot].dt;
  return stage == 1 ? dt / 2.0 : dt;
}

// batched kernels: NP passive scalars per box (one count per batch)
template <int SOLVER, int ORDER, int FLUXCD, typename S, int NP>
__global__ void __launch_bounds__(XY_THREADS) __attribute__((amdgpu_waves_per_eu(NP == 0 ? 4 : 2)))
kb_xy(const Box<S>* __restrict__ boxes, const unsigned* __restrict__ prefix, int nbox,
      const BoxClock* __restrict__ clk, int* err) {
  __shared__ S lds[XYLds<ORDER, 8 + NP>::SIZE];
  int bx;
  unsigned lb;
  batch_locate(prefix, nbox, xcd_logical_block(), bx, lb);
  const Box<S>& b = boxes[bx];
  xy_body<SOLVER, ORDER, FLUXCD, S, NP>(b, lb, stage_dt(clk, b.slot, ORDER), err, lds);
}

// (the FP64 corrector with passive scalars runs at one workgroup per CU,
// as k_slab does)
template <int SOLVER, int ORDER, int FLUXCD, typename S, int NP>
__global__ void __launch_bounds__(SL_THREADS) __attribute__((amdgpu_waves_per_eu(2)))
kb_slab(const Box<S>* __restrict__ boxes, const unsigned* __restrict__ prefix, int nbox,
        const BoxClock* __restrict__ clk, int kchunk, int* err) {
  __shared__ S lds[SlabLds<ORDER, 8 + NP, S>::SIZE];
  int bx;
  unsigned lb;
  batch_locate(prefix, nbox, xcd_logical_block(), bx, lb);
  const Box<S>& b = boxes[bx];
  slab_body<SOLVER, ORDER, FLUXCD, S, NP>(b, lb, stage_dt(clk, b.slot, ORDER), kchunk, err, lds);
}

template <int SOLVER, int ORDER, int FLUXCD, typename S, int NP>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(ORDER == 1 && FLUXCD && NP == 0 ? 3 : 1)))
kb_z(const Box<S>* __restrict__ boxes, const unsigned* __restrict__ prefix,
                                            int nbox, const BoxClock* __restrict__ clk, int kchunk, int* err) {
  int bx;
  unsigned lb;
  batch_locate(prefix, nbox, xcd_logical_block(), bx, lb);
  const Box<S>& b = boxes[bx];
  z_body<SOLVER, ORDER, FLUXCD, S, NP>(b, lb, stage_dt(clk, b.slot, ORDER), kchunk, err);
}

template <typename S>
__global__ void __launch_bounds__(256) kb_bfield(const Box<S>* __restrict__ boxes,
                                                 const unsigned* __restrict__ prefix, int nbox,
                                                 const BoxClock* __restrict__ clk, int stage, int kchunk) {
  int bx;
  unsigned lb;
  batch_locate(prefix, nbox, xcd_logical_block(), bx, lb);
  const Box<S>& b = boxes[bx];
  bfield_body(b, lb, stage_dt(clk, b.slot, stage), stage, kchunk);
}

// the corrector's field update of a batch with each box's next CFL
// minimum folded in (k_bfield_cfl per box): into the box's clock, which
// kb_clock_dt reset to 1e30 when it took the step's dt
template <typename S, int NQ>
__global__ void __launch_bounds__(256) kb_bfield_cfl(const Box<S>* __restrict__ boxes,
                                                     const unsigned* __restrict__ prefix, int nbox,
                                                     BoxClock* __restrict__ clk, int kchunk, int* err) {
  int bx;
  unsigned lb;
  batch_locate(prefix, nbox, xcd_logical_block(), bx, lb);
  const Box<S>& b = boxes[bx];
  double dtp = 1e30;
  bool bad = false;
  bfield_body<S, 1, NQ>(b, lb, stage_dt(clk, b.slot, 2), 2, kchunk, &dtp, &bad);
  cfl_reduce(dtp, bad, &clk[b.slot].dtmin, err);
}

template <typename S>
__global__ void __launch_bounds__(256) kb_sources(const Box<S>* __restrict__ boxes,
                                                  const unsigned* __restrict__ prefix, int nbox,
                                                  const BoxClock* __restrict__ clk, int stage, int kchunk) {
  int bx;
  unsigned lb;
  batch_locate(prefix, nbox, xcd_logical_block(), bx, lb);
  const Box<S>& b = boxes[bx];
  sources_body(b, lb, stage_dt(clk, b.slot, stage), stage, kchunk);
}

// viscous_copy over a batch's column tiles (STORE: the copy back from W).
// A box without viscosity (Box::eta = 0), or whose step has dt = 0 (it has
// completed, or the autotuner is timing shapes), keeps its state: unlike
// the update, viscous_copy does not scale with dt.
template <typename S, bool STORE>
__global__ void __launch_bounds__(256) kb_viscous(const Box<S>* __restrict__ boxes,
                                                  const unsigned* __restrict__ prefix, int nbox,
                                                  const BoxClock* __restrict__ clk, int kchunk) {
  int bx;
  unsigned lb;
  batch_locate(prefix, nbox, xcd_logical_block(), bx, lb);
  const Box<S>& b = boxes[bx];
  if (b.eta == 0.0 || clk[b.slot].dt == 0.0) return;
  const Geom& g = b.g;
  int t, kc;
  if (!tile_item(lb, tiles_of((unsigned)(g.nx * g.ny)), (unsigned)(g.nx * g.ny), t, kc)) return;
  const int i = t % g.nx + 1, j = t / g.nx + 1;
  const int k0 = kc * kchunk + 1, k1 = min(k0 + kchunk - 1, g.nz);
  for (int k = k0; k <= k1; k++) {
    if (STORE) visc_store_cell(b, i, j, k);
    else visc_cell(b, i, j, k, b.eta);
  }
}

template <typename S, int NQ>
__global__ void __launch_bounds__(256) kb_cfl(const Box<S>* __restrict__ boxes, const unsigned* __restrict__ prefix,
                                              int nbox, int kchunk, BoxClock* clk, int* err) {
  int bx;
  unsigned lb;
  batch_locate(prefix, nbox, xcd_logical_block(), bx, lb);
  cfl_body<S, NQ>(boxes[bx], lb, kchunk, &clk[boxes[bx].slot].dtmin, err);
}


#if GM_PART(1)
// Start of a batched step: reset the CFL minimum of every box in the
// launch (slots[0..nbox), the packed table's clock slots).
__global__ void kb_clock_begin(BoxClock* clk, const int* __restrict__ slots, int nbox) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nbox) clk[slots[i]].dtmin = (unsigned long long)__double_as_longlong(1e30);
}

// get_timestep's ramp and clamp on the device (src/hydro_core.f90:677-694):
// dt = cfl * 2^-(n_iter+1-it) * dtp for it <= n_iter, cfl * dtp after;
// clamped to land on t_end, which completes the box.  A box that is done,
// or a timing-only launch (tuning != 0), steps with dt = 0: the fluxes are
// computed in full and the state keeps its values.
__global__ void kb_clock_dt(BoxClock* clk, const int* __restrict__ slots, int nbox, int tuning) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nbox) return;
  const int b = slots[i];
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
  c.dtmin = (unsigned long long)__double_as_longlong(1e30);  // for the corrector's folded minimum
  if (finish || (c.max_steps > 0 && c.steps >= c.max_steps)) c.done = 1;
  clk[b] = c;
}
#endif  // GM_PART(1)

// ---------------------------------------------------------------------------
// ghost fill of one face of the conserved state, serial branches:
//   layers 1, boundaryI (src/boundaries.f90:44-247): ghost 0 (n+1); periodic
//     copies over the full transverse range, closed/outflow over 0..n+1;
//   layers 2, boundaryII (:256-512): ghosts -1, 0 (n+1, n+2) over the full
//     transverse range; periodic from n-1, n (1, 2), closed/outflow mirrored
//     (ghost 0 <- cell 1, ghost -1 <- cell 2).
// Closed walls negate the normal momentum.
// ---------------------------------------------------------------------------
template <typename S>
__global__ void __launch_bounds__(256) k_ghost(Geom g, S* __restrict__ U, int axis, int side, int kind, int layers) {
  const int n0 = axis == 0 ? g.ny : g.nx;  // transverse extents (interior)
  const int n1 = axis == 2 ? g.ny : g.nz;
  const int na = axis == 0 ? g.nx : axis == 1 ? g.ny : g.nz;
  const int lo = (kind == BC_PERIODIC || layers == 2) ? -1 : 0;
  const int e0 = n0 + 2 - 2 * lo, e1 = n1 + 2 - 2 * lo;
  const long long n = (long long)e0 * e1 * layers;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x) {
    const long long plane = (long long)e0 * e1;
    const int layer = (int)(t / plane);  // 0: the layer next to the domain
    const long long r = t - (long long)layer * plane;
    const int a = (int)(r % e0) + lo;
    const int b = (int)(r / e0) + lo;
    const int dst = side == 0 ? -layer : na + 1 + layer;
    int src;
    if (kind == BC_PERIODIC) src = side == 0 ? dst + na : dst - na;
    else src = side == 0 ? 1 - dst : 2 * na + 1 - dst;
    int di, dj, dk, si, sj, sk;
    if (axis == 0) { di = dst; si = src; dj = sj = a; dk = sk = b; }
    else if (axis == 1) { dj = dst; sj = src; di = si = a; dk = sk = b; }
    else { dk = dst; sk = src; di = si = a; dj = sj = b; }
    const long long cd = cidx(g, di, dj, dk), cs = cidx(g, si, sj, sk);
    for (int q = 0; q < 8 + g.npas; q++) {
      const S v = U[q * g.ncell + cs];
      U[q * g.ncell + cd] = (kind == BC_CLOSED && q == 1 + axis) ? -v : v;
    }
  }
}

// ---------------------------------------------------------------------------
// layout conversion between the reference's AoS u(neq,i,j,k) and SoA
// ---------------------------------------------------------------------------
template <typename S>
__global__ void __launch_bounds__(256) k_aos_to_soa(long long ncell, int nq, const double* __restrict__ A,
                                                    S* __restrict__ D) {
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < ncell;
       c += (long long)gridDim.x * blockDim.x) {
    for (int q = 0; q < nq; q++) D[q * ncell + c] = (S)A[c * nq + q];
  }
}

template <typename S>
__global__ void __launch_bounds__(256) k_soa_to_aos(long long ncell, int nq, const S* __restrict__ D,
                                                    double* __restrict__ A) {
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < ncell;
       c += (long long)gridDim.x * blockDim.x) {
    for (int q = 0; q < nq; q++) A[c * nq + q] = (double)D[q * ncell + c];
  }
}

// primitives and temperature of every cell as stored (calcprim over the
// whole array, hydro_core.f90:304-316), AoS output
template <typename S>
__global__ void __launch_bounds__(256) k_prim_all(long long ncell, int nq, Phys ph, const S* __restrict__ U,
                                                  double* __restrict__ Paos, double* __restrict__ T, Geom g,
                                                  const double* __restrict__ pc) {
  for (long long c = blockIdx.x * (long long)blockDim.x + threadIdx.x; c < ncell;
       c += (long long)gridDim.x * blockDim.x) {
    double u[10], p[10];
#pragma unroll
    for (int q = 0; q < 10; q++) u[q] = q < nq ? (double)U[q * ncell + c] : 0.0;
    const double te = u2prim_t<10>(ph, u, p);
    if (pc) {  // the pressure primit carries after table cooling, interior cells (Box::pc)
      const long long i = c % g.NX - 1, r = c / g.NX, j = r % g.NY - 1, k = r / g.NY - 1;
      if (i >= 1 && i <= g.nx && j >= 1 && j <= g.ny && k >= 1 && k <= g.nz) p[4] = pc[c];
    }
    for (int q = 0; q < nq; q++) Paos[c * nq + q] = p[q];
    if (T) T[c] = te;
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
[[maybe_unused]] static inline unsigned grid_for(long long n, int block = 256) {
  long long b = (n + block - 1) / block;
  const long long cap = 256LL * 1024;  // grid-stride loops cover the rest
  if (b > cap) b = cap;
  if (b < 1) b = 1;
  return (unsigned)b;
}

[[maybe_unused]] static inline unsigned ceil_div(long long a, long long b) { return (unsigned)((a + b - 1) / b); }

#if GM_PART(1)
unsigned box_tiles(const Geom& g, int kclass, const Shape& sh) {
  switch (kclass) {
    case KC_XY: return ceil_div(g.nx, TX) * ceil_div(g.ny, TY) * (unsigned)g.nz;
    case KC_COL: return ceil_div((long long)g.nx * g.ny, sh.block) * ceil_div(g.nz, sh.kchunk);
    case KC_SLAB: return sh.slab > 0 ? slab_tiles(g, sh.slab) : 0u;
  }
  return 0;
}
#endif

// solver x order x flux-CD dispatch; CALL(SOLVER, ORDER, FLUXCD)
#define GM_STAGE_DISPATCH(CALL)                                                      \
  switch (solver) {                                                                  \
    case SOLVER_HLLD: GM_STAGE_OF(SOLVER_HLLD, CALL) break;                           \
    case SOLVER_HLLE: GM_STAGE_OF(SOLVER_HLLE, CALL) break;                           \
    case SOLVER_HLL: GM_STAGE_OF(SOLVER_HLL, CALL) break;                             \
    default: GM_STAGE_OF(SOLVER_HLLC, CALL) break;                                    \
  }
#define GM_STAGE_OF(SV, CALL)                                                        \
  if (stage == 1) {                                                                  \
    if (flux_cd) CALL(SV, 1, 1) else CALL(SV, 1, 0)                                  \
  } else {                                                                           \
    if (flux_cd) CALL(SV, 2, 1) else CALL(SV, 2, 0)                                  \
  }

template <typename S>
void launch_xy(hipStream_t s, const Box<S>& b, double dt, int stage, int solver, int flux_cd, int* err) {
  const unsigned n = box_tiles(b.g, KC_XY, Shape{XY_THREADS, 1});
#define GM_CALL(SV, O, F)                                                                                  \
  {                                                                                                        \
    if (b.g.npas == 0) hipLaunchKernelGGL((k_xy<SV, O, F, S, 0>), dim3(n), dim3(XY_THREADS), 0, s, b, dt, err); \
    else if (b.g.npas == 1)                                                                                \
      hipLaunchKernelGGL((k_xy<SV, O, F, S, 1>), dim3(n), dim3(XY_THREADS), 0, s, b, dt, err);           \
    else hipLaunchKernelGGL((k_xy<SV, O, F, S, 2>), dim3(n), dim3(XY_THREADS), 0, s, b, dt, err);       \
  }
  GM_STAGE_DISPATCH(GM_CALL)
#undef GM_CALL
}

template <typename S>
void launch_z(hipStream_t s, const Box<S>& b, double dt, int stage, int solver, int flux_cd, int* err,
              const Shape& sh) {
  const unsigned n = box_tiles(b.g, KC_COL, sh);
#define GM_CALL(SV, O, F)                                                                                  \
  {                                                                                                        \
    if (b.srcin) hipLaunchKernelGGL((k_z<SV, O, F, S, 0, true>), dim3(n), dim3(sh.block), 0, s, b, dt, sh.kchunk, err); \
    else if (b.g.npas == 0) hipLaunchKernelGGL((k_z<SV, O, F, S, 0>), dim3(n), dim3(sh.block), 0, s, b, dt, sh.kchunk, err); \
    else if (b.g.npas == 1)                                                                                \
      hipLaunchKernelGGL((k_z<SV, O, F, S, 1>), dim3(n), dim3(sh.block), 0, s, b, dt, sh.kchunk, err);    \
    else hipLaunchKernelGGL((k_z<SV, O, F, S, 2>), dim3(n), dim3(sh.block), 0, s, b, dt, sh.kchunk, err); \
  }
  GM_STAGE_DISPATCH(GM_CALL)
#undef GM_CALL
}

#if GM_PART(1)
unsigned slab_tiles(const Geom& g, int kchunk) {
  return ceil_div(g.nx, TX) * ceil_div(g.ny, TYS) * ceil_div(g.nz, kchunk);
}

// z-planes per slab.  About 2048 workgroups (eight per CU, so the launch's
// tail is short): at 256^3 that is 32 planes per slab, and 8 to 64 run
// within 1%; at 96^3 to 192^3, 2 to 13 planes, within 2% of the best length
// swept.  A grid of no more than two rounds of resident workgroups at one
// plane per slab (`resident`: the slab workgroups the GPU holds at once,
// two per CU in FP64, three in the mixed form) does not keep the GPU busy
// for long, and there the length k minimises
//   ceil(workgroups(k) / resident) * (k + 1),
// the rounds of resident workgroups times a workgroup's march (k planes
// plus about one plane of window warm-up): one round of one- or two-plane
// marches rather than a second, nearly empty, round.  Per step against the
// former choice (at least 8 planes per slab): 128x128x4 0.077 ms against
// 0.108, 1024x4x4 0.066 against 0.096, 32^3 0.092 against 0.193, 64^3
// 0.137 against 0.209.
int slab_kchunk(const Geom& g, int resident) {
  const long long per_plane = (long long)ceil_div(g.nx, TX) * ceil_div(g.ny, TYS);
  if (per_plane * g.nz > 2LL * resident) {
    long long nslab = (2048 + per_plane - 1) / per_plane;
    if (nslab > g.nz) nslab = g.nz;
    if (nslab < 1) nslab = 1;
    return (int)((g.nz + nslab - 1) / nslab);
  }
  int best = 1;
  long long best_cost = -1;
  for (int k = 1; k <= g.nz; k++) {
    const long long wg = per_plane * ((g.nz + k - 1) / k);
    const long long cost = ((wg + resident - 1) / resident) * (k + 1);
    if (best_cost < 0 || cost < best_cost) {
      best_cost = cost;
      best = k;
    }
  }
  return best;
}
#endif  // GM_PART(1)

template <typename S>
void launch_slab(hipStream_t s, const Box<S>& b, double dt, int stage, int solver, int flux_cd, int* err,
                 int kchunk) {
  const unsigned n = slab_tiles(b.g, kchunk);
#define GM_CALL(SV, O, F)                                                                                        \
  {                                                                                                              \
    if (b.g.npas == 0) hipLaunchKernelGGL((k_slab<SV, O, F, S, 0>), dim3(n), dim3(SL_THREADS), 0, s, b, dt, kchunk, err); \
    else if (b.g.npas == 1)                                                                                      \
      hipLaunchKernelGGL((k_slab<SV, O, F, S, 1>), dim3(n), dim3(SL_THREADS), 0, s, b, dt, kchunk, err);        \
    else hipLaunchKernelGGL((k_slab<SV, O, F, S, 2>), dim3(n), dim3(SL_THREADS), 0, s, b, dt, kchunk, err);    \
  }
  GM_STAGE_DISPATCH(GM_CALL)
#undef GM_CALL
}

template <typename S>
void launch_bfield(hipStream_t s, const Box<S>& b, double dt, int stage, const Shape& sh) {
  hipLaunchKernelGGL(k_bfield<S>, dim3(box_tiles(b.g, KC_COL, sh)), dim3(sh.block), 0, s, b, dt, stage, sh.kchunk);
}

template <typename S>
void launch_bfield_cfl(hipStream_t s, const Box<S>& b, double dt, unsigned long long* dtmin, int* err,
                       const Shape& sh) {
  if (b.g.npas == 0)
    hipLaunchKernelGGL((k_bfield_cfl<S, 8>), dim3(box_tiles(b.g, KC_COL, sh)), dim3(sh.block), 0, s, b, dt,
                       sh.kchunk, dtmin, err);
  else
    hipLaunchKernelGGL((k_bfield_cfl<S, 10>), dim3(box_tiles(b.g, KC_COL, sh)), dim3(sh.block), 0, s, b, dt,
                       sh.kchunk, dtmin, err);
}

template <typename S>
void launch_cfl(hipStream_t s, const Box<S>& b, unsigned long long* dtmin, int* err, const Shape& sh) {
  if (b.g.npas == 0)
    hipLaunchKernelGGL((k_cfl<S, 8>), dim3(box_tiles(b.g, KC_COL, sh)), dim3(sh.block), 0, s, b, sh.kchunk, dtmin, err);
  else  // EOS_H_RATE reads the first passive scalar
    hipLaunchKernelGGL((k_cfl<S, 10>), dim3(box_tiles(b.g, KC_COL, sh)), dim3(sh.block), 0, s, b, sh.kchunk, dtmin, err);
}

template <typename S>
void launch_viscous(hipStream_t s, const Box<S>& b, double eta) {
  const long long n = (long long)b.g.nx * b.g.ny * b.g.nz;
  const dim3 grid((unsigned)((n + 255) / 256));
  hipLaunchKernelGGL(k_viscous<S>, grid, dim3(256), 0, s, b, eta);
  hipLaunchKernelGGL(k_visc_store<S>, grid, dim3(256), 0, s, b);
}

template <typename S>
void launch_cooling(hipStream_t s, const Box<S>& b, const Cooling& cl, double dt) {
  const long long n = (long long)b.g.nx * b.g.ny * b.g.nz;
  if (cl.kind == GM_COOL_H) hipLaunchKernelGGL(k_cooling_h<S>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, b, cl, dt);
  else hipLaunchKernelGGL(k_cooling<S>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, b, cl, dt);
}

template <typename S>
void launch_tc_dtmin(hipStream_t s, const Box<S>& b, const Conduction& tc, unsigned long long* out) {
  const long long n = (long long)b.g.nx * b.g.ny * b.g.nz;
  hipLaunchKernelGGL(k_tc_dtmin<S>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, b, tc, out);
}

template <typename S>
void launch_tc_substep(hipStream_t s, const Box<S>& b, const Conduction& tc, double dts) {
  const Geom& g = b.g;
  const long long nf = (long long)(g.nx + 1) * (g.ny + 1) * (g.nz + 1);
  hipLaunchKernelGGL(k_tc_flux<S>, dim3((unsigned)((nf + 255) / 256)), dim3(256), 0, s, b, tc);
  const long long ni = (long long)g.nx * g.ny * g.nz;
  hipLaunchKernelGGL(k_tc_update<S>, dim3((unsigned)((ni + 255) / 256)), dim3(256), 0, s, b, dts);
  auto bound = [&](int axis, int side, int periodic) {
    const int n0 = axis == 0 ? g.ny : g.nx;
    const int n1 = axis == 2 ? g.ny : g.nz;
    const long long n = (long long)(n0 + 4) * (n1 + 4);
    hipLaunchKernelGGL(k_tc_bound<S>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, b.U, axis, side,
                       periodic);
  };
  for (int a = 0; a < 3; a++)
    if (g.bc[2 * a] == BC_PERIODIC && g.bc[2 * a + 1] == BC_PERIODIC) {
      bound(a, 0, 1);
      bound(a, 1, 1);
    }
  for (int a = 0; a < 3; a++) {
    bound(a, 0, 0);
    bound(a, 1, 0);
  }
}

template <typename S>
void launch_sources(hipStream_t s, const Box<S>& b, double dt, int stage, const Shape& sh) {
  hipLaunchKernelGGL(k_sources<S>, dim3(box_tiles(b.g, KC_COL, sh)), dim3(sh.block), 0, s, b, dt, stage, sh.kchunk);
}

template <typename S>
void launch_sources_batch(hipStream_t s, const Box<S>* boxes, const unsigned* prefix, int nbox, unsigned tiles,
                          const BoxClock* clk, int stage, const Shape& sh) {
  hipLaunchKernelGGL(kb_sources<S>, dim3(tiles), dim3(sh.block), 0, s, boxes, prefix, nbox, clk, stage, sh.kchunk);
}

template <typename S>
void launch_viscous_batch(hipStream_t s, const Box<S>* boxes, const unsigned* prefix, int nbox, unsigned tiles,
                          const BoxClock* clk, const Shape& sh) {
  hipLaunchKernelGGL((kb_viscous<S, false>), dim3(tiles), dim3(sh.block), 0, s, boxes, prefix, nbox, clk, sh.kchunk);
  hipLaunchKernelGGL((kb_viscous<S, true>), dim3(tiles), dim3(sh.block), 0, s, boxes, prefix, nbox, clk, sh.kchunk);
}

template <typename S>
void launch_xy_batch(hipStream_t s, const Box<S>* boxes, const unsigned* prefix, int nbox, unsigned tiles,
                     const BoxClock* clk, int stage, int solver, int flux_cd, int* err, int npas) {
#define GM_CALL(SV, O, F)                                                                                        \
  {                                                                                                               \
    if (npas == 0) hipLaunchKernelGGL((kb_xy<SV, O, F, S, 0>), dim3(tiles), dim3(XY_THREADS), 0, s, boxes, prefix, nbox, clk, err); \
    else if (npas == 1) hipLaunchKernelGGL((kb_xy<SV, O, F, S, 1>), dim3(tiles), dim3(XY_THREADS), 0, s, boxes, prefix, nbox, clk, err); \
    else hipLaunchKernelGGL((kb_xy<SV, O, F, S, 2>), dim3(tiles), dim3(XY_THREADS), 0, s, boxes, prefix, nbox, clk, err); \
  }
  GM_STAGE_DISPATCH(GM_CALL)
#undef GM_CALL
}

template <typename S>
void launch_z_batch(hipStream_t s, const Box<S>* boxes, const unsigned* prefix, int nbox, unsigned tiles,
                    const BoxClock* clk, int stage, int solver, int flux_cd, int* err, const Shape& sh, int npas) {
#define GM_CALL(SV, O, F)                                                                                   \
  {                                                                                                         \
    if (npas == 0) hipLaunchKernelGGL((kb_z<SV, O, F, S, 0>), dim3(tiles), dim3(sh.block), 0, s, boxes, prefix, nbox, clk, sh.kchunk, err); \
    else if (npas == 1) hipLaunchKernelGGL((kb_z<SV, O, F, S, 1>), dim3(tiles), dim3(sh.block), 0, s, boxes, prefix, nbox, clk, sh.kchunk, err); \
    else hipLaunchKernelGGL((kb_z<SV, O, F, S, 2>), dim3(tiles), dim3(sh.block), 0, s, boxes, prefix, nbox, clk, sh.kchunk, err); \
  }
  GM_STAGE_DISPATCH(GM_CALL)
#undef GM_CALL
}

template <typename S>
void launch_slab_batch(hipStream_t s, const Box<S>* boxes, const unsigned* prefix, int nbox, unsigned tiles,
                       const BoxClock* clk, int stage, int solver, int flux_cd, int* err, const Shape& sh, int npas) {
#define GM_CALL(SV, O, F)                                                                                       \
  {                                                                                                              \
    if (npas == 0) hipLaunchKernelGGL((kb_slab<SV, O, F, S, 0>), dim3(tiles), dim3(SL_THREADS), 0, s, boxes, prefix, nbox, clk, sh.slab, err); \
    else if (npas == 1) hipLaunchKernelGGL((kb_slab<SV, O, F, S, 1>), dim3(tiles), dim3(SL_THREADS), 0, s, boxes, prefix, nbox, clk, sh.slab, err); \
    else hipLaunchKernelGGL((kb_slab<SV, O, F, S, 2>), dim3(tiles), dim3(SL_THREADS), 0, s, boxes, prefix, nbox, clk, sh.slab, err); \
  }
  GM_STAGE_DISPATCH(GM_CALL)
#undef GM_CALL
}
#undef GM_STAGE_OF
#undef GM_STAGE_DISPATCH

template <typename S>
void launch_bfield_batch(hipStream_t s, const Box<S>* boxes, const unsigned* prefix, int nbox, unsigned tiles,
                         const BoxClock* clk, int stage, const Shape& sh) {
  hipLaunchKernelGGL(kb_bfield<S>, dim3(tiles), dim3(sh.block), 0, s, boxes, prefix, nbox, clk, stage, sh.kchunk);
}

template <typename S>
void launch_bfield_cfl_batch(hipStream_t s, const Box<S>* boxes, const unsigned* prefix, int nbox, unsigned tiles,
                             BoxClock* clk, int* err, const Shape& sh, int npas) {
  // as launch_bfield_cfl: with passive scalars every variable through NQ = 10
  if (npas == 0) hipLaunchKernelGGL((kb_bfield_cfl<S, 8>), dim3(tiles), dim3(sh.block), 0, s, boxes, prefix, nbox, clk, sh.kchunk, err);
  else hipLaunchKernelGGL((kb_bfield_cfl<S, 10>), dim3(tiles), dim3(sh.block), 0, s, boxes, prefix, nbox, clk, sh.kchunk, err);
}

template <typename S>
void launch_cfl_batch(hipStream_t s, const Box<S>* boxes, const unsigned* prefix, int nbox, unsigned tiles,
                      BoxClock* clk, int* err, const Shape& sh, int npas) {
  if (npas == 0) hipLaunchKernelGGL((kb_cfl<S, 8>), dim3(tiles), dim3(sh.block), 0, s, boxes, prefix, nbox, sh.kchunk, clk, err);
  else hipLaunchKernelGGL((kb_cfl<S, 10>), dim3(tiles), dim3(sh.block), 0, s, boxes, prefix, nbox, sh.kchunk, clk, err);
}

#if GM_PART(1)
void launch_clock_begin(hipStream_t s, BoxClock* clk, const int* slots, int nbox) {
  hipLaunchKernelGGL(kb_clock_begin, dim3(ceil_div(nbox, 64)), dim3(64), 0, s, clk, slots, nbox);
}

void launch_clock_dt(hipStream_t s, BoxClock* clk, const int* slots, int nbox, int tuning) {
  hipLaunchKernelGGL(kb_clock_dt, dim3(ceil_div(nbox, 64)), dim3(64), 0, s, clk, slots, nbox, tuning);
}
#endif  // GM_PART(1)

template <typename S>
void launch_ghost(hipStream_t s, const Geom& g, S* U, int axis, int side, int kind, int layers) {
  const int n0 = axis == 0 ? g.ny : g.nx;
  const int n1 = axis == 2 ? g.ny : g.nz;
  const long long n = (long long)(n0 + 4) * (n1 + 4) * layers;
  hipLaunchKernelGGL(k_ghost<S>, dim3(grid_for(n)), dim3(256), 0, s, g, U, axis, side, kind, layers);
}

template <typename S>
void launch_aos_to_soa(hipStream_t s, long long ncell, const double* A, S* D, int nq) {
  hipLaunchKernelGGL(k_aos_to_soa<S>, dim3(grid_for(ncell)), dim3(256), 0, s, ncell, nq, A, D);
}

template <typename S>
void launch_soa_to_aos(hipStream_t s, long long ncell, const S* D, double* A, int nq) {
  hipLaunchKernelGGL(k_soa_to_aos<S>, dim3(grid_for(ncell)), dim3(256), 0, s, ncell, nq, D, A);
}

template <typename S>
void launch_prim_all(hipStream_t s, long long ncell, const Phys& ph, const S* U, double* Paos, double* T, int nq,
                     const Geom& g, const double* pc) {
  hipLaunchKernelGGL(k_prim_all<S>, dim3(grid_for(ncell)), dim3(256), 0, s, ncell, nq, ph, U, Paos, T, g, pc);
}

template <typename S>
void launch_divb(hipStream_t s, const Geom& g, const S* U, double dx, double dy, double dz, double* D) {
  const long long n = (long long)g.nx * g.ny * g.nz;
  hipLaunchKernelGGL(k_divb<S>, dim3(grid_for(n)), dim3(256), 0, s, g, U, dx, dy, dz, D);
}

// Explicit instantiations, split into parts so the library builds as
// parallel compilations of this one file: GM_KERNELS_PART 1-4 the FP64
// kernels (1: plane tiles and the other single-box kernels, 2: z pencils,
// 3: slabs, 4: batches), 5-8 the same for FP32; 0 (the default) all of
// them.  Part 1 also holds the non-template kernels and host functions.
#define GM_INST_XY(S)                                                                                            \
  template void launch_xy<S>(hipStream_t, const Box<S>&, double, int, int, int, int*);
#define GM_INST_Z(S)                                                                                             \
  template void launch_z<S>(hipStream_t, const Box<S>&, double, int, int, int, int*, const Shape&);
#define GM_INST_SLAB(S)                                                                                          \
  template void launch_slab<S>(hipStream_t, const Box<S>&, double, int, int, int, int*, int);
#define GM_INST_REST(S)                                                                                          \
  template void launch_bfield<S>(hipStream_t, const Box<S>&, double, int, const Shape&);                          \
  template void launch_cfl<S>(hipStream_t, const Box<S>&, unsigned long long*, int*, const Shape&);               \
  template void launch_bfield_cfl<S>(hipStream_t, const Box<S>&, double, unsigned long long*, int*, const Shape&); \
  template void launch_sources<S>(hipStream_t, const Box<S>&, double, int, const Shape&);                         \
  template void launch_cooling<S>(hipStream_t, const Box<S>&, const Cooling&, double);                            \
  template void launch_viscous<S>(hipStream_t, const Box<S>&, double);                                            \
  template void launch_tc_dtmin<S>(hipStream_t, const Box<S>&, const Conduction&, unsigned long long*);           \
  template void launch_tc_substep<S>(hipStream_t, const Box<S>&, const Conduction&, double);                      \
  template void launch_ghost<S>(hipStream_t, const Geom&, S*, int, int, int, int);                               \
  template void launch_aos_to_soa<S>(hipStream_t, long long, const double*, S*, int);                             \
  template void launch_soa_to_aos<S>(hipStream_t, long long, const S*, double*, int);                             \
  template void launch_prim_all<S>(hipStream_t, long long, const Phys&, const S*, double*, double*, int,         \
                                   const Geom&, const double*);                                               \
  template void launch_divb<S>(hipStream_t, const Geom&, const S*, double, double, double, double*);
#define GM_INST_BATCH(S)                                                                                         \
  template void launch_sources_batch<S>(hipStream_t, const Box<S>*, const unsigned*, int, unsigned,               \
                                        const BoxClock*, int, const Shape&);                                      \
  template void launch_viscous_batch<S>(hipStream_t, const Box<S>*, const unsigned*, int, unsigned,               \
                                        const BoxClock*, const Shape&);                                           \
  template void launch_xy_batch<S>(hipStream_t, const Box<S>*, const unsigned*, int, unsigned, const BoxClock*,   \
                                   int, int, int, int*, int);                                                     \
  template void launch_z_batch<S>(hipStream_t, const Box<S>*, const unsigned*, int, unsigned, const BoxClock*,    \
                                  int, int, int, int*, const Shape&, int);                                        \
  template void launch_bfield_batch<S>(hipStream_t, const Box<S>*, const unsigned*, int, unsigned,                \
                                       const BoxClock*, int, const Shape&);                                       \
  template void launch_slab_batch<S>(hipStream_t, const Box<S>*, const unsigned*, int, unsigned, const BoxClock*,   \
                                     int, int, int, int*, const Shape&, int);                                     \
  template void launch_bfield_cfl_batch<S>(hipStream_t, const Box<S>*, const unsigned*, int, unsigned, BoxClock*,  \
                                           int*, const Shape&, int);                                              \
  template void launch_cfl_batch<S>(hipStream_t, const Box<S>*, const unsigned*, int, unsigned, BoxClock*, int*,  \
                                    const Shape&, int);
#if GM_PART(1)
GM_INST_XY(double)
GM_INST_REST(double)
#endif
#if GM_PART(2)
GM_INST_Z(double)
#endif
#if GM_PART(3)
GM_INST_SLAB(double)
#endif
#if GM_PART(4)
GM_INST_BATCH(double)
#endif
#if GM_PART(5)
GM_INST_XY(float)
GM_INST_REST(float)
#endif
#if GM_PART(6)
GM_INST_Z(float)
#endif
#if GM_PART(7)
GM_INST_SLAB(float)
#endif
#if GM_PART(8)
GM_INST_BATCH(float)
#endif
#undef GM_INST_XY
#undef GM_INST_Z
#undef GM_INST_SLAB
#undef GM_INST_REST
#undef GM_INST_BATCH

}  // namespace gm

// Test plugin: the oracle's user boundary hook as a device kernel -- an
// inflow jet through the left x face (ghost layers 1-order..0 where
// (y-1/2)^2 + (z-1/2)^2 < 1/16) with vx = 1 + time, and an obstacle of
// interior cells within 0.15 of (0.7, 0.5, 0.5) reset to rest at density
// 5, pressure 1, keeping B -- the same arithmetic, in the same order, as
// the Fortran impose_user_bc the reference is built with for the parity
// oracle (oracle/ref/user_mod.f90).  Every cell is written from its own
// values or constants only, so the parallel loop equals the serial one.
#include "guacho_amd_user.h"

extern "C" __global__ void __launch_bounds__(256) gm_test_jet(gm_user_args a) {
  int i, j, k;
  if (!gm_user_any_cell(a, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
  const int order = a.stage;
  const int64_t c = gm_user_index(a, i, j, k);
  if (i >= 1 - order && i <= 0) {
    const double yc = ((double)j - 0.5) * a.dy - 0.5;
    const double zc = ((double)k - 0.5) * a.dz - 0.5;
    if (yc * yc + zc * zc < 0.0625) {
      const double rho = 2.0, vx = 1.0 + a.time, pr = 1.0, bx = 0.5;
      gm_user_set(a, c, 0, rho);
      gm_user_set(a, c, 1, rho * vx);
      gm_user_set(a, c, 2, 0.0);
      gm_user_set(a, c, 3, 0.0);
      gm_user_set(a, c, 4, ((0.5 * rho) * (vx * vx) + a.cv * pr) + 0.5 * (bx * bx));
      gm_user_set(a, c, 5, bx);
      gm_user_set(a, c, 6, 0.0);
      gm_user_set(a, c, 7, 0.0);
    }
  }
  if (i >= 1 && i <= a.nx && j >= 1 && j <= a.ny && k >= 1 && k <= a.nz) {
    const double xc = ((double)i - 0.5) * a.dx - 0.7;
    const double yc = ((double)j - 0.5) * a.dy - 0.5;
    const double zc = ((double)k - 0.5) * a.dz - 0.5;
    if ((xc * xc + yc * yc) + zc * zc < 0.0225) {
      const double bx = gm_user_get(a, c, 5), by = gm_user_get(a, c, 6), bz = gm_user_get(a, c, 7);
      gm_user_set(a, c, 0, 5.0);
      gm_user_set(a, c, 1, 0.0);
      gm_user_set(a, c, 2, 0.0);
      gm_user_set(a, c, 3, 0.0);
      gm_user_set(a, c, 4, a.cv * 1.0 + 0.5 * ((bx * bx + by * by) + bz * bz));
    }
  }
}

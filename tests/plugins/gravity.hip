// Test plugin: the oracle's user source term as a device kernel -- uniform
// gravity along -y plus a softened point mass at the box centre, the same
// arithmetic, in the same order, as the Fortran get_user_source_terms the
// reference is built with for the oracle (oracle/ref/user_mod.f90).
// Built by tests/plugins/Makefile into gravity.hsaco (gfx950 code object).
#include "guacho_amd_user.h"

extern "C" __global__ void __launch_bounds__(256) gm_test_gravity(gm_user_args a) {
  int i, j, k;
  if (!gm_user_cell(a, (int64_t)blockIdx.x * blockDim.x + threadIdx.x, i, j, k)) return;
  const int64_t c = gm_user_index(a, i, j, k);
  const int64_t N = a.ncell;
  double u[GM_USER_MAXVAR], pp[GM_USER_MAXVAR];
  gm_user_load(a, c, u);
  gm_user_u2prim(a, u, pp);
  const double g0 = 0.5, gm = 0.02, soft = 0.01;
  const double xc = ((double)i - 0.5) * a.dx - 0.5;
  const double yc = ((double)j - 0.5) * a.dy - 0.5;
  const double zc = ((double)k - 0.5) * a.dz - 0.5;
  const double r2 = ((xc * xc + yc * yc) + zc * zc) + soft;
  const double r3 = r2 * __builtin_sqrt(r2);
  double* s = a.s;
  s[2 * N + c] = s[2 * N + c] - pp[0] * g0;
  s[4 * N + c] = s[4 * N + c] - (pp[0] * pp[2]) * g0;
  s[1 * N + c] = s[1 * N + c] - ((pp[0] * gm) * xc) / r3;
  s[2 * N + c] = s[2 * N + c] - ((pp[0] * gm) * yc) / r3;
  s[3 * N + c] = s[3 * N + c] - ((pp[0] * gm) * zc) / r3;
  s[4 * N + c] = s[4 * N + c] - ((pp[0] * gm) * ((pp[1] * xc + pp[2] * yc) + pp[3] * zc)) / r3;
}

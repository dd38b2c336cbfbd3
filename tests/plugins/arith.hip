// Test kernel: gm_div / gm_sqrt (include/guacho_amd_user.h) beside the
// compiler's IEEE division and square root, on the same operands.
#include "guacho_amd_user.h"

extern "C" __global__ void gm_test_div_sqrt(const double* a, const double* b, double* out, long long n) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double x = a[t], y = b[t];
  out[4 * t + 0] = gm_div(x, y);
  out[4 * t + 1] = x / y;
  out[4 * t + 2] = gm_sqrt(x);
  out[4 * t + 3] = __builtin_sqrt(x);
}

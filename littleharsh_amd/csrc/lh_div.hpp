// lh_div.hpp -- the divergence monitor's per-column maximum (device functions only).
//
// error (littleharsh.f90:323-352) reports max |div u| over the p grid.  One column's maximum
// is shared by the kernel behind lh_max_div and the one behind lh_max_div_noalloc, so the two
// entry points return the same double for the same state.
#pragma once
#include <hip/hip_runtime.h>

namespace lh {

// max that keeps NaN: a NaN operand wins over every number
__device__ __forceinline__ double nanmax(double e, double v) { return (v > e || v != v) ? v : e; }

// max over the p planes r = 0..nyp-1 of |div| in column c (state slot-major, [slot][col]):
// div = i kx u1 + i kz u3 + ddthetavi dthdyu(r+1) (u2(r+1) - u2(r)).  |div| is the square root
// of the squared sum, not hypot, which returns +inf for (inf, NaN): NaN stays NaN.
__device__ __forceinline__ double col_maxdiv(int c, const double2* u1, const double2* u2,
                                             const double2* u3, const double* dthdyu,
                                             const double* k1x, const double* k1z, double ddthetavi,
                                             int ncol, int mx, int nyp) {
  const int i = c % mx, k = c / mx;
  const double kx = k1x[i], kz = k1z[k];
  double e = 0.0;
  for (int r = 0; r < nyp; ++r) {
    const size_t ou = (size_t)(r + 1) * ncol + c;
    const double2 a = u1[ou], b = u3[ou], p = u2[ou], q = u2[ou - ncol];
    const double m = ddthetavi * dthdyu[r + 1];
    const double dx = (-kx * a.y + -kz * b.y) + m * (p.x - q.x);
    const double dy = (kx * a.x + kz * b.x) + m * (p.y - q.y);
    e = nanmax(e, sqrt(dx * dx + dy * dy));
  }
  return e;
}

}  // namespace lh

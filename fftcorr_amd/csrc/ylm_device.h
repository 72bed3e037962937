// Real spherical harmonics Y_lm for l in {0, 2, 4} evaluated per cell on the
// device (make_ylm, cc/correlate/spherical_harmonics.h:32-179), FP64, with the
// reference's operation order and no FMA contraction so the values agree bit
// for bit with the reference's CPU build.
//
//   ylm_eval(l, m, x, y, z, c, D) = D * c * P_lm(x, y, z) / r^l
// where for l = 0 the host passes c = coeff / sqrt(4 pi) and for l > 0
// c = sqrt(1/pi) * coeff (* sqrt(2) if m != 0), exactly as
// spherical_harmonics.h:51 and :79-82 form them.  D is the `mult` grid value
// (1.0 when the reference passes no mult).  m > 0 -> sqrt(2) Re Y_lm,
// m < 0 -> sqrt(2) Im Y_l|m|, signs per the reference's table.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace fc {

// Host: the constant c described above.
inline double ylm_coeff(int ell, int m, double coeff) {
  if (ell == 0) return coeff / std::sqrt(4.0 * M_PI);
  double isqpi = std::sqrt(1.0 / M_PI);
  isqpi *= coeff;
  if (m != 0) isqpi *= std::sqrt(2.0);
  return isqpi;
}

__device__ __forceinline__ double ylm_eval(int ell, int m, double x, double y, double z,
                                           double c, double D) {
#pragma clang fp contract(off)
  if (ell == 0) return D * c;
  const double tiny = 1e-20;
  const double x2 = x * x;
  const double y2 = y * y;
  const double z2 = z * z;
  const double ir2 = 1.0 / (x2 + y2 + z2 + tiny);
  const double b = D * c;  // D * R * isqpi with R = 1 (exponent 0)
  if (ell == 2) {
    switch (m) {
      case 2:
        return b * sqrt(15.0 / 32.0) * (x2 - y2) * ir2;
      case 1:
        return b * sqrt(15.0 / 8.0) * x * z * ir2;
      case 0:
        return b * sqrt(5.0 / 16.0) * (2.0 * z2 - x2 - y2) * ir2;
      case -1:
        return b * sqrt(15.0 / 8.0) * y * z * ir2;
      default:
        return b * sqrt(15.0 / 8.0) * x * y * ir2;
    }
  }
  const double y3 = y2 * y;
  const double y4 = y3 * y;
  switch (m) {
    case 4:
      return b * 3.0 / 16.0 * sqrt(35.0 / 2.0) * (x2 * x2 - 6.0 * x2 * y2 + y4) * ir2 * ir2;
    case 3:
      return b * 3.0 / 8.0 * sqrt(35.0) * (x2 - 3.0 * y2) * z * x * ir2 * ir2;
    case 2:
      return b * 3.0 / 8.0 * sqrt(5.0 / 2.0) * (6.0 * z2 * (x2 - y2) - x2 * x2 + y4) * ir2 *
             ir2;
    case 1:
      return b * 3.0 / 8.0 * sqrt(5.0) * 3.0 * (4.0 / 3.0 * z2 - x2 - y2) * x * z * ir2 * ir2;
    case 0: {
      const double z4 = z2 * z * z;
      return b * 3.0 / 16.0 * 8.0 *
             (z4 - 3.0 * z2 * (x2 + y2) + 3.0 / 8.0 * (x2 * x2 + 2.0 * x2 * y2 + y4)) * ir2 *
             ir2;
    }
    case -1:
      return b * 3.0 / 8.0 * sqrt(5.0) * 3.0 * (4.0 / 3.0 * z2 - x2 - y2) * y * z * ir2 * ir2;
    case -2:
      return b * 3.0 / 8.0 * sqrt(5.0 / 2.0) * (2.0 * x * y * (6.0 * z2 - x2 - y2)) * ir2 * ir2;
    case -3:
      return b * 3.0 / 8.0 * sqrt(35.0) * (3.0 * x2 * y - y3) * z * ir2 * ir2;
    default:
      return b * 3.0 / 16.0 * sqrt(35.0 / 2.0) * (4.0 * x * (x2 * y - y3)) * ir2 * ir2;
  }
}

// |(x, y, z)| as the reference's subgrid setup computes it
// (correlator.h:133, :185): sqrt((x*x + y*y) + z*z), no contraction.
__device__ __forceinline__ double norm3(double x, double y, double z) {
#pragma clang fp contract(off)
  return sqrt(x * x + y * y + z * z);
}

// HistogramList bin (histogram_list.h:57): floor((v - minval) / binsize).
__device__ __forceinline__ int hist_bin(double v, double minval, double binsize, int nbins) {
#pragma clang fp contract(off)
  const double f = floor((v - minval) / binsize);
  if (!(f >= 0.0) || f >= (double)nbins) return -1;
  return (int)f;
}

}  // namespace fc

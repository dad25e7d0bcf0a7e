// pf_math.h -- FP64 reciprocal and natural logarithm for the fused sweeps.
//
// The fused sweeps evaluate two logarithms and about a dozen divisions per
// cell; the library forms cost ~100 VALU instructions per log and ~10 per
// division on gfx950.  These are shorter forms with an error of about one
// ulp.  Written __host__ __device__ so that the host build can check the
// algorithm's accuracy against a long-double logarithm (the device form
// differs only in the reciprocal's starting estimate).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace pf {

// 1/x: the hardware estimate (device) refined by two Newton steps
// r <- r + r (1 - x r).  For x = 0, +-inf or NaN the estimate itself is
// returned (inf, 0, NaN), as a division would give.
__host__ __device__ __forceinline__ double rcp_refined(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double r0 = __builtin_amdgcn_rcp(x);
#else
  const double r0 = 1.0 / x;
#endif
  double e = fma(-x, r0, 1.0);
  double r = fma(r0, e, r0);
  e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  return r != r ? r0 : r;
}

// 1/x with one Newton step: measured on gfx950 the estimate is good to about
// 2^-24 and one step leaves at most 11 ulp (2.4e-15 relative; two steps are
// correctly rounded).  For the stencil's divisions, whose operands already
// carry rounding noise far above that.
__host__ __device__ __forceinline__ double rcp_fast(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double r0 = __builtin_amdgcn_rcp(x);
#else
  const double r0 = 1.0 / x;
#endif
  const double r = fma(r0, fma(-x, r0, 1.0), r0);
  return r != r ? r0 : r;
}

// ln 2 split so that e * kLn2Hi is exact for |e| < 2^21 (kLn2Hi keeps 32
// significant bits; kLn2Lo = ln 2 - kLn2Hi rounded).
constexpr double kLn2Hi = 0x1.62e42fee00000p-1;
constexpr double kLn2Lo = 0x1.a39ef35793c76p-33;

// Natural logarithm.  x = 2^e m with m in [sqrt(1/2), sqrt(2)), f = m - 1,
// s = f/(2+f); then log m = 2 atanh(s) = f - (f^2/2 - s (f^2/2 + R)) with
// R = sum_{k>=1} 2 s^(2k)/(2k+1), the atanh Taylor series, which is exact
// algebra (2s = f - s f) that keeps the exact f as the leading term.  With
// |s| <= 0.1716, ten terms leave a truncation below 1e-17 relative.  Then
// log x = e kLn2Hi + (log m + e kLn2Lo).  x = 0 gives -inf, x < 0 and NaN
// give NaN, +inf gives +inf, as log() does.
__host__ __device__ __forceinline__ double log_fast(double x) {
  int e;
  double m = frexp(x, &e);  // [0.5, 1) for finite x != 0
  if (m < 0.70710678118654752440) {
    m = m + m;
    e -= 1;
  }
  const double f = m - 1.0;
  const double s = f * rcp_refined(2.0 + f);
  const double z = s * s;
  double R = 2.0 / 21.0;
  R = fma(R, z, 2.0 / 19.0);
  R = fma(R, z, 2.0 / 17.0);
  R = fma(R, z, 2.0 / 15.0);
  R = fma(R, z, 2.0 / 13.0);
  R = fma(R, z, 2.0 / 11.0);
  R = fma(R, z, 2.0 / 9.0);
  R = fma(R, z, 2.0 / 7.0);
  R = fma(R, z, 2.0 / 5.0);
  R = fma(R, z, 2.0 / 3.0);
  R = R * z;
  const double hf = 0.5 * f * f;
  const double lm = f - fma(-s, hf + R, hf);
  const double de = (double)e;
  double r = fma(de, kLn2Hi, fma(de, kLn2Lo, lm));
  if (!(x > 0.0)) r = x == 0.0 ? -__builtin_huge_val() : __builtin_nan("");
  if (x == __builtin_huge_val()) r = x;
  return r;
}

}  // namespace pf

// pf_math.h -- FP64 reciprocal and natural logarithm for the fused sweeps.
//
// The fused sweeps evaluate two logarithms and about a dozen divisions per
// cell; the library forms cost ~100 VALU instructions per log and ~10 per
// division on gfx950.  These are shorter forms with an error of about one
// ulp.  Written __host__ __device__ so that a host build checks the
// logarithm's accuracy against a long-double logarithm (the same expression
// runs on the device).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "pf_log_table.h"

namespace pf {

// 1/x with one Newton step: measured on gfx950 the estimate is good to about
// 2^-24 and one step leaves at most 11 ulp (2.4e-15 relative; two steps are
// correctly rounded).  For the stencil's divisions, whose operands already
// carry rounding noise far above that.  For x = 0, +-inf or NaN the estimate
// itself is returned (inf, 0, NaN), as a division would give.
__host__ __device__ __forceinline__ double rcp_fast(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double r0 = __builtin_amdgcn_rcp(x);
#else
  const double r0 = 1.0 / x;
#endif
  const double r = fma(r0, fma(-x, r0, 1.0), r0);
  return r != r ? r0 : r;
}

// 1/x with one Newton step and no special-case guard, for a denominator that
// is positive and finite whenever the inputs are finite (the node's squared
// gradient plus a positive floor): 0 * inf never arises there, and
// non-finite inputs give NaN, as the division does in that expression.
__host__ __device__ __forceinline__ double rcp_pos(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  const double r0 = __builtin_amdgcn_rcp(x);
#else
  const double r0 = 1.0 / x;
#endif
  return fma(r0, fma(-x, r0, 1.0), r0);
}

// ln 2 split so that e * kLn2Hi is exact for |e| < 2^21 (kLn2Hi keeps 32
// significant bits; kLn2Lo = ln 2 - kLn2Hi rounded).
constexpr double kLn2Hi = 0x1.62e42fee00000p-1;
constexpr double kLn2Lo = 0x1.a39ef35793c76p-33;

// One entry of the logarithm's table (pf_log_table.h): inv = 1/c_j rounded,
// L = -ln(inv) rounded, for c_j = 1 + j/128.  16-byte aligned so that an
// entry is one 128-bit LDS read.
struct alignas(16) LogEnt {
  double inv, L;
};

// Natural logarithm by table lookup (P. T. P. Tang, "Table-driven
// implementation of the logarithm function in IEEE floating-point
// arithmetic", ACM TOMS 16(4), 1990, in the variant with a reciprocal
// table):  x = 2^e m with m in [sqrt(1/2), sqrt(2)); j = trunc(128 (m - 1))
// picks c_j = 1 + j/128 between 1 and m (so ln m, ln c_j and r below share a
// sign and never cancel); r = m inv_j - 1 in one FMA rounding, |r| < 2^-6.5,
// and ln m = L_j + ln(1 + r) exactly, whatever the rounding of inv_j;
// ln(1 + r) is its Taylor series to r^8 (the next term is below 2^-55 |r|).
// Then ln x = (e kLn2Hi + L_j) + (e kLn2Lo + ln(1 + r)).  About one ulp;
// x = 0 gives -inf, x < 0 and NaN give NaN, +inf gives +inf, as log() does.
// `tab` holds the kLogTabN entries of pf_log_table.h (an LDS copy on the
// device).
__host__ __device__ __forceinline__ double log_tab(double x, const LogEnt *tab) {
  int e;
  double m = frexp(x, &e);  // [0.5, 1) for finite x != 0
  if (m < 0.70710678118654752440) {
    m = m + m;
    e -= 1;
  }
  // the clamp only matters for x <= 0, inf and NaN, whose result is
  // replaced below; it keeps the index (and the conversion) in range
  const double u = fmin(fmax((m - 1.0) * 128.0, (double)kLogTabLo),
                        (double)(kLogTabLo + kLogTabN - 1));
  const LogEnt t = tab[(int)u - kLogTabLo];
  const double r = fma(m, t.inv, -1.0);
  double q = fma(r, -1.0 / 8.0, 1.0 / 7.0);
  q = fma(r, q, -1.0 / 6.0);
  q = fma(r, q, 1.0 / 5.0);
  q = fma(r, q, -1.0 / 4.0);
  q = fma(r, q, 1.0 / 3.0);
  q = fma(r, q, -0.5);
  const double p = fma(r * r, q, r);
  const double de = (double)e;
  double res = fma(de, kLn2Hi, fma(de, kLn2Lo, t.L + p));
  if (!(x > 0.0)) res = x == 0.0 ? -__builtin_huge_val() : __builtin_nan("");
  if (x == __builtin_huge_val()) res = x;
  return res;
}

// ln|A/B| for a ratio close to 1, FP64.  With t = (A - B)/(A + B) the ratio
// is (1 + t)/(1 - t) and ln(A/B) = 2 atanh(t) = 2t (1 + t^2/3 + ... + t^12/13)
// + O(t^15), truncated below 2^-56 relative for |t| <= 1/16 (A/B in
// [15/17, 17/15]).  A - B is formed from the operands, so t keeps its full
// relative precision however close the ratio is to 1, where the logarithm
// of a rounded ratio carries an absolute error of about one ulp of 1.
// Every other lane value (a ratio further from 1, zero, negative or not
// finite, and t = NaN from A + B = 0) takes log_tab(|A/B|), the
// reference's ln|ratio|; on the device only waves that hold such a value
// evaluate it.  Which form a lane uses depends on its own t alone.
__host__ __device__ __forceinline__ double ln_ratio_near1(double A, double B,
                                                          const LogEnt *tab) {
  const double t = (A - B) * rcp_pos(A + B);
  const double t2 = t * t;
  double p = fma(t2, 2.0 / 13.0, 2.0 / 11.0);
  p = fma(t2, p, 2.0 / 9.0);
  p = fma(t2, p, 2.0 / 7.0);
  p = fma(t2, p, 2.0 / 5.0);
  p = fma(t2, p, 2.0 / 3.0);
  double r = t * fma(t2, p, 2.0);
  const bool far = !(fabs(t) <= 0.0625);
#if defined(__HIP_DEVICE_COMPILE__)
  if (__builtin_amdgcn_ballot_w64(far) != 0) {
    const double lg = log_tab(fabs(A * rcp_fast(B)), tab);
    r = far ? lg : r;
  }
#else
  if (far) r = log_tab(fabs(A * rcp_fast(B)), tab);
#endif
  return r;
}

// ln|1 + x| in FP32 for an x formed without cancellation: a 6-term series
// x (1 - x (1/2 - x (1/3 - x (1/4 - x (1/5 - x/6))))) below |x| = 1/16
// (truncation < 1e-8 relative), the hardware logarithm of |1 + x| above --
// evaluated only by waves that hold such an x (most waves hold none); which
// form a lane uses depends on x alone.
__device__ __forceinline__ float ln1p_abs_f32(float x) {
  float r = x * fmaf(-x, fmaf(-x, fmaf(-x, fmaf(-x, fmaf(-x, 1.0f / 6.0f, 0.2f), 0.25f),
                                      1.0f / 3.0f),
                              0.5f),
                     1.0f);
  const bool big = !(fabsf(x) < 0.0625f);  // NaN included
  if (__builtin_amdgcn_ballot_w64(big) != 0) {
    // (an instruction with side effects, so that the compiler keeps the
    // branch instead of evaluating the logarithm in every wave and
    // selecting)
    __builtin_amdgcn_sched_barrier(0);
    const float lg = __logf(fabsf(1.0f + x));
    r = big ? lg : r;
  }
  return r;
}

}  // namespace pf

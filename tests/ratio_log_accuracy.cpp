// Host check of pf::ln_ratio_near1 (csrc/pf_math.h), the FP64 logarithm of
// the phase equation's free-energy ratio, against the long-double logarithm
// of the exact quotient of its two double operands.  The device code runs the
// same expression; on the host 1/x is a correctly rounded division, on the
// device a reciprocal with one Newton step (<= 11 ulp), so the device error
// is that much larger in the series range.
// Prints "near_max_ulp far_max_abs_ulp1 n" and exits non-zero when a special
// value differs from log|A/B|.
#include <cfloat>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>

#include "pf_math.h"

static double ulp_of(double v) {
  const double a = std::fabs(v);
  if (a < DBL_MIN) return std::ldexp(1.0, -1074);
  int e;
  std::frexp(a, &e);
  return std::ldexp(1.0, e - 53);
}

int main(int argc, char **argv) {
  const long n = argc > 1 ? std::atol(argv[1]) : 1000000;
  pf::LogEnt tab[pf::kLogTabN];
  for (int t = 0; t < pf::kLogTabN; ++t) tab[t] = pf::LogEnt{pf::kLogTab[t][0], pf::kLogTab[t][1]};
  uint64_t s = 0x2545f4914f6cdd1dull;
  auto next = [&]() {
    s ^= s << 13; s ^= s >> 7; s ^= s << 17;
    return s;
  };
  auto unit = [&]() { return (double)(next() >> 11) * 0x1p-53; };
  // near: ratios with |t| <= 1/16, error in ulps of the result;
  // far: the fallback, error in ulps of 1 (the logarithm of a rounded ratio)
  double near_max = 0.0, far_max = 0.0;
  long cnt = 0;
  auto check = [&](double A, double B) {
    const double got = pf::ln_ratio_near1(A, B, tab);
    const long double t = ((long double)A - B) / ((long double)A + B);
    if (fabsl(t) <= 0.0625L) {
      // ln(1 + (A-B)/B): A - B is exact, so the reference keeps its relative
      // accuracy next to 1, where logl of a rounded quotient would not
      const long double ref = log1pl(((long double)A - B) / B);
      const double u = (double)(fabsl((long double)got - ref) / ulp_of((double)ref));
      if (u > near_max) near_max = u;
    } else {
      const long double ref = logl(fabsl((long double)A / (long double)B));
      const double u = (double)(fabsl((long double)got - ref) / ulp_of(1.0));
      if (u > far_max) far_max = u;
    }
    ++cnt;
  };
  // the phase equation's operands: A = K (den - c), B = den - Ke c with
  // den = Ke h + 1 - h, h in [0,1], c up to 0.3, K = (1-Cs_eq)/(1-Cl_eq)
  const double Ke = (double)0.14f;
  for (long i = 0; i < n; ++i) {
    switch (i % 3) {
      case 0: {
        const double h = unit(), c = 0.3 * unit(), K = 1.0 + 0.05 * (unit() - 0.5);
        const double den = Ke * h + (1.0 - h);
        check(K * (den - c), den - Ke * c);
        break;
      }
      case 1: {  // ratios 1 +- 2^-k, k = 4 .. 52, both orders
        const double B = 0.5 + unit();
        const double d = std::ldexp(unit(), -4 - (int)(next() % 49));
        check(B * (1.0 + d), B);
        check(B * (1.0 - d), B);
        break;
      }
      default: {  // any ratio in [1/64, 64], either sign
        const double B = 0.25 + unit();
        const double r = std::ldexp(1.0 + unit(), (int)(next() % 13) - 6);
        check((next() & 1) ? B * r : -B * r, B);
      }
    }
  }
  int bad = 0;
  const double sp[][2] = {{0.0, 1.0}, {1.0, 0.0}, {-1.0, 1.0}, {1.0, -1.0}, {NAN, 1.0},
                          {1.0, NAN}, {INFINITY, 1.0}, {1.0, 1.0}, {0.5, 0.5}};
  for (const auto &ab : sp) {
    const double got = pf::ln_ratio_near1(ab[0], ab[1], tab);
    const double ref = std::log(std::fabs(ab[0] / ab[1]));
    const bool same = (std::isnan(got) && std::isnan(ref)) || got == ref;
    if (!same) {
      std::printf("special %a/%a: %a vs %a\n", ab[0], ab[1], got, ref);
      bad = 1;
    }
  }
  std::printf("%.4f %.4f %ld\n", near_max, far_max, cnt);
  return bad;
}

// Host check of the FP64 logarithm of the fused sweeps (pf::log_tab in
// csrc/pf_math.h) against the long-double logarithm.  The device code runs
// the same expression (frexp and fma are exact IEEE operations on both), so
// this bounds the device results too.  Prints "max_ulp mean_ulp n" and exits
// non-zero when a special value differs from log().
#include <cfloat>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "pf_math.h"

static double ulp_of(double v) {
  const double a = std::fabs(v);
  if (a < DBL_MIN) return std::ldexp(1.0, -1074);
  int e;
  std::frexp(a, &e);
  return std::ldexp(1.0, e - 53);
}

int main(int argc, char **argv) {
  const long n = argc > 1 ? std::atol(argv[1]) : 2000000;
  pf::LogEnt tab[pf::kLogTabN];
  for (int t = 0; t < pf::kLogTabN; ++t) tab[t] = pf::LogEnt{pf::kLogTab[t][0], pf::kLogTab[t][1]};
  uint64_t s = 0x9e3779b97f4a7c15ull;
  auto next = [&]() {
    s ^= s << 13; s ^= s >> 7; s ^= s << 17;
    return s;
  };
  double maxu = 0.0, sumu = 0.0, worst = 0.0;
  long cnt = 0;
  auto check = [&](double x) {
    const double got = pf::log_tab(x, tab);
    const long double ref = logl((long double)x);
    const double u = (double)fabsl((long double)got - ref) / ulp_of((double)ref);
    if (u > maxu) { maxu = u; worst = x; }
    sumu += u;
    ++cnt;
  };
  for (long t = 0; t < n; ++t) {
    const uint64_t b = next();
    switch (t % 4) {
      case 0: {  // any positive normal or subnormal double
        uint64_t bits = b & 0x7fefffffffffffffull;
        double x;
        std::memcpy(&x, &bits, 8);
        if (x > 0.0) check(x);
        break;
      }
      case 1:  // the sweeps' range: (0, 8)
        check(std::ldexp((double)(b >> 11) + 1.0, -50));
        break;
      case 2: {  // close to 1, where ln x is tiny and must keep its relative accuracy
        const double d = std::ldexp((double)(b >> 11), -53 - (int)(b & 31));
        check(1.0 + d);
        check(1.0 - d);
        break;
      }
      default: {  // each table interval, at its edges and inside
        const int j = (int)(b % pf::kLogTabN) + pf::kLogTabLo;
        const double c = 1.0 + j / 128.0;
        check(c);
        check(std::nextafter(c, 0.0));
        check(std::nextafter(c, 2.0));
        check(c + std::ldexp((double)(b >> 40), -31) / 128.0);
      }
    }
  }
  int bad = 0;
  const double sp[] = {0.0, -0.0, -1.0, INFINITY, -INFINITY, NAN, DBL_MIN, DBL_TRUE_MIN, DBL_MAX};
  for (double x : sp) {
    const double got = pf::log_tab(x, tab), ref = std::log(x);
    const bool same = (std::isnan(got) && std::isnan(ref)) || got == ref ||
                      (std::isfinite(ref) && std::fabs(got - ref) <= 2 * ulp_of(ref));
    if (!same) {
      std::printf("special %a: %a vs %a\n", x, got, ref);
      bad = 1;
    }
  }
  std::printf("%.4f %.4f %ld worst %a\n", maxu, sumu / (double)cnt, cnt, worst);
  return bad;
}

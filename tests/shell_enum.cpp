// Host check of pf::shell_flat (csrc/pf_common.h), the enumeration of a
// grid's boundary shell that the decoration kernels run over: for every
// grid of a set of small and ragged extents, both shell widths and both
// x-face options, the 32-bit and the 64-bit index arithmetic give the same
// cell for every index, every cell comes once, and the set is exactly the
// shell (w planes at each z face; the w rows at each y face of the other
// planes; with xf, the w cells at each x face of their remaining rows).
// Prints "grids cells" and exits non-zero at the first mismatch.
#include <cstdio>
#include <vector>

#include "pf_common.h"

int main() {
  const int ext[] = {6, 7, 9, 16, 17, 33};
  long long grids = 0, cells = 0;
  for (int nx : ext)
    for (int ny : ext)
      for (int nz : ext)
        for (int w = 1; w <= 2; ++w)
          for (int xf = 0; xf <= 1; ++xf) {
            const long long n = pf::shell_count(nx, ny, nz, w, xf != 0);
            std::vector<int> seen((size_t)nx * ny * nz, 0);
            for (long long t = 0; t < n; ++t) {
              int i, j, k, i2, j2, k2;
              if (!pf::shell_flat(t, nx, ny, nz, w, i, j, k, xf != 0)) {
                std::printf("index %lld of %lld rejected (%d %d %d w%d xf%d)\n", t, n, nx, ny, nz, w, xf);
                return 1;
              }
              pf::shell_flat_in<long long>(t, nx, ny, nz, w, i2, j2, k2, xf != 0);
              if (i != i2 || j != j2 || k != k2 || i < 1 || i > nx || j < 1 || j > ny || k < 1 || k > nz) {
                std::printf("mismatch at %lld: (%d %d %d) vs (%d %d %d)\n", t, i, j, k, i2, j2, k2);
                return 1;
              }
              ++seen[pf::cidx(i, j, k, nx, ny)];
            }
            int i, j, k;
            if (pf::shell_flat(n, nx, ny, nz, w, i, j, k, xf != 0)) {
              std::printf("index past the shell accepted\n");
              return 1;
            }
            for (int kk = 1; kk <= nz; ++kk)
              for (int jj = 1; jj <= ny; ++jj)
                for (int ii = 1; ii <= nx; ++ii) {
                  const bool zf = kk <= w || kk > nz - w, yf = jj <= w || jj > ny - w;
                  const bool xfc = ii <= w || ii > nx - w;
                  const int want = (zf || yf || (xf && xfc)) ? 1 : 0;
                  if (seen[pf::cidx(ii, jj, kk, nx, ny)] != want) {
                    std::printf("cell (%d %d %d) seen %d times, want %d (%d %d %d w%d xf%d)\n", ii, jj,
                                kk, seen[pf::cidx(ii, jj, kk, nx, ny)], want, nx, ny, nz, w, xf);
                    return 1;
                  }
                }
            ++grids;
            cells += n;
          }
  std::printf("%lld %lld\n", grids, cells);
  return 0;
}

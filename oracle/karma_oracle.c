/* CPU restatement of the quantitative Karma model (karma/karma3d.f90), the
 * checker of the HIP Karma step.  Test infrastructure only: the product never
 * links it.
 *
 * Fields are padded cubes (n+2)^3, index i + (n+2)(j + (n+2)k) for
 * i, j, k in [0, n+1]; cells 1..n are the domain, 0 and n+1 the mirror
 * ghosts (karma3d.f90:24-25, 313-333).
 *
 * Pinned by tests/test_karma.py against the reference program compiled from
 * karma/karma3d.f90 (its Tecplot output, 8 significant digits), and by the
 * symmetry and closed-form checks listed there. */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "karma_oracle.h"

static inline size_t kidx(int n, int i, int j, int k) {
  const size_t s = (size_t)n + 2;
  return (size_t)i + s * ((size_t)j + s * (size_t)k);
}

static inline int clampc(int v, int n) { return v < 1 ? 1 : (v > n ? n : v); }

/* espelho (karma3d.f90:313-333): the x, y and z passes in that order compose
 * to ghost(i,j,k) = f(clamp(i), clamp(j), clamp(k)) with clamp into [1,n]. */
void pfo_karma_mirror(int n, double *f) {
  for (int k = 0; k <= n + 1; ++k)
    for (int j = 0; j <= n + 1; ++j)
      for (int i = 0; i <= n + 1; ++i) {
        if (i >= 1 && i <= n && j >= 1 && j <= n && k >= 1 && k <= n) continue;
        f[kidx(n, i, j, k)] = f[kidx(n, clampc(i, n), clampc(j, n), clampc(k, n))];
      }
}

/* Initial 1/8 seed at the corner (karma3d.f90:131-144), ghosts included. */
void pfo_karma_init(int n, double dx, double R0, double Omega, double *phi, double *U) {
  for (int k = 0; k <= n + 1; ++k)
    for (int j = 0; j <= n + 1; ++j)
      for (int i = 0; i <= n + 1; ++i) {
        const double x = (i - 0.5) * dx, y = (j - 0.5) * dx, z = (k - 0.5) * dx;
        const double r = sqrt(x * x + y * y + z * z);
        phi[kidx(n, i, j, k)] = -tanh((r - R0) / sqrt(2.0));
        U[kidx(n, i, j, k)] = -Omega;
      }
}

/* fluxn (karma3d.f90:36-47): normal component of the anisotropic flux J at a
 * face from the face gradient (g1 normal, g2 g3 tangential). */
static double fluxn(double g1, double g2, double g3, double eps4) {
  const double g2s = g1 * g1 + g2 * g2 + g3 * g3;
  const double a0 = 1.0 - 3.0 * eps4;
  if (g2s < 1e-20) return a0 * a0 * g1;
  const double q1 = g1 * g1, q2 = g2 * g2, q3 = g3 * g3;
  const double s4 = q1 * q1 + q2 * q2 + q3 * q3;
  const double as = a0 + 4.0 * eps4 * s4 / (g2s * g2s);
  return as * (as * g1 + 16.0 * eps4 * (q1 * g1 - g1 * s4 / g2s) / g2s);
}

/* ruido (karma3d.f90:50-64): uniform in [-1,1) per (i,j,k,it), the same
 * 64-bit mix as the main model's pseudo_random without the salt term. */
double pfo_karma_noise(int i, int j, int k, int it) {
  uint64_t h = (uint64_t)(int64_t)i * 2654435761ULL + (uint64_t)(int64_t)j * 2246822519ULL +
               (uint64_t)(int64_t)k * 3266489917ULL + (uint64_t)(int64_t)it * 668265263ULL + 1ULL;
  h ^= h >> 33;
  h *= 0xff51afd7ed558ccdULL;
  h ^= h >> 33;
  h *= 0xc4ceb9fe1a85ec53ULL;
  h ^= h >> 33;
  const uint64_t m63 = 0x7fffffffffffffffULL;
  return 2.0 * (double)(h & m63) / (double)m63 - 1.0;
}

/* Face gradient of phi at the face between cell c and its neighbour along
 * axis ax on side sd (+1/-1): the normal difference and the two tangential
 * centred differences averaged over the face's two cells
 * (karma3d.f90:158-175, 200-260). */
static void face_grad(int n, const double *phi, int i, int j, int k, int ax, int sd, double idx,
                      double g[3]) {
  int o[3] = {0, 0, 0};
  o[ax] = sd;
  const double pc = phi[kidx(n, i, j, k)];
  const double pn = phi[kidx(n, i + o[0], j + o[1], k + o[2])];
  g[ax] = sd > 0 ? (pn - pc) * idx : (pc - pn) * idx;
  for (int t = 0; t < 3; ++t) {
    if (t == ax) continue;
    int e[3] = {0, 0, 0};
    e[t] = 1;
    const double a = phi[kidx(n, i + e[0], j + e[1], k + e[2])];
    const double b = phi[kidx(n, i + o[0] + e[0], j + o[1] + e[1], k + o[2] + e[2])];
    const double c = phi[kidx(n, i - e[0], j - e[1], k - e[2])];
    const double d = phi[kidx(n, i + o[0] - e[0], j + o[1] - e[1], k + o[2] - e[2])];
    g[t] = 0.25 * idx * (a + b - c - d);
  }
}

/* One step (karma3d.f90:146-296): mirror phi and U, phase equation into
 * phin and dphi, mirror dphi, solute equation with the anti-trapping current
 * into Un, then phi = clamp(phin, -1, 1) and U = Un on the cells 1..n. */
void pfo_karma_step(int n, const pfo_karma_par *p, int it, double *phi, double *U) {
  const size_t tot = kidx(n, n + 1, n + 1, n + 1) + 1;
  double *phin = (double *)calloc(tot, sizeof(double));
  double *dphi = (double *)calloc(tot, sizeof(double));
  double *Un = (double *)calloc(tot, sizeof(double));
  const double idx = 1.0 / p->dx;
  const double D = p->a2 * p->lambda;
  const double at = 0.35355339059327373;
  const double k1 = 1.0 - p->k;
  pfo_karma_mirror(n, phi);
  pfo_karma_mirror(n, U);
  for (int k = 1; k <= n; ++k)
    for (int j = 1; j <= n; ++j)
      for (int i = 1; i <= n; ++i) {
        const size_t c = kidx(n, i, j, k);
        const double pp = phi[c];
        /* J at the faces x+, x-, y+, y-, z+, z-: normal component first, then
         * the tangential ones in x, y, z order */
        double J[6];
        for (int ax = 0; ax < 3; ++ax)
          for (int s = 0; s < 2; ++s) {
            double g[3];
            face_grad(n, phi, i, j, k, ax, s == 0 ? 1 : -1, idx, g);
            const int t1 = ax == 0 ? 1 : 0, t2 = ax == 2 ? 1 : 2;
            J[2 * ax + s] = fluxn(g[ax], g[t1], g[t2], p->eps4);
          }
        const double divJ = (J[0] - J[1] + J[2] - J[3] + J[4] - J[5]) * idx;
        const double g1 = 0.5 * idx * (phi[kidx(n, i + 1, j, k)] - phi[kidx(n, i - 1, j, k)]);
        const double g2 = 0.5 * idx * (phi[kidx(n, i, j + 1, k)] - phi[kidx(n, i, j - 1, k)]);
        const double g3 = 0.5 * idx * (phi[kidx(n, i, j, k + 1)] - phi[kidx(n, i, j, k - 1)]);
        const double gq = g1 * g1 + g2 * g2 + g3 * g3;
        const double a0 = 1.0 - 3.0 * p->eps4;
        double as2;
        if (gq > 1e-20) {
          const double s4 =
              ((g1 * g1) * (g1 * g1) + (g2 * g2) * (g2 * g2) + (g3 * g3) * (g3 * g3)) / (gq * gq);
          const double as = a0 + 4.0 * p->eps4 * s4;
          as2 = as * as;
        } else {
          as2 = a0 * a0;
        }
        const double om = 1.0 - pp * pp;
        double dp = (divJ + pp - pp * pp * pp - p->lambda * (om * om) * U[c]) / as2;
        if (p->Fnoise != 0.0) dp += p->Fnoise * pfo_karma_noise(i, j, k, it) * om / sqrt(p->dt);
        dphi[c] = dp;
        phin[c] = pp + p->dt * dp;
      }
  pfo_karma_mirror(n, dphi);
  for (int k = 1; k <= n; ++k)
    for (int j = 1; j <= n; ++j)
      for (int i = 1; i <= n; ++i) {
        const size_t c = kidx(n, i, j, k);
        const double pp = phi[c], uc = U[c], cc = (1.0 + k1 * uc) * dphi[c];
        double F6[6];
        for (int ax = 0; ax < 3; ++ax)
          for (int sd = 1; sd >= -1; sd -= 2) {
            int o[3] = {0, 0, 0};
            o[ax] = sd;
            const size_t nb = kidx(n, i + o[0], j + o[1], k + o[2]);
            double g[3];
            face_grad(n, phi, i, j, k, ax, sd, idx, g);
            const double gm = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
            const double qf = 0.25 * (2.0 - pp - phi[nb]);
            const double cf = 0.5 * (cc + (1.0 + k1 * U[nb]) * dphi[nb]);
            double F = sd > 0 ? D * qf * (U[nb] - uc) * idx : D * qf * (uc - U[nb]) * idx;
            if (gm > 1e-10) F += at * cf * g[ax] / gm;
            F6[2 * ax + (sd > 0 ? 0 : 1)] = F;
          }
        const double Au = 0.5 * (1.0 + p->k) - 0.5 * k1 * pp;
        const double rhs = (F6[0] - F6[1] + F6[2] - F6[3] + F6[4] - F6[5]) * idx +
                           0.5 * (1.0 + k1 * uc) * dphi[c];
        Un[c] = uc + p->dt * rhs / Au;
      }
  for (int k = 1; k <= n; ++k)
    for (int j = 1; j <= n; ++j)
      for (int i = 1; i <= n; ++i) {
        const size_t c = kidx(n, i, j, k);
        const double v = phin[c];
        phi[c] = v < -1.0 ? -1.0 : (v > 1.0 ? 1.0 : v);
        U[c] = Un[c];
      }
  free(phin);
  free(dphi);
  free(Un);
}

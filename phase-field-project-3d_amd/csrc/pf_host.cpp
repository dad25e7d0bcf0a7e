// pf_host.cpp -- host scalar physics of the C-ABI: the pieces of entrada,
// W0_E0, auxiliar and equacao_campo_fase that run on the host in the
// reference, evaluated in its operation order (REAL(4) where it uses REAL(4))
// so that they are bit-identical to it.
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>

#include "../../include/pf_capi.h"
#include "pf_common.h"
#include "pf_mass.h"

#pragma clang fp contract(off)

extern "C" {

double pf_pseudo_random(int32_t i, int32_t j, int32_t k, int32_t it, int32_t salt,
                        uint64_t seed) {
  return pf::pseudo_random(i, j, k, it, salt, seed);
}

double pf_ruido_suave(int32_t i, int32_t j, int32_t k, int32_t it, uint64_t seed) {
  return pf::ruido_suave(i, j, k, it, seed);
}

// source/W0_E0.f90:10-28.  The DO variable is REAL(8) with a REAL(4) step: it
// starts at 0 and is incremented by (double)1e-7f after each pass (trip count
// fixed up front), and the first try with |E0 - try| < 1e-5 is kept.  Each
// try's |E0 - try| goes to diff[n] while n < cap (the reference prints it).
static int32_t w0_e0_search(double lambda, float ajuste, float sigma, double *W0, double *E0,
                            double *diff, int32_t cap) {
  const double step = (double)0.0000001f;
  const long long trips = (long long)((1.0 - 0.0 + step) / step);
  const float as = ajuste * sigma;  // REAL(4) product
  double t = 0.0, w = 0.0, e = 0.0;
  for (long long n = 0; n < trips; ++n) {
    const double a = (1.5556349186104 / lambda) * t;
    w = a * a;
    e = (double)as * 4.24264068711929 / sqrt(w);
    if (diff && n < cap) diff[n] = fabs(e - t);
    if (fabs(e - t) < 1e-05) {
      if (W0) *W0 = w;
      if (E0) *E0 = e;
      return (int32_t)n;
    }
    t = t + step;
  }
  if (W0) *W0 = w;
  if (E0) *E0 = e;
  return -1;
}

int32_t pf_w0_e0(double lambda, float ajuste, float sigma, double *W0, double *E0) {
  return w0_e0_search(lambda, ajuste, sigma, W0, E0, nullptr, 0);
}

int32_t pf_w0_e0_tries(double lambda, float ajuste, float sigma, double *diff, int32_t cap) {
  const int32_t idx = w0_e0_search(lambda, ajuste, sigma, nullptr, nullptr, diff, cap);
  return idx < 0 ? -1 : idx + 1;
}

// source/auxiliar.f90:14-42 (pf::auxiliar_m0 in pf_common.h).
double pf_auxiliar_m0(double TP, double Cl_eq, double Cs_eq, double R, double Vm, double E0,
                      double W0, float Sigma, float Beta, float Ke, double Me, double D_liq,
                      float ajuste) {
  return pf::auxiliar_m0(TP, Cl_eq, Cs_eq, R, Vm, E0, W0, Sigma, Beta, Ke, Me, D_liq, ajuste);
}

void pf_equilibrium(double TP, double Tm, float Ke, double Me, double *Cs_eq, double *Cl_eq) {
  double cs, cl;
  pf::equilibrium(TP, Tm, Ke, Me, cs, cl);
  if (Cs_eq) *Cs_eq = cs;
  if (Cl_eq) *Cl_eq = cl;
}

double pf_time_step(double dx, double D_liq) { return (dx * dx) / (7.5 * D_liq); }

double pf_noise_scale(double AAA, double dt, double dx, double D_liq, int32_t ruido_tipo) {
  const double ratio = dt / ((dx * dx) / (5.0 * D_liq));
  return ruido_tipo == 0 ? AAA * ratio : AAA * sqrt(ratio);
}

static uint64_t splitmix64(uint64_t *s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

int pf_initial_fields(int32_t nx, int32_t ny, int32_t nz, int32_t tipo, double R_semente,
                      double T0, double C0, float Ke, int32_t n_seeds, uint64_t seed_rng,
                      double *phi_old, double *phi_new, double *c_old, double *c_new,
                      double *T_old, double *T_new, double *dpp) {
  if (nx < 1 || ny < 1 || nz < 1 || !phi_old || !phi_new || !c_old || !c_new || !T_old || !T_new ||
      !dpp)
    return PF_E_ARG;
  if (tipo == 6 && (n_seeds < 0 || nx < 9 || ny < 9)) return PF_E_ARG;
  int ic = nx / 2, jc = ny / 2, kc = nz / 2;
  if (tipo == 2) ic = jc = kc = 3;
  if (tipo == 5) kc = 1;
  const int ns = tipo == 6 ? n_seeds : 0;
  int *sij = ns ? (int *)malloc(sizeof(int) * 2 * (size_t)ns) : nullptr;
  uint64_t st = seed_rng;
  for (int s = 0; s < ns; ++s) {
    sij[2 * s] = 5 + (int)(splitmix64(&st) % (uint64_t)(nx - 8));
    sij[2 * s + 1] = 5 + (int)(splitmix64(&st) % (uint64_t)(ny - 8));
  }
  const double r2 = R_semente * R_semente;
  const double solid_c = (double)Ke * C0;
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        double d2 = DBL_MAX;
        if (tipo == 1 || tipo == 2 || tipo == 5)
          d2 = (double)(i - ic) * (double)(i - ic) + (double)(j - jc) * (double)(j - jc) +
               (double)(k - kc) * (double)(k - kc);
        for (int s = 0; s < ns; ++s) {
          const double di = (double)(i - sij[2 * s]), dj = (double)(j - sij[2 * s + 1]);
          const double e = di * di + dj * dj + (double)(k - 1) * (double)(k - 1);
          if (e < d2) d2 = e;
        }
        const size_t c = (size_t)(i - 1) + (size_t)nx * ((size_t)(j - 1) + (size_t)ny * (k - 1));
        const bool solid = (tipo == 0 && j <= 4) || (tipo == 3 && k <= 4) || d2 <= r2;
        phi_old[c] = solid ? 1.0 : 0.0;
        c_old[c] = solid ? solid_c : C0;
        phi_new[c] = 0.0;
        c_new[c] = C0;
        T_old[c] = T0;
        T_new[c] = T0;
        dpp[c] = 0.0;
      }
  free(sij);
  return PF_OK;
}

}  // extern "C"

// FP64 interior ([2,n-1]^3) sum of a host field, k, j, i in order.
double pf::host_interior_mass(const double *c, int nx, int ny, int nz) {
  double s = 0.0;
  for (int k = 2; k <= nz - 1; ++k)
    for (int j = 2; j <= ny - 1; ++j) {
      const double *row = c + (size_t)nx * ((size_t)(j - 1) + (size_t)ny * (size_t)(k - 1));
      for (int i = 2; i <= nx - 1; ++i) s += row[i - 1];
    }
  return s;
}

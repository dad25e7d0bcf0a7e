/*
 * pf_oracle.c -- CPU oracle for the coupled phase-field time step.
 *
 * TEST INFRASTRUCTURE ONLY (see pf_oracle.h).  A restatement of the
 * reference's FP64 CPU path, function by function, in its operation order:
 *   Fortran a*b*c    is evaluated ((a*b)*c); a+b+c as ((a+b)+c);
 *   x**2 / x**2.d0   are x*x; x**4 is x*(x*(x*x)); x**3 is (x*x)*x;
 *   REAL(4) variables/literals are float, promoted to double where the
 *   Fortran expression mixes kinds.
 * Build with -O2 -ffp-contract=off (no FMA contraction), like the x86-64
 * flang build of the reference it is checked against.
 */
#include "pf_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PFO_SEED_MUL 0x9E3779B97F4A7C15ull

static int32_t g_threads = 1;
void pfo_set_threads(int32_t n) { g_threads = n < 1 ? 1 : n; }
int32_t pfo_get_threads(void) { return g_threads; }

#define ID(i, j, k)                                                            \
  ((size_t)((i)-1) + (size_t)nx * ((size_t)((j)-1) + (size_t)ny * (size_t)((k)-1)))

/* ------------------------------------------------------------------------ */
/* source/variaveis.f90:73-92 -- counter hash (MurmurHash3 fmix64 finaliser). */
double pfo_pseudo_random(int32_t i, int32_t j, int32_t k, int32_t it, int32_t salt,
                         uint64_t seed) {
  uint64_t h = (uint64_t)(int64_t)i * 2654435761ull + (uint64_t)(int64_t)j * 2246822519ull +
               (uint64_t)(int64_t)k * 3266489917ull + (uint64_t)(int64_t)it * 668265263ull +
               (uint64_t)(int64_t)salt * 374761393ull + 1ull;
  h += seed * PFO_SEED_MUL; /* build extension: vanishes for seed 0 */
  h ^= h >> 33;             /* :84  ishft is a logical shift */
  h *= 0xff51afd7ed558ccdull;
  h ^= h >> 33;
  h *= 0xc4ceb9fe1a85ec53ull;
  h ^= h >> 33;
  /* :90  real(iand(h,MASK63),8)/real(MASK63,8) */
  return (double)(int64_t)(h & 0x7fffffffffffffffull) / (double)(int64_t)0x7fffffffffffffffll;
}

/* source/variaveis.f90:98-111 -- trilinear noise on a pitch-6 lattice,
 * renewed every 20 steps (LR_RUIDO=6, TR_RUIDO=20 at :62). */
double pfo_ruido_suave(int32_t i, int32_t j, int32_t k, int32_t it, uint64_t seed) {
  const int L = 6, TR = 20;
  int i0 = i / L, j0 = j / L, k0 = k / L, tb = it / TR;
  double fx = (double)(i % L) / (double)L;
  double fy = (double)(j % L) / (double)L;
  double fz = (double)(k % L) / (double)L;
#define PR(a, b, c) pfo_pseudo_random((a), (b), (c), tb, 3, seed)
  double r = (1.0 - fz) * ((1.0 - fy) * ((1.0 - fx) * PR(i0, j0, k0) + fx * PR(i0 + 1, j0, k0)) +
                           fy * ((1.0 - fx) * PR(i0, j0 + 1, k0) + fx * PR(i0 + 1, j0 + 1, k0))) +
             fz * ((1.0 - fy) * ((1.0 - fx) * PR(i0, j0, k0 + 1) + fx * PR(i0 + 1, j0, k0 + 1)) +
                   fy * ((1.0 - fx) * PR(i0, j0 + 1, k0 + 1) + fx * PR(i0 + 1, j0 + 1, k0 + 1)));
#undef PR
  return r;
}

/* source/W0_E0.f90:10-28 -- DO EPSLON_Try=0,1,0.0000001 with a REAL(4) step:
 * the DO variable starts at 0 and is incremented by (double)1e-7f after each
 * pass; the first try with |E0-try| < 1e-5 is taken. */
int32_t pfo_w0_e0(double lambda, float ajuste, float sigma, double *W0, double *E0) {
  const double step = (double)0.0000001f;
  const double m1 = 0.0, m2 = 1.0;
  long long trips = (long long)((m2 - m1 + step) / step);
  double t = m1;
  float as = ajuste * sigma; /* :16  REAL(4)*REAL(4) */
  double w = 0.0, e = 0.0;
  for (long long n = 0; n < trips; ++n) {
    w = ((1.5556349186104 / lambda) * t) * ((1.5556349186104 / lambda) * t);
    e = (double)as * 4.24264068711929 / sqrt(w);
    if (fabs(e - t) < 1e-05) {
      *W0 = w;
      *E0 = e;
      return (int32_t)n;
    }
    t = t + step;
  }
  *W0 = w;
  *E0 = e;
  return -1;
}

/* source/auxiliar.f90:14-42 -- thin-interface mobility. */
double pfo_auxiliar(double TP, double Cl_eq, double Cs_eq, double R, double Vm,
                    double E0, double W0, float Sigma, float Beta, float Ke,
                    double Me, double D_liq, float ajuste) {
  double acc = 0.0;
  double phi_i = 0.11;
  do { /* :21-34 */
    double h_phi = phi_i * phi_i * (3.0 - 2.0 * phi_i);
    double function_1 = phi_i * (1.0 - phi_i);
    double numerador = h_phi * (1.0 - h_phi);
    double denominador =
        ((1.0 - h_phi) * Cl_eq * (1.0 - Cl_eq) + h_phi * Cs_eq * (1.0 - Cs_eq)) * function_1;
    acc = acc + numerador / denominador;
    phi_i = phi_i + 0.01;
  } while (phi_i <= 0.8999);
  /* :31  0.01 and 54.39180136 are REAL(4) literals */
  double integral = ((double)0.01f / 2.0) * ((double)54.39180136f + 2.0 * acc + 234.073053);
  double qsi = R * TP * (1.0 / Vm) * (Cl_eq - Cs_eq) * (Cl_eq - Cs_eq) * integral; /* :36 */
  float one_minus_ke = 1.0f - Ke; /* :39  (1-Ke) in REAL(4) */
  double inv_m0 = (E0 * E0 / (double)Sigma) *
                  (((double)Beta * R * TP * (double)one_minus_ke / (Vm * Me)) +
                   (E0 * qsi / (D_liq * sqrt(2.0 * W0))));
  return (double)ajuste * 1.0 / inv_m0; /* :42 */
}

/* source/entrada.f90:173 */
double pfo_delta_t(double dx, double D_liq) { return (dx * dx) / (7.5 * D_liq); }

/* source/entrada.f90:179-185 */
double pfo_noise_scale(double AAA, double dt, double dx, double D_liq, int32_t ruido_tipo) {
  if (ruido_tipo == 0) return AAA * (dt / ((dx * dx) / (5.0 * D_liq)));
  return AAA * sqrt(dt / ((dx * dx) / (5.0 * D_liq)));
}

static uint64_t splitmix64(uint64_t *s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* source/entrada.f90:226-262 for tipo 0..3.  Mode 5 (bottom-centre sphere at
 * (nx/2,ny/2,1)) and mode 6 (n_seeds spheres centred on k=1 at SplitMix64
 * positions in [5,n-4]) are build extensions used by the workload configs. */
int pfo_initial_fields(int32_t nx, int32_t ny, int32_t nz, int32_t tipo,
                       double R_semente, double T0, double C0, float Ke,
                       int32_t n_seeds, uint64_t seed_rng, double *phi_old,
                       double *phi_new, double *c_old, double *c_new,
                       double *T_old, double *T_new, double *dpp) {
  int ic = nx / 2, jc = ny / 2, kc = nz / 2;
  if (tipo == 2) ic = jc = kc = 3;
  if (tipo == 5) kc = 1;
  int ns = (tipo == 6) ? n_seeds : 0;
  int *si = NULL, *sj = NULL;
  if (ns > 0) {
    si = (int *)malloc(sizeof(int) * (size_t)ns);
    sj = (int *)malloc(sizeof(int) * (size_t)ns);
    uint64_t st = seed_rng;
    for (int s = 0; s < ns; ++s) {
      si[s] = 5 + (int)(splitmix64(&st) % (uint64_t)(nx - 8));
      sj[s] = 5 + (int)(splitmix64(&st) % (uint64_t)(ny - 8));
    }
  }
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        double dist2;
        if (tipo == 1 || tipo == 2 || tipo == 5)
          dist2 = (double)(i - ic) * (double)(i - ic) + (double)(j - jc) * (double)(j - jc) +
                  (double)(k - kc) * (double)(k - kc);
        else
          dist2 = DBL_MAX;
        for (int s = 0; s < ns; ++s) {
          double d2 = (double)(i - si[s]) * (double)(i - si[s]) +
                      (double)(j - sj[s]) * (double)(j - sj[s]) + (double)(k - 1) * (double)(k - 1);
          if (d2 < dist2) dist2 = d2;
        }
        size_t c = ID(i, j, k);
        if ((tipo == 0 && j <= 4) || (tipo == 3 && k <= 4) || dist2 <= R_semente * R_semente) {
          phi_old[c] = 1.0;
          c_old[c] = (double)Ke * C0;
        } else {
          phi_old[c] = 0.0;
          c_old[c] = C0;
        }
        phi_new[c] = 0.0;
        T_new[c] = T0;
        T_old[c] = T0;
        c_new[c] = C0;
        dpp[c] = 0.0;
      }
  free(si);
  free(sj);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* source/temperatura.f90:55-70 + temp_body.inc:11-45 (CPU path). */
void pfo_temperatura(const pf_params *p, const double *Told, const double *phi_old,
                     const double *dpp, double *Tnew) {
  const int nx = p->nx, ny = p->ny, nz = p->nz;
  const double cp = p->cp_s; /* temperatura.f90:9 */
  const double D = p->D, dt = p->dt, dH = p->delta_H, TP = p->TP;
  const double dx = p->dx, dy = p->dy, dz = p->dz;
  const int hq = p->h_quintico;
  memcpy(Tnew, Told, sizeof(double) * (size_t)nx * ny * nz); /* :55 TNEW=TOLD */
#pragma omp parallel for collapse(2) schedule(static) num_threads(g_threads)
  for (int k = 2; k <= nz - 1; ++k)
    for (int j = 2; j <= ny - 1; ++j)
      for (int i = 2; i <= nx - 1; ++i) {
        double TE = Told[ID(i + 1, j, k)], TW = Told[ID(i - 1, j, k)];
        double TN = Told[ID(i, j + 1, k)], TS = Told[ID(i, j - 1, k)];
        double TT = Told[ID(i, j, k + 1)], TB = Told[ID(i, j, k - 1)];
        double PP = phi_old[ID(i, j, k)];
        double h_linha;
        if (hq == 1)
          h_linha = 30.0 * PP * PP * (1.0 - PP) * (1.0 - PP);
        else
          h_linha = 2.0 * PP * (3.0 - 2.0 * PP) - 2.0 * PP * PP;
        double FONTE = (dH / cp) * h_linha * dpp[ID(i, j, k)];
        double TNEW1 = D * dt *
                           (((TE - 2.0 * TP + TW) / (dx * dx)) + ((TN - 2.0 * TP + TS) / (dy * dy)) +
                            ((TT - 2.0 * TP + TB) / (dz * dz))) +
                       FONTE + TP;
        Tnew[ID(i, j, k)] = TNEW1;
      }
}

/* source/phase_body.inc:14-198 -- one interior cell of a phase sweep. */
static inline void phase_cell(const double *po, const double *co, int nx, int ny, int i, int j,
                              int k, const pf_params *p, double Cs_eq, double Cl_eq,
                              double *pnew, double *dpp) {
#define P(a, b, c) po[ID((a), (b), (c))]
  const double dx = p->dx, dy = p->dy, dz = p->dz, E0 = p->E0, DE = p->delta_eps;
  double PP = P(i, j, k);
  double PN = P(i, j + 1, k), PS = P(i, j - 1, k), PE = P(i + 1, j, k), PW = P(i - 1, j, k);
  double PT = P(i, j, k + 1), PB = P(i, j, k - 1);
  double PNE = P(i + 1, j + 1, k), PSE = P(i + 1, j - 1, k), PNW = P(i - 1, j + 1, k),
         PSW = P(i - 1, j - 1, k);
  double PNT = P(i, j + 1, k + 1), PNB = P(i, j + 1, k - 1), PST = P(i, j - 1, k + 1),
         PSB = P(i, j - 1, k - 1);
  double PET = P(i + 1, j, k + 1), PEB = P(i + 1, j, k - 1), PWT = P(i - 1, j, k + 1),
         PWB = P(i - 1, j, k - 1);
  const double small = 1.0e-23;          /* :47 */
  const double K = 16.0 * E0 * DE;       /* :49 */
#define SQ(x) ((x) * (x))
#define P4(x) ((x) * ((x) * ((x) * (x)))) /* powi(x,4) as flang expands it */
  /* E: node (i+1,j,k)  :55-66 */
  double gxE = (P(i + 2, j, k) - PP) * 0.5 / dx;
  double gyE = (P(i + 1, j + 1, k) - P(i + 1, j - 1, k)) * 0.5 / dy;
  double gzE = (P(i + 1, j, k + 1) - P(i + 1, j, k - 1)) * 0.5 / dz;
  double QE = SQ(gxE) + SQ(gyE) + SQ(gzE) + small;
  double S4E = P4(gxE) + P4(gyE) + P4(gzE);
  double epsE = E0 * (1.0 + DE * (4.0 * S4E / SQ(QE) - 3.0));
  double exyE = K * epsE * gxE * gyE * (SQ(gxE) - SQ(gyE)) / SQ(QE);
  double ezxE = K * epsE * gzE * gxE * (SQ(gzE) - SQ(gxE)) / SQ(QE);
  /* W: node (i-1,j,k)  :68-78 */
  double gxW = (PP - P(i - 2, j, k)) * 0.5 / dx;
  double gyW = (P(i - 1, j + 1, k) - P(i - 1, j - 1, k)) * 0.5 / dy;
  double gzW = (P(i - 1, j, k + 1) - P(i - 1, j, k - 1)) * 0.5 / dz;
  double QW = SQ(gxW) + SQ(gyW) + SQ(gzW) + small;
  double S4W = P4(gxW) + P4(gyW) + P4(gzW);
  double epsW = E0 * (1.0 + DE * (4.0 * S4W / SQ(QW) - 3.0));
  double exyW = K * epsW * gxW * gyW * (SQ(gxW) - SQ(gyW)) / SQ(QW);
  double ezxW = K * epsW * gzW * gxW * (SQ(gzW) - SQ(gxW)) / SQ(QW);
  /* N: node (i,j+1,k)  :81-92 */
  double gxN = (P(i + 1, j + 1, k) - P(i - 1, j + 1, k)) * 0.5 / dx;
  double gyN = (P(i, j + 2, k) - PP) * 0.5 / dy;
  double gzN = (P(i, j + 1, k + 1) - P(i, j + 1, k - 1)) * 0.5 / dz;
  double QN = SQ(gxN) + SQ(gyN) + SQ(gzN) + small;
  double S4N = P4(gxN) + P4(gyN) + P4(gzN);
  double epsN = E0 * (1.0 + DE * (4.0 * S4N / SQ(QN) - 3.0));
  double exyN = K * epsN * gxN * gyN * (SQ(gxN) - SQ(gyN)) / SQ(QN);
  double eyzN = K * epsN * gyN * gzN * (SQ(gyN) - SQ(gzN)) / SQ(QN);
  /* S: node (i,j-1,k)  :94-103 */
  double gxS = (P(i + 1, j - 1, k) - P(i - 1, j - 1, k)) * 0.5 / dx;
  double gyS = (PP - P(i, j - 2, k)) * 0.5 / dy;
  double gzS = (P(i, j - 1, k + 1) - P(i, j - 1, k - 1)) * 0.5 / dz;
  double QS = SQ(gxS) + SQ(gyS) + SQ(gzS) + small;
  double S4S = P4(gxS) + P4(gyS) + P4(gzS);
  double epsS = E0 * (1.0 + DE * (4.0 * S4S / SQ(QS) - 3.0));
  double exyS = K * epsS * gxS * gyS * (SQ(gxS) - SQ(gyS)) / SQ(QS);
  double eyzS = K * epsS * gyS * gzS * (SQ(gyS) - SQ(gzS)) / SQ(QS);
  /* T: node (i,j,k+1)  :107-116 */
  double gxT = (P(i + 1, j, k + 1) - P(i - 1, j, k + 1)) * 0.5 / dx;
  double gyT = (P(i, j + 1, k + 1) - P(i, j - 1, k + 1)) * 0.5 / dy;
  double gzT = (P(i, j, k + 2) - PP) * 0.5 / dz;
  double QT = SQ(gxT) + SQ(gyT) + SQ(gzT) + small;
  double S4T = P4(gxT) + P4(gyT) + P4(gzT);
  double epsT = E0 * (1.0 + DE * (4.0 * S4T / SQ(QT) - 3.0));
  double eyzT = K * epsT * gyT * gzT * (SQ(gyT) - SQ(gzT)) / SQ(QT);
  double ezxT = K * epsT * gzT * gxT * (SQ(gzT) - SQ(gxT)) / SQ(QT);
  /* B: node (i,j,k-1)  :120-129 */
  double gxB = (P(i + 1, j, k - 1) - P(i - 1, j, k - 1)) * 0.5 / dx;
  double gyB = (P(i, j + 1, k - 1) - P(i, j - 1, k - 1)) * 0.5 / dy;
  double gzB = (PP - P(i, j, k - 2)) * 0.5 / dz;
  double QB = SQ(gxB) + SQ(gyB) + SQ(gzB) + small;
  double S4B = P4(gxB) + P4(gyB) + P4(gzB);
  double epsB = E0 * (1.0 + DE * (4.0 * S4B / SQ(QB) - 3.0));
  double eyzB = K * epsB * gyB * gzB * (SQ(gyB) - SQ(gzB)) / SQ(QB);
  double ezxB = K * epsB * gzB * gxB * (SQ(gzB) - SQ(gxB)) / SQ(QB);
  /* :133-149 */
  double AA = (SQ(epsE) * (PE - PP) + SQ(epsW) * (PW - PP)) / (dx * dx) +
              (SQ(epsN) * (PN - PP) + SQ(epsS) * (PS - PP)) / (dy * dy) +
              (SQ(epsT) * (PT - PP) + SQ(epsB) * (PB - PP)) / (dz * dz);
  const double cxy = 0.25 / (dx * dy), cyz = 0.25 / (dy * dz), czx = 0.25 / (dz * dx);
  double Cxy = cxy * (PNE * (-exyN - (-exyE)) + PSE * (-exyE - (-exyS)) + PNW * (-exyW - (-exyN)) +
                      PSW * (-exyS - (-exyW))) +
               PN * cxy * (-exyW - (-exyE)) + PS * cxy * (-exyE - (-exyW)) +
               PE * cxy * (-exyN - (-exyS)) + PW * cxy * (-exyS - (-exyN));
  double Cyz = cyz * (PNT * (-eyzT - (-eyzN)) + PNB * (-eyzN - (-eyzB)) + PST * (-eyzS - (-eyzT)) +
                      PSB * (-eyzB - (-eyzS))) +
               PT * cyz * (-eyzS - (-eyzN)) + PB * cyz * (-eyzN - (-eyzS)) +
               PN * cyz * (-eyzT - (-eyzB)) + PS * cyz * (-eyzB - (-eyzT));
  double Czx = czx * (PET * (-ezxE - (-ezxT)) + PWT * (-ezxT - (-ezxW)) + PEB * (-ezxB - (-ezxE)) +
                      PWB * (-ezxW - (-ezxB))) +
               PE * czx * (-ezxB - (-ezxT)) + PW * czx * (-ezxT - (-ezxB)) +
               PT * czx * (-ezxE - (-ezxW)) + PB * czx * (-ezxW - (-ezxE));
  AA = AA + Cxy + Cyz + Czx;
  /* :153-174 */
  double g_linha = 2.0 * PP * (1.0 - PP) * (1.0 - PP) - 2.0 * PP * PP * (1.0 - PP);
  double CPP = co[ID(i, j, k)];
  double h_linha, h;
  if (p->h_quintico == 1) {
    h_linha = 30.0 * PP * PP * (1.0 - PP) * (1.0 - PP);
    h = PP * PP * PP * (10.0 - 15.0 * PP + 6.0 * PP * PP);
  } else {
    h_linha = 2.0 * PP * (3.0 - 2.0 * PP) - 2.0 * PP * PP;
    h = PP * PP * (3.0 - 2.0 * PP);
  }
  const double Ke = (double)p->Ke;
  double CLP = CPP / (Ke * h + (1.0 - h));
  double CSP = Ke * CPP / (Ke * h + (1.0 - h));
  double GG = (-(p->W0 * g_linha)) +
              (p->R * p->TP / p->Vm) * h_linha *
                  log(fabs(((1.0 - Cs_eq) * (1.0 - CLP)) / ((1.0 - Cl_eq) * (1.0 - CSP))));
  /* :181-190  ran is REAL(4) */
  float ran;
  if (p->ruido_tipo == 2)
    ran = (float)pfo_ruido_suave(i, j, k, p->it, p->seed);
  else
    ran = (float)pfo_pseudo_random(i, j, k, p->it, 1, p->seed);
  double ruidos;
  if (p->ruido_tipo == 0)
    ruidos = (double)ran * p->AAA * 16.0 * (PP * PP) * ((1.0 - PP) * (1.0 - PP));
  else
    ruidos = (2.0 * (double)ran - 1.0) * p->AAA * 16.0 * (PP * PP) * ((1.0 - PP) * (1.0 - PP));
  /* :194-198 */
  double PPNEW = PP + (p->M0 * p->dt) * (AA + GG) + ruidos;
  *dpp = PPNEW - PP;
  /* dmax1(0,dmin1(1,x)) as compiled: NaN propagates (C fmin/fmax would not) */
  double cl = PPNEW > 1.0 ? 1.0 : PPNEW;
  *pnew = cl < 0.0 ? 0.0 : cl;
#undef P
#undef SQ
#undef P4
}

/* source/equacao_campo_fase.f90:21-22, 110-174 (CPU path). */
void pfo_equacao_campo_fase(pf_params *p, const double *phi_old, const double *c_old,
                            double *phi_new, double *dpp, double probe[3]) {
  const int nx = p->nx, ny = p->ny, nz = p->nz;
  const double Ke = (double)p->Ke;
  p->Cs_eq = ((p->Tm - p->TP) * Ke) / p->Me; /* :21 */
  p->Cl_eq = (p->Tm - p->TP) / p->Me;        /* :22 */
  const double Cs_eq = p->Cs_eq, Cl_eq = p->Cl_eq;
  memcpy(phi_new, phi_old, sizeof(double) * (size_t)nx * ny * nz); /* :110 */
#pragma omp parallel for collapse(2) schedule(static) num_threads(g_threads)
  for (int k = 3; k <= nz - 2; ++k)
    for (int j = 3; j <= ny - 2; ++j)
      for (int i = 3; i <= nx - 2; ++i)
        phase_cell(phi_old, c_old, nx, ny, i, j, k, p, Cs_eq, Cl_eq, &phi_new[ID(i, j, k)],
                   &dpp[ID(i, j, k)]);
  /* :151-167  boundary decoration, sequential array sections, y then x then z */
  double *f = phi_new;
  for (int k = 1; k <= nz; ++k)
    for (int i = 1; i <= nx; ++i) f[ID(i, 1, k)] = f[ID(i, 2, k)];
  for (int k = 1; k <= nz; ++k)
    for (int i = 1; i <= nx; ++i) f[ID(i, 2, k)] = f[ID(i, 3, k)];
  for (int k = 1; k <= nz; ++k)
    for (int i = 1; i <= nx; ++i) f[ID(i, ny, k)] = f[ID(i, ny - 1, k)];
  for (int k = 1; k <= nz; ++k)
    for (int i = 1; i <= nx; ++i) f[ID(i, ny - 1, k)] = f[ID(i, ny - 2, k)];
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j) f[ID(1, j, k)] = f[ID(2, j, k)];
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j) f[ID(2, j, k)] = f[ID(3, j, k)];
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j) f[ID(nx, j, k)] = f[ID(nx - 1, j, k)];
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j) f[ID(nx - 1, j, k)] = f[ID(nx - 2, j, k)];
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) f[ID(i, j, 1)] = f[ID(i, j, 2)];
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) f[ID(i, j, 2)] = f[ID(i, j, 3)];
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) f[ID(i, j, nz)] = f[ID(i, j, nz - 1)];
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) f[ID(i, j, nz - 1)] = f[ID(i, j, nz - 2)];
  if (probe) { /* :174 */
    int kz = nz / 2 > 1 ? nz / 2 : 1;
    int j2 = ny - 1 < 100 ? ny - 1 : 100;
    probe[0] = f[ID(nx / 2, 10, kz)];
    probe[1] = p->TP;
    probe[2] = f[ID(nx / 2, j2, kz)];
  }
}

/* source/concentracao_soluto.f90:9-11, 86-124 + conc_body.inc:11-149. */
void pfo_concentracao_soluto(const pf_params *p, const double *phi_old, double *c_old,
                             double *c_new) {
  const int nx = p->nx, ny = p->ny, nz = p->nz;
  const double InvDX = 1.0 / (p->dx * p->dx), InvDY = 1.0 / (p->dy * p->dy),
               InvDZ = 1.0 / (p->dz * p->dz);
  const double Ke = (double)p->Ke, dt = p->dt, D_liq = p->D_liq, D_sol = p->D_sol;
  const int hq = p->h_quintico;
  const size_t n = (size_t)nx * ny * nz;
  memcpy(c_new, c_old, sizeof(double) * n); /* :86 */
#pragma omp parallel for collapse(2) schedule(static) num_threads(g_threads)
  for (int k = 2; k <= nz - 1; ++k)
    for (int j = 2; j <= ny - 1; ++j)
      for (int i = 2; i <= nx - 1; ++i) {
        const double *po = phi_old, *co = c_old;
        double PP = po[ID(i, j, k)], PN = po[ID(i, j + 1, k)], PS = po[ID(i, j - 1, k)];
        double PW = po[ID(i - 1, j, k)], PE = po[ID(i + 1, j, k)];
        double PT = po[ID(i, j, k + 1)], PB = po[ID(i, j, k - 1)];
        double CPP = co[ID(i, j, k)], CPS = co[ID(i, j - 1, k)], CPN = co[ID(i, j + 1, k)];
        double CPW = co[ID(i - 1, j, k)], CPE = co[ID(i + 1, j, k)];
        double CPT = co[ID(i, j, k + 1)], CPB = co[ID(i, j, k - 1)];
        double h, hE, hW, hN, hS, hT, hB;
        if (hq == 1) { /* :42-49  x**3 = (x*x)*x, x**2 = x*x */
#define HQ(x) (((x) * (x) * (x)) * (10.0 - (15.0 * (x)) + (6.0 * ((x) * (x)))))
          h = HQ(PP); hE = HQ(PE); hW = HQ(PW); hN = HQ(PN);
          hS = HQ(PS); hT = HQ(PT); hB = HQ(PB);
#undef HQ
        } else { /* :51-57 */
#define HC(x) ((x) * (x) * (3.0 - 2.0 * (x)))
          h = HC(PP); hE = HC(PE); hW = HC(PW); hN = HC(PN);
          hS = HC(PS); hT = HC(PT); hB = HC(PB);
#undef HC
        }
#define CS(c, hh) ((Ke * (c)) / ((Ke * (hh)) + (1.0 - (hh))))
#define CL(c, hh) ((c) / (Ke * (hh) + (1.0 - (hh))))
        double CSP = CS(CPP, h), CSN = CS(CPN, hN), CSS = CS(CPS, hS), CSE = CS(CPE, hE);
        double CSW = CS(CPW, hW), CST = CS(CPT, hT), CSB = CS(CPB, hB);
        double CLP = CL(CPP, h), CLN = CL(CPN, hN), CLS = CL(CPS, hS), CLE = CL(CPE, hE);
        double CLW = CL(CPW, hW), CLT = CL(CPT, hT), CLB = CL(CPB, hB);
#undef CS
#undef CL
        double DC = (PP <= 0.95) ? D_liq : D_sol; /* :92-100 */
#define GA(hh, cl, cs) (DC * ((1.0 - (hh)) * (cl) * (1.0 - (cl)) + (hh) * (cs) * (1.0 - (cs))))
        double GP = GA(h, CLP, CSP), GS = GA(hS, CLS, CSS), GN = GA(hN, CLN, CSN);
        double GW = GA(hW, CLW, CSW), GE = GA(hE, CLE, CSE), GT = GA(hT, CLT, CST);
        double GB = GA(hB, CLB, CSB);
#undef GA
        double Gs = 2.0 / ((1.0 / GP) + (1.0 / GS)); /* :121-131 */
        double Gn = 2.0 / ((1.0 / GN) + (1.0 / GP));
        double Gw = 2.0 / ((1.0 / GP) + (1.0 / GW));
        double Ge = 2.0 / ((1.0 / GE) + (1.0 / GP));
        double Gt = 2.0 / ((1.0 / GT) + (1.0 / GP));
        double Gb = 2.0 / ((1.0 / GP) + (1.0 / GB));
#define MU(cl) log((cl) / (1.0 - (cl)))
        double cn = (Ge * InvDX * (MU(CLE) - MU(CLP)) - Gw * InvDX * (MU(CLP) - MU(CLW)) +
                     Gn * InvDY * (MU(CLN) - MU(CLP)) - Gs * InvDY * (MU(CLP) - MU(CLS)) +
                     Gt * InvDZ * (MU(CLT) - MU(CLP)) - Gb * InvDZ * (MU(CLP) - MU(CLB))) *
                        dt +
                    CPP; /* :139-144 */
#undef MU
        c_new[ID(i, j, k)] = cn;
      }
  memcpy(c_old, c_new, sizeof(double) * n); /* :116 */
  /* :119-124 */
  for (int k = 1; k <= nz; ++k)
    for (int i = 1; i <= nx; ++i) c_new[ID(i, 1, k)] = c_old[ID(i, 2, k)];
  for (int k = 1; k <= nz; ++k)
    for (int i = 1; i <= nx; ++i) c_new[ID(i, ny, k)] = c_old[ID(i, ny - 1, k)];
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j) c_new[ID(1, j, k)] = c_old[ID(2, j, k)];
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j) c_new[ID(nx, j, k)] = c_old[ID(nx - 1, j, k)];
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) c_new[ID(i, j, 1)] = c_old[ID(i, j, 2)];
  for (int j = 1; j <= ny; ++j)
    for (int i = 1; i <= nx; ++i) c_new[ID(i, j, nz)] = c_old[ID(i, j, nz - 1)];
}

/* source/diagnosticos.f90:10-56 + diag_body.inc:8-22.  Serial k,j,i order
 * (the reference's single-thread order). */
void pfo_diagnosticos(const pf_params *p, const double *phi_new, const double *c_new,
                      double out[5]) {
  const int nx = p->nx, ny = p->ny, nz = p->nz;
  const double dx = p->dx, dy = p->dy, dz = p->dz, E0 = p->E0, w0 = p->W0;
  double L2 = 0.0, pmax = -DBL_MAX, pmin = DBL_MAX, sc = 0.0, fe = 0.0;
  for (int k = 1; k <= nz; ++k)
    for (int j = 1; j <= ny; ++j)
      for (int i = 1; i <= nx; ++i) {
        double ph = phi_new[ID(i, j, k)];
        L2 = L2 + ph * ph;
        if (ph > pmax) pmax = ph;
        if (ph < pmin) pmin = ph;
        sc = sc + c_new[ID(i, j, k)];
        double w = w0 * ph * ph * (1.0 - ph) * (1.0 - ph);
        double g2;
        if (i > 1 && i < nx && j > 1 && j < ny && k > 1 && k < nz) {
          double a = (phi_new[ID(i + 1, j, k)] - phi_new[ID(i - 1, j, k)]) * 0.5 / dx;
          double b = (phi_new[ID(i, j + 1, k)] - phi_new[ID(i, j - 1, k)]) * 0.5 / dy;
          double c = (phi_new[ID(i, j, k + 1)] - phi_new[ID(i, j, k - 1)]) * 0.5 / dz;
          g2 = a * a + b * b + c * c;
        } else {
          g2 = 0.0;
        }
        fe = fe + (0.5 * E0 * E0 * g2 + w) * dx * dy * dz;
      }
  out[0] = sqrt(L2);
  out[1] = pmax;
  out[2] = pmin;
  out[3] = sc;
  out[4] = fe;
}

/* source/principal.f90:93-226 without the output event (a stop on
 * stoptime is the caller's business). */
int pfo_run_steps(pf_params *p, pf_loop *L, int32_t nsteps, double *phi_old,
                  double *phi_new, double *c_old, double *c_new, double *T_old,
                  double *T_new, double *dpp, double *probes) {
  const size_t n = (size_t)p->nx * p->ny * p->nz;
  for (int s = 0; s < nsteps; ++s) {
    p->it = p->it + 1;
    const int it = p->it;
    L->time = (float)(p->dt + (double)L->time);          /* :96 */
    pfo_temperatura(p, T_old, phi_old, dpp, T_new);       /* :98 */
    if (L->taxa_resfr > 0.0) p->TP = L->T0 - L->taxa_resfr * p->dt * (double)it; /* :104 */
    int nv = p->n_varreduras < 1 ? 1 : p->n_varreduras;
    for (int sw = 1; sw <= nv; ++sw) {                    /* :127-152 */
      if (sw > 1) memcpy(phi_old, phi_new, sizeof(double) * n);
      double *pr = probes ? probes + 3 * ((size_t)s * nv + (sw - 1)) : NULL;
      pfo_equacao_campo_fase(p, phi_old, c_old, phi_new, dpp, pr);
    }
    pfo_concentracao_soluto(p, phi_old, c_old, c_new);   /* :154 */
    p->M0 = pfo_auxiliar(p->TP, p->Cl_eq, p->Cs_eq, p->R, p->Vm, p->E0, p->W0, L->Sigma,
                         L->Beta, p->Ke, p->Me, p->D_liq, L->ajuste); /* :157 */
    memcpy(phi_old, phi_new, sizeof(double) * n);        /* :200-226 */
    memcpy(T_old, T_new, sizeof(double) * n);
    memcpy(c_old, c_new, sizeof(double) * n);
  }
  return 0;
}

// pf_common.h -- scalar physics shared by the host code and the gfx950
// kernels: the counter-based noise, h(phi) and friends, the kernel parameter
// block and the boundary-decoration index maps.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pf_capi.h"

#define PF_HD __host__ __device__ __forceinline__

namespace pf {

// Multiplier of the seed term added to the reference hash (0 for seed 0, so
// seed 0 reproduces the reference stream bit for bit).
constexpr uint64_t kSeedMul = 0x9E3779B97F4A7C15ull;

// source/variaveis.f90:82-83: the linear pre-mix of the counter hash.
PF_HD uint64_t hash_base(int32_t i, int32_t j, int32_t k, int32_t it, int32_t salt,
                         uint64_t seed) {
  return (uint64_t)(int64_t)i * 2654435761ull + (uint64_t)(int64_t)j * 2246822519ull +
         (uint64_t)(int64_t)k * 3266489917ull + (uint64_t)(int64_t)it * 668265263ull +
         (uint64_t)(int64_t)salt * 374761393ull + 1ull + seed * kSeedMul;
}
constexpr uint64_t kHashStepK = 3266489917ull;  // hash_base(k+1) - hash_base(k)

// source/variaveis.f90:84-90: fmix64 finaliser and the map to [0,1].
PF_HD double hash_unit(uint64_t h) {
  h ^= h >> 33;
  h *= 0xff51afd7ed558ccdull;
  h ^= h >> 33;
  h *= 0xc4ceb9fe1a85ec53ull;
  h ^= h >> 33;
  // (h & MASK63) / real(MASK63): real(MASK63,8) rounds to 2^63, so the
  // division is an exact scaling.
  return (double)(int64_t)(h & 0x7fffffffffffffffull) * 0x1p-63;
}

PF_HD double pseudo_random(int32_t i, int32_t j, int32_t k, int32_t it, int32_t salt,
                           uint64_t seed) {
  return hash_unit(hash_base(i, j, k, it, salt, seed));
}

// source/variaveis.f90:98-111 (LR_RUIDO=6, TR_RUIDO=20).  Evaluated without
// FMA contraction so that the REAL(4) `ran` it feeds is bit-identical.
PF_HD double ruido_suave(int32_t i, int32_t j, int32_t k, int32_t it, uint64_t seed) {
#pragma clang fp contract(off)
  const int L = 6;
  const int i0 = i / L, j0 = j / L, k0 = k / L, tb = it / 20;
  const double fx = (double)(i % L) / (double)L;
  const double fy = (double)(j % L) / (double)L;
  const double fz = (double)(k % L) / (double)L;
  const double p000 = pseudo_random(i0, j0, k0, tb, 3, seed);
  const double p100 = pseudo_random(i0 + 1, j0, k0, tb, 3, seed);
  const double p010 = pseudo_random(i0, j0 + 1, k0, tb, 3, seed);
  const double p110 = pseudo_random(i0 + 1, j0 + 1, k0, tb, 3, seed);
  const double p001 = pseudo_random(i0, j0, k0 + 1, tb, 3, seed);
  const double p101 = pseudo_random(i0 + 1, j0, k0 + 1, tb, 3, seed);
  const double p011 = pseudo_random(i0, j0 + 1, k0 + 1, tb, 3, seed);
  const double p111 = pseudo_random(i0 + 1, j0 + 1, k0 + 1, tb, 3, seed);
  return (1.0 - fz) * ((1.0 - fy) * ((1.0 - fx) * p000 + fx * p100) +
                       fy * ((1.0 - fx) * p010 + fx * p110)) +
         fz * ((1.0 - fy) * ((1.0 - fx) * p001 + fx * p101) +
               fy * ((1.0 - fx) * p011 + fx * p111));
}

// The REAL(4) `ran` of phase_body.inc:181-186.
PF_HD float ran_real4(int32_t i, int32_t j, int32_t k, int32_t it, int32_t ruido,
                      uint64_t seed) {
  return ruido == PF_RUIDO_SMOOTH ? (float)ruido_suave(i, j, k, it, seed)
                                  : (float)pseudo_random(i, j, k, it, 1, seed);
}

// Boundary decoration of a phase sweep (equacao_campo_fase.f90:151-167): the
// sequential "plane1<-plane2, plane2<-plane3" copies in y, then x, then z
// compose, for n >= 5, into phi(i,j,k) = F(m(i),m(j),m(k)) with
// m(1)=2, m(2)=3, m(n-1)=n-2, m(n)=n-1 and m = identity on [3,n-2], where F is
// the swept field inside [3,n-2]^3 and the pre-sweep field elsewhere.
PF_HD int phase_map(int i, int n) {
  return i <= 2 ? i + 1 : (i >= n - 1 ? i - 1 : i);
}

// Solute boundary (concentracao_soluto.f90:116-124): the faces copy their
// inner neighbour from the post-sweep snapshot; z writes last, then x, then
// y.  Returns the source cell of a face cell (1-based, in place).
PF_HD void solute_source(int &i, int &j, int &k, int nx, int ny, int nz) {
  if (k == 1 || k == nz) {
    k = (k == 1) ? 2 : nz - 1;
  } else if (i == 1 || i == nx) {
    i = (i == 1) ? 2 : nx - 1;
  } else if (j == 1 || j == ny) {
    j = (j == 1) ? 2 : ny - 1;
  }
}

// source/equacao_campo_fase.f90:21-22.
PF_HD void equilibrium(double TP, double Tm, float Ke, double Me, double &Cs_eq, double &Cl_eq) {
#pragma clang fp contract(off)
  Cs_eq = ((Tm - TP) * (double)Ke) / Me;
  Cl_eq = (Tm - TP) / Me;
}

// Terms of auxiliar's integral (below): phi_i runs from 0.11 in rounded
// steps of 0.01 while phi_i <= 0.8999 -- the same additions, hence the same
// count, every time.
constexpr int m0_terms() {
  double phi_i = 0.11;
  int n = 0;
  do {
    ++n;
    phi_i = phi_i + 0.01;
  } while (phi_i <= 0.8999);
  return n;
}
constexpr int kM0Terms = m0_terms();
static_assert(kM0Terms == 79, "auxiliar's integral has 79 terms");

// The integral's abscissae: phi_i from 0.11 in rounded steps of 0.01, the
// same sequential additions as auxiliar.f90's loop, evaluated at compile
// time (IEEE double, round to nearest, as at run time).
struct M0Abscissae {
  double v[kM0Terms];
};
constexpr M0Abscissae m0_abscissae() {
  M0Abscissae t{};
  double phi_i = 0.11;
  for (int n = 0; n < kM0Terms; ++n) {
    t.v[n] = phi_i;
    phi_i = phi_i + 0.01;
  }
  return t;
}

// source/auxiliar.f90:14-29: one term of the integral's sum at phi_i.
PF_HD double m0_term(double phi_i, double Cl_eq, double Cs_eq) {
#pragma clang fp contract(off)
  const double h = phi_i * phi_i * (3.0 - 2.0 * phi_i);
  const double f1 = phi_i * (1.0 - phi_i);
  const double num = h * (1.0 - h);
  const double den = ((1.0 - h) * Cl_eq * (1.0 - Cl_eq) + h * Cs_eq * (1.0 - Cs_eq)) * f1;
  return num / den;
}

// source/auxiliar.f90:31-42: M0 from the sum of the terms (acc, summed in
// the loop's order).
PF_HD double m0_from_sum(double acc, double TP, double Cl_eq, double Cs_eq, double R, double Vm,
                         double E0, double W0, float Sigma, float Beta, float Ke, double Me,
                         double D_liq, float ajuste) {
#pragma clang fp contract(off)
  const double integral = ((double)0.01f / 2.0) * ((double)54.39180136f + 2.0 * acc + 234.073053);
  const double qsi = R * TP * (1.0 / Vm) * (Cl_eq - Cs_eq) * (Cl_eq - Cs_eq) * integral;
  const float omk = 1.0f - Ke;
  const double inv = (E0 * E0 / (double)Sigma) *
                     (((double)Beta * R * TP * (double)omk / (Vm * Me)) +
                      (E0 * qsi / (D_liq * sqrt(2.0 * W0))));
  return (double)ajuste * 1.0 / inv;
}

// source/auxiliar.f90:14-42: trapezoid rule over phi_i = 0.11..0.89 with the
// hard-coded end values, then 1/M0 from the thin-interface relation, in the
// reference's operation order with its REAL(4) literals and (1-Ke).  Shared
// by the host loop and the batched form's per-instance device update (which
// evaluates the terms in parallel, one lane each, and sums them in this
// order), so both give the same bits.
PF_HD double auxiliar_m0(double TP, double Cl_eq, double Cs_eq, double R, double Vm, double E0,
                         double W0, float Sigma, float Beta, float Ke, double Me, double D_liq,
                         float ajuste) {
#pragma clang fp contract(off)
  double acc = 0.0, phi_i = 0.11;
  // with a fixed trip count the compiler overlaps the terms' divisions (the
  // terms are independent; the sum keeps its order)
#pragma unroll 8
  for (int n = 0; n < kM0Terms; ++n) {
    acc = acc + m0_term(phi_i, Cl_eq, Cs_eq);
    phi_i = phi_i + 0.01;
  }
  return m0_from_sum(acc, TP, Cl_eq, Cs_eq, R, Vm, E0, W0, Sigma, Beta, Ke, Me, D_liq, ajuste);
}

PF_HD size_t cidx(int i, int j, int k, int nx, int ny) {
  return (size_t)(i - 1) + (size_t)nx * ((size_t)(j - 1) + (size_t)ny * (size_t)(k - 1));
}
__host__ __device__ inline long long shell_count(int nx, int ny, int nz, int w, bool xf = true) {
  return 2LL * w * nx * ny + (2LL * w * nx + (xf ? 2LL * w * (ny - 2 * w) : 0LL)) * (nz - 2 * w);
}

// shell_flat's index arithmetic in the integer type I (the same values in
// either type when they fit it).
template <typename I>
__host__ __device__ __forceinline__ void shell_flat_in(I t, I nx, I ny, I nz, I w, int &i, int &j,
                                                      int &k, bool xf) {
  const I plane = nx * ny;
  const I full = w * plane;  // the w planes at one z face
  const I per = 2 * w * nx + (xf ? 2 * w * (ny - 2 * w) : 0);  // one middle plane
  const I mid = per * (nz - 2 * w);
  if (t < full || t >= full + mid) {
    const I u = t < full ? t : t - full - mid;
    const I kk = u / plane;
    k = (int)(t < full ? kk + 1 : nz - w + 1 + kk);
    const I r = u - kk * plane;
    const I jj = r / nx;
    i = (int)(r - jj * nx) + 1;
    j = (int)jj + 1;
    return;
  }
  const I u = t - full;
  const I km = u / per;
  k = (int)(w + 1 + km);
  const I r = u - km * per;
  const I rows = 2 * w * nx;
  if (r < rows) {
    const I q = r / nx;
    i = (int)(r - q * nx) + 1;
    j = (int)(q < w ? q + 1 : ny - (2 * w - 1 - q));
    return;
  }
  const I q = r - rows;
  const I qq = q / (2 * w);
  const I col = q - qq * (2 * w);
  j = (int)(qq + w + 1);
  i = (int)(col < w ? col + 1 : nx - (2 * w - 1 - col));
}

// Cell t (0-based) of the boundary shell of thickness w (2 for the phase
// decoration, 1 for the solute faces), enumerated plane by plane: the w
// planes at each z face are whole; every other plane contributes its w rows
// at each y face (x-contiguous) and, if xf, the w cells at each x face of its
// other rows (without them: the cells the fused sweeps write themselves).
// Returns false when t is past the shell.  In 32-bit unsigned arithmetic
// whenever the shell has fewer than 2^31 cells (a 64-bit division is a
// long software sequence on the GPU; the indices are the same either way).
__host__ __device__ __forceinline__ bool shell_flat(long long t, int nx, int ny, int nz, int w, int &i,
                                           int &j, int &k, bool xf = true) {
  const long long total = shell_count(nx, ny, nz, w, xf);
  if (t < 0 || t >= total) return false;
  if (total < (1LL << 31))
    shell_flat_in<uint32_t>((uint32_t)t, (uint32_t)nx, (uint32_t)ny, (uint32_t)nz, (uint32_t)w, i,
                            j, k, xf);
  else
    shell_flat_in<long long>(t, nx, ny, nz, w, i, j, k, xf);
  return true;
}

// One cell of the phase decoration (equacao_campo_fase.f90:151-167 in
// closed form): the value of its mapped source cell, from the swept field
// inside [3,n-2]^3 and from the pre-sweep field elsewhere.
template <typename R>
PF_HD void phase_deco_cell(int i, int j, int k, int nx, int ny, int nz, const R *po, R *pn) {
  const int si = phase_map(i, nx), sj = phase_map(j, ny), sk = phase_map(k, nz);
  const bool inner = si >= 3 && si <= nx - 2 && sj >= 3 && sj <= ny - 2 && sk >= 3 && sk <= nz - 2;
  const size_t s = cidx(si, sj, sk, nx, ny);
  pn[cidx(i, j, k, nx, ny)] = inner ? pn[s] : po[s];
}

// One cell of the solute face copy (concentracao_soluto.f90:116-124): the
// inner neighbour's new value, or the snapshot's where the source is itself
// on the boundary (edges, z > x > y priority).
template <typename R>
PF_HD void solute_deco_cell(int i, int j, int k, int nx, int ny, int nz, const R *co, R *cn) {
  int si = i, sj = j, sk = k;
  solute_source(si, sj, sk, nx, ny, nz);
  const bool inner = si > 1 && si < nx && sj > 1 && sj < ny && sk > 1 && sk < nz;
  const size_t s = cidx(si, sj, sk, nx, ny);
  cn[cidx(i, j, k, nx, ny)] = inner ? cn[s] : co[s];
}

// Kernel parameter block, passed by value.  The scalars each launch needs,
// already combined the way the reference combines them.
struct KParams {
  int32_t nx, ny, nz, it;
  int32_t h_quintico, ruido_tipo;
  uint64_t seed;
  double dx, dy, dz;
  double dt, D, delta_H, cp_s, TP;
  double R, Vm, E0, W0, M0, delta_eps, AAA, D_liq, D_sol;
  double Ke;  // (double)(float)Ke
  double Cs_eq, Cl_eq;
  // The FP32-mixed form stores c as c - c_off (c_off: a liquid value of
  // the initial field), which keeps the solute increments of the near-
  // uniform liquid that an FP32 c would round away; 0 for FP64 storage.
  double c_off;
};

// The kernel scalars of a pf_params (Cs_eq/Cl_eq recomputed from TP, as
// equacao_campo_fase.f90:21-22 does for every sweep).
inline KParams kparams_from(const pf_params &p) {
  KParams k;
  k.nx = p.nx; k.ny = p.ny; k.nz = p.nz; k.it = p.it;
  k.h_quintico = p.h_quintico; k.ruido_tipo = p.ruido_tipo; k.seed = p.seed;
  k.dx = p.dx; k.dy = p.dy; k.dz = p.dz;
  k.dt = p.dt; k.D = p.D; k.delta_H = p.delta_H; k.cp_s = p.cp_s; k.TP = p.TP;
  k.R = p.R; k.Vm = p.Vm; k.E0 = p.E0; k.W0 = p.W0; k.M0 = p.M0; k.delta_eps = p.delta_eps;
  k.AAA = p.AAA; k.D_liq = p.D_liq; k.D_sol = p.D_sol;
  k.Ke = (double)p.Ke;
  equilibrium(p.TP, p.Tm, p.Ke, p.Me, k.Cs_eq, k.Cl_eq);
  k.c_off = 0.0;  // set by the owner of the stored field
  return k;
}

// Records the calling thread's last error message (pf_last_error) and
// returns code.  Defined in pf_runtime.hip.
int fail_msg(int code, const char *fmt, ...);

}  // namespace pf

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
};

}  // namespace pf

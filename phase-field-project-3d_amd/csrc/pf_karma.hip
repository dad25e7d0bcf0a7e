// The quantitative Karma model (karma/karma3d.f90) on gfx950: phi and U in
// HBM as padded (n+2)^3 FP64 cubes (cells 1..n, mirror ghosts 0 and n+1),
// four launches per step on the model's own stream:
//   ghosts of phi and U -> phase equation (dphi, clamped phi of the next
//   step) -> ghosts of dphi -> solute equation with the anti-trapping current
//   (U of the next step), then the two buffer pairs swap.
// One thread per cell, x fastest (32x8 threads per work-group).  Every
// expression keeps the reference's operation order with FP contraction off
// (division and square root correctly rounded), so each cell rounds as a
// plain IEEE-754 C build of the same expressions does.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include "pf_capi.h"
#include "pf_common.h"
#include "pf_karma.h"

namespace {

constexpr int kBX = 32, kBY = 8;
constexpr double kA1 = 0.8839, kA2 = 0.6267;          // karma3d.f90:91
constexpr double kAt = 0.35355339059327373;          // 1/(2 sqrt 2), karma3d.f90:92
constexpr int kMaxN = 1200;                          // (n+2)^3 < 2^31: 32-bit indices

struct KStep {
  int n, s, ss;  // cells per edge, padded edge, padded plane
  double idx, eps4, lambda, dt, Fnoise, k1, kp, D, sqdt;
  int it;
};

__device__ __forceinline__ int clampc(int v, int n) { return v < 1 ? 1 : (v > n ? n : v); }

// espelho (karma3d.f90:318-339) for one or two fields: the x, y, z passes
// in that order compose to ghost(i,j,k) = f(clamp(i), clamp(j), clamp(k)),
// so every ghost cell reads interior cells only and the axes need no order
// (cells on ghost edges are written with the same value by each axis).
__global__ void __launch_bounds__(kBX *kBY)
    k_karma_mirror(KStep g, double *__restrict__ f1, double *__restrict__ f2) {
  const int u = blockIdx.x * kBX + threadIdx.x, v = blockIdx.y * kBY + threadIdx.y;
  const int n = g.n;
  if (u > n + 1 || v > n + 1) return;
  const int ax = blockIdx.z;
  for (int side = 0; side < 2; ++side) {
    const int w = side ? n + 1 : 0;
    int i, j, k;
    if (ax == 0) { i = w; j = u; k = v; }
    else if (ax == 1) { i = u; j = w; k = v; }
    else { i = u; j = v; k = w; }
    const int dst = i + g.s * j + g.ss * k;
    const int src = clampc(i, n) + g.s * clampc(j, n) + g.ss * clampc(k, n);
    f1[dst] = f1[src];
    if (f2) f2[dst] = f2[src];
  }
}

// fluxn (karma3d.f90:36-47)
__device__ __forceinline__ double fluxn(double g1, double g2, double g3, double eps4) {
#pragma clang fp contract(off)
  const double g2s = g1 * g1 + g2 * g2 + g3 * g3;
  const double a0 = 1.0 - 3.0 * eps4;
  if (g2s < 1e-20) return a0 * a0 * g1;
  const double q1 = g1 * g1, q2 = g2 * g2, q3 = g3 * g3;
  const double s4 = q1 * q1 + q2 * q2 + q3 * q3;
  const double as = a0 + 4.0 * eps4 * s4 / (g2s * g2s);
  return as * (as * g1 + 16.0 * eps4 * (q1 * g1 - g1 * s4 / g2s) / g2s);
}

// Face gradient between cell c and its neighbour c + o (o = +-stride of the
// face's axis): the normal difference and the two tangential centred
// differences averaged over the face's cells, along e1 then e2 (the other
// two axes in x, y, z order) (karma3d.f90:170-187, 215-268).
__device__ __forceinline__ void face(const double *__restrict__ P, int c, int o, bool plus, int e1,
                                     int e2, double idx, double &gn, double &t1, double &t2) {
#pragma clang fp contract(off)
  const double pc = P[c], pn = P[c + o];
  gn = plus ? (pn - pc) * idx : (pc - pn) * idx;
  t1 = 0.25 * idx * (P[c + e1] + P[c + o + e1] - P[c - e1] - P[c + o - e1]);
  t2 = 0.25 * idx * (P[c + e2] + P[c + o + e2] - P[c - e2] - P[c + o - e2]);
}

// ruido (karma3d.f90:50-64)
__device__ __forceinline__ double ruido(int i, int j, int k, int it) {
#pragma clang fp contract(off)
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

// Phase equation (karma3d.f90:164-207): dphi and the clamped phi of the next
// step (karma3d.f90:280-289; the unclamped value is not read elsewhere).
__global__ void __launch_bounds__(kBX *kBY)
    k_karma_phase(KStep g, const double *__restrict__ P, const double *__restrict__ U,
                  double *__restrict__ dphi, double *__restrict__ Pn) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * kBX + threadIdx.x + 1, j = blockIdx.y * kBY + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  if (i > g.n || j > g.n) return;
  const int sx = 1, sy = g.s, sz = g.ss;
  const int c = i + sy * j + sz * k;
  const double idx = g.idx, pp = P[c];
  double gn, t1, t2;
  face(P, c, sx, true, sy, sz, idx, gn, t1, t2);
  const double Jxp = fluxn(gn, t1, t2, g.eps4);
  face(P, c, -sx, false, sy, sz, idx, gn, t1, t2);
  const double Jxm = fluxn(gn, t1, t2, g.eps4);
  face(P, c, sy, true, sx, sz, idx, gn, t1, t2);
  const double Jyp = fluxn(gn, t1, t2, g.eps4);
  face(P, c, -sy, false, sx, sz, idx, gn, t1, t2);
  const double Jym = fluxn(gn, t1, t2, g.eps4);
  face(P, c, sz, true, sx, sy, idx, gn, t1, t2);
  const double Jzp = fluxn(gn, t1, t2, g.eps4);
  face(P, c, -sz, false, sx, sy, idx, gn, t1, t2);
  const double Jzm = fluxn(gn, t1, t2, g.eps4);
  const double divJ = (Jxp - Jxm + Jyp - Jym + Jzp - Jzm) * idx;
  const double g1 = 0.5 * idx * (P[c + sx] - P[c - sx]);
  const double g2 = 0.5 * idx * (P[c + sy] - P[c - sy]);
  const double g3 = 0.5 * idx * (P[c + sz] - P[c - sz]);
  const double gq = g1 * g1 + g2 * g2 + g3 * g3;
  const double a0 = 1.0 - 3.0 * g.eps4;
  double as2;
  if (gq > 1e-20) {
    const double s4 = ((g1 * g1) * (g1 * g1) + (g2 * g2) * (g2 * g2) + (g3 * g3) * (g3 * g3)) /
                      (gq * gq);
    const double as = a0 + 4.0 * g.eps4 * s4;
    as2 = as * as;
  } else {
    as2 = a0 * a0;
  }
  const double om = 1.0 - pp * pp;
  double dp = (divJ + pp - pp * pp * pp - g.lambda * (om * om) * U[c]) / as2;
  if (g.Fnoise != 0.0) dp += g.Fnoise * ruido(i, j, k, g.it) * om / g.sqdt;
  dphi[c] = dp;
  const double v = pp + g.dt * dp;
  Pn[c] = v < -1.0 ? -1.0 : (v > 1.0 ? 1.0 : v);
}

// One face's solute flux (karma3d.f90:215-268): diffusion D q (U_X - U_P)
// plus the anti-trapping current a_t cf n.
__device__ __forceinline__ double uflux(const double *__restrict__ P, const double *__restrict__ U,
                                        const double *__restrict__ dphi, int c, int o, bool plus,
                                        int ax, int e1, int e2, double pp, double uc, double cc,
                                        const KStep &g) {
#pragma clang fp contract(off)
  double gn, t1, t2;
  face(P, c, o, plus, e1, e2, g.idx, gn, t1, t2);
  // |grad phi| over x, y, z in that order
  double gx, gy, gz;
  if (ax == 0) { gx = gn; gy = t1; gz = t2; }
  else if (ax == 1) { gx = t1; gy = gn; gz = t2; }
  else { gx = t1; gy = t2; gz = gn; }
  const double gm = sqrt(gx * gx + gy * gy + gz * gz);
  const int nb = c + o;
  const double qf = 0.25 * (2.0 - pp - P[nb]);
  const double cf = 0.5 * (cc + (1.0 + g.k1 * U[nb]) * dphi[nb]);
  double F = plus ? g.D * qf * (U[nb] - uc) * g.idx : g.D * qf * (uc - U[nb]) * g.idx;
  if (gm > 1e-10) F = F + kAt * cf * gn / gm;
  return F;
}

// Solute equation (karma3d.f90:211-278): U of the next step.
__global__ void __launch_bounds__(kBX *kBY)
    k_karma_solute(KStep g, const double *__restrict__ P, const double *__restrict__ U,
                   const double *__restrict__ dphi, double *__restrict__ Un) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * kBX + threadIdx.x + 1, j = blockIdx.y * kBY + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  if (i > g.n || j > g.n) return;
  const int sx = 1, sy = g.s, sz = g.ss;
  const int c = i + sy * j + sz * k;
  const double pp = P[c], uc = U[c];
  const double cc = (1.0 + g.k1 * uc) * dphi[c];
  const double Fxp = uflux(P, U, dphi, c, sx, true, 0, sy, sz, pp, uc, cc, g);
  const double Fxm = uflux(P, U, dphi, c, -sx, false, 0, sy, sz, pp, uc, cc, g);
  const double Fyp = uflux(P, U, dphi, c, sy, true, 1, sx, sz, pp, uc, cc, g);
  const double Fym = uflux(P, U, dphi, c, -sy, false, 1, sx, sz, pp, uc, cc, g);
  const double Fzp = uflux(P, U, dphi, c, sz, true, 2, sx, sy, pp, uc, cc, g);
  const double Fzm = uflux(P, U, dphi, c, -sz, false, 2, sx, sy, pp, uc, cc, g);
  const double Au = 0.5 * (1.0 + g.kp) - 0.5 * g.k1 * pp;
  const double rhs =
      (Fxp - Fxm + Fyp - Fym + Fzp - Fzm) * g.idx + 0.5 * (1.0 + g.k1 * uc) * dphi[c];
  Un[c] = uc + g.dt * rhs / Au;
}

}  // namespace

struct pf_karma {
  pf_karma_params p;
  int n = 0, s = 0, ss = 0;
  size_t cells = 0;  // padded
  double D = 0, Omega = 0;
  double *phi[2] = {nullptr, nullptr}, *U[2] = {nullptr, nullptr}, *dphi = nullptr;
  int cur = 0, it = 0;
  hipStream_t st = nullptr;
};

namespace {

void karma_free(pf_karma *s) {
  for (int b = 0; b < 2; ++b) {
    if (s->phi[b]) (void)hipFree(s->phi[b]);
    if (s->U[b]) (void)hipFree(s->U[b]);
  }
  if (s->dphi) (void)hipFree(s->dphi);
  if (s->st) (void)hipStreamDestroy(s->st);
  delete s;
}

KStep kstep(const pf_karma *s) {
  KStep g;
  g.n = s->n; g.s = s->s; g.ss = s->ss;
  g.idx = 1.0 / s->p.dx;
  g.eps4 = s->p.eps4; g.lambda = s->p.lambda; g.dt = s->p.dt; g.Fnoise = s->p.Fnoise;
  g.k1 = 1.0 - s->p.k; g.kp = s->p.k; g.D = s->D;
  g.sqdt = std::sqrt(s->p.dt);
  g.it = s->it;
  return g;
}

}  // namespace

extern "C" int pf_karma_create(const pf_karma_params *p, pf_karma **out) {
  if (!p || !out) return pf::fail_msg(PF_E_ARG, "pf_karma_create: null argument");
  *out = nullptr;
  if (p->n < 2 || p->n > kMaxN)
    return pf::fail_msg(PF_E_ARG, "pf_karma_create: n = %d outside [2, %d]", p->n, kMaxN);
  if (!(p->dx > 0) || !(p->dt > 0) || !(p->mliq != 0) || !(p->Tm != p->TP) || !(p->k != 1.0))
    return pf::fail_msg(PF_E_ARG, "pf_karma_create: dx, dt must be > 0; mliq != 0, Tm != TP, k != 1");
  pf_karma *s = new pf_karma;
  s->p = *p;
  s->n = p->n; s->s = p->n + 2; s->ss = s->s * s->s;
  s->cells = (size_t)s->ss * s->s;
  // karma3d.f90:110-112
  s->D = kA2 * p->lambda;
  const double cl0 = (p->Tm - p->TP) / p->mliq;
  s->Omega = (cl0 - p->C0) / ((1.0 - p->k) * cl0);
  (void)kA1;
  const size_t bytes = s->cells * sizeof(double);
  if (hipStreamCreate(&s->st) != hipSuccess) {
    karma_free(s);
    return pf::fail_msg(PF_E_HIP, "pf_karma_create: hipStreamCreate failed");
  }
  for (int b = 0; b < 2; ++b)
    if (hipMalloc((void **)&s->phi[b], bytes) != hipSuccess ||
        hipMalloc((void **)&s->U[b], bytes) != hipSuccess) {
      karma_free(s);
      return pf::fail_msg(PF_E_NOMEM, "pf_karma_create: %zu bytes per field", bytes);
    }
  if (hipMalloc((void **)&s->dphi, bytes) != hipSuccess) {
    karma_free(s);
    return pf::fail_msg(PF_E_NOMEM, "pf_karma_create: %zu bytes per field", bytes);
  }
  // seed (karma3d.f90:130-144), ghosts included, on the host
  std::vector<double> ph(s->cells), u(s->cells, -s->Omega);
  const int n = s->n;
  for (int k = 0; k <= n + 1; ++k)
    for (int j = 0; j <= n + 1; ++j)
      for (int i = 0; i <= n + 1; ++i) {
        const double x = (i - 0.5) * p->dx, y = (j - 0.5) * p->dx, z = (k - 0.5) * p->dx;
        const double r = std::sqrt(x * x + y * y + z * z);
        ph[(size_t)i + (size_t)s->s * j + (size_t)s->ss * k] = -std::tanh((r - p->R0) / std::sqrt(2.0));
      }
  if (hipMemcpy(s->phi[0], ph.data(), bytes, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(s->U[0], u.data(), bytes, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(s->dphi, 0, bytes) != hipSuccess) {
    karma_free(s);
    return pf::fail_msg(PF_E_HIP, "pf_karma_create: upload failed");
  }
  *out = s;
  return PF_OK;
}

extern "C" int pf_karma_run(pf_karma *s, int32_t nsteps) {
  if (!s || nsteps < 0) return pf::fail_msg(PF_E_ARG, "pf_karma_run: bad argument");
  const int n = s->n;
  const dim3 blk(kBX, kBY, 1);
  const dim3 gm((n + 2 + kBX - 1) / kBX, (n + 2 + kBY - 1) / kBY, 3);
  const dim3 gc((n + kBX - 1) / kBX, (n + kBY - 1) / kBY, n);
  for (int t = 0; t < nsteps; ++t) {
    s->it += 1;
    const KStep g = kstep(s);
    const int a = s->cur, b = 1 - s->cur;
    hipLaunchKernelGGL(k_karma_mirror, gm, blk, 0, s->st, g, s->phi[a], s->U[a]);
    hipLaunchKernelGGL(k_karma_phase, gc, blk, 0, s->st, g, s->phi[a], s->U[a], s->dphi, s->phi[b]);
    hipLaunchKernelGGL(k_karma_mirror, gm, blk, 0, s->st, g, s->dphi, (double *)nullptr);
    hipLaunchKernelGGL(k_karma_solute, gc, blk, 0, s->st, g, s->phi[a], s->U[a], s->dphi, s->U[b]);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess)
      return pf::fail_msg(PF_E_HIP, "pf_karma_run: launch failed: %s", hipGetErrorString(e));
    s->cur = b;
  }
  return PF_OK;
}

extern "C" int pf_karma_download(pf_karma *s, double *phi, double *U, int32_t *it) {
  if (!s) return pf::fail_msg(PF_E_ARG, "pf_karma_download: null context");
  const int n = s->n;
  std::vector<double> tmp(s->cells);
  double *dst[2] = {phi, U};
  double *src[2] = {s->phi[s->cur], s->U[s->cur]};
  for (int f = 0; f < 2; ++f) {
    if (!dst[f]) continue;
    if (hipMemcpyAsync(tmp.data(), src[f], s->cells * sizeof(double), hipMemcpyDeviceToHost,
                       s->st) != hipSuccess ||
        hipStreamSynchronize(s->st) != hipSuccess)
      return pf::fail_msg(PF_E_HIP, "pf_karma_download: copy failed");
    size_t q = 0;
    for (int k = 1; k <= n; ++k)
      for (int j = 1; j <= n; ++j)
        for (int i = 1; i <= n; ++i) dst[f][q++] = tmp[(size_t)i + (size_t)s->s * j + (size_t)s->ss * k];
  }
  if (it) *it = s->it;
  return PF_OK;
}

extern "C" int pf_karma_synchronize(pf_karma *s) {
  if (!s) return pf::fail_msg(PF_E_ARG, "pf_karma_synchronize: null context");
  const hipError_t e = hipStreamSynchronize(s->st);
  if (e != hipSuccess) return pf::fail_msg(PF_E_HIP, "pf_karma_synchronize: %s", hipGetErrorString(e));
  return PF_OK;
}

extern "C" int pf_karma_destroy(pf_karma *s) {
  if (!s) return PF_OK;
  const hipError_t e = hipStreamSynchronize(s->st);
  karma_free(s);
  if (e != hipSuccess) return pf::fail_msg(PF_E_HIP, "pf_karma_destroy: %s", hipGetErrorString(e));
  return PF_OK;
}

// The quantitative Karma model (karma/karma3d.f90) on gfx950: phi and U in
// HBM as padded (n+2)^3 FP64 cubes (cells 1..n, mirror ghosts 0 and n+1),
// two launches per step on the model's own stream:
//   phase equation (dphi and the clamped phi of the next step, ghosts
//   included) -> solute equation with the anti-trapping current (U of the
//   next step, ghosts included), then the two buffer pairs swap.
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

// espelho (karma3d.f90:318-339): its x, y, z passes compose to
// ghost(i,j,k) = f(clamp(i), clamp(j), clamp(k)), so the thread that writes a
// boundary cell's new value also writes it to every ghost position that
// clamps to that cell (up to 7 for a corner).  The fields of the next step
// therefore arrive with their ghosts set and no separate mirror pass runs;
// no kernel reads the buffer it writes.
__device__ __forceinline__ void put_ghosts(double *__restrict__ f, const KStep &g, int i, int j,
                                           int k, int c, double v) {
  const int di = i == 1 ? -1 : (i == g.n ? 1 : 0);
  const int dj = j == 1 ? -g.s : (j == g.n ? g.s : 0);
  const int dk = k == 1 ? -g.ss : (k == g.n ? g.ss : 0);
  if ((di | dj | dk) == 0) return;
  for (int a = 0; a <= (di != 0); ++a)
    for (int b = 0; b <= (dj != 0); ++b)
      for (int e = 0; e <= (dk != 0); ++e)
        if (a + b + e) f[c + a * di + b * dj + e * dk] = v;
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

// A field around one cell: f(dx, dy, dz) is its value at the neighbour
// (i+dx, j+dy, k+dz), |dx|, |dy|, |dz| <= 1 (the offsets are compile-time
// constants once inlined).  GlobalNb reads the padded cube in HBM; PlaneNb
// reads three LDS planes k-1, k, k+1 of a tile with a one-cell halo.
struct GlobalNb {
  const double *__restrict__ f;
  int c, s, ss;
  __device__ __forceinline__ double operator()(int dx, int dy, int dz) const {
    return f[c + dx + s * dy + ss * dz];
  }
};

constexpr int kTX = 32, kTY = 8;                      // z-march tile (cells)
constexpr int kLX = kTX + 2, kLY = kTY + 2, kLP = kLX * kLY;  // with the halo ring

struct PlaneNb {
  const double *lo, *mid, *hi;  // planes k-1, k, k+1
  int q;                        // the cell's index in a plane
  __device__ __forceinline__ double operator()(int dx, int dy, int dz) const {
    const double *p = dz < 0 ? lo : (dz > 0 ? hi : mid);
    return p[q + dx + kLX * dy];
  }
};

// Face gradient on the face between the cell and its neighbour along axis
// AX (PLUS: +1, else -1): the normal difference and the two tangential
// centred differences averaged over the face's two cells, along the other
// two axes in x, y, z order (karma3d.f90:170-187, 215-268).
template <int AX, bool PLUS, class F>
__device__ __forceinline__ void face(const F &P, double idx, double &gn, double &t1, double &t2) {
#pragma clang fp contract(off)
  constexpr int sg = PLUS ? 1 : -1;
  constexpr int A1 = AX == 0 ? 1 : 0, A2 = AX == 2 ? 1 : 2;
  constexpr int nx = (AX == 0) * sg, ny = (AX == 1) * sg, nz = (AX == 2) * sg;
  constexpr int x1 = A1 == 0, y1 = A1 == 1, z1 = A1 == 2;
  constexpr int x2 = A2 == 0, y2 = A2 == 1, z2 = A2 == 2;
  const double pc = P(0, 0, 0), pn = P(nx, ny, nz);
  gn = PLUS ? (pn - pc) * idx : (pc - pn) * idx;
  t1 = 0.25 * idx *
       (P(x1, y1, z1) + P(nx + x1, ny + y1, nz + z1) - P(-x1, -y1, -z1) - P(nx - x1, ny - y1, nz - z1));
  t2 = 0.25 * idx *
       (P(x2, y2, z2) + P(nx + x2, ny + y2, nz + z2) - P(-x2, -y2, -z2) - P(nx - x2, ny - y2, nz - z2));
}

template <int AX, bool PLUS, class F>
__device__ __forceinline__ double flux_phi(const F &P, const KStep &g) {
  double gn, t1, t2;
  face<AX, PLUS>(P, g.idx, gn, t1, t2);
  return fluxn(gn, t1, t2, g.eps4);
}

// Phase equation at one cell (karma3d.f90:164-207): dphi.
template <class F>
__device__ __forceinline__ double phase_cell(const F &P, double uc, int i, int j, int k,
                                             const KStep &g) {
#pragma clang fp contract(off)
  const double idx = g.idx, pp = P(0, 0, 0);
  const double Jxp = flux_phi<0, true>(P, g), Jxm = flux_phi<0, false>(P, g);
  const double Jyp = flux_phi<1, true>(P, g), Jym = flux_phi<1, false>(P, g);
  const double Jzp = flux_phi<2, true>(P, g), Jzm = flux_phi<2, false>(P, g);
  const double divJ = (Jxp - Jxm + Jyp - Jym + Jzp - Jzm) * idx;
  const double g1 = 0.5 * idx * (P(1, 0, 0) - P(-1, 0, 0));
  const double g2 = 0.5 * idx * (P(0, 1, 0) - P(0, -1, 0));
  const double g3 = 0.5 * idx * (P(0, 0, 1) - P(0, 0, -1));
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
  double dp = (divJ + pp - pp * pp * pp - g.lambda * (om * om) * uc) / as2;
  if (g.Fnoise != 0.0) dp += g.Fnoise * ruido(i, j, k, g.it) * om / g.sqdt;
  return dp;
}

// One face's solute flux (karma3d.f90:215-268): diffusion D q (U_X - U_P)
// plus the anti-trapping current a_t cf n.
template <int AX, bool PLUS, class F>
__device__ __forceinline__ double uflux(const F &P, const F &U, const F &Dp, double pp, double uc,
                                        double cc, const KStep &g) {
#pragma clang fp contract(off)
  constexpr int sg = PLUS ? 1 : -1;
  constexpr int nx = (AX == 0) * sg, ny = (AX == 1) * sg, nz = (AX == 2) * sg;
  double gn, t1, t2;
  face<AX, PLUS>(P, g.idx, gn, t1, t2);
  // |grad phi| over x, y, z in that order
  double gx, gy, gz;
  if (AX == 0) { gx = gn; gy = t1; gz = t2; }
  else if (AX == 1) { gx = t1; gy = gn; gz = t2; }
  else { gx = t1; gy = t2; gz = gn; }
  const double gm = sqrt(gx * gx + gy * gy + gz * gz);
  const double un = U(nx, ny, nz);
  const double qf = 0.25 * (2.0 - pp - P(nx, ny, nz));
  const double cf = 0.5 * (cc + (1.0 + g.k1 * un) * Dp(nx, ny, nz));
  double Fl = PLUS ? g.D * qf * (un - uc) * g.idx : g.D * qf * (uc - un) * g.idx;
  if (gm > 1e-10) Fl = Fl + kAt * cf * gn / gm;
  return Fl;
}

// Solute equation at one cell (karma3d.f90:211-278): U of the next step.
template <class F>
__device__ __forceinline__ double solute_cell(const F &P, const F &U, const F &Dp, const KStep &g) {
#pragma clang fp contract(off)
  const double pp = P(0, 0, 0), uc = U(0, 0, 0), dc = Dp(0, 0, 0);
  const double cc = (1.0 + g.k1 * uc) * dc;
  const double Fxp = uflux<0, true>(P, U, Dp, pp, uc, cc, g);
  const double Fxm = uflux<0, false>(P, U, Dp, pp, uc, cc, g);
  const double Fyp = uflux<1, true>(P, U, Dp, pp, uc, cc, g);
  const double Fym = uflux<1, false>(P, U, Dp, pp, uc, cc, g);
  const double Fzp = uflux<2, true>(P, U, Dp, pp, uc, cc, g);
  const double Fzm = uflux<2, false>(P, U, Dp, pp, uc, cc, g);
  const double Au = 0.5 * (1.0 + g.kp) - 0.5 * g.k1 * pp;
  const double rhs = (Fxp - Fxm + Fyp - Fym + Fzp - Fzm) * g.idx + 0.5 * (1.0 + g.k1 * uc) * dc;
  return uc + g.dt * rhs / Au;
}

// Stores a new phi (clamped, karma3d.f90:280-289) and dphi of cell c.
__device__ __forceinline__ void put_phase(const KStep &g, int i, int j, int k, int c, double pp,
                                          double dp, double *__restrict__ dphi,
                                          double *__restrict__ Pn) {
#pragma clang fp contract(off)
  const double v = pp + g.dt * dp;
  const double vc = v < -1.0 ? -1.0 : (v > 1.0 ? 1.0 : v);
  dphi[c] = dp;
  Pn[c] = vc;
  put_ghosts(dphi, g, i, j, k, c, dp);
  put_ghosts(Pn, g, i, j, k, c, vc);
}

// One thread per cell, neighbours read from HBM through the caches.
__global__ void __launch_bounds__(kBX *kBY)
    k_karma_phase(KStep g, const double *__restrict__ P, const double *__restrict__ U,
                  double *__restrict__ dphi, double *__restrict__ Pn) {
  const int i = blockIdx.x * kBX + threadIdx.x + 1, j = blockIdx.y * kBY + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  if (i > g.n || j > g.n) return;
  const int c = i + g.s * j + g.ss * k;
  const GlobalNb Pv{P, c, g.s, g.ss};
  put_phase(g, i, j, k, c, P[c], phase_cell(Pv, U[c], i, j, k, g), dphi, Pn);
}

__global__ void __launch_bounds__(kBX *kBY)
    k_karma_solute(KStep g, const double *__restrict__ P, const double *__restrict__ U,
                   const double *__restrict__ dphi, double *__restrict__ Un) {
  const int i = blockIdx.x * kBX + threadIdx.x + 1, j = blockIdx.y * kBY + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  if (i > g.n || j > g.n) return;
  const int c = i + g.s * j + g.ss * k;
  const double un = solute_cell(GlobalNb{P, c, g.s, g.ss}, GlobalNb{U, c, g.s, g.ss},
                                GlobalNb{dphi, c, g.s, g.ss}, g);
  Un[c] = un;
  put_ghosts(Un, g, i, j, k, c, un);
}

// z-march forms: a work-group owns a kTX x kTY column of cells over kKC
// planes and keeps the planes k-1, k, k+1 of its inputs (tile plus a
// one-cell halo ring, ghosts included) in LDS, loading one new plane per
// step of the march, so every input value is read from HBM about once.
constexpr int kKC = 8;

// Loads padded plane kk of the tile's halo region of up to three fields.
__device__ __forceinline__ void load_plane(const KStep &g, int i0, int j0, int kk,
                                           const double *__restrict__ a, double *sa,
                                           const double *__restrict__ b, double *sb,
                                           const double *__restrict__ d, double *sd) {
  const int t = threadIdx.x + kTX * threadIdx.y;
  for (int q = t; q < kLP; q += kTX * kTY) {
    const int i = i0 + q % kLX, j = j0 + q / kLX;  // padded coordinates
    if (i <= g.n + 1 && j <= g.n + 1) {
      const int c = i + g.s * j + g.ss * kk;
      sa[q] = a[c];
      if (b) sb[q] = b[c];
      if (d) sd[q] = d[c];
    }
  }
}

__global__ void __launch_bounds__(kTX *kTY)
    k_karma_phase_zm(KStep g, const double *__restrict__ P, const double *__restrict__ U,
                     double *__restrict__ dphi, double *__restrict__ Pn) {
  __shared__ double sP[3][kLP];
  const int i0 = blockIdx.x * kTX, j0 = blockIdx.y * kTY;  // padded origin of the halo region
  const int k0 = blockIdx.z * kKC + 1, k1 = min(k0 + kKC - 1, g.n);
  const int i = i0 + threadIdx.x + 1, j = j0 + threadIdx.y + 1;
  const bool own = i <= g.n && j <= g.n;
  const int q = threadIdx.x + 1 + kLX * (threadIdx.y + 1);
  load_plane(g, i0, j0, k0 - 1, P, sP[0], nullptr, nullptr, nullptr, nullptr);
  load_plane(g, i0, j0, k0, P, sP[1], nullptr, nullptr, nullptr, nullptr);
  int lo = 0, mid = 1, hi = 2;
  for (int k = k0; k <= k1; ++k) {
    load_plane(g, i0, j0, k + 1, P, sP[hi], nullptr, nullptr, nullptr, nullptr);
    __syncthreads();
    if (own) {
      const int c = i + g.s * j + g.ss * k;
      const PlaneNb Pv{sP[lo], sP[mid], sP[hi], q};
      put_phase(g, i, j, k, c, Pv(0, 0, 0), phase_cell(Pv, U[c], i, j, k, g), dphi, Pn);
    }
    __syncthreads();  // plane lo is overwritten next
    const int t = lo; lo = mid; mid = hi; hi = t;
  }
}

__global__ void __launch_bounds__(kTX *kTY)
    k_karma_solute_zm(KStep g, const double *__restrict__ P, const double *__restrict__ U,
                      const double *__restrict__ dphi, double *__restrict__ Un) {
  __shared__ double sP[3][kLP], sU[3][kLP], sD[3][kLP];
  const int i0 = blockIdx.x * kTX, j0 = blockIdx.y * kTY;
  const int k0 = blockIdx.z * kKC + 1, k1 = min(k0 + kKC - 1, g.n);
  const int i = i0 + threadIdx.x + 1, j = j0 + threadIdx.y + 1;
  const bool own = i <= g.n && j <= g.n;
  const int q = threadIdx.x + 1 + kLX * (threadIdx.y + 1);
  load_plane(g, i0, j0, k0 - 1, P, sP[0], U, sU[0], dphi, sD[0]);
  load_plane(g, i0, j0, k0, P, sP[1], U, sU[1], dphi, sD[1]);
  int lo = 0, mid = 1, hi = 2;
  for (int k = k0; k <= k1; ++k) {
    load_plane(g, i0, j0, k + 1, P, sP[hi], U, sU[hi], dphi, sD[hi]);
    __syncthreads();
    if (own) {
      const int c = i + g.s * j + g.ss * k;
      const double un = solute_cell(PlaneNb{sP[lo], sP[mid], sP[hi], q},
                                    PlaneNb{sU[lo], sU[mid], sU[hi], q},
                                    PlaneNb{sD[lo], sD[mid], sD[hi], q}, g);
      Un[c] = un;
      put_ghosts(Un, g, i, j, k, c, un);
    }
    __syncthreads();
    const int t = lo; lo = mid; mid = hi; hi = t;
  }
}

}  // namespace

struct pf_karma {
  pf_karma_params p;
  int n = 0, s = 0, ss = 0;
  size_t cells = 0;  // padded
  double D = 0, Omega = 0;
  double *phi[2] = {nullptr, nullptr}, *U[2] = {nullptr, nullptr}, *dphi = nullptr;
  int cur = 0, it = 0;
  int zmarch = 1;  // z-march kernels (1) or one thread per cell (0)
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
  // seed (karma3d.f90:130-144) on the cells 1..n, then the mirror ghosts
  // the first step's espelho sets (U is uniform, its ghosts included)
  std::vector<double> ph(s->cells), u(s->cells, -s->Omega);
  const int n = s->n;
  auto at = [&](int i, int j, int k) { return (size_t)i + (size_t)s->s * j + (size_t)s->ss * k; };
  auto cl = [n](int v) { return v < 1 ? 1 : (v > n ? n : v); };
  for (int k = 1; k <= n; ++k)
    for (int j = 1; j <= n; ++j)
      for (int i = 1; i <= n; ++i) {
        const double x = (i - 0.5) * p->dx, y = (j - 0.5) * p->dx, z = (k - 0.5) * p->dx;
        const double r = std::sqrt(x * x + y * y + z * z);
        ph[at(i, j, k)] = -std::tanh((r - p->R0) / std::sqrt(2.0));
      }
  for (int k = 0; k <= n + 1; ++k)
    for (int j = 0; j <= n + 1; ++j)
      for (int i = 0; i <= n + 1; ++i)
        if (i != cl(i) || j != cl(j) || k != cl(k)) ph[at(i, j, k)] = ph[at(cl(i), cl(j), cl(k))];
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
  const dim3 gc((n + kBX - 1) / kBX, (n + kBY - 1) / kBY, n);
  const dim3 bz(kTX, kTY, 1);
  const dim3 gz((n + kTX - 1) / kTX, (n + kTY - 1) / kTY, (n + kKC - 1) / kKC);
  for (int t = 0; t < nsteps; ++t) {
    s->it += 1;
    const KStep g = kstep(s);
    const int a = s->cur, b = 1 - s->cur;
    if (s->zmarch) {
      hipLaunchKernelGGL(k_karma_phase_zm, gz, bz, 0, s->st, g, s->phi[a], s->U[a], s->dphi,
                         s->phi[b]);
      hipLaunchKernelGGL(k_karma_solute_zm, gz, bz, 0, s->st, g, s->phi[a], s->U[a], s->dphi,
                         s->U[b]);
    } else {
      hipLaunchKernelGGL(k_karma_phase, gc, blk, 0, s->st, g, s->phi[a], s->U[a], s->dphi,
                         s->phi[b]);
      hipLaunchKernelGGL(k_karma_solute, gc, blk, 0, s->st, g, s->phi[a], s->U[a], s->dphi,
                         s->U[b]);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess)
      return pf::fail_msg(PF_E_HIP, "pf_karma_run: launch failed: %s", hipGetErrorString(e));
    s->cur = b;
  }
  return PF_OK;
}

extern "C" int pf_karma_set_zmarch(pf_karma *s, int32_t on) {
  if (!s) return pf::fail_msg(PF_E_ARG, "pf_karma_set_zmarch: null context");
  s->zmarch = on ? 1 : 0;
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

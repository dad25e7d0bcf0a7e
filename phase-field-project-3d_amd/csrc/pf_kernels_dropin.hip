// pf_kernels_dropin.hip -- one gfx950 kernel per reference subroutine.
//
// These are the kernels behind the drop-in calls (pf_temperatura,
// pf_equacao_campo_fase, pf_concentracao_soluto, pf_diagnosticos).  Each
// thread owns one cell and evaluates the reference's per-cell body in its
// operation order with FMA contraction off, so that the FP64 form matches the
// reference's CPU path bit for bit except where the device `log` rounds
// differently from the host libm.  Buffer swaps replace the reference's
// whole-array copies; every cell a copy used to fill is written explicitly.
//
// Template parameter R is the storage and arithmetic type (double for the
// FP64 form, float for the FP32-mixed form).
#include <hip/hip_runtime.h>

#include "pf_common.h"
#include "pf_kernels.h"
#include "pf_mass.h"
#include "pf_math.h"

#pragma clang fp contract(off)

namespace pf {

template <typename R>
__device__ __forceinline__ R rlog(R x) { return log(x); }
template <>
__device__ __forceinline__ float rlog<float>(float x) { return logf(x); }



// ---------------------------------------------------------------------------
// temperatura.f90:55-70 + temp_body.inc:11-45.  Boundary cells keep TOLD.
template <typename R>
__global__ void __launch_bounds__(256) k_temperatura(KParams p, const R *__restrict__ Told,
                                                     const R *__restrict__ phi,
                                                     const R *__restrict__ dpp,
                                                     R *__restrict__ Tnew) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  const int nx = p.nx, ny = p.ny, nz = p.nz;
  if (i > nx || j > ny) return;
  const size_t c = cidx(i, j, k, nx, ny);
  if (i == 1 || i == nx || j == 1 || j == ny || k == 1 || k == nz) {
    Tnew[c] = Told[c];
    return;
  }
  const R TP = (R)p.TP, dx = (R)p.dx, dy = (R)p.dy, dz = (R)p.dz;
  const R PP = phi[c];
  R h_linha;
  if (p.h_quintico == 1)
    h_linha = (R)30 * PP * PP * ((R)1 - PP) * ((R)1 - PP);
  else
    h_linha = (R)2 * PP * ((R)3 - (R)2 * PP) - (R)2 * PP * PP;
  const R fonte = ((R)p.delta_H / (R)p.cp_s) * h_linha * dpp[c];
  const size_t sx = 1, sy = (size_t)nx, sz = (size_t)nx * ny;
  // The Laplacian is evaluated even for D == 0 (the reference's D is never
  // assigned): a non-finite TOLD then propagates exactly as it does there.
  const R TE = Told[c + sx], TW = Told[c - sx], TN = Told[c + sy], TS = Told[c - sy];
  const R TT = Told[c + sz], TB = Told[c - sz];
  const R lap = (((TE - (R)2 * TP + TW) / (dx * dx)) + ((TN - (R)2 * TP + TS) / (dy * dy)) +
                 ((TT - (R)2 * TP + TB) / (dz * dz)));
  Tnew[c] = (R)p.D * (R)p.dt * lap + fonte + TP;
}

// ---------------------------------------------------------------------------
// One "face" point of phase_body.inc:55-129: eps and the three cross
// coefficients from the centred gradient at a node.  FP64 keeps the
// reference's expressions and order; FP32 uses the same quantities with the
// gradient normalised by Q (Q^2 = 1e-46 in a flat region is below the FP32
// range).
template <typename R>
__device__ __forceinline__ void node_terms(R gx, R gy, R gz, R E0, R DE, R K, R small, R &eps,
                                           R &exy, R &eyz, R &ezx) {
  if constexpr (sizeof(R) == 8) {
    const R Q = gx * gx + gy * gy + gz * gz + small;
    const R S4 = (gx * (gx * (gx * gx))) + (gy * (gy * (gy * gy))) + (gz * (gz * (gz * gz)));
    eps = E0 * ((R)1 + DE * ((R)4 * S4 / (Q * Q) - (R)3));
    exy = K * eps * gx * gy * (gx * gx - gy * gy) / (Q * Q);
    eyz = K * eps * gy * gz * (gy * gy - gz * gz) / (Q * Q);
    ezx = K * eps * gz * gx * (gz * gz - gx * gx) / (Q * Q);
  } else {
    const R rq = (R)1 / (gx * gx + gy * gy + gz * gz + small);
    const R a = gx * gx * rq, b = gy * gy * rq, c = gz * gz * rq;
    eps = E0 * ((R)1 + DE * ((R)4 * (a * a + b * b + c * c) - (R)3));
    exy = K * eps * (gx * gy * rq) * (a - b);
    eyz = K * eps * (gy * gz * rq) * (b - c);
    ezx = K * eps * (gz * gx * rq) * (c - a);
  }
}

// phase_body.inc:14-198 for one interior cell, reading phi_old from memory.
template <typename R>
__device__ __forceinline__ void phase_body(const KParams &p, const R *__restrict__ po,
                                           const R *__restrict__ co, int i, int j, int k,
                                           R &pnew, R &dpp) {
  const int nx = p.nx, ny = p.ny;
  const size_t sx = 1, sy = (size_t)nx, sz = (size_t)nx * ny;
  const size_t c = cidx(i, j, k, nx, ny);
#define P(o) po[c + (o)]
  const R dx = (R)p.dx, dy = (R)p.dy, dz = (R)p.dz, E0 = (R)p.E0, DE = (R)p.delta_eps;
  const R h5 = (R)0.5;
  const R PP = P(0);
  const R PN = P(sy), PS = P(-sy), PE = P(sx), PW = P(-sx), PT = P(sz), PB = P(-sz);
  const R PNE = P(sx + sy), PSE = P(sx - sy), PNW = P(sy - sx), PSW = P(-sx - sy);
  const R PNT = P(sy + sz), PNB = P(sy - sz), PST = P(sz - sy), PSB = P(-sy - sz);
  const R PET = P(sx + sz), PEB = P(sx - sz), PWT = P(sz - sx), PWB = P(-sx - sz);
  const R small = (R)1.0e-23;
  const R K = (R)16 * E0 * DE;
#define SQ(x) ((x) * (x))
#define P4(x) ((x) * ((x) * ((x) * (x))))
  // E: node (i+1,j,k)
  const R gxE = (P(2 * sx) - PP) * h5 / dx, gyE = (PNE - PSE) * h5 / dy, gzE = (PET - PEB) * h5 / dz;
  R epsE, exyE, eyzE, ezxE;
  node_terms<R>(gxE, gyE, gzE, E0, DE, K, small, epsE, exyE, eyzE, ezxE);
  // W: node (i-1,j,k)
  const R gxW = (PP - P(-2 * sx)) * h5 / dx, gyW = (PNW - PSW) * h5 / dy, gzW = (PWT - PWB) * h5 / dz;
  R epsW, exyW, eyzW, ezxW;
  node_terms<R>(gxW, gyW, gzW, E0, DE, K, small, epsW, exyW, eyzW, ezxW);
  // N: node (i,j+1,k)
  const R gxN = (PNE - PNW) * h5 / dx, gyN = (P(2 * sy) - PP) * h5 / dy, gzN = (PNT - PNB) * h5 / dz;
  R epsN, exyN, eyzN, ezxN;
  node_terms<R>(gxN, gyN, gzN, E0, DE, K, small, epsN, exyN, eyzN, ezxN);
  // S: node (i,j-1,k)
  const R gxS = (PSE - PSW) * h5 / dx, gyS = (PP - P(-2 * sy)) * h5 / dy, gzS = (PST - PSB) * h5 / dz;
  R epsS, exyS, eyzS, ezxS;
  node_terms<R>(gxS, gyS, gzS, E0, DE, K, small, epsS, exyS, eyzS, ezxS);
  // T: node (i,j,k+1)
  const R gxT = (PET - PWT) * h5 / dx, gyT = (PNT - PST) * h5 / dy, gzT = (P(2 * sz) - PP) * h5 / dz;
  R epsT, exyT, eyzT, ezxT;
  node_terms<R>(gxT, gyT, gzT, E0, DE, K, small, epsT, exyT, eyzT, ezxT);
  // B: node (i,j,k-1)
  const R gxB = (PEB - PWB) * h5 / dx, gyB = (PNB - PSB) * h5 / dy, gzB = (PP - P(-2 * sz)) * h5 / dz;
  R epsB, exyB, eyzB, ezxB;
  node_terms<R>(gxB, gyB, gzB, E0, DE, K, small, epsB, exyB, eyzB, ezxB);
  (void)eyzE; (void)eyzW; (void)ezxN; (void)ezxS; (void)exyT; (void)exyB;
  R AA = (SQ(epsE) * (PE - PP) + SQ(epsW) * (PW - PP)) / (dx * dx) +
         (SQ(epsN) * (PN - PP) + SQ(epsS) * (PS - PP)) / (dy * dy) +
         (SQ(epsT) * (PT - PP) + SQ(epsB) * (PB - PP)) / (dz * dz);
  const R cxy = (R)0.25 / (dx * dy), cyz = (R)0.25 / (dy * dz), czx = (R)0.25 / (dz * dx);
  const R Cxy = cxy * (PNE * (exyE - exyN) + PSE * (exyS - exyE) + PNW * (exyN - exyW) +
                       PSW * (exyW - exyS)) +
                PN * cxy * (exyE - exyW) + PS * cxy * (exyW - exyE) + PE * cxy * (exyS - exyN) +
                PW * cxy * (exyN - exyS);
  const R Cyz = cyz * (PNT * (eyzN - eyzT) + PNB * (eyzB - eyzN) + PST * (eyzT - eyzS) +
                       PSB * (eyzS - eyzB)) +
                PT * cyz * (eyzN - eyzS) + PB * cyz * (eyzS - eyzN) + PN * cyz * (eyzB - eyzT) +
                PS * cyz * (eyzT - eyzB);
  const R Czx = czx * (PET * (ezxT - ezxE) + PWT * (ezxW - ezxT) + PEB * (ezxE - ezxB) +
                       PWB * (ezxB - ezxW)) +
                PE * czx * (ezxT - ezxB) + PW * czx * (ezxB - ezxT) + PT * czx * (ezxW - ezxE) +
                PB * czx * (ezxE - ezxW);
  AA = AA + Cxy + Cyz + Czx;
#undef SQ
#undef P4
#undef P
  const R one = (R)1;
  const R g_linha = (R)2 * PP * (one - PP) * (one - PP) - (R)2 * PP * PP * (one - PP);
  // the FP32-mixed form stores c - c_off
  R CPP = co[c];
  if constexpr (sizeof(R) == 4) CPP = (R)p.c_off + CPP;
  R h_linha, h;
  if (p.h_quintico == 1) {
    h_linha = (R)30 * PP * PP * (one - PP) * (one - PP);
    h = PP * PP * PP * ((R)10 - (R)15 * PP + (R)6 * PP * PP);
  } else {
    h_linha = (R)2 * PP * ((R)3 - (R)2 * PP) - (R)2 * PP * PP;
    h = PP * PP * ((R)3 - (R)2 * PP);
  }
  const R Ke = (R)p.Ke;
  R lnr;
  if constexpr (sizeof(R) == 4) {
    // the ratio minus 1 formed without cancellation, as the fused FP32
    // sweeps do (ratio_ln_f32 in pf_march.h): (den A - c B) / (den - Ke c)
    // with A = K - 1, B = K - Ke, K = (1-Cs)/(1-Cl), and den A - c B =
    // (A - c_off B) - h (1 - Ke) A - dc B for the stored dc = c - c_off (the
    // constants in FP64, then rounded)
    const double K = (1.0 - p.Cs_eq) / (1.0 - p.Cl_eq);
    const double A = K - 1.0, B = K - p.Ke;
    const R den = Ke * h + (one - h);
    const R x = ((R)(A - p.c_off * B) - h * (R)((1.0 - p.Ke) * A) - co[c] * (R)B) /
                (den - Ke * CPP);
    lnr = ln1p_abs_f32(x);
  } else {
    const R CLP = CPP / (Ke * h + (one - h));
    const R CSP = Ke * CPP / (Ke * h + (one - h));
    lnr = rlog<R>(fabs(((one - (R)p.Cs_eq) * (one - CLP)) / ((one - (R)p.Cl_eq) * (one - CSP))));
  }
  const R GG = (-((R)p.W0 * g_linha)) + ((R)p.R * (R)p.TP / (R)p.Vm) * h_linha * lnr;
  const float ran = ran_real4(i, j, k, p.it, p.ruido_tipo, p.seed);
  const R amp = (p.ruido_tipo == PF_RUIDO_CELL_POSITIVE) ? (R)ran * (R)p.AAA
                                                         : ((R)2 * (R)ran - one) * (R)p.AAA;
  const R ruidos = amp * (R)16 * (PP * PP) * ((one - PP) * (one - PP));
  R PPNEW;
  if constexpr (sizeof(R) == 4)  // stochastically rounded store (csum, pf_mass.h)
    PPNEW = csum<R>(PP, ((R)p.M0 * (R)p.dt) * (AA + GG) + ruidos,
                    cell_hash_col(i, j, p.it) ^ kSaltPhiFirst, k);
  else
    PPNEW = PP + ((R)p.M0 * (R)p.dt) * (AA + GG) + ruidos;
  dpp = PPNEW - PP;
  const R cl = PPNEW > one ? one : PPNEW;  // dmax1(0,dmin1(1,x)), NaN kept
  pnew = cl < (R)0 ? (R)0 : cl;
}

// Interior [3,n-2]^3 of one phase sweep: phi_new and DPP.
template <typename R>
__global__ void __launch_bounds__(256) k_phase_interior(KParams p, const R *__restrict__ po,
                                                        const R *__restrict__ co,
                                                        R *__restrict__ pn,
                                                        R *__restrict__ dpp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 3;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 3;
  const int k = blockIdx.z + 3;
  if (i > p.nx - 2 || j > p.ny - 2) return;
  R v, d;
  phase_body<R>(p, po, co, i, j, k, v, d);
  const size_t c = cidx(i, j, k, p.nx, p.ny);
  pn[c] = v;
  dpp[c] = d;
}

// The 12-plane decoration for the cells outside [3,n-2]^3.  Reads phi_new
// only inside [3,n-2]^3 (never written here) and phi_old elsewhere.
// xf = 0 leaves out the x-face cells of the rows inside [3,n-2] in y and z,
// which the fused sweeps write in their own launch.
template <typename R>
__global__ void __launch_bounds__(256) k_phase_boundary(KParams p, const R *__restrict__ po,
                                                        R *__restrict__ pn, int xf) {
  const int nx = p.nx, ny = p.ny, nz = p.nz;
  int i, j, k;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
       shell_flat(t, nx, ny, nz, 2, i, j, k, xf != 0); t += (long long)gridDim.x * blockDim.x) {
    phase_deco_cell(i, j, k, nx, ny, nz, po, pn);
  }
}

// ---------------------------------------------------------------------------
// conc_body.inc:11-149 for one interior cell [2,n-1]^3.
template <typename R>
__global__ void __launch_bounds__(256) k_solute_interior(KParams p, const R *__restrict__ po,
                                                         const R *__restrict__ co,
                                                         R *__restrict__ cn) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 2;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 2;
  const int k = blockIdx.z + 2;
  const int nx = p.nx, ny = p.ny;
  if (i > nx - 1 || j > ny - 1) return;
  const size_t sx = 1, sy = (size_t)nx, sz = (size_t)nx * ny;
  const size_t c = cidx(i, j, k, nx, ny);
  const R one = (R)1;
  const R PP = po[c], PN = po[c + sy], PS = po[c - sy], PW = po[c - sx], PE = po[c + sx];
  const R PT = po[c + sz], PB = po[c - sz];
  // the FP32-mixed form stores c - c_off
  auto cat = [&](size_t q) -> R {
    if constexpr (sizeof(R) == 8) return co[q];
    else return (R)p.c_off + co[q];
  };
  const R CPP = cat(c), CPS = cat(c - sy), CPN = cat(c + sy), CPW = cat(c - sx), CPE = cat(c + sx);
  const R CPT = cat(c + sz), CPB = cat(c - sz);
  R h, hE, hW, hN, hS, hT, hB;
  if (p.h_quintico == 1) {
#define HQ(x) (((x) * (x) * (x)) * ((R)10 - ((R)15 * (x)) + ((R)6 * ((x) * (x)))))
    h = HQ(PP); hE = HQ(PE); hW = HQ(PW); hN = HQ(PN); hS = HQ(PS); hT = HQ(PT); hB = HQ(PB);
#undef HQ
  } else {
#define HC(x) ((x) * (x) * ((R)3 - (R)2 * (x)))
    h = HC(PP); hE = HC(PE); hW = HC(PW); hN = HC(PN); hS = HC(PS); hT = HC(PT); hB = HC(PB);
#undef HC
  }
  const R Ke = (R)p.Ke;
#define CS(cc, hh) ((Ke * (cc)) / ((Ke * (hh)) + (one - (hh))))
#define CL(cc, hh) ((cc) / (Ke * (hh) + (one - (hh))))
  const R CSP = CS(CPP, h), CSN = CS(CPN, hN), CSS = CS(CPS, hS), CSE = CS(CPE, hE);
  const R CSW = CS(CPW, hW), CST = CS(CPT, hT), CSB = CS(CPB, hB);
  const R CLP = CL(CPP, h), CLN = CL(CPN, hN), CLS = CL(CPS, hS), CLE = CL(CPE, hE);
  const R CLW = CL(CPW, hW), CLT = CL(CPT, hT), CLB = CL(CPB, hB);
#undef CS
#undef CL
  const R DC = (PP <= (R)0.95) ? (R)p.D_liq : (R)p.D_sol;
#define GA(hh, cl, cs) (DC * ((one - (hh)) * (cl) * (one - (cl)) + (hh) * (cs) * (one - (cs))))
  const R GP = GA(h, CLP, CSP), GS = GA(hS, CLS, CSS), GN = GA(hN, CLN, CSN);
  const R GW = GA(hW, CLW, CSW), GE = GA(hE, CLE, CSE), GT = GA(hT, CLT, CST);
  const R GB = GA(hB, CLB, CSB);
#undef GA
  const R Gs = (R)2 / ((one / GP) + (one / GS));
  const R Gn = (R)2 / ((one / GN) + (one / GP));
  const R Gw = (R)2 / ((one / GP) + (one / GW));
  const R Ge = (R)2 / ((one / GE) + (one / GP));
  const R Gt = (R)2 / ((one / GT) + (one / GP));
  const R Gb = (R)2 / ((one / GP) + (one / GB));
  const R IX = (R)(1.0 / (p.dx * p.dx)), IY = (R)(1.0 / (p.dy * p.dy)), IZ = (R)(1.0 / (p.dz * p.dz));
#define MU(cl) rlog<R>((cl) / (one - (cl)))
  const R muP = MU(CLP);
  // the increment onto the stored value (= CPP for FP64; FP32: unbiased
  // stochastic rounding, as in the fused sweeps)
  const R inc = (Ge * IX * (MU(CLE) - muP) - Gw * IX * (muP - MU(CLW)) + Gn * IY * (MU(CLN) - muP) -
                 Gs * IY * (muP - MU(CLS)) + Gt * IZ * (MU(CLT) - muP) - Gb * IZ * (muP - MU(CLB))) *
                (R)p.dt;
  cn[c] = csum<R>(co[c], inc, cell_hash_col(i, j, p.it), k);
#undef MU
}

// Solute faces: c_new(face) = snapshot(inner neighbour), with the snapshot
// equal to the swept c_new inside [2,n-1]^3 and to c_old on the faces.
// xf = 0 leaves out the x-face cells of the rows inside [2,n-1] in y and z,
// which the fused solute sweeps write in their own launch.
template <typename R>
__global__ void __launch_bounds__(256) k_solute_boundary(KParams p, const R *__restrict__ co,
                                                         R *__restrict__ cn, int xf) {
  const int nx = p.nx, ny = p.ny, nz = p.nz;
  int i, j, k;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
       shell_flat(t, nx, ny, nz, 1, i, j, k, xf != 0); t += (long long)gridDim.x * blockDim.x) {
    solute_deco_cell(i, j, k, nx, ny, nz, co, cn);
  }
}

// Boundary cells whose solute source is an interior cell take that cell's
// value again (in place: interior cells are only read).  Used after the
// interior of c has been rescaled, so that the face planes stay copies of
// their interior neighbours; edge and corner cells, whose sources are other
// boundary cells, border no interior cell and are left as they are.
template <typename R>
__global__ void __launch_bounds__(256) k_solute_refresh(KParams p, R *__restrict__ c) {
  const int nx = p.nx, ny = p.ny, nz = p.nz;
  int i, j, k;
  for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x;
       shell_flat(t, nx, ny, nz, 1, i, j, k); t += (long long)gridDim.x * blockDim.x) {
    int si = i, sj = j, sk = k;
    solute_source(si, sj, sk, nx, ny, nz);
    if (si > 1 && si < nx && sj > 1 && sj < ny && sk > 1 && sk < nz)
      c[cidx(i, j, k, nx, ny)] = c[cidx(si, sj, sk, nx, ny)];
  }
}

// ---------------------------------------------------------------------------
// diag_body.inc:8-22: block b folds the rows (j,k) = b, b+G, ... (threads
// striding over i) into a partial {sum phi^2, max, min, sum c, F}; the second
// stage folds the partials in a fixed order, so the result is the same for
// every run and does not depend on scheduling.
template <typename R>
__global__ void __launch_bounds__(256) k_diag_partial(KParams p, const R *__restrict__ ph,
                                                      const R *__restrict__ cc,
                                                      double *__restrict__ part) {
  const int nx = p.nx, ny = p.ny, nz = p.nz;
  const double dx = p.dx, dy = p.dy, dz = p.dz, E0 = p.E0, w0 = p.W0;
  const double dV = dx * dy * dz;
  const size_t sy = (size_t)nx, sz = (size_t)nx * ny;
  double s2 = 0.0, mx = -1.7976931348623157e308, mn = 1.7976931348623157e308, sc = 0.0, fe = 0.0;
  for (int r = blockIdx.x; r < ny * nz; r += gridDim.x) {
    const int j = r % ny + 1, k = r / ny + 1;
    const bool rin = j > 1 && j < ny && k > 1 && k < nz;
    const size_t row = (size_t)(j - 1) * sy + (size_t)(k - 1) * sz;
    for (int i = threadIdx.x + 1; i <= nx; i += blockDim.x) {
      const size_t t = row + (size_t)(i - 1);
      const double v = (double)ph[t];
      s2 += v * v;
      mx = v > mx ? v : mx;
      mn = v < mn ? v : mn;
      if constexpr (sizeof(R) == 8) sc += (double)cc[t];
      else sc += (double)cc[t] + p.c_off;
      const double w = w0 * v * v * (1.0 - v) * (1.0 - v);
      double g2 = 0.0;
      if (rin && i > 1 && i < nx) {
        const double a = ((double)ph[t + 1] - (double)ph[t - 1]) * 0.5 / dx;
        const double b = ((double)ph[t + sy] - (double)ph[t - sy]) * 0.5 / dy;
        const double c = ((double)ph[t + sz] - (double)ph[t - sz]) * 0.5 / dz;
        g2 = a * a + b * b + c * c;
      }
      fe += (0.5 * E0 * E0 * g2 + w) * dV;
    }
  }
  __shared__ double sh[5][256];
  const int tid = threadIdx.x;
  sh[0][tid] = s2; sh[1][tid] = mx; sh[2][tid] = mn; sh[3][tid] = sc; sh[4][tid] = fe;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) {
      sh[0][tid] += sh[0][tid + w];
      sh[1][tid] = sh[1][tid + w] > sh[1][tid] ? sh[1][tid + w] : sh[1][tid];
      sh[2][tid] = sh[2][tid + w] < sh[2][tid] ? sh[2][tid + w] : sh[2][tid];
      sh[3][tid] += sh[3][tid + w];
      sh[4][tid] += sh[4][tid + w];
    }
    __syncthreads();
  }
  if (tid == 0)
    for (int q = 0; q < 5; ++q) part[(size_t)blockIdx.x * 5 + q] = sh[q][0];
}

__global__ void __launch_bounds__(256) k_diag_final(const double *__restrict__ part, int nblk,
                                                    double *__restrict__ out) {
  __shared__ double sh[5][256];
  const int tid = threadIdx.x;
  double s2 = 0.0, mx = -1.7976931348623157e308, mn = 1.7976931348623157e308, sc = 0.0, fe = 0.0;
  for (int b = tid; b < nblk; b += 256) {
    const double *q = part + (size_t)b * 5;
    s2 += q[0];
    mx = q[1] > mx ? q[1] : mx;
    mn = q[2] < mn ? q[2] : mn;
    sc += q[3];
    fe += q[4];
  }
  sh[0][tid] = s2; sh[1][tid] = mx; sh[2][tid] = mn; sh[3][tid] = sc; sh[4][tid] = fe;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) {
      sh[0][tid] += sh[0][tid + w];
      sh[1][tid] = sh[1][tid + w] > sh[1][tid] ? sh[1][tid + w] : sh[1][tid];
      sh[2][tid] = sh[2][tid + w] < sh[2][tid] ? sh[2][tid + w] : sh[2][tid];
      sh[3][tid] += sh[3][tid + w];
      sh[4][tid] += sh[4][tid + w];
    }
    __syncthreads();
  }
  if (tid == 0) {
    out[0] = sqrt(sh[0][0]);
    out[1] = sh[1][0];
    out[2] = sh[2][0];
    out[3] = sh[3][0];
    out[4] = sh[4][0];
  }
}

// ---------------------------------------------------------------------------
// Launchers.
static inline dim3 plane_grid(int ni, int nj, int nk, dim3 blk) {
  return dim3((unsigned)((ni + blk.x - 1) / blk.x), (unsigned)((nj + blk.y - 1) / blk.y),
              (unsigned)nk);
}

template <typename R>
hipError_t launch_temperatura(const KParams &p, const R *Told, const R *phi, const R *dpp, R *Tnew,
                              hipStream_t s) {
  const dim3 blk(64, 4, 1);
  hipLaunchKernelGGL(k_temperatura<R>, plane_grid(p.nx, p.ny, p.nz, blk), blk, 0, s, p, Told, phi,
                     dpp, Tnew);
  return hipGetLastError();
}

static unsigned shell_blocks(const KParams &p, int w, bool xf = true) {
  const long long n = shell_count(p.nx, p.ny, p.nz, w, xf);
  const long long b = (n + 255) / 256;
  return (unsigned)(b < 1 ? 1 : (b > 16384 ? 16384 : b));
}

template <typename R>
hipError_t launch_phase_boundary(const KParams &p, const R *po, R *pn, hipStream_t s, bool xf) {
  hipLaunchKernelGGL(k_phase_boundary<R>, dim3(shell_blocks(p, 2, xf)), dim3(256), 0, s, p, po, pn,
                     xf ? 1 : 0);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_solute_boundary(const KParams &p, const R *co, R *cn, hipStream_t s, bool xf) {
  hipLaunchKernelGGL(k_solute_boundary<R>, dim3(shell_blocks(p, 1, xf)), dim3(256), 0, s, p, co,
                     cn, xf ? 1 : 0);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_solute_refresh(const KParams &p, R *c, hipStream_t s) {
  hipLaunchKernelGGL(k_solute_refresh<R>, dim3(shell_blocks(p, 1)), dim3(256), 0, s, p, c);
  return hipGetLastError();
}

template <typename R>
__global__ void k_probe(const R *__restrict__ f, size_t a, size_t b, double *__restrict__ out) {
  if (threadIdx.x == 0) {
    out[0] = (double)f[a];
    out[1] = (double)f[b];
  }
}

template <typename R>
hipError_t launch_probe(const R *f, size_t a, size_t b, double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_probe<R>, dim3(1), dim3(64), 0, s, f, a, b, out);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_phase(const KParams &p, const R *po, const R *co, R *pn, R *dpp, hipStream_t s) {
  const dim3 blk(64, 4, 1);
  hipLaunchKernelGGL(k_phase_interior<R>, plane_grid(p.nx - 4, p.ny - 4, p.nz - 4, blk), blk, 0, s,
                     p, po, co, pn, dpp);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_phase_boundary<R>(p, po, pn, s, true);
}

template <typename R>
hipError_t launch_solute(const KParams &p, const R *po, const R *co, R *cn, hipStream_t s) {
  const dim3 blk(64, 4, 1);
  hipLaunchKernelGGL(k_solute_interior<R>, plane_grid(p.nx - 2, p.ny - 2, p.nz - 2, blk), blk, 0, s,
                     p, po, co, cn);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  return launch_solute_boundary<R>(p, co, cn, s, true);
}

template <typename R>
hipError_t launch_diag(const KParams &p, const R *ph, const R *cc, double *part, int nblk,
                       double *out, hipStream_t s) {
  hipLaunchKernelGGL(k_diag_partial<R>, dim3(nblk), dim3(256), 0, s, p, ph, cc, part);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(256), 0, s, part, nblk, out);
  return hipGetLastError();
}

#define PF_INST(R)                                                                              \
  template hipError_t launch_temperatura<R>(const KParams &, const R *, const R *, const R *, R *, \
                                            hipStream_t);                                       \
  template hipError_t launch_phase<R>(const KParams &, const R *, const R *, R *, R *,           \
                                      hipStream_t);                                             \
  template hipError_t launch_solute<R>(const KParams &, const R *, const R *, R *, hipStream_t);  \
  template hipError_t launch_diag<R>(const KParams &, const R *, const R *, double *, int,       \
                                     double *, hipStream_t);                                    \
  template hipError_t launch_phase_boundary<R>(const KParams &, const R *, R *, hipStream_t,    \
                                               bool);                                           \
  template hipError_t launch_solute_boundary<R>(const KParams &, const R *, R *, hipStream_t,   \
                                                bool);                                          \
  template hipError_t launch_solute_refresh<R>(const KParams &, R *, hipStream_t);               \
  template hipError_t launch_probe<R>(const R *, size_t, size_t, double *, hipStream_t);
PF_INST(double)
PF_INST(float)
#undef PF_INST

}  // namespace pf

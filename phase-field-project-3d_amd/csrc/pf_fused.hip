// pf_fused.hip -- fused 2.5D z-march sweeps for gfx950.
//
// A work-group owns a TX x TY column tile (x fastest, 64 = one wavefront per
// row) and a segment of z-planes, and marches up the segment:
//   - phi planes (tile + 2-cell halo) rotate through 5 LDS slots, loaded once
//     from HBM with x-contiguous (coalesced) reads;
//   - the anisotropy quantities of a phase cell's six neighbours
//     (phase_body.inc:55-129) are the centred-gradient quantities of those
//     neighbour nodes, so each node's eps^2, exy, eyz, ezx is computed once
//     (tile + 1-cell ring), shared in-plane through LDS and along z in
//     registers -- instead of six times per cell;
//   - sweep B also evaluates the solute (conc_body.inc) with the per-node
//     1/gamma and mu = ln(CL/(1-CL)) computed once per node, and the
//     temperature of the next step (temp_body.inc) from the fresh DPP, which
//     therefore never goes to HBM except on the last step of a call.
// Compulsory HBM traffic per cell: sweep A reads phi, c and writes phi;
// sweep B reads phi, c and writes phi, c, T.
//
// Arithmetic follows the reference's formulas with algebraic reuse
// (reciprocals of dx^2, node reuse, the harmonic mean as 2/(1/g+1/g'), the
// free-energy ratio with CL = c/den folded in, divisions as a refined
// reciprocal, the ~1 ulp logarithm of pf_math.h), so results agree with the
// reference to rounding, not bit for bit; the drop-in kernels keep the
// bit-faithful form.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pf_common.h"
#include "pf_fused.h"
#include "pf_kernels.h"
#include "pf_math.h"

namespace pf {
namespace {

constexpr int NSLOT = 5;

// Tile shape: TX x TY cells per work-group (one thread per column).
template <int TX_, int TY_>
struct Tile {
  static constexpr int TX = TX_, TY = TY_, NT = TX * TY;
  static constexpr int PX = TX + 4, PY = TY + 4, PPL = PX * PY;  // phi plane, halo 2
  static constexpr int QX = TX + 2, QY = TY + 2, QPL = QX * QY;  // node plane, halo 1
  static constexpr int NRING = 2 * QX + 2 * TY;                  // ring nodes per plane
  static constexpr int NLD = (PPL + NT - 1) / NT;                 // phi loads per thread
  static_assert(NRING <= NT, "one ring node per thread at most");
  static_assert(TX % 16 == 0 && NT % 64 == 0, "whole wavefronts, 128-byte row segments");
};

// log and 1/x in the arithmetic type.  FP64: the ~1 ulp forms of
// pf_math.h.  FP32: the hardware log2 / reciprocal (about 1 ulp).
template <typename R>
__device__ __forceinline__ R dlog(R x);
template <>
__device__ __forceinline__ double dlog<double>(double x) { return log_fast(x); }
template <>
__device__ __forceinline__ float dlog<float>(float x) { return __logf(x); }

template <typename R>
__device__ __forceinline__ R rcpn(R x);
template <>
__device__ __forceinline__ double rcpn<double>(double x) { return rcp_refined(x); }
template <>
__device__ __forceinline__ float rcpn<float>(float x) { return __builtin_amdgcn_rcpf(x); }

// Scalars of one sweep, in the arithmetic type.
template <typename R>
struct SC {
  R hx, hy, hz;        // 0.5/dx ...
  R idx2, idy2, idz2;  // 1/dx^2 ...
  R cxy, cyz, czx;     // 0.25/(dx dy) ...
  R E0a, E0b, K;       // eps = E0a + E0b*S4/Q^2, K = 16 E0 delta
  R small;
  R Ke, W0, RTV;       // RTV = R*TP/Vm
  R omCsCl;            // (1-Cs_eq)/(1-Cl_eq)
  R M0dt, AAA;
  R Dliq2dt, Dsol2dt;  // 2*D*dt for the solute faces
  R fonte, TP, Ddt;    // delta_H/cp, TP, D*dt
};

template <typename R>
SC<R> make_sc(const KParams &p) {
  SC<R> s;
  s.hx = (R)(0.5 / p.dx); s.hy = (R)(0.5 / p.dy); s.hz = (R)(0.5 / p.dz);
  s.idx2 = (R)(1.0 / (p.dx * p.dx)); s.idy2 = (R)(1.0 / (p.dy * p.dy));
  s.idz2 = (R)(1.0 / (p.dz * p.dz));
  s.cxy = (R)(0.25 / (p.dx * p.dy)); s.cyz = (R)(0.25 / (p.dy * p.dz));
  s.czx = (R)(0.25 / (p.dz * p.dx));
  s.E0a = (R)(p.E0 * (1.0 - 3.0 * p.delta_eps));
  s.E0b = (R)(p.E0 * 4.0 * p.delta_eps);
  s.K = (R)(16.0 * p.E0 * p.delta_eps);
  s.small = (R)1.0e-23;
  s.Ke = (R)p.Ke; s.W0 = (R)p.W0; s.RTV = (R)(p.R * p.TP / p.Vm);
  s.omCsCl = (R)((1.0 - p.Cs_eq) / (1.0 - p.Cl_eq));
  s.M0dt = (R)(p.M0 * p.dt); s.AAA = (R)p.AAA;
  s.Dliq2dt = (R)(2.0 * p.D_liq * p.dt); s.Dsol2dt = (R)(2.0 * p.D_sol * p.dt);
  s.fonte = (R)(p.delta_H / p.cp_s); s.TP = (R)p.TP; s.Ddt = (R)(p.D * p.dt);
  return s;
}

template <typename R>
struct Node {
  R e2, xy, yz, zx;
};

// Node quantities from the centred gradient at a node (phase_body.inc:55-129),
// written with the gradient normalised by Q = |g|^2 + small:
//   S4/Q^2 = a^2+b^2+c^2 with a = gx^2/Q ...,
//   gx gy (gx^2-gy^2)/Q^2 = (gx gy/Q)(a-b) ...
// which is the same quantity but never forms Q^2 (1e-46 for a flat region,
// below the FP32 range).
template <typename R>
__device__ __forceinline__ Node<R> node_from(R gx, R gy, R gz, const SC<R> &s) {
  const R gx2 = gx * gx, gy2 = gy * gy, gz2 = gz * gz;
  const R rq = rcpn<R>(gx2 + gy2 + gz2 + s.small);
  const R a = gx2 * rq, b = gy2 * rq, c = gz2 * rq;
  const R eps = s.E0a + s.E0b * (a * a + b * b + c * c);
  const R f = s.K * eps;
  Node<R> v;
  v.e2 = eps * eps;
  v.xy = f * (gx * gy * rq) * (a - b);
  v.yz = f * (gy * gz * rq) * (b - c);
  v.zx = f * (gz * gx * rq) * (c - a);
  return v;
}

template <typename R>
__device__ __forceinline__ void h_of(R p, int quint, R &h, R &hl) {
  if (quint) {
    h = p * p * p * ((R)10 - (R)15 * p + (R)6 * p * p);
    hl = (R)30 * p * p * ((R)1 - p) * ((R)1 - p);
  } else {
    h = p * p * ((R)3 - (R)2 * p);
    hl = (R)6 * p * ((R)1 - p);
  }
}

// Solute node quantities (conc_body.inc:61-118, 139-144): 1/gamma (gamma
// without the centre's D_C) and q, the node's share of the chemical
// potential difference taken across each face by dmu():
//   FP64: q = mu = ln(CL/(1-CL)) = ln(c/(den-c)), dmu = mu_X - mu_P;
//   FP32: q = CL, dmu = ln((CL_X (1-CL_P)) / (CL_P (1-CL_X)))
//         = log1p((CL_X - CL_P) / (CL_P (1-CL_X)))   (same quantity).
// In FP32 the difference of two ~-4 logarithms is noise of about one FP32
// ulp of mu, which in near-uniform liquid is as large as the true
// difference and would random-walk c by ~1 ulp per step; the log1p form
// takes the exact difference CL_X - CL_P first (c_X - c_P where den = 1).
template <typename R>
__device__ __forceinline__ void solute_node(R p, R cc, int quint, const SC<R> &s, R &ig, R &q) {
  R h, hl;
  h_of(p, quint, h, hl);
  const R den = s.Ke * h + ((R)1 - h);
  const R CL = cc * rcpn<R>(den);
  const R CS = s.Ke * CL;
  const R g = ((R)1 - h) * CL * ((R)1 - CL) + h * CS * ((R)1 - CS);
  ig = rcpn<R>(g);
  if constexpr (sizeof(R) == 8) q = dlog<R>(cc * rcpn<R>(den - cc));
  else q = CL;
}

template <typename R>
__device__ __forceinline__ R dmu(R qX, R qP);
template <>
__device__ __forceinline__ double dmu<double>(double qX, double qP) { return qX - qP; }
template <>
__device__ __forceinline__ float dmu<float>(float qX, float qP) {
  const float d = (qX - qP) * rcpn<float>(qP * (1.0f - qX));
  // log1p(d): Taylor series to d^6 for |d| < 1/32 (truncation < 2e-10
  // relative), else log(1 + d)
  const float ser =
      d * (1.0f + d * (-0.5f + d * (1.0f / 3.0f + d * (-0.25f + d * (0.2f + d * (-1.0f / 6.0f))))));
  return fabsf(d) < 0.03125f ? ser : dlog<float>(1.0f + d);
}

struct Geo {
  int nx, ny, nz;
  int nbx, nby, nseg, seg, k_lo, k_hi;  // cell planes [k_lo,k_hi] split in segments
  int ntiles, per_xcd;
};

// XCD-aware tile order: consecutive work-group ids go round-robin over the
// 8 XCDs, so give each XCD a contiguous run of tiles (neighbours share halo
// lines in that XCD's L2).
__device__ __forceinline__ bool tile_of(const Geo &g, int &bx, int &by, int &sg) {
  const int b = blockIdx.x;
  const int t = (b & 7) * g.per_xcd + (b >> 3);
  if (t >= g.ntiles) return false;
  bx = t % g.nbx;
  by = (t / g.nbx) % g.nby;
  sg = t / (g.nbx * g.nby);
  return true;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// SWEEP_B == false: sweep A (phase only).  SWEEP_B == true: phase + solute +
// (write_T ? next-step T : DPP); if mslot is non-null, the work-group adds
// the FP64 sum of its cells' solute increments to mslot[blockIdx.x] (the
// solute-mass ledger; one slot per work-group, so no atomics and a fixed
// summation order).
// FP32 sweeps fit four waves per SIMD (<= 128 VGPRs); the hint keeps the
// compiler there.  FP64 sweeps need more registers than three waves allow.
template <typename R, bool SWEEP_B, typename TL>
__global__ void __launch_bounds__(TL::NT)
    __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? 4 : 1))) k_march(Geo g, KParams kp, SC<R> s, const R *__restrict__ po,
                                              const R *__restrict__ co, R *__restrict__ pn,
                                              R *__restrict__ cn, const R *__restrict__ Tcur,
                                              R *__restrict__ Tnext, R *__restrict__ dpp,
                                              int write_T, double *__restrict__ mslot) {
  constexpr int TX = TL::TX, NT = TL::NT, PX = TL::PX, PPL = TL::PPL;
  constexpr int QX = TL::QX, QY = TL::QY, QPL = TL::QPL, NRING = TL::NRING, NLD = TL::NLD;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  R *sphi = reinterpret_cast<R *>(smem_raw);  // NSLOT * PPL
  R *snode = sphi + NSLOT * PPL;              // 4 * QPL
  R *sgm = snode + 4 * QPL;                   // 2 * QPL (sweep B)

  int bx, by, sg;
  if (!tile_of(g, bx, by, sg)) return;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TX + tx;
  const int i0 = 2 + bx * TX, j0 = 2 + by * TL::TY;  // first cell of the tile
  const int i = i0 + tx, j = j0 + ty;
  const bool own = (i <= nx - 1) && (j <= ny - 1);
  const bool ph_col = own && i >= 3 && i <= nx - 2 && j >= 3 && j <= ny - 2;
  const int ks = g.k_lo + sg * g.seg;
  const int ke = min(ks + g.seg - 1, g.k_hi);
  const size_t sz = (size_t)nx * ny;
  const int quint = kp.h_quintico;
  const bool noise = s.AAA != (R)0;

  // ring node handled by this thread (tid < NRING): its node-frame coords
  int rlx = 0, rly = 0;
  if (tid < NRING) {
    if (tid < QX) { rlx = tid; rly = 0; }
    else if (tid < 2 * QX) { rlx = tid - QX; rly = QY - 1; }
    else { const int r = tid - 2 * QX; rly = 1 + (r >> 1); rlx = (r & 1) ? QX - 1 : 0; }
  }
  const int olx = tx + 1, oly = ty + 1;  // own node-frame coords

  // In-plane offsets (32 bit; nx*ny < 2^31) of the phi window elements this
  // thread loads, of its own column and of its ring column; clamped so every
  // load stays inside the grid (clamped values feed only unused nodes).
  int loff[NLD];
#pragma unroll
  for (int u = 0; u < NLD; ++u) {
    const int e = tid + u * NT;
    const int gi = clampi(i0 - 2 + (e % PX), 1, nx), gj = clampi(j0 - 2 + (e / PX), 1, ny);
    loff[u] = (gi - 1) + (gj - 1) * nx;
  }
  const int ooff = (clampi(i, 1, nx) - 1) + (clampi(j, 1, ny) - 1) * nx;
  const int roff = (clampi(i0 - 1 + rlx, 1, nx) - 1) + (clampi(j0 - 1 + rly, 1, ny) - 1) * nx;
  auto plane = [&](int q) -> size_t { return (size_t)(clampi(q, 1, nz) - 1) * sz; };

  R pre[NLD];
  auto gload = [&](int q) {
    const R *src = po + plane(q);
#pragma unroll
    for (int u = 0; u < NLD; ++u)
      if (tid + u * NT < PPL) pre[u] = src[loff[u]];
  };
  auto slot = [&](int q) -> R * { return sphi + (q % NSLOT) * PPL; };  // q >= 0
  auto lstore = [&](int q) {
    R *dst = slot(q);
#pragma unroll
    for (int u = 0; u < NLD; ++u)
      if (tid + u * NT < PPL) dst[tid + u * NT] = pre[u];
  };

  // node at node-frame (lx,ly) of plane q-1 from phi planes q-2, q-1, q
  auto node_at = [&](int q, int lx, int ly) -> Node<R> {
    const R *Pm = slot(q - 2), *P0 = slot(q - 1), *Pp = slot(q);
    const int c = (ly + 1) * PX + (lx + 1);
    const R gx = (P0[c + 1] - P0[c - 1]) * s.hx;
    const R gy = (P0[c + PX] - P0[c - PX]) * s.hy;
    const R gz = (Pp[c] - Pm[c]) * s.hz;
    return node_from(gx, gy, gz, s);
  };

  // own-column registers: nodes of planes k-1 (M), k (Z), k+1 (P) and the
  // solute quantities of planes k-1, k, k+1
  Node<R> nM{}, nZ{}, nP{};
  R igM = 0, muM = 0, igZ = 0, muZ = 0, igP = 0, muP = 0;
  R cZ = 0;
  uint64_t hbase = hash_base(i, j, ks, kp.it, 1, kp.seed);
  // solute increments of this column for the ledger, summed in the
  // arithmetic type over the column's z-segment (a few dozen cells) and in
  // FP64 from there on (FP64 here would cost the FP32 form a wave per SIMD)
  R macc = (R)0;

  const int q_first = ks - 2, q_last = ke + 2;
  gload(q_first);
  for (int q = q_first; q <= q_last; ++q) {
    lstore(q);
    if (q < q_last) gload(q + 1);  // prefetch; lands while this plane computes
    __syncthreads();
    const int kn = q - 1;  // plane whose node / solute quantities are built now
    Node<R> nr{};
    R rig = 0, rmu = 0, cnew_col = 0;
    if (q >= ks) {
      if (kn >= 2 && kn <= nz - 1) {  // nodes are only used for planes [2, nz-1]
        nP = node_at(q, olx, oly);
        if (tid < NRING) nr = node_at(q, rlx, rly);
      }
      if constexpr (SWEEP_B) {
        if (kn >= 1 && kn <= nz) {
          const R *P0 = slot(q - 1);
          const size_t pk = (size_t)(kn - 1) * sz;
          cnew_col = co[pk + ooff];
          solute_node(P0[(oly + 1) * PX + olx + 1], cnew_col, quint, s, igP, muP);
          if (tid < NRING)
            solute_node(P0[(rly + 1) * PX + rlx + 1], co[pk + roff], quint, s, rig, rmu);
        }
      }
    }
    const int k = q - 2;  // cell plane updated now
    if (q >= ks + 2 && own) {
      const R *Pm = slot(k - 1), *P0 = slot(k), *Pp = slot(k + 1);
      const int c = (ty + 2) * PX + (tx + 2);
      const R PP = P0[c];
      const size_t gidx = (size_t)(k - 1) * sz + ooff;
      R dppv = (R)0, pnew = PP;
      R cPP;
      if constexpr (SWEEP_B) cPP = cZ; else cPP = co[gidx];
      if (ph_col && k >= 3 && k <= nz - 2) {
        const R PE = P0[c + 1], PW = P0[c - 1], PN = P0[c + PX], PS = P0[c - PX];
        const R PNE = P0[c + PX + 1], PSE = P0[c - PX + 1], PNW = P0[c + PX - 1],
                PSW = P0[c - PX - 1];
        const R PT = Pp[c], PB = Pm[c];
        const R PNT = Pp[c + PX], PST = Pp[c - PX], PET = Pp[c + 1], PWT = Pp[c - 1];
        const R PNB = Pm[c + PX], PSB = Pm[c - PX], PEB = Pm[c + 1], PWB = Pm[c - 1];
        const int nc = oly * QX + olx;
        const R *e2 = snode, *xy = snode + QPL, *yz = snode + 2 * QPL, *zx = snode + 3 * QPL;
        const R e2E = e2[nc + 1], e2W = e2[nc - 1], e2N = e2[nc + QX], e2S = e2[nc - QX];
        const R xyE = xy[nc + 1], xyW = xy[nc - 1], xyN = xy[nc + QX], xyS = xy[nc - QX];
        const R yzN = yz[nc + QX], yzS = yz[nc - QX];
        const R zxE = zx[nc + 1], zxW = zx[nc - 1];
        const R yzT = nP.yz, yzB = nM.yz, zxT = nP.zx, zxB = nM.zx;
        R AA = (e2E * (PE - PP) + e2W * (PW - PP)) * s.idx2 +
               (e2N * (PN - PP) + e2S * (PS - PP)) * s.idy2 +
               (nP.e2 * (PT - PP) + nM.e2 * (PB - PP)) * s.idz2;
        const R Cxy = PNE * (xyE - xyN) + PSE * (xyS - xyE) + PNW * (xyN - xyW) +
                      PSW * (xyW - xyS) + PN * (xyE - xyW) + PS * (xyW - xyE) +
                      PE * (xyS - xyN) + PW * (xyN - xyS);
        const R Cyz = PNT * (yzN - yzT) + PNB * (yzB - yzN) + PST * (yzT - yzS) +
                      PSB * (yzS - yzB) + PT * (yzN - yzS) + PB * (yzS - yzN) +
                      PN * (yzB - yzT) + PS * (yzT - yzB);
        const R Czx = PET * (zxT - zxE) + PWT * (zxW - zxT) + PEB * (zxE - zxB) +
                      PWB * (zxB - zxW) + PE * (zxT - zxB) + PW * (zxB - zxT) +
                      PT * (zxW - zxE) + PB * (zxE - zxW);
        AA = AA + s.cxy * Cxy + s.cyz * Cyz + s.czx * Czx;
        const R om = (R)1 - PP;
        const R gl = (R)2 * PP * om * (om - PP);  // 2p(1-p)^2 - 2p^2(1-p)
        R h, hl;
        h_of(PP, quint, h, hl);
        // (1-Cs)(1-CLP)/((1-Cl)(1-CSP)) with CLP = c/den, CSP = Ke CLP
        const R den = s.Ke * h + ((R)1 - h);
        const R arg = s.omCsCl * (den - cPP) * rcpn<R>(den - s.Ke * cPP);
        const R GG = hl * s.RTV * dlog<R>(fabs(arg)) - s.W0 * gl;
        R ruidos = (R)0;
        if (noise) {
          const float ran = kp.ruido_tipo == PF_RUIDO_SMOOTH
                                ? (float)ruido_suave(i, j, k, kp.it, kp.seed)
                                : (float)hash_unit(hbase);
          const R amp = kp.ruido_tipo == PF_RUIDO_CELL_POSITIVE ? (R)ran * s.AAA
                                                                : ((R)2 * (R)ran - (R)1) * s.AAA;
          ruidos = amp * (R)16 * (PP * PP) * (om * om);
        }
        const R pstar = PP + s.M0dt * (AA + GG) + ruidos;
        dppv = pstar - PP;
        const R cl = pstar > (R)1 ? (R)1 : pstar;
        pnew = cl < (R)0 ? (R)0 : cl;
        pn[gidx] = pnew;
      }
      if constexpr (SWEEP_B) {
        // solute at (i,j,k), k in [2,nz-1] (conc_body.inc:92-144): face
        // mobility 2 D_C/(1/g_P + 1/g_X), flux along mu differences
        const int nc = oly * QX + olx;
        const R *ig = sgm, *mu = sgm + QPL;
        const R D2dt = PP <= (R)0.95 ? s.Dliq2dt : s.Dsol2dt;
        const R igc = igZ, muc = muZ;
        const R fx = dmu<R>(mu[nc + 1], muc) * rcpn<R>(ig[nc + 1] + igc) +
                     dmu<R>(mu[nc - 1], muc) * rcpn<R>(ig[nc - 1] + igc);
        const R fy = dmu<R>(mu[nc + QX], muc) * rcpn<R>(ig[nc + QX] + igc) +
                     dmu<R>(mu[nc - QX], muc) * rcpn<R>(ig[nc - QX] + igc);
        const R fz = dmu<R>(muP, muc) * rcpn<R>(igP + igc) + dmu<R>(muM, muc) * rcpn<R>(igM + igc);
        const R inc = D2dt * (fx * s.idx2 + fy * s.idy2 + fz * s.idz2);
        cn[gidx] = cPP + inc;
        macc += inc;
        if (write_T) {
          // temp_body.inc:31-41 of the next step: TP + D dt lap(T) + fonte
          R hh, hl2;
          h_of(pnew, quint, hh, hl2);
          R lap = (R)0;
          if (s.Ddt != (R)0) {
            const R TP2 = (R)2 * s.TP;
            lap = (Tcur[gidx + 1] - TP2 + Tcur[gidx - 1]) * s.idx2 +
                  (Tcur[gidx + nx] - TP2 + Tcur[gidx - nx]) * s.idy2 +
                  (Tcur[gidx + sz] - TP2 + Tcur[gidx - sz]) * s.idz2;
          }
          Tnext[gidx] = s.Ddt * lap + s.fonte * hl2 * dppv + s.TP;
        } else if (ph_col && k >= 3 && k <= nz - 2) {
          dpp[gidx] = dppv;
        }
      } else {
        (void)Tcur; (void)Tnext; (void)dpp; (void)write_T;
      }
      hbase += kHashStepK;
    }
    __syncthreads();
    // publish plane kn's node / solute quantities and rotate the column
    if (q >= ks) {
      const int oc = oly * QX + olx, rc = rly * QX + rlx;
      if (kn >= 2 && kn <= nz - 1) {
        snode[oc] = nP.e2; snode[QPL + oc] = nP.xy; snode[2 * QPL + oc] = nP.yz;
        snode[3 * QPL + oc] = nP.zx;
        if (tid < NRING) {
          snode[rc] = nr.e2; snode[QPL + rc] = nr.xy; snode[2 * QPL + rc] = nr.yz;
          snode[3 * QPL + rc] = nr.zx;
        }
      }
      if constexpr (SWEEP_B) {
        sgm[oc] = igP; sgm[QPL + oc] = muP;
        if (tid < NRING) {
          sgm[rc] = rig; sgm[QPL + rc] = rmu;
        }
      }
    }
    nM = nZ; nZ = nP;
    if constexpr (SWEEP_B) {
      igM = igZ; muM = muZ; igZ = igP; muZ = muP;
      cZ = cnew_col;
    }
  }
  if constexpr (SWEEP_B) {
    if (mslot) {  // uniform over the work-group
      __shared__ double wsum[NT / 64];
      double v = (double)macc;
      for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
      if ((tid & 63) == 0) wsum[tid >> 6] = v;
      __syncthreads();
      if (tid == 0) {
        double t = 0.0;
        for (int w = 0; w < NT / 64; ++w) t += wsum[w];
        mslot[blockIdx.x] += t;
      }
    }
  } else {
    (void)mslot; (void)macc;
  }
}

template <typename R, typename TL>
size_t march_lds(bool sweep_b) {
  return sizeof(R) * (size_t)(NSLOT * TL::PPL + 4 * TL::QPL + (sweep_b ? 2 * TL::QPL : 0));
}

using Tile64x8 = Tile<64, 8>;
using Tile32x8 = Tile<32, 8>;
using Tile32x16 = Tile<32, 16>;
constexpr int kTileX[PF_NVARIANTS] = {64, 32, 32};
constexpr int kTileY[PF_NVARIANTS] = {8, 8, 16};

Geo make_geo(int nx, int ny, int nz, int k_lo, int k_hi, int seg, int tile) {
  Geo g;
  g.nx = nx; g.ny = ny; g.nz = nz;
  g.nbx = (nx - 2 + kTileX[tile] - 1) / kTileX[tile];
  g.nby = (ny - 2 + kTileY[tile] - 1) / kTileY[tile];
  g.k_lo = k_lo; g.k_hi = k_hi;
  const int ncell = k_hi - k_lo + 1;
  g.seg = seg < 1 ? 1 : (seg > ncell ? ncell : seg);
  g.nseg = (ncell + g.seg - 1) / g.seg;
  g.ntiles = g.nbx * g.nby * g.nseg;
  g.per_xcd = (g.ntiles + 7) / 8;
  return g;
}

// Segment length: enough work-groups to fill 256 CUs a few times over while
// keeping the z warm-up (4 extra planes per segment) small.
int pick_seg(int nx, int ny, int ncell, int tile, int target) {
  const int tiles = ((nx - 2 + kTileX[tile] - 1) / kTileX[tile]) *
                    ((ny - 2 + kTileY[tile] - 1) / kTileY[tile]);
  int nseg = (target + tiles - 1) / tiles;
  if (nseg < 1) nseg = 1;
  int seg = (ncell + nseg - 1) / nseg;
  if (seg < 16) seg = 16;
  return seg;
}

template <typename R, typename TL>
hipError_t launch_march_t(bool sweep_b, const Geo &g, const KParams &kp, const SC<R> &s,
                          const R *po, const R *co, R *pn, R *cn, const R *Tcur, R *Tnext, R *dpp,
                          int write_T, double *mslot, hipStream_t st) {
  const dim3 blk(TL::TX, TL::TY, 1), grd(8 * g.per_xcd, 1, 1);
  const size_t lds = march_lds<R, TL>(sweep_b);
  if (sweep_b)
    hipLaunchKernelGGL((k_march<R, true, TL>), grd, blk, lds, st, g, kp, s, po, co, pn, cn, Tcur,
                       Tnext, dpp, write_T, mslot);
  else
    hipLaunchKernelGGL((k_march<R, false, TL>), grd, blk, lds, st, g, kp, s, po, co, pn, cn, Tcur,
                       Tnext, dpp, write_T, nullptr);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_march(bool sweep_b, int tile, const Geo &g, const KParams &kp, const R *po,
                        const R *co, R *pn, R *cn, const R *Tcur, R *Tnext, R *dpp, int write_T,
                        double *mslot, hipStream_t st) {
  const SC<R> s = make_sc<R>(kp);
#define PF_MARCH(TL) \
  return launch_march_t<R, TL>(sweep_b, g, kp, s, po, co, pn, cn, Tcur, Tnext, dpp, write_T, \
                               mslot, st)
  switch (tile) {
    case 1: PF_MARCH(Tile32x8);
    case 2: PF_MARCH(Tile32x16);
    default: PF_MARCH(Tile64x8);
  }
#undef PF_MARCH
}

// The ledger is kept while it is valid and has a slot for every sweep-B
// work-group (fused_mass_prepare sizes the slots before a run).
bool mass_on(const FusedState &S, const Geo &gb) {
  return S.mass_every > 0 && S.ledger_valid && S.mslot && 8 * gb.per_xcd <= S.nslot;
}

template <typename R>
int step_impl(FusedState &S, const KParams &kt, const KParams &k, const FusedStep &fs,
              void *phi[2], int *phi_old, bool *phi_alias, void *c[2], int *c_old, bool *c_alias,
              void *T[2], int *T_old, bool *T_alias, void *dpp, double *probe, hipStream_t st,
              KTimer &tm) {
  const int nx = S.nx, ny = S.ny, nz = S.nz;
  hipError_t e;
  auto NEWW = [](int old, bool &alias) { alias = false; return 1 - old; };
  // temperatura for this step (later steps get it from the previous sweep B)
  if (fs.first) {
    const int tn = NEWW(*T_old, *T_alias);
    tm.begin(PF_KT_TEMPERATURA, st);
    e = launch_temperatura<R>(kt, (const R *)T[*T_old], (const R *)phi[*phi_old], (const R *)dpp,
                              (R *)T[tn], st);
    tm.end(st);
    if (e != hipSuccess) return PF_E_HIP;
  }
  const Geo ga = make_geo(nx, ny, nz, 3, nz - 2, S.seg_a, S.tile_a);
  const Geo gb = make_geo(nx, ny, nz, 2, nz - 1, S.seg_b, S.tile_b);
  size_t pa, pb;
  probe_cells(nx, ny, nz, pa, pb);
  for (int sw = 1; sw <= fs.nv; ++sw) {
    if (sw > 1) {  // principal.f90:127-150
      if (!*phi_alias) *phi_old = 1 - *phi_old;
      *phi_alias = true;
    }
    const int pnw = NEWW(*phi_old, *phi_alias);
    const R *po = (const R *)phi[*phi_old];
    R *pn = (R *)phi[pnw];
    if (sw < fs.nv) {
      tm.begin(PF_KT_SWEEP_A, st);
      e = launch_march<R>(false, S.tile_a, ga, k, po, (const R *)c[*c_old], pn, nullptr, nullptr,
                          nullptr, nullptr, 0, nullptr, st);
      tm.end(st);
      if (e != hipSuccess) return PF_E_HIP;
      tm.begin(PF_KT_BOUNDARY, st);
      e = launch_phase_boundary<R>(k, po, pn, st);
      tm.end(st);
      if (e != hipSuccess) return PF_E_HIP;
    } else {
      const int cnw = NEWW(*c_old, *c_alias);
      // T of the next step goes into the buffer of this step's TOLD (dead now)
      const int t_cur = *T_alias ? *T_old : 1 - *T_old;
      const int t_nxt = 1 - t_cur;
      tm.begin(PF_KT_SWEEP_B, st);
      e = launch_march<R>(true, S.tile_b, gb, k, po, (const R *)c[*c_old], pn, (R *)c[cnw],
                          (const R *)T[t_cur], (R *)T[t_nxt], (R *)dpp, fs.last ? 0 : 1,
                          mass_on(S, gb) ? S.mslot : nullptr, st);
      tm.end(st);
      if (e != hipSuccess) return PF_E_HIP;
      tm.begin(PF_KT_BOUNDARY, st);
      e = launch_phase_boundary<R>(k, po, pn, st);
      if (e == hipSuccess) e = launch_solute_boundary<R>(k, (const R *)c[*c_old], (R *)c[cnw], st);
      tm.end(st);
      if (e != hipSuccess) return PF_E_HIP;
    }
    if (probe) {
      e = launch_probe<R>((const R *)pn, pa, pb, probe + 2 * (sw - 1), st);
      if (e != hipSuccess) return PF_E_HIP;
    }
  }
  // end of step (principal.f90:200-226)
  if (!*phi_alias) *phi_old = 1 - *phi_old;
  *phi_alias = true;
  if (!*c_alias) *c_old = 1 - *c_old;
  *c_alias = true;
  {
    const int t_cur = *T_alias ? *T_old : 1 - *T_old;  // holds T of this step
    *T_old = t_cur;
    // not last: the other buffer holds the next step's T (its temperatura)
    *T_alias = fs.last;
  }
  return PF_OK;
}

}  // namespace

int fused_init(FusedState &s, int nx, int ny, int nz, int prec) {
  s.nx = nx; s.ny = ny; s.nz = nz; s.prec = prec;
  s.enabled = nx >= 6 && ny >= 6 && nz >= 6;
  // defaults until pf_autotune measures: the shapes that were fastest for
  // both precisions on 256^2 x 528 (32x8 for the phase-only sweep, 32x16
  // for the phase + solute sweep, about 1024 work-groups)
  fused_set_tile(s, 1, 2, 1024);
  s.mass_every = prec == PF_FP32_MIXED ? 100 : 0;
  return PF_OK;
}

int fused_set_tile(FusedState &s, int tile_a, int tile_b, int target_groups) {
  if (tile_a < 0 || tile_a >= PF_NVARIANTS || tile_b < 0 || tile_b >= PF_NVARIANTS ||
      target_groups < 1)
    return PF_E_ARG;
  s.tile_a = tile_a;
  s.tile_b = tile_b;
  s.target = target_groups;
  s.seg_a = pick_seg(s.nx, s.ny, s.nz - 4, tile_a, target_groups);
  s.seg_b = pick_seg(s.nx, s.ny, s.nz - 2, tile_b, target_groups);
  return PF_OK;
}

namespace {

template <typename R>
int autotune_impl(FusedState &S, const KParams &k, void *phi_old, void *c_old, void *T_old,
                  int reps, hipStream_t st, AutotuneResult *res) {
  const int nx = S.nx, ny = S.ny, nz = S.nz;
  const size_t n = (size_t)nx * ny * nz;
  R *out = nullptr;  // scratch outputs: phi, c, T (the context's fields are not touched)
  if (hipMalloc((void **)&out, 3 * n * sizeof(R)) != hipSuccess) return PF_E_NOMEM;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = PF_OK;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = PF_E_HIP;
  static const int targets[] = {1024, 2048, 4096, 8192};
  const R *po = (const R *)phi_old, *co = (const R *)c_old, *Tc = (const R *)T_old;
  float best[2] = {1e30f, 1e30f};
  int cand = 0;
  for (int sweep_b = 0; sweep_b < 2 && rc == PF_OK; ++sweep_b) {
    for (int tile = 0; tile < PF_NVARIANTS && rc == PF_OK; ++tile) {
      for (int target : targets) {
        const int seg = pick_seg(nx, ny, sweep_b ? nz - 2 : nz - 4, tile, target);
        const Geo g = sweep_b ? make_geo(nx, ny, nz, 2, nz - 1, seg, tile)
                              : make_geo(nx, ny, nz, 3, nz - 2, seg, tile);
        hipError_t e = hipSuccess;
        for (int r = 0; r <= reps && e == hipSuccess; ++r) {  // r = 0 is the warm-up
          if (r == 1) hipEventRecord(e0, st);
          e = launch_march<R>(sweep_b != 0, tile, g, k, po, co, out, out + n, Tc, out + 2 * n,
                              nullptr, 1, nullptr, st);
        }
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventRecord(e1, st);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) { rc = PF_E_HIP; break; }
        ms /= (float)reps;
        if (res && cand < PF_AUTOTUNE_MAX) {
          res->cand[cand].sweep = sweep_b;
          res->cand[cand].tile = tile;
          res->cand[cand].target = target;
          res->cand[cand].ms = ms;
        }
        ++cand;
        if (ms < best[sweep_b]) {
          best[sweep_b] = ms;
          if (sweep_b) { S.tile_b = tile; S.seg_b = seg; }
          else { S.tile_a = tile; S.seg_a = seg; }
        }
      }
    }
  }
  if (res) {
    res->n = cand < PF_AUTOTUNE_MAX ? cand : PF_AUTOTUNE_MAX;
    res->best_ms[0] = best[0];
    res->best_ms[1] = best[1];
  }
  if (e0) hipEventDestroy(e0);
  if (e1) hipEventDestroy(e1);
  hipStreamSynchronize(st);
  hipFree(out);
  return rc;
}

}  // namespace

int fused_autotune(FusedState &s, const KParams &k, void *phi_old, void *c_old, void *T_old,
                   int reps, hipStream_t stream, AutotuneResult *res) {
  if (!s.enabled) return PF_E_STATE;
  if (reps < 1) return PF_E_ARG;
  if (s.prec == PF_FP64)
    return autotune_impl<double>(s, k, phi_old, c_old, T_old, reps, stream, res);
  return autotune_impl<float>(s, k, phi_old, c_old, T_old, reps, stream, res);
}

// ---- solute-mass ledger (FP32-mixed form) -----------------------------------
// The ledger is the FP64 interior solute mass the scheme's own increments
// imply: its initial FP64 sum plus the FP64 sum of every FP32 increment
// computed by sweep B (before the increment is rounded into the FP32 field).
// Every mass_every steps the FP64 sum of the stored field over the same
// interior [2,n-1]^3 is compared with it and the interior is rescaled by
// ledger/sum (FP64 product, unbiased stochastic rounding to FP32), which
// removes the drift that FP32 storage rounding accumulates.  The face planes
// are then refreshed from their interior sources: left ~1 ulp above the
// rescaled interior (the factor is mostly < 1), they would feed solute back
// through the domain faces at every correction.

namespace {

constexpr int kMassBlocks = 1024;

__device__ __forceinline__ double block_sum256(double v, double *sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) sh[t] += sh[t + w];
    __syncthreads();
  }
  return sh[0];
}

template <typename R>
__global__ void __launch_bounds__(256) k_mass_interior(const R *__restrict__ c, int nx, int ny,
                                                       int nz, double *__restrict__ part) {
  __shared__ double sh[256];
  const int nrow = (ny - 2) * (nz - 2);
  const size_t sz = (size_t)nx * ny;
  double acc = 0.0;
  for (int r = blockIdx.x; r < nrow; r += gridDim.x) {
    const int j = 2 + r % (ny - 2), k = 2 + r / (ny - 2);
    const size_t row = (size_t)(j - 1) * nx + (size_t)(k - 1) * sz;
    for (int i = 2 + (int)threadIdx.x; i <= nx - 1; i += 256) acc += (double)c[row + i - 1];
  }
  const double t = block_sum256(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// mode 0: ledger := interior sum (slots discarded); 1: ledger += slots;
// 2: ledger += slots, mass[1] := interior sum, mass[2] := ledger / sum.
// Slots are zeroed in every mode.
__global__ void __launch_bounds__(256) k_mass_update(const double *__restrict__ part, int npart,
                                                     double *__restrict__ mslot, int nslot,
                                                     double *__restrict__ mass, int mode) {
  __shared__ double sh[256];
  double a = 0.0, b = 0.0;
  for (int q = threadIdx.x; q < npart; q += 256) a += part[q];
  for (int q = threadIdx.x; q < nslot; q += 256) {
    b += mslot[q];
    mslot[q] = 0.0;
  }
  const double S = block_sum256(a, sh);
  __syncthreads();
  const double L = block_sum256(b, sh);
  if (threadIdx.x == 0) {
    if (mode == 0) {
      mass[0] = S;
    } else {
      mass[0] += L;
      if (mode == 2) {
        mass[1] = S;
        mass[2] = mass[0] / S;
      }
    }
  }
}

// Rescales the interior cells by mass[2].  FP32: the FP64 product is
// rounded stochastically to one of its two neighbouring floats with the
// probability that makes the expectation exact (a counter hash of cell and
// step, salt 4, draws the choice), because the factor is typically 1 -+ 1e-9,
// far below one FP32 ulp, and round-to-nearest would leave every cell as it
// was.  The summed rounding error is then a zero-mean random walk of about
// sqrt(n) * ulp / 2.
template <typename R>
__global__ void __launch_bounds__(256) k_mass_scale(R *__restrict__ c, int nx, int ny, int nz,
                                                    int it, uint64_t seed,
                                                    const double *__restrict__ mass) {
  const double f = mass[2];
  const int nrow = (ny - 2) * (nz - 2);
  const size_t sz = (size_t)nx * ny;
  for (int r = blockIdx.x; r < nrow; r += gridDim.x) {
    const int j = 2 + r % (ny - 2), k = 2 + r / (ny - 2);
    const size_t row = (size_t)(j - 1) * nx + (size_t)(k - 1) * sz;
    for (int i = 2 + (int)threadIdx.x; i <= nx - 1; i += 256) {
      const size_t t = row + i - 1;
      const double x = (double)c[t] * f;
      if constexpr (sizeof(R) == 8) {
        c[t] = (R)x;
      } else {
        const float nearest = (float)x;
        float lo = nearest, hi = nearest;
        if ((double)nearest > x) lo = nextafterf(nearest, -__builtin_huge_valf());
        else if ((double)nearest < x) hi = nextafterf(nearest, __builtin_huge_valf());
        float out = lo;
        if (hi != lo) {
          const double p = (x - (double)lo) / ((double)hi - (double)lo);
          out = hash_unit(hash_base(i, j, k, it, 4, seed)) < p ? hi : lo;
        }
        c[t] = out;
      }
    }
  }
}

template <typename R>
hipError_t mass_interior(const FusedState &S, const void *c, hipStream_t st) {
  hipLaunchKernelGGL(k_mass_interior<R>, dim3(kMassBlocks), dim3(256), 0, st, (const R *)c, S.nx,
                     S.ny, S.nz, S.mpart);
  return hipGetLastError();
}

}  // namespace

int fused_mass_prepare(FusedState &S, const void *c_old, hipStream_t st) {
  if (S.mass_every <= 0 || !S.enabled) return PF_OK;
  const Geo gb = make_geo(S.nx, S.ny, S.nz, 2, S.nz - 1, S.seg_b, S.tile_b);
  const int need = 8 * gb.per_xcd;
  if (!S.mass) {
    if (hipMalloc((void **)&S.mass, 4 * sizeof(double)) != hipSuccess) return PF_E_NOMEM;
    if (hipMalloc((void **)&S.mpart, kMassBlocks * sizeof(double)) != hipSuccess)
      return PF_E_NOMEM;
    if (hipMemsetAsync(S.mass, 0, 4 * sizeof(double), st) != hipSuccess) return PF_E_HIP;
  }
  if (need > S.nslot) {  // slots are folded (zero) between runs: safe to replace
    if (S.mslot) {
      if (hipStreamSynchronize(st) != hipSuccess) return PF_E_HIP;
      hipFree(S.mslot);
      S.mslot = nullptr;
      S.nslot = 0;
    }
    if (hipMalloc((void **)&S.mslot, (size_t)need * sizeof(double)) != hipSuccess)
      return PF_E_NOMEM;
    if (hipMemsetAsync(S.mslot, 0, (size_t)need * sizeof(double), st) != hipSuccess)
      return PF_E_HIP;
    S.nslot = need;
  }
  if (!S.ledger_valid) {
    hipError_t e = S.prec == PF_FP64 ? mass_interior<double>(S, c_old, st)
                                     : mass_interior<float>(S, c_old, st);
    if (e != hipSuccess) return PF_E_HIP;
    hipLaunchKernelGGL(k_mass_update, dim3(1), dim3(256), 0, st, S.mpart, kMassBlocks, S.mslot,
                       S.nslot, S.mass, 0);
    if (hipGetLastError() != hipSuccess) return PF_E_HIP;
    S.ledger_valid = true;
  }
  return PF_OK;
}

int fused_mass_fold(FusedState &S, hipStream_t st) {
  if (!S.ledger_valid || !S.mslot) return PF_OK;
  hipLaunchKernelGGL(k_mass_update, dim3(1), dim3(256), 0, st, S.mpart, 0, S.mslot, S.nslot,
                     S.mass, 1);
  return hipGetLastError() == hipSuccess ? PF_OK : PF_E_HIP;
}

int fused_mass_correct(FusedState &S, const KParams &k, void *c, hipStream_t st) {
  if (!S.ledger_valid || !S.mslot) return PF_OK;
  const int it = k.it;
  const uint64_t seed = k.seed;
  hipError_t e = S.prec == PF_FP64 ? mass_interior<double>(S, c, st)
                                   : mass_interior<float>(S, c, st);
  if (e != hipSuccess) return PF_E_HIP;
  hipLaunchKernelGGL(k_mass_update, dim3(1), dim3(256), 0, st, S.mpart, kMassBlocks, S.mslot,
                     S.nslot, S.mass, 2);
  if (hipGetLastError() != hipSuccess) return PF_E_HIP;
  if (S.prec == PF_FP64) {
    hipLaunchKernelGGL(k_mass_scale<double>, dim3(kMassBlocks), dim3(256), 0, st, (double *)c, S.nx,
                       S.ny, S.nz, it, seed, S.mass);
    e = hipGetLastError();
    if (e == hipSuccess) e = launch_solute_refresh<double>(k, (double *)c, st);
  } else {
    hipLaunchKernelGGL(k_mass_scale<float>, dim3(kMassBlocks), dim3(256), 0, st, (float *)c, S.nx,
                       S.ny, S.nz, it, seed, S.mass);
    e = hipGetLastError();
    if (e == hipSuccess) e = launch_solute_refresh<float>(k, (float *)c, st);
  }
  return e == hipSuccess ? PF_OK : PF_E_HIP;
}

int fused_mass_read(FusedState &S, double out[3], hipStream_t st) {
  out[0] = out[1] = out[2] = 0.0;
  if (!S.mass) return PF_OK;
  if (hipMemcpyAsync(out, S.mass, 3 * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return PF_E_HIP;
  return PF_OK;
}

void fused_free(FusedState &s) {
  s.enabled = false;
  if (s.mass) hipFree(s.mass);
  if (s.mpart) hipFree(s.mpart);
  if (s.mslot) hipFree(s.mslot);
  s.mass = s.mpart = s.mslot = nullptr;
  s.nslot = 0;
  s.ledger_valid = false;
}
size_t fused_bytes(const FusedState &s) {
  return s.mass ? (4 + kMassBlocks + (size_t)s.nslot) * sizeof(double) : 0;
}

bool fused_supported(const FusedState &s, const pf_params &p) {
  (void)p;
  return s.enabled;
}

int fused_step(FusedState &s, const KParams &kt, const KParams &k, const FusedStep &fs,
               void *phi[2], int *phi_old, bool *phi_alias, void *c[2], int *c_old,
               bool *c_alias, void *T[2], int *T_old, bool *T_alias, void *dpp, double *probe,
               hipStream_t stream, KTimer &timer) {
  if (s.prec == PF_FP64)
    return step_impl<double>(s, kt, k, fs, phi, phi_old, phi_alias, c, c_old, c_alias, T, T_old,
                             T_alias, dpp, probe, stream, timer);
  return step_impl<float>(s, kt, k, fs, phi, phi_old, phi_alias, c, c_old, c_alias, T, T_old,
                          T_alias, dpp, probe, stream, timer);
}

}  // namespace pf

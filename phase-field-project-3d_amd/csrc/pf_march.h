// pf_march.h -- the fused z-march sweep body shared by the single-instance
// kernels (pf_fused.hip) and the batched kernels (pf_batch.hip).
//
// A work-group owns a TX x TY column tile (x fastest) and a segment of
// z-planes and marches up the segment:
//   - phi planes (tile + 2-cell halo) rotate through 5 LDS slots, loaded once
//     from HBM with x-contiguous (coalesced) reads;
//   - the anisotropy quantities of a phase cell's six neighbours
//     (phase_body.inc:55-129) are the centred-gradient quantities of those
//     neighbour nodes, so each node's eps^2, exy, eyz, ezx is computed once
//     (tile + 1-cell ring), shared in-plane through LDS and along z in
//     registers -- instead of six times per cell;
//   - sweep B also evaluates the solute (conc_body.inc) with the per-node
//     1/gamma and mu computed once per node, and the temperature of the next
//     step (temp_body.inc) from the fresh DPP, which therefore never goes to
//     HBM except on the last step of a call.
// Compulsory HBM traffic per cell: sweep A reads phi, c and writes phi;
// sweep B reads phi, c and writes phi, c, T.
//
// Arithmetic follows the reference's formulas with algebraic reuse
// (reciprocals of dx^2, node reuse, the harmonic mean as 2/(1/g+1/g'), the
// free-energy ratio with CL = c/den folded in, divisions as a refined
// reciprocal, the ~1 ulp logarithm of pf_math.h), so results agree with the
// reference to rounding, not bit for bit; the drop-in kernels keep the
// bit-faithful form.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

#include "pf_common.h"
#include "pf_mass.h"
#include "pf_math.h"

namespace pf {

constexpr int NSLOT = 5;

// Tile shape: TX x TY cells per work-group (one thread per column).
template <int TX_, int TY_>
struct Tile {
  static constexpr int TX = TX_, TY = TY_, NT = TX * TY;
  static constexpr int PX = TX + 4, PY = TY + 4, PPL = PX * PY;  // phi plane, halo 2
  static constexpr int QX = TX + 2, QY = TY + 2, QPL = QX * QY;  // node plane, halo 1
  // ring nodes per plane: the halo nodes east/west/north/south of the tile's
  // edge cells (the four corner nodes of the halo are no cell's neighbour)
  static constexpr int NRING = 2 * TX + 2 * TY;
  static constexpr int NLD = (PPL + NT - 1) / NT;                 // phi loads per thread
  static_assert(NRING <= NT, "one ring node per thread at most");
  static_assert(TX % 16 == 0 && NT % 64 == 0, "whole wavefronts, 128-byte row segments");
};

// Bytes of a march_tile work-group's LDS arrays (before the FP64 log
// table): NSLOT phi planes and four node planes of the arithmetic type R
// and, in sweep B, two solute node planes (1/gamma and q).
template <typename R, typename TL>
constexpr size_t march_arrays_bytes(bool sweep_b) {
  return (size_t)(NSLOT * TL::PPL + 4 * TL::QPL + (sweep_b ? 2 * TL::QPL : 0)) * sizeof(R);
}

// log and 1/x in the arithmetic type.  FP64: the ~1 ulp table-driven
// logarithm (its table in LDS, `lt`) and the one-Newton-step reciprocal
// (<= 11 ulp) of pf_math.h; FP32: the hardware log2 / reciprocal (about
// 1 ulp).
template <typename R>
__device__ __forceinline__ R dlog(R x, const LogEnt *lt);
template <>
__device__ __forceinline__ double dlog<double>(double x, const LogEnt *lt) {
  return log_tab(x, lt);
}
template <>
__device__ __forceinline__ float dlog<float>(float x, const LogEnt *) { return __logf(x); }

template <typename R>
__device__ __forceinline__ R rcpn(R x);
template <>
__device__ __forceinline__ double rcpn<double>(double x) { return rcp_fast(x); }
template <>
__device__ __forceinline__ float rcpn<float>(float x) { return __builtin_amdgcn_rcpf(x); }

// 1/x for a denominator that is positive and finite for finite inputs.
template <typename R>
__device__ __forceinline__ R rcpn_pos(R x);
template <>
__device__ __forceinline__ double rcpn_pos<double>(double x) { return rcp_pos(x); }
template <>
__device__ __forceinline__ float rcpn_pos<float>(float x) { return __builtin_amdgcn_rcpf(x); }

// The FP64 logarithm's table in LDS: kLogTabN 16-byte entries placed after
// a kernel's other LDS arrays (`bytes` from the start), filled by the
// work-group before its first barrier.  The FP32 form has none.
template <typename R>
constexpr size_t log_tab_bytes() {
  return sizeof(R) == 8 ? sizeof(LogEnt) * (size_t)kLogTabN : 0;
}
constexpr size_t log_tab_offset(size_t bytes) { return (bytes + 15) & ~(size_t)15; }
template <typename R>
__device__ __forceinline__ const LogEnt *log_tab_fill(unsigned char *smem, size_t bytes,
                                                      int tid, int nt) {
  LogEnt *lt = reinterpret_cast<LogEnt *>(smem + log_tab_offset(bytes));
  if constexpr (sizeof(R) == 8) {
    for (int t = tid; t < kLogTabN; t += nt) lt[t] = LogEnt{kLogTab[t][0], kLogTab[t][1]};
  }
  return lt;
}

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
  R coff;              // offset of the stored c (KParams::c_off; 0 for FP64)
  // the free-energy ratio K (1-CL)/(1-Ke CL) minus 1 (FP32 form, see
  // ratio_ln_f32): A = K - 1, B = K - Ke, K1 = A - c_off B, gAh = (1 - Ke) A
  R gAh, gB, gK1;
};

// Computed on the host for a single instance and on the device for batched
// instances; contraction off so that both give the same bits.
template <typename R>
__host__ __device__ inline SC<R> make_sc(const KParams &p) {
#pragma clang fp contract(off)
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
  s.coff = (R)p.c_off;
  {
    const double K = (1.0 - p.Cs_eq) / (1.0 - p.Cl_eq);
    const double A = K - 1.0, B = K - p.Ke;
    s.gAh = (R)((1.0 - p.Ke) * A);
    s.gB = (R)B;
    s.gK1 = (R)(A - p.c_off * B);
  }
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
  const R rq = rcpn_pos<R>(gx2 + gy2 + gz2 + s.small);
  const R a = gx2 * rq, b = gy2 * rq, c = gz2 * rq;
  const R eps = s.E0a + s.E0b * (a * a + b * b + c * c);
  // K eps (g_x g_y / |g|^2)(a - b) etc., with K eps / |g|^2 formed once
  const R frq = s.K * eps * rq;
  Node<R> v;
  v.e2 = eps * eps;
  v.xy = frq * (gx * gy) * (a - b);
  v.yz = frq * (gy * gz) * (b - c);
  v.zx = frq * (gz * gx) * (c - a);
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

// ln of the free-energy ratio K (den - c)/(den - Ke c) of the phase
// equation (phase_body.inc:159-170, with CL = c/den folded in), FP32 form.
// Near equilibrium the ratio is within a few 1e-3 of 1, and forming it in
// FP32 before the logarithm leaves an absolute error of ~1 ulp of 1
// (6e-8) in ln, i.e. ~50 J/m^3 of driving force per cell, which pushed phi
// off the FP64 trajectory by ~1e-5 within 1,000 steps and, through the
// reference's non-conservative solute faces, the solute mass by up to 1e-6.
// Instead the ratio minus 1 is formed directly,
//   (den A - c B) / (den - Ke c),  A = K - 1, B = K - Ke,
// with den A - c B = K1 - h (1 - Ke) A - dc B for the stored dc = c - c_off
// (den - 1 = -h (1 - Ke); K1 = A - c_off B), whose terms are all small, and
// its logarithm is ln|1 + x| (pf_math.h; the reference takes ln of |ratio|).
__device__ __forceinline__ float ratio_ln_f32(float h, float den, float dc, float cfull,
                                              const SC<float> &s) {
  const float x = (s.gK1 - h * s.gAh - dc * s.gB) * rcpn<float>(den - s.Ke * cfull);
  return ln1p_abs_f32(x);
}

// The full c of a stored value (FP32: stored as c - c_off).
template <typename R>
__device__ __forceinline__ R cfull(R stored, const SC<R> &s) {
  if constexpr (sizeof(R) == 8) return stored;
  else return s.coff + stored;
}

// Solute node quantities (conc_body.inc:61-118, 139-144): 1/gamma (gamma
// without the centre's D_C) and q, the node's share of the chemical
// potential difference taken across each face by dmu():
//   FP64: q = mu = ln(CL/(1-CL)) = ln(c/(den-c)), dmu = mu_X - mu_P;
//   FP32: q = CL - c_off = (dc + c_off (1-Ke) h) / den for the stored
//         dc = c - c_off (since 1 - den = (1-Ke) h), dmu = ln((CL_X (1-CL_P))
//         / (CL_P (1-CL_X))) (the same quantity) from the difference
//         q_X - q_P = CL_X - CL_P (see dmu<float>).
// In FP32 the difference of two ~-4 logarithms is noise of about one FP32
// ulp of mu, which in near-uniform liquid is as large as the true
// difference and would random-walk c by ~1 ulp per step; the FP32 form takes
// the difference CL_X - CL_P first, which in the liquid (h = 0, den = 1)
// is dc_X - dc_P: the increments of the near-uniform liquid keep the
// precision of the stored offsets, far below an FP32 ulp of c.
// GUARD selects the reciprocal: rcpn (a zero, infinite or NaN denominator
// gives what a division gives) or rcpn_pos (NaN there, otherwise the same
// bits).  The FP64 callers take the unguarded form and, in a wave where a
// result is NaN, evaluate the guarded one (solute_node_w, solute_inc_w):
// exact for every input at the cost of one compare in the usual case.
template <typename R, bool GUARD>
__device__ __forceinline__ R rcp_g(R x) {
  if constexpr (GUARD) return rcpn<R>(x);
  else return rcpn_pos<R>(x);
}

template <typename R, bool GUARD = true>
__device__ __forceinline__ void solute_node(R p, R cst, int quint, const SC<R> &s,
                                            const LogEnt *lt, R &ig, R &q) {
  R h, hl;
  h_of(p, quint, h, hl);
  const R den = s.Ke * h + ((R)1 - h);
  const R cc = cfull<R>(cst, s);
  const R rden = rcp_g<R, GUARD>(den);
  const R CL = cc * rden;
  const R CS = s.Ke * CL;
  const R g = ((R)1 - h) * CL * ((R)1 - CL) + h * CS * ((R)1 - CS);
  ig = rcp_g<R, GUARD>(g);
  if constexpr (sizeof(R) == 8) q = dlog<R>(cc * rcp_g<R, GUARD>(den - cc), lt);
  else q = (cst + s.coff * ((R)1 - s.Ke) * h) * rden;
}

// solute_node with the FP64 reciprocals checked per wave (see rcp_g).
template <typename R>
__device__ __forceinline__ void solute_node_w(R p, R cst, int quint, const SC<R> &s,
                                              const LogEnt *lt, R &ig, R &q) {
  if constexpr (sizeof(R) == 8) {
    solute_node<R, false>(p, cst, quint, s, lt, ig, q);
    if (__builtin_amdgcn_ballot_w64(ig != ig || q != q) != 0)
      solute_node<R, true>(p, cst, quint, s, lt, ig, q);
  } else {
    solute_node<R, true>(p, cst, quint, s, lt, ig, q);
  }
}

template <typename R>
__device__ __forceinline__ R dmu(R qX, R qP, R coff);
template <>
__device__ __forceinline__ double dmu<double>(double qX, double qP, double) { return qX - qP; }
template <>
__device__ __forceinline__ float dmu<float>(float qX, float qP, float coff) {
  // ln r, r = CL_X (1-CL_P) / (CL_P (1-CL_X)), as 2 atanh(t) with
  // t = (r-1)/(r+1) = (CL_X - CL_P) / ((CL_X + CL_P) - 2 CL_X CL_P), where
  // CL = c_off + q.  The denominator is symmetric as computed and the
  // numerator flips sign exactly, so t(X,P) = -t(P,X); the odd series and
  // the log difference below keep dmu(X,P) = -dmu(P,X) bit for bit: a
  // face's flux is the same from both sides (and carrying it down the
  // column changes nothing).
  const float cx = coff + qX, cp = coff + qP;
  const float sum = cx + cp, prod = cx * cp;  // each symmetric
  const float t = (qX - qP) * rcpn<float>(sum - 2.0f * prod);
  const float a = fabsf(t), t2 = t * t;
  // 2 atanh(t) = 2t (1 + t^2/3 + t^4/5 + ...): truncation < 1e-10 relative
  // for |t| < 1/8; a polynomial in t^2 (even) times 2t (odd), with explicit
  // FMAs so that every kernel rounds it alike
  const float ser =
      (2.0f * t) *
      fmaf(t2, fmaf(t2, fmaf(t2, fmaf(t2, 1.0f / 9.0f, 1.0f / 7.0f), 0.2f), 1.0f / 3.0f), 1.0f);
  // large jumps of CL: ln 2 (log2(CL_X (1-CL_P)) - log2(CL_P (1-CL_X))), the
  // same two products in swapped roles, so also exactly odd (and ln r >= 0.25
  // here, where the difference of logarithms loses nothing that matters).
  // The base-2 logarithms are subtracted before the one scaling by ln 2: with
  // two natural logarithms the compiler may fuse one scaling into the
  // subtraction (an FMA), which breaks the exact sign flip.
  // Only waves with such a face evaluate it (near-uniform regions, most of
  // the grid, have none); which form a lane uses depends on |t| alone.
  float r = ser;
  if (__builtin_amdgcn_ballot_w64(a >= 0.125f) != 0) {
    const float lg = (__builtin_amdgcn_logf(cx * (1.0f - cp)) -
                      __builtin_amdgcn_logf(cp * (1.0f - cx))) * 0.693147180559945309f;
    r = a < 0.125f ? ser : lg;
  }
  return r;
}

// Solute increment of the cell at LDS index nc (conc_body.inc:92-144): face
// mobility 2 D_C/(1/g_P + 1/g_X), flux along the q differences.  The flux
// through the face below is the one through the face above the cell beneath,
// negated (exact: dmu is odd and a + b = b + a), except for the first cell
// of a z-segment (fresh): the carried flux is a product rounded on its own
// (fzu, a statement of its own, is never fused into a sum).
template <typename R, bool GUARD = true>
__device__ __forceinline__ R solute_inc(const R *ig, const R *mu, int nc, int qx, R igc, R muc,
                                        R igP, R muP, R igM, R muM, bool fresh, R fzu_prev,
                                        R &fzu, R D2dt, const SC<R> &s) {
  const R co = s.coff;
  const R fx = dmu<R>(mu[nc + 1], muc, co) * rcp_g<R, GUARD>(ig[nc + 1] + igc) +
               dmu<R>(mu[nc - 1], muc, co) * rcp_g<R, GUARD>(ig[nc - 1] + igc);
  const R fy = dmu<R>(mu[nc + qx], muc, co) * rcp_g<R, GUARD>(ig[nc + qx] + igc) +
               dmu<R>(mu[nc - qx], muc, co) * rcp_g<R, GUARD>(ig[nc - qx] + igc);
  fzu = dmu<R>(muP, muc, co) * rcp_g<R, GUARD>(igP + igc);
  const R fzd = fresh ? dmu<R>(muM, muc, co) * rcp_g<R, GUARD>(igM + igc) : -fzu_prev;
  const R fz = fzu + fzd;
  return D2dt * (fx * s.idx2 + fy * s.idy2 + fz * s.idz2);
}

// solute_inc with the FP64 reciprocals checked per wave (see rcp_g); the
// flux through the cell's up face goes to fzu_prev for the next cell.
template <typename R>
__device__ __forceinline__ R solute_inc_w(const R *ig, const R *mu, int nc, int qx, R igc,
                                          R muc, R igP, R muP, R igM, R muM, bool fresh,
                                          R &fzu_prev, R D2dt, const SC<R> &s) {
  R fzu;
  R inc;
  if constexpr (sizeof(R) == 8) {
    inc = solute_inc<R, false>(ig, mu, nc, qx, igc, muc, igP, muP, igM, muM, fresh, fzu_prev,
                               fzu, D2dt, s);
    if (__builtin_amdgcn_ballot_w64(inc != inc || fzu != fzu) != 0)
      inc = solute_inc<R, true>(ig, mu, nc, qx, igc, muc, igP, muP, igM, muM, fresh, fzu_prev,
                                fzu, D2dt, s);
  } else {
    inc = solute_inc<R, true>(ig, mu, nc, qx, igc, muc, igP, muP, igM, muM, fresh, fzu_prev, fzu,
                              D2dt, s);
  }
  fzu_prev = fzu;
  return inc;
}

// The element at byte offset boff from base.  With a uniform base (a field
// pointer plus a whole number of planes, kept in scalar registers) and a
// 32-bit per-lane byte offset, the code generator addresses global memory
// with an SGPR base and a VGPR offset instead of 64-bit vector arithmetic per
// access.  A plane of a field must be under 4 GiB (checked at launch).
template <typename T>
__device__ __forceinline__ T &at(T *base, uint32_t boff) {
  return *reinterpret_cast<T *>(reinterpret_cast<char *>(base) + boff);
}
template <typename T>
__device__ __forceinline__ const T &at(const T *base, uint32_t boff) {
  return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + boff);
}

// base + a byte offset (a uniform plane offset: see plane_bytes)
template <typename T>
__device__ __forceinline__ T *boff(T *base, uint64_t b) {
  return reinterpret_cast<T *>(reinterpret_cast<char *>(base) + b);
}
template <typename T>
__device__ __forceinline__ const T *boff(const T *base, uint64_t b) {
  return reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + b);
}
// Byte offset of plane kk (1-based, kk >= 1) of a field with planes of szb
// bytes: a 32 x 32 -> 64-bit product (a plane of a field is under 4 GiB,
// checked at launch), two scalar instructions instead of a 64 x 64 one.
__device__ __forceinline__ uint64_t plane_bytes(int kk, uint32_t szb) {
  return (uint64_t)(uint32_t)(kk - 1) * (uint64_t)szb;
}

// Store of one element at a per-lane byte offset from a uniform plane base,
// without a branch: a lane that must not store passes kNoStore, an offset
// beyond the range of the buffer resource built here, and the hardware drops
// its write (raw buffer accesses are range-checked against num_records; a
// plane is under 4 GiB, so every real offset is in range).  With every
// vector memory instruction of a plane unconditional, the compiler waits for
// exactly the loads it needs; a store behind a branch makes it wait for all
// outstanding memory operations at the next plane, stores included.
constexpr uint32_t kNoStore = 0xffffffffu;
typedef unsigned int pf_u32x2 __attribute__((ext_vector_type(2)));
template <typename T>
__device__ __forceinline__ void bstore(T *base, uint32_t boff, T v) {
  // word 3 of a raw (untyped, stride 0) buffer resource on gfx9-family parts
  const __amdgpu_buffer_rsrc_t r =
      __builtin_amdgcn_make_buffer_rsrc(base, (short)0, (int)0xffffffffu, 0x00020000);
  if constexpr (sizeof(T) == 8)
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(pf_u32x2, v), r, boff, 0, 0);
  else
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, boff, 0, 0);
}

struct Geo {
  int nx, ny, nz;
  int nbx, nby, nseg, seg, k_lo, k_hi;  // cell planes [k_lo,k_hi] split in segments
  int ntiles, per_xcd;
};

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Sweep kinds of march_tile.
enum : int {
  kSweepA = 0,   // phase sweep only
  kSweepB = 1,   // phase + solute + (write_T ? next-step T : DPP)
  kSweepPT = 2,  // phase + (write_T ? next-step T : DPP); the solute runs in solute_tile
};

// One work-group's tile (bx, by) and z-segment sg of a sweep of kind MODE.
// With the solute (kSweepB) and macc non-null, the work-group adds its
// cells' solute increments, in ledger units (pf_mass.h), to *macc.
// Kernel variants (bits of VAR): paths that few decks take but whose
// register demand would burden every other path are compiled only into the
// kernels that need them; launches pick the variant from the deck (a kernel
// must not be launched for a deck that needs a path it lacks).
//   kVarSmooth: the smooth noise (ruido_tipo 2: eight hashes and an
//               interpolation per cell);
//   kVarTDiff:  the thermal Laplacian of the next step's T (D != 0; the
//               reference's decks have D = 0).
enum : int { kVarSmooth = 1, kVarTDiff = 2, kVarAll = 3 };

// Occupancy hints for the kernels around march_tile (waves per SIMD the
// register allocation must allow): without the optional paths, FP32 sweep A
// and phase + T fit six (<= 80 VGPRs), FP64 sweep A and phase + T three
// (<= 168); the fused FP32 sweep B (the solute with its ledger units and the
// stochastically rounded store) four (<= 128; at five it spilled 16
// registers), three with the smooth noise; with the optional paths the
// other FP32 sweeps fit four; FP64 sweep B takes what it needs.
template <typename R, int MODE, int VAR>
constexpr int march_waves() {
  if (sizeof(R) == 4) {
    if (MODE == kSweepB) return (VAR & kVarSmooth) ? 3 : 4;
    return VAR != 0 ? 4 : 6;
  }
  return MODE != kSweepB && VAR == 0 ? 3 : 1;
}

// q % NSLOT for the first plane a work-group's march loads, q = k_lo +
// sg * seg - 2 (>= 0).
__device__ __forceinline__ int q_first_slot(const Geo &g, int sg) {
  return (g.k_lo + sg * g.seg - 2) % NSLOT;
}

// Whether the tile column bx holds a column whose sweep writes x-face cells
// (the phase update's i = 3 and nx-2, the solute's i = 2 and nx-1).  The
// tiles without one run a loop body without the x-face stores (XF false).
__device__ __forceinline__ bool tile_has_xface(int bx, int tx, int nx) {
  const int ilo = 2 + bx * tx;
  return ilo <= 3 || ilo + tx - 1 >= nx - 2;
}

template <typename R, int MODE, typename TL, int VAR, bool XF>
__device__ __forceinline__ void march_tile_body(const Geo &g, const KParams &kp, const SC<R> &s,
                                                const R *__restrict__ po, const R *__restrict__ co,
                                                R *__restrict__ pn, R *__restrict__ cn,
                                                const R *__restrict__ Tcur, R *__restrict__ Tnext,
                                                R *__restrict__ dpp, int write_T,
                                                unsigned long long *__restrict__ macc_dst, int bx,
                                                int by, int sg) {
  // FMAs only within one source expression, decided by the front end, so the
  // single-instance and batched kernels that inline this body round alike
#pragma clang fp contract(on)
  constexpr bool SWEEP_B = MODE == kSweepB;  // with the solute
  constexpr bool T_OUT = MODE != kSweepA;     // writes T (or DPP)
  constexpr int TX = TL::TX, NT = TL::NT, PX = TL::PX, PPL = TL::PPL;
  constexpr int QX = TL::QX, QY = TL::QY, QPL = TL::QPL, NRING = TL::NRING, NLD = TL::NLD;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  R *sphi = reinterpret_cast<R *>(smem_raw);  // NSLOT * PPL
  R *snode = sphi + NSLOT * PPL;              // 4 * QPL
  R *sig = snode + 4 * QPL;                   // QPL (sweep B): 1/gamma
  R *smu = sig + QPL;                         // QPL (sweep B): q

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
  // visible to the whole work-group after the first plane's barrier, before
  // any logarithm
  const LogEnt *lt =
      log_tab_fill<R>(smem_raw, march_arrays_bytes<R, TL>(MODE == kSweepB), tid, NT);

  // ring node handled by this thread (tid < NRING): its node-frame coords
  int rlx = 0, rly = 0;
  if (tid < NRING) {
    if (tid < TX) { rlx = tid + 1; rly = 0; }
    else if (tid < 2 * TX) { rlx = tid - TX + 1; rly = QY - 1; }
    else { const int r = tid - 2 * TX; rly = 1 + (r >> 1); rlx = (r & 1) ? QX - 1 : 0; }
  }
  const int olx = tx + 1, oly = ty + 1;  // own node-frame coords

  // In-plane offsets (32 bit; nx*ny < 2^31) of the phi window elements this
  // thread loads, of its own column and of its ring column; clamped so every
  // load stays inside the grid (clamped values feed only unused nodes).
  uint32_t loff[NLD];  // byte offsets
#pragma unroll
  for (int u = 0; u < NLD; ++u) {
    const int e = min(tid + u * NT, PPL - 1);  // lanes past the window reload its last element
    const int gi = clampi(i0 - 2 + (e % PX), 1, nx), gj = clampi(j0 - 2 + (e / PX), 1, ny);
    loff[u] = (uint32_t)((gi - 1) + (gj - 1) * nx) * (uint32_t)sizeof(R);
  }
  const uint32_t ooff =
      (uint32_t)((clampi(i, 1, nx) - 1) + (clampi(j, 1, ny) - 1) * nx) * (uint32_t)sizeof(R);
  const uint32_t roff =
      (uint32_t)((clampi(i0 - 1 + rlx, 1, nx) - 1) + (clampi(j0 - 1 + rly, 1, ny) - 1) * nx) *
      (uint32_t)sizeof(R);
  const uint32_t szb = (uint32_t)(sz * sizeof(R));  // bytes per plane of a field
  auto plane = [&](int q) -> uint64_t { return plane_bytes(clampi(q, 1, nz), szb); };

  R pre[NLD];
  auto gload = [&](int q) {
    const R *src = boff(po, plane(q));
#pragma unroll
    for (int u = 0; u < NLD; ++u)
      pre[u] = at(src, loff[u]);
  };
  // LDS slots of phi planes q-3 .. q+1 for the plane q being marched
  // (slot of plane q = q % NSLOT), rotated from plane to plane: no index
  // arithmetic per plane
  R *sl[NSLOT];
  {
    const int s0 = q_first_slot(g, sg);
#pragma unroll
    for (int d = 0; d < NSLOT; ++d) sl[d] = sphi + ((s0 + d + NSLOT - 3) % NSLOT) * PPL;
  }
  auto slotr = [&](int d) -> R * { return sl[d + 3]; };  // d in [-3, 1]
  auto lstore = [&](R *dst) {
#pragma unroll
    for (int u = 0; u < NLD; ++u)
      if (tid + u * NT < PPL) dst[tid + u * NT] = pre[u];
  };

  // node at node-frame (lx,ly) of plane q-1 from phi planes q-2, q-1, q
  auto node_at = [&](int lx, int ly) -> Node<R> {
    const R *Pm = slotr(-2), *P0 = slotr(-1), *Pp = slotr(0);
    const int c = (ly + 1) * PX + (lx + 1);
    const R gx = (P0[c + 1] - P0[c - 1]) * s.hx;
    const R gy = (P0[c + PX] - P0[c - PX]) * s.hy;
    const R gz = (Pp[c] - Pm[c]) * s.hz;
    return node_from(gx, gy, gz, s);
  };

  // own-column registers: nodes of planes k-1 (M), k (Z), k+1 (P) and the
  // solute quantities of planes k-1, k, k+1
  Node<R> nM{}, nZ{}, nP{};
  // phi of the own column in planes q-1, q-2, q-3 (read from LDS once, at
  // plane q, and carried down the column in registers)
  R a1 = 0, a2 = 0, a3 = 0;
  R igM = 0, muM = 0, igZ = 0, muZ = 0, igP = 0, muP = 0;
  R cZ = 0;
  R cPPcur = 0;  // c of the plane updated next (sweeps A and phase + T)
  uint64_t hbase = hash_base(i, j, ks, kp.it, 1, kp.seed);
  const uint32_t hcol = cell_hash_col(i, j, kp.it);  // csum's draws (FP32 stores)
  (void)hcol;
  // solute increments of this column for the ledger, in ledger units
  const bool acc_on = SWEEP_B && macc_dst != nullptr;  // uniform
  long long macc = 0;
  R fzu_prev = (R)0;  // solute flux through the up face of the previous cell

  const int q_first = ks - 2, q_last = ke + 2;
  // One z-plane of the march.  The first four planes (ST false) only load
  // phi and build nodes.  From plane ks + 2 on (ST true) each iteration builds
  // the nodes of plane q-1 and updates cell plane q-2 with no branch between
  // the two: cells outside the update range (edge columns, planes 2 and nz-1,
  // columns beyond the grid) are computed on clamped, in-bounds data and only
  // their stores are masked, so the long dependency chain of a node (the
  // reciprocal) and the independent arithmetic of the update are one
  // scheduling region.
  auto plane_step = [&](int q, auto steady) __attribute__((always_inline)) {
    constexpr bool ST = decltype(steady)::value;
    const int kn = q - 1;  // plane whose node / solute quantities are built now
    // (steady planes: kn in [ks+1, ke+1], inside [3, nz]; a node of plane nz
    // is built from clamped planes and never used)
    // This plane's loads of c go first, the prefetch of phi plane q + 1
    // second: waiting for c then never waits for the prefetch, and the wait
    // for the prefetch (before its LDS store at the end of the plane) only for
    // what was issued after it in this same plane, not for the stores.
    // FP64: the c of the next plane's update (sweeps A and phase + T) is
    // loaded one plane ahead, before this plane's stores are issued, so that
    // using it never waits for them (measured 3% faster phase + T; the FP32
    // sweep A, with twice the waves to hide latency, measured 3% slower).
    constexpr bool C_AHEAD = !SWEEP_B && sizeof(R) == 8;
    const bool cnode = SWEEP_B && (ST || (q >= ks && kn >= 1 && kn <= nz));
    R cnew_col = 0, cring = 0, cPPnext = 0, cPPnow = 0;
    if (cnode) {
      const uint64_t pk = plane_bytes(kn, szb);
      cnew_col = at(boff(co, pk), ooff);
      cring = at(boff(co, pk), roff);  // every lane: no branch around the load
    }
    if constexpr (C_AHEAD) cPPnext = at(boff(co, plane(q - 1)), ooff);  // plane q - 1 (clamped)
    if constexpr (ST && !SWEEP_B && !C_AHEAD)
      cPPnow = at(boff(co, plane_bytes(q - 2, szb)), ooff);
    gload(min(q + 1, q_last));  // prefetch; lands while this plane computes
    __syncthreads();            // phi plane q (stored at the end of plane q - 1) in LDS
    const bool build = ST || (q >= ks && kn >= 2 && kn <= nz - 1);
    const R aq = slotr(0)[(oly + 1) * PX + (olx + 1)];  // phi of plane q, own column
    Node<R> nr{};
    R rig = 0, rmu = 0;
    if (build) {
      // own node: phi of plane q and q-2 at the own column from the carry
      const R *P0 = slotr(-1);
      const int c = (oly + 1) * PX + (olx + 1);
      nP = node_from((P0[c + 1] - P0[c - 1]) * s.hx, (P0[c + PX] - P0[c - PX]) * s.hy,
                     (aq - a2) * s.hz, s);
      if (tid < NRING) nr = node_at(rlx, rly);
    }
    if constexpr (SWEEP_B) {
      if (cnode) {
        const R *P0 = slotr(-1);
        solute_node_w(a1, cnew_col, quint, s, lt, igP, muP);  // plane q - 1, own column
        if (tid < NRING)
          solute_node_w(P0[(rly + 1) * PX + rlx + 1], cring, quint, s, lt, rig, rmu);
      }
    }
    if constexpr (ST) {
      const int k = q - 2;  // cell plane updated now
      // the noise draw first, before the update's values are loaded: the
      // smooth noise's eight hashes and interpolation then need no
      // registers beside them (phase_body.inc:181-189; the draw depends on
      // the step, not the sweep -- passing it from the first sweep to the
      // last through HBM measured slower than hashing again)
      // The draw enters only as amp 16 phi^2 (1-phi)^2: where phi^2 or
      // (1-phi)^2 is zero (the bulk liquid and solid, phi exactly 0 or 1)
      // the term is +-0 whatever the draw, and adding it changes no sum (at
      // most the sign of a zero phi).  A wave with no other cell skips the
      // hash.
      float ran = 0.f;
      if (noise && __builtin_amdgcn_ballot_w64(a2 * a2 != (R)0 &&
                                               ((R)1 - a2) * ((R)1 - a2) != (R)0) != 0) {
        __builtin_amdgcn_sched_barrier(0);  // keeps the branch (not select)
        ran = (VAR & kVarSmooth) && kp.ruido_tipo == PF_RUIDO_SMOOTH
                  ? (float)ruido_suave(i, j, k, kp.it, kp.seed)
                  : (float)hash_unit(hbase);
      }
      const R *Pm = slotr(-3), *P0 = slotr(-2), *Pp = slotr(-1);  // planes k-1, k, k+1
      const int c = (ty + 2) * PX + (tx + 2);
      const R PP = a2;  // planes k, k+1, k-1 of the own column

      const uint64_t pk = plane_bytes(k, szb);  // plane k of a field
      const bool upd = ph_col && k >= 3 && k <= nz - 2;  // phase cells (phase_body.inc)
      R cPP;
      if constexpr (SWEEP_B) cPP = cZ;
      else if constexpr (C_AHEAD) cPP = cPPcur;
      else cPP = cPPnow;  // plane q - 2
      // the Laplacian-like term, then the three cross terms, each reading
      // its own phi and node values
      const int nc = oly * QX + olx;
      const R *e2 = snode, *xy = snode + QPL, *yz = snode + 2 * QPL, *zx = snode + 3 * QPL;
      const R PE = P0[c + 1], PW = P0[c - 1], PN = P0[c + PX], PS = P0[c - PX];
      const R PT = a1, PB = a3;
      const R e2E = e2[nc + 1], e2W = e2[nc - 1], e2N = e2[nc + QX], e2S = e2[nc - QX];
      R AA = (e2E * (PE - PP) + e2W * (PW - PP)) * s.idx2 +
             (e2N * (PN - PP) + e2S * (PS - PP)) * s.idy2 +
             (nP.e2 * (PT - PP) + nM.e2 * (PB - PP)) * s.idz2;
      // the cross-derivative terms (phase_body.inc:137-149), each a sum of
      // eight phi x node-difference products, collected by node quantity:
      // every node value multiplies the sum of two phi differences across it
      // (the same sum; 39 operations instead of 48)
      const R dNS0 = PN - PS, dEW0 = PE - PW, dTB0 = PT - PB;
      const R PNE = P0[c + PX + 1], PSE = P0[c - PX + 1], PNW = P0[c + PX - 1],
              PSW = P0[c - PX - 1];
      const R xyE = xy[nc + 1], xyW = xy[nc - 1], xyN = xy[nc + QX], xyS = xy[nc - QX];
      const R Cxy = xyE * ((PNE - PSE) + dNS0) - xyW * ((PNW - PSW) + dNS0) -
                    xyN * ((PNE - PNW) + dEW0) + xyS * ((PSE - PSW) + dEW0);
      const R PNT = Pp[c + PX], PST = Pp[c - PX], PNB = Pm[c + PX], PSB = Pm[c - PX];
      const R yzN = yz[nc + QX], yzS = yz[nc - QX], yzT = nP.yz, yzB = nM.yz;
      const R Cyz = yzN * ((PNT - PNB) + dTB0) - yzS * ((PST - PSB) + dTB0) -
                    yzT * ((PNT - PST) + dNS0) + yzB * ((PNB - PSB) + dNS0);
      const R PET = Pp[c + 1], PWT = Pp[c - 1], PEB = Pm[c + 1], PWB = Pm[c - 1];
      const R zxE = zx[nc + 1], zxW = zx[nc - 1], zxT = nP.zx, zxB = nM.zx;
      const R Czx = zxT * ((PET - PWT) + dEW0) - zxE * ((PET - PEB) + dTB0) +
                    zxW * ((PWT - PWB) + dTB0) - zxB * ((PEB - PWB) + dEW0);
      AA = AA + s.cxy * Cxy + s.cyz * Cyz + s.czx * Czx;
      const R om = (R)1 - PP;
      const R gl = (R)2 * PP * om * (om - PP);  // 2p(1-p)^2 - 2p^2(1-p)
      R h, hl;
      h_of(PP, quint, h, hl);
      // (1-Cs)(1-CLP)/((1-Cl)(1-CSP)) with CLP = c/den, CSP = Ke CLP
      const R den = s.Ke * h + ((R)1 - h);
      const R cfP = cfull<R>(cPP, s);
      R lnr;
      if constexpr (sizeof(R) == 4) {
        lnr = ratio_ln_f32(h, den, cPP, cfP, s);
      } else {
        // ln|(1-Cs_eq)(1-CLP)/((1-Cl_eq)(1-CSP))| = ln|A/B| with
        // A = (1-Cs_eq)/(1-Cl_eq) (den - c), B = den - Ke c: a ratio within
        // a few 1e-3 of 1 near equilibrium
        lnr = ln_ratio_near1(s.omCsCl * (den - cfP), den - s.Ke * cfP, lt);
      }
      const R GG = hl * s.RTV * lnr - s.W0 * gl;
      R ruidos = (R)0;
      if (noise) {
        const R amp = kp.ruido_tipo == PF_RUIDO_CELL_POSITIVE ? (R)ran * s.AAA
                                                              : ((R)2 * (R)ran - (R)1) * s.AAA;
        ruidos = amp * (R)16 * (PP * PP) * (om * om);
      }
      R pstar;
      if constexpr (sizeof(R) == 4) {
        // stochastically rounded store of phi (csum in pf_mass.h)
        const R incp = s.M0dt * (AA + GG) + ruidos;
        pstar = csum<R>(PP, incp, hcol ^ (T_OUT ? kSaltPhiLast : kSaltPhiFirst), k);
      } else {
        pstar = PP + s.M0dt * (AA + GG) + ruidos;
      }
      const R cl = pstar > (R)1 ? (R)1 : pstar;
      const R pnew = upd ? (cl < (R)0 ? (R)0 : cl) : PP;
      const R dppv = upd ? pstar - PP : (R)0;
      bstore(boff(pn, pk), upd ? ooff : kNoStore, pnew);
      if constexpr (XF) {
        // the x faces of an interior row (equacao_campo_fase.f90:151-167 in
        // closed form: the face cells of rows inside [3,n-2] in y and z):
        // cell 2 (nx-1) takes the swept value of cell 3 (nx-2), cell 1 (nx)
        // the pre-sweep value of cell 2 (nx-1), i.e. PW (PE) of cell 3 (nx-2);
        // the shell kernel after the sweep leaves these cells out
        constexpr uint32_t E = (uint32_t)sizeof(R);
        const bool xw = upd && i == 3, xe = upd && i == nx - 2;
        bstore(boff(pn, pk), xw ? ooff - E : xe ? ooff + E : kNoStore, pnew);
        bstore(boff(pn, pk), xw ? ooff - 2 * E : xe ? ooff + 2 * E : kNoStore, xw ? PW : PE);
      }
      if constexpr (SWEEP_B) {
        // solute at (i,j,k), k in [2,nz-1]
        const R *ig = sig, *mu = smu;
        const R D2dt = PP <= (R)0.95 ? s.Dliq2dt : s.Dsol2dt;
        const R inc = solute_inc_w<R>(ig, mu, nc, QX, igZ, muZ, igP, muP, igM, muM, q == ks + 2,
                                    fzu_prev, D2dt, s);
        const R cnew = csum<R>(cPP, inc, hcol, k);
        bstore(boff(cn, pk), own ? ooff : kNoStore, cnew);
        // x faces of the row (concentracao_soluto.f90:116-124: a face cell of
        // a row inside [2,n-1] in y and z copies its inner neighbour's new
        // value); the face kernel after the sweep leaves them out
        if constexpr (XF) {
          constexpr uint32_t E = (uint32_t)sizeof(R);
          bstore(boff(cn, pk), own && i == 2 ? ooff - E : own && i == nx - 1 ? ooff + E : kNoStore,
                 cnew);
        }
        if (acc_on) macc += mass_units<R>(own ? inc : (R)0);
      }
      if constexpr (T_OUT) {
        if (write_T) {
          // temp_body.inc:31-41 of the next step: TP + D dt lap(T) + fonte
          R hh, hl2;
          h_of(pnew, quint, hh, hl2);
          R lap = (R)0;
          if ((VAR & kVarTDiff) && s.Ddt != (R)0 && own) {
            const R TP2 = (R)2 * s.TP;
            const R *tc = &at(boff(Tcur, pk), ooff);
            lap = (tc[1] - TP2 + tc[-1]) * s.idx2 + (tc[nx] - TP2 + tc[-nx]) * s.idy2 +
                  (tc[sz] - TP2 + tc[-(ptrdiff_t)sz]) * s.idz2;
          }
          bstore(boff(Tnext, pk), own ? ooff : kNoStore, s.Ddt * lap + s.fonte * hl2 * dppv + s.TP);
        } else {
          bstore(boff(dpp, pk), upd ? ooff : kNoStore, dppv);
        }
      } else {
        (void)Tcur; (void)Tnext; (void)dpp; (void)write_T; (void)dppv;
      }
      hbase += kHashStepK;
    }
    __syncthreads();
    // publish plane kn's node / solute quantities and rotate the column
    const int oc = oly * QX + olx, rc = rly * QX + rlx;
    if (build) {
      snode[oc] = nP.e2; snode[QPL + oc] = nP.xy; snode[2 * QPL + oc] = nP.yz;
      snode[3 * QPL + oc] = nP.zx;
      if (tid < NRING) {
        snode[rc] = nr.e2; snode[QPL + rc] = nr.xy; snode[2 * QPL + rc] = nr.yz;
        snode[3 * QPL + rc] = nr.zx;
      }
    }
    if constexpr (SWEEP_B) {
      if (ST || q >= ks) {
        sig[oc] = igP; smu[oc] = muP;
        if (tid < NRING) {
          sig[rc] = rig; smu[rc] = rmu;
        }
      }
    }
    nM = nZ; nZ = nP;
    a3 = a2; a2 = a1; a1 = aq;
    if constexpr (SWEEP_B) {
      igM = igZ; muM = muZ; igZ = igP; muZ = muP;
      cZ = cnew_col;
    } else if constexpr (C_AHEAD) {
      cPPcur = cPPnext;
    }
    // phi plane q + 1 into its slot, (q + 1) % NSLOT = (q - 4) % NSLOT, which
    // no thread reads in this plane (nodes read slots q-2..q, the update
    // q-3..q-1) and every thread finished reading in plane q - 1; the next
    // plane's first barrier publishes it
    lstore(slotr(1));
    R *const t = sl[0];
#pragma unroll
    for (int d = 0; d < NSLOT - 1; ++d) sl[d] = sl[d + 1];
    sl[NSLOT - 1] = t;
  };
  gload(q_first);
  lstore(slotr(0));
  for (int q = q_first; q < ks + 2; ++q) plane_step(q, std::false_type{});
  for (int q = ks + 2; q <= q_last; ++q) plane_step(q, std::true_type{});
  if constexpr (SWEEP_B) {
    if (acc_on) mass_flush<NT>(macc, tid, macc_dst);
  } else {
    (void)macc_dst; (void)macc; (void)cn;
  }
}

// One work-group's tile (bx, by) and z-segment sg of a sweep of kind MODE;
// the x-face stores are compiled only into the body that tiles touching an
// x face run.
template <typename R, int MODE, typename TL, int VAR = kVarAll>
__device__ __forceinline__ void march_tile(const Geo &g, const KParams &kp, const SC<R> &s,
                                           const R *__restrict__ po, const R *__restrict__ co,
                                           R *__restrict__ pn, R *__restrict__ cn,
                                           const R *__restrict__ Tcur, R *__restrict__ Tnext,
                                           R *__restrict__ dpp, int write_T,
                                           unsigned long long *__restrict__ macc_dst, int bx,
                                           int by, int sg) {
  if (tile_has_xface(bx, TL::TX, g.nx))
    march_tile_body<R, MODE, TL, VAR, true>(g, kp, s, po, co, pn, cn, Tcur, Tnext, dpp, write_T,
                                            macc_dst, bx, by, sg);
  else
    march_tile_body<R, MODE, TL, VAR, false>(g, kp, s, po, co, pn, cn, Tcur, Tnext, dpp,
                                             write_T, macc_dst, bx, by, sg);
}

// LDS bytes of a march_tile work-group.
template <typename R, typename TL>
constexpr size_t march_lds_bytes(int mode) {
  return log_tab_offset(march_arrays_bytes<R, TL>(mode == kSweepB)) + log_tab_bytes<R>();
}

// LDS of a solute_tile work-group: three node planes of 1/gamma and three
// of q (arithmetic type), then the FP64 log table.
template <typename R, typename TL>
constexpr size_t solute_arrays_bytes() {
  return (size_t)(6 * TL::QPL) * sizeof(R);
}

// One work-group's tile and z-segment of the solute update alone
// (conc_body.inc:61-144), for the split form of sweep B: the same per-node
// 1/gamma and mu and the same face fluxes as march_tile's kSweepB, in the
// same expressions (same bits), but it needs only phi and c at the nodes --
// no phi planes, gradients or anisotropy -- so it runs with few registers
// and many waves.  Node planes go through three LDS buffers (plane q is
// published before the barrier of iteration q and read after the barrier of
// iteration q+1; its buffer is rewritten in iteration q+3, after a barrier
// all its readers have passed), so one barrier per plane suffices.
template <typename R, typename TL, bool XF, int LEDGER>
__device__ __forceinline__ void solute_tile_body(const Geo &g, const KParams &kp,
                                                 const SC<R> &s, const R *__restrict__ po,
                                                 const R *__restrict__ co, R *__restrict__ cn,
                                                 unsigned long long *__restrict__ macc_dst,
                                                 int bx, int by, int sg) {
#pragma clang fp contract(on)
  constexpr int TX = TL::TX, NT = TL::NT, QX = TL::QX, QY = TL::QY, QPL = TL::QPL;
  constexpr int NRING = TL::NRING;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  // three node planes of 1/gamma, then three of q
  R *sig = reinterpret_cast<R *>(smem_raw);
  R *smu = sig + 3 * QPL;

  const int nx = g.nx, ny = g.ny;
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TX + tx;
  const int i0 = 2 + bx * TX, j0 = 2 + by * TL::TY;
  const int i = i0 + tx, j = j0 + ty;
  const bool own = (i <= nx - 1) && (j <= ny - 1);
  const int ks = g.k_lo + sg * g.seg;
  const int ke = min(ks + g.seg - 1, g.k_hi);
  const size_t sz = (size_t)nx * ny;
  const uint32_t szb = (uint32_t)(sz * sizeof(R));  // bytes per plane of a field
  const int quint = kp.h_quintico;
  const LogEnt *lt = log_tab_fill<R>(smem_raw, solute_arrays_bytes<R, TL>(), tid, NT);
  if constexpr (sizeof(R) == 8) __syncthreads();  // the first node's log
  int rlx = 0, rly = 0;
  if (tid < NRING) {
    if (tid < TX) { rlx = tid + 1; rly = 0; }
    else if (tid < 2 * TX) { rlx = tid - TX + 1; rly = QY - 1; }
    else { const int r = tid - 2 * TX; rly = 1 + (r >> 1); rlx = (r & 1) ? QX - 1 : 0; }
  }
  const int olx = tx + 1, oly = ty + 1;
  const uint32_t ooff =
      (uint32_t)((clampi(i, 1, nx) - 1) + (clampi(j, 1, ny) - 1) * nx) * (uint32_t)sizeof(R);
  const uint32_t roff =
      (uint32_t)((clampi(i0 - 1 + rlx, 1, nx) - 1) + (clampi(j0 - 1 + rly, 1, ny) - 1) * nx) *
      (uint32_t)sizeof(R);
  const int oc = oly * QX + olx, rc = rly * QX + rlx;
  const uint32_t hcol = cell_hash_col(i, j, kp.it);  // csum's draw (FP32)
  (void)hcol;

  R igM = 0, muM = 0, igZ = 0, muZ = 0, igP = 0, muP = 0;
  R pZ = 0, cZ = 0;  // phi and c of the own cell of plane k
  // LEDGER 0 / 1: a body without / with the ledger's increments; 2: decided
  // per work-group (macc_dst null or not)
  const bool acc_on = LEDGER == 2 ? macc_dst != nullptr : LEDGER == 1;  // uniform
  long long macc = 0;
  R fzu_prev = (R)0;
  // node plane q (all in [1, nz]); prefetched one plane ahead
  const int q_first = ks - 1, q_last = ke + 1;
  R pq = 0, cq = 0, rp = 0, rcv = 0;
  // every lane loads (the ring values of non-ring lanes are clamped in-grid
  // duplicates, unused): no branch around a load
  auto fetch = [&](int q, R &p, R &c, R &p2, R &c2) {
    const uint64_t pk = plane_bytes(q, szb);
    p = at(boff(po, pk), ooff);
    c = at(boff(co, pk), ooff);
    p2 = at(boff(po, pk), roff);
    c2 = at(boff(co, pk), roff);
  };
  fetch(q_first, pq, cq, rp, rcv);
  // node-plane buffers of planes q - 1 (bm) and q (bq), rotated over three
  // from plane to plane (buffer of plane q = q % 3)
  int bq = q_first % 3, bm = (q_first + 2) % 3;
  // One node plane.  The first two (ST false) only build and publish nodes;
  // from plane ks + 1 on each also updates cell plane q - 1, every lane
  // computing (lanes beyond the grid on clamped data) and only the stores
  // masked, so the plane's memory instructions are the same on every path.
  auto step = [&](int q, auto steady) __attribute__((always_inline)) {
    constexpr bool ST = decltype(steady)::value;
    const R pcur = pq, ccur = cq, rpc = rp, rcc = rcv;
    fetch(min(q + 1, q_last), pq, cq, rp, rcv);
    R *gw = sig + bq * QPL;
    R *qw = smu + bq * QPL;
    solute_node_w(pcur, ccur, quint, s, lt, igP, muP);
    gw[oc] = igP;
    qw[oc] = muP;
    if (tid < NRING) {
      R rig, rmu;
      solute_node_w(rpc, rcc, quint, s, lt, rig, rmu);
      gw[rc] = rig;
      qw[rc] = rmu;
    }
    __syncthreads();
    if constexpr (ST) {
      const int k = q - 1;  // cell plane updated now
      const uint64_t pk = plane_bytes(k, szb);
      const R PP = pZ, cPP = cZ;
      const int nc = oc;
      const R *ig = sig + bm * QPL, *mu = smu + bm * QPL;  // plane k = q - 1
      const R D2dt = PP <= (R)0.95 ? s.Dliq2dt : s.Dsol2dt;
      const R igc = igZ, muc = muZ;
      const R inc = solute_inc_w<R>(ig, mu, nc, QX, igc, muc, igP, muP, igM, muM, q == ks + 1,
                                  fzu_prev, D2dt, s);
      const R cnew = csum<R>(cPP, inc, hcol, k);
      bstore(boff(cn, pk), own ? ooff : kNoStore, cnew);
      // x faces of the row, as in march_tile's sweep B
      if constexpr (XF) {
        constexpr uint32_t E = (uint32_t)sizeof(R);
        bstore(boff(cn, pk), own && i == 2 ? ooff - E : own && i == nx - 1 ? ooff + E : kNoStore,
               cnew);
      }
      if (acc_on) macc += mass_units<R>(own ? inc : (R)0);
    }
    igM = igZ; muM = muZ; igZ = igP; muZ = muP;
    pZ = pcur; cZ = ccur;
    bm = bq;
    bq = bq == 2 ? 0 : bq + 1;
  };
  for (int q = q_first; q < ks + 1; ++q) step(q, std::false_type{});
  for (int q = ks + 1; q <= q_last; ++q) step(q, std::true_type{});
  if (acc_on) mass_flush<NT>(macc, tid, macc_dst);
}

// The loop bodies: with or without the x-face stores; FP64 also with or
// without the ledger's increments (FP64 runs keep no ledger by default:
// its per-cell conversion to ledger units is then not computed), FP32 with
// the ledger decided per work-group (on by default).
template <typename R, typename TL>
__device__ __forceinline__ void solute_tile(const Geo &g, const KParams &kp, const SC<R> &s,
                                            const R *__restrict__ po, const R *__restrict__ co,
                                            R *__restrict__ cn,
                                            unsigned long long *__restrict__ macc_dst, int bx,
                                            int by, int sg) {
  const bool xf = tile_has_xface(bx, TL::TX, g.nx);
  if constexpr (sizeof(R) == 8) {
    if (macc_dst) {
      if (xf) solute_tile_body<R, TL, true, 1>(g, kp, s, po, co, cn, macc_dst, bx, by, sg);
      else solute_tile_body<R, TL, false, 1>(g, kp, s, po, co, cn, macc_dst, bx, by, sg);
    } else {
      if (xf) solute_tile_body<R, TL, true, 0>(g, kp, s, po, co, cn, macc_dst, bx, by, sg);
      else solute_tile_body<R, TL, false, 0>(g, kp, s, po, co, cn, macc_dst, bx, by, sg);
    }
  } else {
    if (xf) solute_tile_body<R, TL, true, 2>(g, kp, s, po, co, cn, macc_dst, bx, by, sg);
    else solute_tile_body<R, TL, false, 2>(g, kp, s, po, co, cn, macc_dst, bx, by, sg);
  }
}

// (its log table after the arrays)
template <typename TL, typename R>
constexpr size_t solute_lds_bytes() {
  return log_tab_offset(solute_arrays_bytes<R, TL>()) + log_tab_bytes<R>();
}


}  // namespace pf

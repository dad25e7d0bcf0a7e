// pf_batch.hip -- the batched form: many independent instances of uneven
// grid sizes advanced together, one launch per sweep, precision and tile
// shape for all of them.
//
// Every instance keeps its fields and its loop scalars (it, TP, Cs/Cl, M0,
// the REAL(4) time, buffer roles, steps left) on the device.  A step is
//   begin (k_batch_turn)   per instance: principal.f90:96-104 + the kernel
//                          scalars of the sweeps (one thread per instance)
// then, per precision (the FP32 chain on a second stream, forked and
// joined within the step),
//   (nv-1) x [sweep A, phase decoration]      one launch per tile shape
//   sweep B: phase + T, then the solute       (work-group -> tile table)
//     with the phase decoration
//   solute faces (with a due mass correction's interior sums), and for the
//   instances due, the correction's rescaling and face refresh
// and
//   end (k_batch_turn)     per instance: auxiliar M0 (principal.f90:157),
//                          buffer roles, steps left
// with the end of a step and the beginning of the next in one launch, so
// K steps need no host work and are captured once in a hipGraph that is
// replayed.  An instance that has run its steps goes inactive on the device
// (its work-groups exit at once) until the host removes it.
//
// The sweeps are the single-instance z-march (pf_march.h) with the same
// arithmetic, and the per-instance scalars are computed with the host loop's
// functions (pf_common.h, contraction off), so an instance's result does
// not depend on what else is in the batch, nor on the tiling.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cstring>
#include <map>
#include <vector>

#include "../../include/pf_capi.h"
#include "pf_common.h"
#include "pf_kernels.h"
#include "pf_march.h"
#include "pf_mass.h"

namespace pf {
namespace {

using Tile64x8 = Tile<64, 8>;
using Tile32x8 = Tile<32, 8>;
using Tile32x16 = Tile<32, 16>;
using Tile16x16 = Tile<16, 16>;
constexpr int kNShapes = 4;
constexpr int kShapeX[kNShapes] = {64, 32, 32, 16};
constexpr int kShapeY[kNShapes] = {8, 8, 16, 16};

// Device-resident state of one instance (arithmetic type R).
template <typename R>
struct InstDev {
  Geo ga, gb, gs;       // geometry (k ranges, tiles, z segments) of sweep A,
                        // sweep B and the split sweep B's solute kernel
  KParams kt, k;        // temperatura scalars (TP of the previous step), sweep scalars
  SC<R> s;              // derived sweep scalars of this step
  R *phi[2], *c[2], *T[2];
  int po, co, to;       // buffer of phi_old / c_old / T of the current step
  int nv;               // phase sweeps per step
  int active, steps_left, steps_done;
  // host-loop scalars (principal.f90:96-157)
  double T0, taxa, Tm, Me;
  float Sigma, Beta, ajuste, time, Ke;
  int mass_every;       // solute-mass correction period (0 = off), as pf_set_mass_correction
  MassLedger m;         // the instance's solute-mass ledger (pf_mass.h)
};

// the host writes the three geometries with one copy
static_assert(offsetof(InstDev<double>, gs) == offsetof(InstDev<double>, ga) + 2 * sizeof(Geo) &&
                  offsetof(InstDev<float>, gs) == offsetof(InstDev<float>, ga) + 2 * sizeof(Geo),
              "ga, gb, gs adjacent");

// The instance corrects its solute mass after the current step.
template <typename R>
__device__ __forceinline__ bool mass_due(const InstDev<R> &I) {
  return I.active && I.mass_every > 0 && I.k.it % I.mass_every == 0;
}

struct TileRef {
  int inst, bx, by, sg;
};

// A field pointer read from an instance record, typed as a pointer to
// global memory: the compiler then accesses the field with global memory
// instructions (a plain pointer loaded from memory is generic, and flat
// loads share their wait counter with the LDS accesses of the z-march).
template <typename T>
__device__ __forceinline__ T *gfield(T *const &field) {
  __attribute__((address_space(1))) T *g;
  __builtin_memcpy(&g, &field, sizeof g);
  return (T *)g;
}
struct Chunk {
  int inst, prec;
  long long first;
};
constexpr int kChunkCells = 1024;

// ---- per-step scalar kernels ------------------------------------------------

// due[q] (when non-null): instance q corrects its solute mass after this
// step -- a compact array the correction kernels read instead of every
// instance's record
template <typename R>
__device__ __forceinline__ void batch_begin(InstDev<R> &I, int *due) {
#pragma clang fp contract(off)
  if (!I.active) {
    if (due) *due = 0;
    return;
  }
  I.k.it += 1;
  const int it = I.k.it;
  I.time = (float)(I.k.dt + (double)I.time);  // principal.f90:96
  I.kt = I.k;                                 // temperatura sees TP of step it-1
  if (I.taxa > 0.0) I.k.TP = I.T0 - I.taxa * I.k.dt * (double)it;  // :104
  equilibrium(I.k.TP, I.Tm, I.Ke, I.Me, I.k.Cs_eq, I.k.Cl_eq);
  I.s = make_sc<R>(I.k);
  if (due) *due = mass_due(I) ? 1 : 0;
}

// The end of a step after M0: buffer roles, steps left.
template <typename R>
__device__ __forceinline__ void batch_end_roles(InstDev<R> &I) {
  if (!I.active) return;
  I.po ^= (I.nv & 1);  // principal.f90:127-150, 200-226 as buffer roles
  I.co ^= 1;
  I.to ^= 1;
  I.steps_done += 1;
  I.steps_left -= 1;
  if (I.steps_left <= 0) I.active = 0;
}

// auxiliar's abscissae in constant memory (the lanes index them)
__constant__ const M0Abscissae kM0Phi = m0_abscissae();

// One 64-lane work-group per instance.  The end of one step (do_end) and
// the beginning of the next (do_begin) are one launch.  The end's M0
// (auxiliar.f90, principal.f90:157) evaluates the integral's terms one per
// lane and sums them on lane 0 in auxiliar_m0's order (the same terms and
// order, so the same bits): its chain of 79 divisions is the longest latency
// of the step's end.
template <typename R>
__device__ __forceinline__ void batch_turn(InstDev<R> &I, int *due, int do_end, int do_begin,
                                           double *terms) {
  const int lane = threadIdx.x;
  const bool m0 = do_end && I.active;  // uniform; read before lane 0 writes the record
  if (m0)
    for (int n = lane; n < kM0Terms; n += 64) terms[n] = m0_term(kM0Phi.v[n], I.k.Cl_eq, I.k.Cs_eq);
  __syncthreads();
  if (lane != 0) return;
  if (m0) {
    double acc = 0.0;
    for (int n = 0; n < kM0Terms; ++n) acc = acc + terms[n];
    I.k.M0 = m0_from_sum(acc, I.k.TP, I.k.Cl_eq, I.k.Cs_eq, I.k.R, I.k.Vm, I.k.E0, I.k.W0,
                         I.Sigma, I.Beta, I.Ke, I.Me, I.k.D_liq, I.ajuste);
  }
  if (do_end) batch_end_roles<R>(I);
  if (do_begin) batch_begin<R>(I, due);
}

// Work-group b < n64: FP64 instance b; the rest: the FP32 instances.
__global__ void __launch_bounds__(64)
    k_batch_turn(InstDev<double> *__restrict__ inst64, int n64, int *__restrict__ due64,
                 InstDev<float> *__restrict__ inst32, int n32, int *__restrict__ due32,
                 int do_end, int do_begin) {
  __shared__ double terms[kM0Terms];
  const int q = blockIdx.x;
  if (q < n64)
    batch_turn<double>(inst64[q], due64 ? due64 + q : nullptr, do_end, do_begin, terms);
  else if (q - n64 < n32)
    batch_turn<float>(inst32[q - n64], due32 ? due32 + (q - n64) : nullptr, do_end, do_begin,
                      terms);
}

// ---- batched sweeps and decorations ----------------------------------------

// One batched sweep of kind MODE (kSweepA: a phase sweep; kSweepB: the last
// phase sweep + solute + next T; kSweepPT: the last phase sweep + next T,
// the solute then running in k_solute_batch over the same tiles).
template <typename R, int MODE, typename TL, int VAR>
__global__ void __launch_bounds__(TL::NT)
    __attribute__((amdgpu_waves_per_eu(march_waves<R, MODE, VAR>())))
    k_march_batch(const TileRef *__restrict__ tiles, int ntiles, int per_xcd,
                  InstDev<R> *__restrict__ inst, int sweep, int ledger) {
  const int b = blockIdx.x;
  const int t = (b & 7) * per_xcd + (b >> 3);  // each XCD gets a contiguous run
  if (t >= ntiles) return;
  const TileRef d = tiles[t];
  const InstDev<R> &I = inst[d.inst];
  if (!I.active) return;
  const int pin = I.po ^ ((sweep - 1) & 1), pout = pin ^ 1;
  // the instance's scalars as values (not references into the record,
  // which the compiler would read again plane after plane)
  const KParams kp = I.k;
  const SC<R> sc = I.s;
  const R *po = gfield(I.phi[pin]), *co = gfield(I.c[I.co]);
  R *pn = gfield(I.phi[pout]);
  if constexpr (MODE == kSweepB) {
    const Geo g = I.gb;
    march_tile<R, kSweepB, TL, VAR>(g, kp, sc, po, co, pn, gfield(I.c[I.co ^ 1]),
                                       gfield(I.T[I.to]), gfield(I.T[I.to ^ 1]), nullptr, 1,
                                       ledger && I.mass_every > 0 ? &inst[d.inst].m.acc : nullptr,
                                       d.bx, d.by, d.sg);
  } else if constexpr (MODE == kSweepPT) {
    const Geo g = I.gb;
    march_tile<R, kSweepPT, TL, VAR>(g, kp, sc, po, co, pn, nullptr, gfield(I.T[I.to]),
                                        gfield(I.T[I.to ^ 1]), nullptr, 1, nullptr, d.bx, d.by,
                                        d.sg);
  } else {
    const Geo g = I.ga;
    march_tile<R, kSweepA, TL, VAR>(g, kp, sc, po, co, pn, nullptr, nullptr, nullptr, nullptr, 0,
                                       nullptr, d.bx, d.by, d.sg);
  }
}

// Phase decoration (which = 0, shell of width 2, from the swept phi) or
// solute faces (which = 1, width 1) of one chunk of an instance's shell; the
// x-face cells of the interior rows are written by the sweeps.
template <typename R>
__device__ __forceinline__ void deco_chunk(const Chunk &ch, const InstDev<R> &I, int which,
                                           int sweep, int tid, int nt) {
  const int nx = I.k.nx, ny = I.k.ny, nz = I.k.nz;
  const int w = which == 0 ? 2 : 1;
  const long long cnt = shell_count(nx, ny, nz, w, false);
  const long long end = ch.first + kChunkCells < cnt ? ch.first + kChunkCells : cnt;
  const int pin = I.po ^ ((sweep - 1) & 1);
  for (long long t = ch.first + tid; t < end; t += nt) {
    int i, j, k;
    shell_flat(t, nx, ny, nz, w, i, j, k, false);
    if (which == 0)
      phase_deco_cell(i, j, k, nx, ny, nz, (const R *)gfield(I.phi[pin]), gfield(I.phi[pin ^ 1]));
    else
      solute_deco_cell(i, j, k, nx, ny, nz, (const R *)gfield(I.c[I.co]), gfield(I.c[I.co ^ 1]));
  }
}

// The solute update of the split sweep B over its own tile table: reads the
// phi and c sweep B reads, writes c (solute_tile, the single-instance split
// form's body).  The ndeco work-groups after the tiles' decorate the swept
// phi of the instances of this precision (chunks, which = 0): the solute
// reads only the pre-sweep phi and c, so the two are independent and the
// decoration needs no launch of its own.
template <typename R, typename TL>
__global__ void __launch_bounds__(TL::NT)
    k_solute_batch(const TileRef *__restrict__ tiles, int ntiles, int per_xcd,
                   InstDev<R> *__restrict__ inst, int sweep, int ledger,
                   const Chunk *__restrict__ chunks, int ndeco) {
  const int b = blockIdx.x;
  if (b >= 8 * per_xcd) {
    const int q = b - 8 * per_xcd;
    if (q >= ndeco) return;
    const Chunk ch = chunks[q];
    const InstDev<R> &I = inst[ch.inst];
    if (I.active) deco_chunk<R>(ch, I, 0, sweep, threadIdx.y * TL::TX + threadIdx.x, TL::NT);
    return;
  }
  const int t = (b & 7) * per_xcd + (b >> 3);
  if (t >= ntiles) return;
  const TileRef d = tiles[t];
  const InstDev<R> &I = inst[d.inst];
  if (!I.active) return;
  const int pin = I.po ^ ((sweep - 1) & 1);
  const Geo g = I.gs;
  const KParams kp = I.k;
  const SC<R> sc = I.s;
  solute_tile<R, TL>(g, kp, sc, gfield(I.phi[pin]), gfield(I.c[I.co]), gfield(I.c[I.co ^ 1]),
                     ledger && I.mass_every > 0 ? &inst[d.inst].m.acc : nullptr, d.bx, d.by,
                     d.sg);
}

// which = 0: phase decoration of sweep `sweep` (shell of width 2);
// which = 1: solute faces of the step (width 1); a chunk's precision picks
// the instance array.  With nmass > 0 the kMassBlocks work-groups after the
// chunks' compute partial b of the interior sum of the new c of every
// instance of precision mass_prec due for a solute-mass correction (into
// part; the face copy writes only faces, the sum reads only the interior).
__global__ void __launch_bounds__(256) k_deco_batch(const Chunk *__restrict__ chunks, int nchunk,
                                                    const InstDev<double> *__restrict__ inst64,
                                                    const InstDev<float> *__restrict__ inst32,
                                                    int which, int sweep, int nmass, int mass_prec,
                                                    const int *__restrict__ due,
                                                    double *__restrict__ part) {
  if ((int)blockIdx.x >= nchunk) {
    const int mb = (int)blockIdx.x - nchunk;
    if (mb >= kMassBlocks) return;
    __shared__ double sh[256];
    for (int q = 0; q < nmass; ++q) {
      if (!due[q]) continue;  // uniform over the block
      double t;
      if (mass_prec == PF_FP64) {
        const InstDev<double> &I = inst64[q];
        t = block_sum256(mass_rows(gfield(I.c[I.co ^ 1]), I.k.nx, I.k.ny, I.k.nz, mb, I.k.c_off),
                         sh);
      } else {
        const InstDev<float> &I = inst32[q];
        t = block_sum256(mass_rows(gfield(I.c[I.co ^ 1]), I.k.nx, I.k.ny, I.k.nz, mb, I.k.c_off),
                         sh);
      }
      if (threadIdx.x == 0) part[(size_t)q * kMassBlocks + mb] = t;
    }
    return;
  }
  const Chunk ch = chunks[blockIdx.x];
  if (ch.prec == PF_FP64) {
    const InstDev<double> &I = inst64[ch.inst];
    if (I.active) deco_chunk<double>(ch, I, which, sweep, threadIdx.x, 256);
  } else {
    const InstDev<float> &I = inst32[ch.inst];
    if (I.active) deco_chunk<float>(ch, I, which, sweep, threadIdx.x, 256);
  }
}

// sweep kinds of a batched launch: kSweepA, kSweepB, kSweepPT, and
// kSweepSolute (k_solute_batch)
constexpr int kSweepSolute = 3;

template <typename R, typename TL, int VAR>
void launch_batch_v(int mode, const TileRef *tiles, int ntiles, InstDev<R> *inst, int sweep,
                    int ledger, const Chunk *chunks, int ndeco, hipStream_t st) {
  const int per_xcd = (ntiles + 7) / 8;
  const dim3 blk(TL::TX, TL::TY, 1), grd(8 * per_xcd, 1, 1);
  const size_t lds_b = march_lds_bytes<R, TL>(kSweepB), lds_pt = march_lds_bytes<R, TL>(kSweepPT);
  const size_t lds_a = march_lds_bytes<R, TL>(kSweepA), lds_s = solute_lds_bytes<TL, R>();
  if (mode == kSweepB)
    hipLaunchKernelGGL((k_march_batch<R, kSweepB, TL, VAR>), grd, blk, lds_b, st, tiles, ntiles,
                       per_xcd, inst, sweep, ledger);
  else if (mode == kSweepPT)
    hipLaunchKernelGGL((k_march_batch<R, kSweepPT, TL, VAR>), grd, blk, lds_pt, st, tiles, ntiles,
                       per_xcd, inst, sweep, 0);
  else if (mode == kSweepSolute)
    hipLaunchKernelGGL((k_solute_batch<R, TL>), dim3(8 * per_xcd + ndeco), blk, lds_s, st, tiles,
                       ntiles, per_xcd, inst, sweep, ledger, chunks, ndeco);
  else
    hipLaunchKernelGGL((k_march_batch<R, kSweepA, TL, VAR & kVarSmooth>), grd, blk, lds_a, st,
                       tiles, ntiles, per_xcd, inst, sweep, 0);
}

// One launch of kernel kind `mode` (kSweepA, kSweepB, kSweepPT or
// kSweepSolute) over a tile table of one tile shape.  var: the optional
// paths (kVar*) some instance of the group needs (the kernels without a path
// must not see an instance that needs it); ledger: sweep B (or the split
// form's solute) adds its solute increments to the instances' ledgers (off
// for the autotuner's timing launches); the solute kernel also decorates
// `ndeco` chunks of swept phi.
template <typename R>
hipError_t launch_batch_group(int shape, int mode, int var, const TileRef *tiles, int ntiles,
                              InstDev<R> *inst, int sweep, int ledger, const Chunk *chunks,
                              int ndeco, hipStream_t st) {
  if (ntiles == 0) return hipSuccess;
  auto go = [&](auto tl) {
    using TL = decltype(tl);
    switch (var) {
      case 0: launch_batch_v<R, TL, 0>(mode, tiles, ntiles, inst, sweep, ledger, chunks, ndeco, st); break;
      case kVarSmooth:
        launch_batch_v<R, TL, kVarSmooth>(mode, tiles, ntiles, inst, sweep, ledger, chunks, ndeco, st);
        break;
      case kVarTDiff:
        launch_batch_v<R, TL, kVarTDiff>(mode, tiles, ntiles, inst, sweep, ledger, chunks, ndeco, st);
        break;
      default:
        launch_batch_v<R, TL, kVarAll>(mode, tiles, ntiles, inst, sweep, ledger, chunks, ndeco, st);
    }
  };
  switch (shape) {
    case 1: go(Tile32x8{}); break;
    case 2: go(Tile32x16{}); break;
    case 3: go(Tile16x16{}); break;
    default: go(Tile64x8{});
  }
  return hipGetLastError();
}

// ---- solute-mass correction of the instances due this step (pf_mass.h) -----
// The same decomposition and trees as the single-instance correction
// (pf_fused.hip), per instance, so a corrected instance gives the same bits
// batched and alone.  Each kernel covers every instance and skips those not
// due, so a captured step needs no host decision.

// Rescales block b's rows of every due instance's new c by f = ledger / S,
// with S the fold of the instance's kMassBlocks interior partials: every
// work-group folds them itself, in the same order, so all hold the same S
// and f (work-group 0 records both in the ledger).
template <typename R>
__global__ void __launch_bounds__(256) k_bmass_scale(InstDev<R> *__restrict__ inst, int n,
                                                     const int *__restrict__ due,
                                                     const double *__restrict__ part) {
  __shared__ double sh[256];
  for (int q = 0; q < n; ++q) {
    if (!due[q]) continue;  // uniform over the block
    InstDev<R> &I = inst[q];
    const double S = mass_fold(part + (size_t)q * kMassBlocks, sh);
    const double f = ledger_value(I.m) / S;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      I.m.S = S;
      I.m.f = f;
    }
    mass_scale_rows(gfield(I.c[I.co ^ 1]), I.k.nx, I.k.ny, I.k.nz, I.k.it, I.k.seed, f,
                    (int)blockIdx.x, I.k.c_off);
  }
}

// the face planes of a corrected instance from their interior sources
// (concentracao_soluto.f90:116-124, as after the single-instance correction)
constexpr int kRefreshBlocks = 512;
template <typename R>
__global__ void __launch_bounds__(256) k_bmass_refresh(const InstDev<R> *__restrict__ inst, int n,
                                                       const int *__restrict__ due) {
  for (int q = 0; q < n; ++q) {
    if (!due[q]) continue;
    const InstDev<R> &I = inst[q];
    const int nx = I.k.nx, ny = I.k.ny, nz = I.k.nz;
    R *c = gfield(I.c[I.co ^ 1]);
    int i, j, k;
    for (long long t = blockIdx.x * 256LL + threadIdx.x; shell_flat(t, nx, ny, nz, 1, i, j, k);
         t += (long long)kRefreshBlocks * 256) {
      int si = i, sj = j, sk = k;
      solute_source(si, sj, sk, nx, ny, nz);
      if (si > 1 && si < nx && sj > 1 && sj < ny && sk > 1 && sk < nz)
        c[cidx(i, j, k, nx, ny)] = c[cidx(si, sj, sk, nx, ny)];
    }
  }
}

// Geometry of a sweep for a tile shape and segment length.
Geo batch_geo(int nx, int ny, int nz, bool sweep_b, int shape, int seg) {
  Geo g;
  g.nx = nx; g.ny = ny; g.nz = nz;
  g.nbx = (nx - 2 + kShapeX[shape] - 1) / kShapeX[shape];
  g.nby = (ny - 2 + kShapeY[shape] - 1) / kShapeY[shape];
  g.k_lo = sweep_b ? 2 : 3;
  g.k_hi = sweep_b ? nz - 1 : nz - 2;
  const int ncell = g.k_hi - g.k_lo + 1;
  g.seg = std::max(1, std::min(seg, ncell));
  g.nseg = (ncell + g.seg - 1) / g.seg;
  g.ntiles = g.nbx * g.nby * g.nseg;
  g.per_xcd = (g.ntiles + 7) / 8;
  return g;
}

// Size class of an instance: its precision and the bucket of its xy plane
// (tile shape and z-march depth are chosen per class).
int size_class(int prec, int nx, int ny) {
  const long long a = (long long)nx * ny;
  const int bucket = a <= 48 * 48 ? 0 : a <= 96 * 96 ? 1 : a <= 160 * 160 ? 2 : 3;
  return prec * 4 + bucket;
}
constexpr int kNClasses = 8;
// z segments of the batched sweeps: between kSegMin and kSegMax planes
// (each segment reads 4 planes before its first update).  With the default
// group target a batch of ~30M cells marches 16-plane segments and one of
// < 10M cells 8-plane ones: measured on MI355X, the fastest uniform segment
// length is 16 for a mix of 18 instances / 30M cells (8: +6%, 24: +3%,
// 32: +6% step time) and 8 for mixes of 4-8 instances / 6-9M cells (16:
// +4%, 32: +18-38%).
constexpr int kSegMin = 8, kSegMax = 16;
// kernel kinds with a tile shape of their own
enum : int { kKindA = 0, kKindB = 1, kKindS = 2, kNKinds = 3 };

}  // namespace
}  // namespace pf

// ---- host side ----------------------------------------------------------------

struct pf_batch_slot {
  bool used = false;
  int prec = PF_FP64;
  int idx = -1;  // index in the per-precision device array
  pf_params p{};
  pf_loop L{};
  int nx = 0, ny = 0, nz = 0, nv = 2;
  size_t n = 0, bytes = 0;
  void *buf[6] = {};  // phi0, phi1, c0, c1, T0, T1
  int steps_total = 0, steps_left = 0;
  int mass_every = 0;      // solute-mass correction period
  double host_mass = 0.0;  // FP64 interior mass of the admitted c field
  double c_off = 0.0;      // FP32-mixed: c stored as c - c_off (as pf_init)
};

struct pf_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  // the FP32 instances' chain of a step (forked from stream, joined back)
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  size_t budget = 0, used = 0;
  int nv = -1;  // shared n_varreduras of the instances
  std::vector<pf_batch_slot> slots;
  // device arrays of instance state, per precision
  void *inst_dev[2] = {nullptr, nullptr};
  int inst_cap[2] = {0, 0};
  std::vector<int> slot_of[2];  // per precision: device index -> slot
  // tile shape per size class and kernel kind (kKindA: sweep A, kKindB:
  // sweep B or the split form's phase + T kernel, kKindS: the split form's
  // solute kernel); work-groups per sweep the z split aims for over the
  // whole batch (each instance gets its share by cell count), with segments
  // of at least kSegMin planes
  int cls_shape[pf::kNClasses][pf::kNKinds];
  int group_target = 7168;
  // sweep B split into the phase + T kernel and the solute kernel (as the
  // single-instance split form; same bits)
  bool split_b = true;
  size_t total_cells = 0;
  // tables (rebuilt when instances change)
  bool dirty = true;
  pf::TileRef *tiles_dev = nullptr;
  size_t tiles_cap = 0;
  pf::Chunk *chunks_dev = nullptr;
  size_t chunks_cap = 0;
  struct Group {
    int prec, shape, kind;
    int var;  // kernel variant: the optional paths an instance of the group needs
    size_t off;
    int n;
    int ndeco;  // solute groups: phase-decoration chunks its launch also runs
    size_t deco_off;
  };
  std::vector<Group> groups;
  // chunk tables of the shell work: [which] = 0 phase decoration, 1 solute
  // faces; the FP64 instances' chunks first, then the FP32 ones
  size_t deco_off[2] = {0, 0};
  int deco_n[2] = {0, 0}, deco_n64[2] = {0, 0};
  // captured graph of graph_steps steps
  int graph_steps = 16;
  hipGraphExec_t graph = nullptr;
  bool graph_valid = false;
  double *diag_part = nullptr;
  int diag_cap = 0;
  double *pinned = nullptr;
  // per precision: partial interior sums of the solute-mass correction
  // (kMassBlocks per instance) and the instances due for a correction after
  // the current step
  double *mass_part[2] = {nullptr, nullptr};
  size_t mass_part_cap[2] = {0, 0};
  int *mass_due[2] = {nullptr, nullptr};
  int mass_due_cap[2] = {0, 0};
};

namespace {

using pf::fail_msg;

#define HIP_OK(x)                                                                          \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess) return fail_msg(PF_E_HIP, "%s: %s", #x, hipGetErrorString(e_)); \
  } while (0)

size_t inst_size(int prec) {
  return prec == PF_FP64 ? sizeof(pf::InstDev<double>) : sizeof(pf::InstDev<float>);
}

int check_batch(const pf_batch *b) {
  if (!b) return fail_msg(PF_E_ARG, "null batch");
  return PF_OK;
}

int check_slot(const pf_batch *b, int32_t slot) {
  int rc = check_batch(b);
  if (rc != PF_OK) return rc;
  if (slot < 0 || slot >= (int)b->slots.size() || !b->slots[slot].used)
    return fail_msg(PF_E_ARG, "no instance in slot %d", slot);
  return PF_OK;
}

template <typename R>
pf::InstDev<R> make_inst(const pf_batch_slot &s) {
  pf::InstDev<R> I;
  memset(&I, 0, sizeof I);
  I.k = pf::kparams_from(s.p);
  I.k.c_off = s.c_off;
  I.kt = I.k;
  I.s = pf::make_sc<R>(I.k);
  for (int f = 0; f < 2; ++f) {
    I.phi[f] = (R *)s.buf[f];
    I.c[f] = (R *)s.buf[2 + f];
    I.T[f] = (R *)s.buf[4 + f];
  }
  I.po = 0;
  I.co = 0;
  I.to = 1;  // T of step 1 (temperatura at admission) goes to T[1]
  I.nv = s.nv;
  I.active = s.steps_left > 0;
  I.steps_left = s.steps_left;
  I.steps_done = 0;
  I.T0 = s.L.T0;
  I.taxa = s.L.taxa_resfr;
  I.Tm = s.p.Tm;
  I.Me = s.p.Me;
  I.Sigma = s.L.Sigma;
  I.Beta = s.L.Beta;
  I.ajuste = s.L.ajuste;
  I.time = s.L.time;
  I.Ke = s.p.Ke;
  I.mass_every = s.mass_every;
  I.m.L0 = s.host_mass;  // the ledger starts from the caller's FP64 field
  return I;
}

// Rebuilds the device instance arrays' geometry and the tile / chunk tables
// after instances were added or removed or the tiling changed.
template <typename R>
void set_geo(pf::InstDev<R> &I, const pf_batch &b, const pf_batch_slot &s) {
  const int cls = pf::size_class(s.prec, s.nx, s.ny);
  // z segments: the instance's share (by cells) of group_target work-groups
  // over the whole batch; a batch with many cells marches longer segments,
  // which saves warm-up planes, a small batch shorter ones, which shortens
  // its launches
  const double share = b.total_cells > s.n ? (double)s.n / (double)b.total_cells : 1.0;
  const int want = std::max(1, (int)(share * b.group_target + 0.5));
  auto geo = [&](int kind, bool sweep_b) {
    const int shape = b.cls_shape[cls][kind];
    const int plane_tiles = ((s.nx - 2 + pf::kShapeX[shape] - 1) / pf::kShapeX[shape]) *
                            ((s.ny - 2 + pf::kShapeY[shape] - 1) / pf::kShapeY[shape]);
    const int nseg = std::max(1, (want + plane_tiles - 1) / plane_tiles);
    const int seg = std::min(pf::kSegMax, std::max(pf::kSegMin, (s.nz - 2 + nseg - 1) / nseg));
    return pf::batch_geo(s.nx, s.ny, s.nz, sweep_b, shape, seg);
  };
  I.ga = geo(pf::kKindA, false);
  I.gb = geo(pf::kKindB, true);
  I.gs = geo(pf::kKindS, true);
}

int rebuild(pf_batch *b) {
  b->total_cells = 0;
  for (const auto &s : b->slots)
    if (s.used) b->total_cells += s.n;
  // tile tables: one group per (precision, kernel kind, shape); the
  // solute kind only when sweep B is split
  std::vector<pf::TileRef> tiles;
  std::vector<pf::Chunk> chunks;
  b->groups.clear();
  const int nkinds = b->split_b ? pf::kNKinds : pf::kKindS;
  for (int prec = 0; prec < 2; ++prec) {
    const auto &ids = b->slot_of[prec];
    if (ids.empty()) continue;
    // read geometry from the slots (host mirror) in device-index order
    for (int kind = 0; kind < nkinds; ++kind)
      for (int shape = 0; shape < pf::kNShapes; ++shape) {
        const size_t off = tiles.size();
        int var = 0;
        for (size_t q = 0; q < ids.size(); ++q) {
          const pf_batch_slot &s = b->slots[ids[q]];
          const int cls = pf::size_class(s.prec, s.nx, s.ny);
          if (b->cls_shape[cls][kind] != shape) continue;
          if (s.p.AAA != 0.0 && s.p.ruido_tipo == PF_RUIDO_SMOOTH) var |= pf::kVarSmooth;
          if (s.p.D != 0.0) var |= pf::kVarTDiff;
          pf::InstDev<double> tmp;  // geometry only
          set_geo(tmp, *b, s);
          const pf::Geo &g = kind == pf::kKindA ? tmp.ga : kind == pf::kKindB ? tmp.gb : tmp.gs;
          for (int sg = 0; sg < g.nseg; ++sg)
            for (int by = 0; by < g.nby; ++by)
              for (int bx = 0; bx < g.nbx; ++bx) tiles.push_back({(int)q, bx, by, sg});
        }
        const int n = (int)(tiles.size() - off);
        if (n) b->groups.push_back({prec, shape, kind, var, off, n, 0, 0});
      }
  }
  for (int which = 0; which < 2; ++which) {
    b->deco_off[which] = chunks.size();
    for (int prec = 0; prec < 2; ++prec) {
      const auto &ids = b->slot_of[prec];
      const size_t off = chunks.size();
      for (size_t q = 0; q < ids.size(); ++q) {
        const pf_batch_slot &s = b->slots[ids[q]];
        const long long cnt = pf::shell_count(s.nx, s.ny, s.nz, which == 0 ? 2 : 1, false);
        for (long long f = 0; f < cnt; f += pf::kChunkCells) chunks.push_back({(int)q, prec, f});
      }
      if (prec == PF_FP64) b->deco_n64[which] = (int)(chunks.size() - off);
    }
    b->deco_n[which] = (int)(chunks.size() - b->deco_off[which]);
  }
  // the split form's last phase decoration runs in the first solute launch
  // of each precision
  for (int prec = 0; prec < 2 && b->split_b; ++prec)
    for (auto &g : b->groups)
      if (g.prec == prec && g.kind == pf::kKindS) {
        g.ndeco = prec == PF_FP64 ? b->deco_n64[0] : b->deco_n[0] - b->deco_n64[0];
        g.deco_off = b->deco_off[0] + (prec == PF_FP64 ? 0 : b->deco_n64[0]);
        break;
      }
  if (tiles.size() > b->tiles_cap) {
    if (b->tiles_dev) HIP_OK(hipFree(b->tiles_dev));
    b->tiles_dev = nullptr;
    b->tiles_cap = 0;
    HIP_OK(hipMalloc((void **)&b->tiles_dev, tiles.size() * sizeof(pf::TileRef)));
    b->tiles_cap = tiles.size();
  }
  if (chunks.size() > b->chunks_cap) {
    if (b->chunks_dev) HIP_OK(hipFree(b->chunks_dev));
    b->chunks_dev = nullptr;
    b->chunks_cap = 0;
    HIP_OK(hipMalloc((void **)&b->chunks_dev, chunks.size() * sizeof(pf::Chunk)));
    b->chunks_cap = chunks.size();
  }
  for (int prec = 0; prec < 2; ++prec) {
    const int need = (int)b->slot_of[prec].size();
    if (need > b->mass_due_cap[prec]) {
      if (b->mass_due[prec]) HIP_OK(hipFree(b->mass_due[prec]));
      b->mass_due[prec] = nullptr;
      b->mass_due_cap[prec] = 0;
      HIP_OK(hipMalloc((void **)&b->mass_due[prec], (size_t)need * sizeof(int)));
      b->mass_due_cap[prec] = need;
    }
  }
  for (int prec = 0; prec < 2; ++prec) {
    const size_t nmp = b->slot_of[prec].size() * pf::kMassBlocks;
    if (nmp > b->mass_part_cap[prec]) {
      if (b->mass_part[prec]) HIP_OK(hipFree(b->mass_part[prec]));
      b->mass_part[prec] = nullptr;
      b->mass_part_cap[prec] = 0;
      HIP_OK(hipMalloc((void **)&b->mass_part[prec], nmp * sizeof(double)));
      b->mass_part_cap[prec] = nmp;
    }
  }
  if (!tiles.empty())
    HIP_OK(hipMemcpyAsync(b->tiles_dev, tiles.data(), tiles.size() * sizeof(pf::TileRef),
                          hipMemcpyHostToDevice, b->stream));
  if (!chunks.empty())
    HIP_OK(hipMemcpyAsync(b->chunks_dev, chunks.data(), chunks.size() * sizeof(pf::Chunk),
                          hipMemcpyHostToDevice, b->stream));
  // the geometry of every instance on the device
  for (int prec = 0; prec < 2; ++prec)
    for (size_t q = 0; q < b->slot_of[prec].size(); ++q) {
      const pf_batch_slot &s = b->slots[b->slot_of[prec][q]];
      if (prec == PF_FP64) {
        pf::InstDev<double> tmp;
        set_geo(tmp, *b, s);
        char *base = (char *)b->inst_dev[prec] + q * sizeof(pf::InstDev<double>);
        HIP_OK(hipMemcpyAsync(base + offsetof(pf::InstDev<double>, ga), &tmp.ga, 3 * sizeof(pf::Geo),
                              hipMemcpyHostToDevice, b->stream));
      } else {
        pf::InstDev<float> tmp;
        set_geo(tmp, *b, s);
        char *base = (char *)b->inst_dev[prec] + q * sizeof(pf::InstDev<float>);
        HIP_OK(hipMemcpyAsync(base + offsetof(pf::InstDev<float>, ga), &tmp.ga, 3 * sizeof(pf::Geo),
                              hipMemcpyHostToDevice, b->stream));
      }
    }
  HIP_OK(hipStreamSynchronize(b->stream));
  b->dirty = false;
  if (b->graph) {
    (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
  }
  b->graph_valid = false;
  return PF_OK;
}

// One group's launch (sweep `sweep` of the step); `deco`: the solute
// launch also runs its phase-decoration chunks.
hipError_t launch_group(pf_batch *b, const pf_batch::Group &g, int sweep, int ledger, bool deco,
                        hipStream_t st) {
  const int mode = g.kind == pf::kKindA   ? pf::kSweepA
                   : g.kind == pf::kKindS ? pf::kSweepSolute
                   : b->split_b           ? pf::kSweepPT
                                          : pf::kSweepB;
  const pf::Chunk *ch = b->chunks_dev + g.deco_off;
  const int nd = deco ? g.ndeco : 0;
  if (g.prec == PF_FP64)
    return pf::launch_batch_group<double>(g.shape, mode, g.var, b->tiles_dev + g.off, g.n,
                                          (pf::InstDev<double> *)b->inst_dev[0], sweep, ledger,
                                          ch, nd, st);
  return pf::launch_batch_group<float>(g.shape, mode, g.var, b->tiles_dev + g.off, g.n,
                                       (pf::InstDev<float> *)b->inst_dev[1], sweep, ledger, ch,
                                       nd, st);
}

// Enqueues the sweeps, decorations and solute-mass corrections of one
// precision's instances for one step (between its beginning and its end) on
// stream st.
int enqueue_chain(pf_batch *b, int prec, hipStream_t st) {
  auto *i64 = (pf::InstDev<double> *)b->inst_dev[0];
  auto *i32 = (pf::InstDev<float> *)b->inst_dev[1];
  auto sweeps = [&](int kind, int sweep) -> int {
    for (const auto &g : b->groups) {
      if (g.kind != kind || g.prec != prec) continue;
      const hipError_t e = launch_group(b, g, sweep, 1, true, st);
      if (e != hipSuccess) return fail_msg(PF_E_HIP, "batched sweep: %s", hipGetErrorString(e));
    }
    return PF_OK;
  };
  // instances of this precision keeping a solute-mass ledger: the
  // correction's interior sums ride on the solute-face launch
  const int ninst = (int)b->slot_of[prec].size();
  bool mass = false;
  for (int id : b->slot_of[prec]) mass = mass || b->slots[id].mass_every > 0;
  const int *due = b->mass_due[prec];
  double *part = b->mass_part[prec];
  auto decos = [&](int which, int sweep) -> int {
    const int n = prec == PF_FP64 ? b->deco_n64[which] : b->deco_n[which] - b->deco_n64[which];
    const int nm = which == 1 && mass ? ninst : 0;
    const unsigned grid = (unsigned)n + (nm ? pf::kMassBlocks : 0);
    if (grid == 0) return PF_OK;
    const size_t off = b->deco_off[which] + (prec == PF_FP64 ? 0 : b->deco_n64[which]);
    hipLaunchKernelGGL(pf::k_deco_batch, dim3(grid), dim3(256), 0, st, b->chunks_dev + off, n,
                       i64, i32, which, sweep, nm, prec, due, part);
    HIP_OK(hipGetLastError());
    return PF_OK;
  };
  int rc;
  const int nv = b->nv < 1 ? 1 : b->nv;
  for (int sw = 1; sw < nv; ++sw) {
    if ((rc = sweeps(pf::kKindA, sw)) != PF_OK) return rc;
    if ((rc = decos(0, sw)) != PF_OK) return rc;
  }
  // sweep B: the phase + T kernels before the solute kernels (which read
  // the same phi and c, partly still in the caches, and whose launches also
  // decorate the swept phi)
  if ((rc = sweeps(pf::kKindB, nv)) != PF_OK) return rc;
  if (b->split_b) {
    if ((rc = sweeps(pf::kKindS, nv)) != PF_OK) return rc;
  } else if ((rc = decos(0, nv)) != PF_OK) {
    return rc;
  }
  if ((rc = decos(1, nv)) != PF_OK) return rc;
  // solute-mass corrections of the instances due this step (after the
  // step's face copy, as in the single-instance loop; the interior sums ran
  // with the face copy)
  if (!ninst || !mass) return PF_OK;
  if (prec == PF_FP64) {
    hipLaunchKernelGGL(pf::k_bmass_scale<double>, dim3(pf::kMassBlocks), dim3(256), 0, st, i64,
                       ninst, due, (const double *)part);
    hipLaunchKernelGGL(pf::k_bmass_refresh<double>, dim3(pf::kRefreshBlocks), dim3(256), 0, st,
                       i64, ninst, due);
  } else {
    hipLaunchKernelGGL(pf::k_bmass_scale<float>, dim3(pf::kMassBlocks), dim3(256), 0, st, i32,
                       ninst, due, (const double *)part);
    hipLaunchKernelGGL(pf::k_bmass_refresh<float>, dim3(pf::kRefreshBlocks), dim3(256), 0, st,
                       i32, ninst, due);
  }
  HIP_OK(hipGetLastError());
  return PF_OK;
}

// One step's sweeps, decorations and corrections: with instances of both
// precisions, the FP32 chain runs on a second stream beside the FP64 chain
// (forked from and joined back to the batch's stream; a captured graph
// keeps the two as parallel branches), so that each chain's launches fill
// the GPU the other leaves idle.
int enqueue_step_body(pf_batch *b) {
  const bool both = !b->slot_of[0].empty() && !b->slot_of[1].empty();
  int rc;
  if (!both) return enqueue_chain(b, b->slot_of[0].empty() ? PF_FP32_MIXED : PF_FP64, b->stream);
  HIP_OK(hipEventRecord(b->ev_fork, b->stream));
  HIP_OK(hipStreamWaitEvent(b->stream2, b->ev_fork, 0));
  rc = enqueue_chain(b, PF_FP64, b->stream);
  if (rc == PF_OK) rc = enqueue_chain(b, PF_FP32_MIXED, b->stream2);
  // joined even after a failed launch, so that a graph capture can end
  HIP_OK(hipEventRecord(b->ev_join, b->stream2));
  HIP_OK(hipStreamWaitEvent(b->stream, b->ev_join, 0));
  return rc;
}

// Enqueues k steps of every active instance (k >= 1); one launch ends a
// step and begins the next.
int enqueue_steps(pf_batch *b, int k) {
  hipStream_t st = b->stream;
  auto *i64 = (pf::InstDev<double> *)b->inst_dev[0];
  auto *i32 = (pf::InstDev<float> *)b->inst_dev[1];
  const int n64 = (int)b->slot_of[0].size(), n32 = (int)b->slot_of[1].size();
  auto turn = [&](int do_end, int do_begin) -> int {
    hipLaunchKernelGGL(pf::k_batch_turn, dim3((unsigned)(n64 + n32)), dim3(64), 0, st, i64, n64,
                       b->mass_due[0], i32, n32, b->mass_due[1], do_end, do_begin);
    HIP_OK(hipGetLastError());
    return PF_OK;
  };
  int rc;
  if ((rc = turn(0, 1)) != PF_OK) return rc;
  for (int q = 0; q < k; ++q) {
    if ((rc = enqueue_step_body(b)) != PF_OK) return rc;
    if ((rc = turn(1, q + 1 < k)) != PF_OK) return rc;
  }
  return PF_OK;
}

int capture_graph(pf_batch *b) {
  if (b->graph) {
    (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
  }
  b->graph_valid = false;
  hipGraph_t g = nullptr;
  HIP_OK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int rc = PF_OK;
  rc = enqueue_steps(b, b->graph_steps);
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc != PF_OK) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (e != hipSuccess) return fail_msg(PF_E_HIP, "graph capture: %s", hipGetErrorString(e));
  e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (e != hipSuccess) {
    b->graph = nullptr;
    return fail_msg(PF_E_HIP, "graph instantiate: %s", hipGetErrorString(e));
  }
  b->graph_valid = true;
  return PF_OK;
}

template <typename R>
int download_field(pf_batch *b, const void *src, size_t n, double *dst, double off = 0.0) {
  if (!dst) return PF_OK;
  if (sizeof(R) == 8) {
    HIP_OK(hipMemcpyAsync(dst, src, n * 8, hipMemcpyDeviceToHost, b->stream));
    HIP_OK(hipStreamSynchronize(b->stream));
    return PF_OK;
  }
  std::vector<float> tmp(n);
  HIP_OK(hipMemcpyAsync(tmp.data(), src, n * 4, hipMemcpyDeviceToHost, b->stream));
  HIP_OK(hipStreamSynchronize(b->stream));
  for (size_t q = 0; q < n; ++q) dst[q] = (double)tmp[q] + off;
  return PF_OK;
}

int upload_field(pf_batch *b, void *dst, size_t n, int prec, const double *src,
                 double off = 0.0) {
  if (prec == PF_FP64) {
    HIP_OK(hipMemcpyAsync(dst, src, n * 8, hipMemcpyHostToDevice, b->stream));
    HIP_OK(hipStreamSynchronize(b->stream));
    return PF_OK;
  }
  std::vector<float> tmp(n);
  for (size_t q = 0; q < n; ++q) tmp[q] = (float)(src[q] - off);
  HIP_OK(hipMemcpyAsync(dst, tmp.data(), n * 4, hipMemcpyHostToDevice, b->stream));
  HIP_OK(hipStreamSynchronize(b->stream));
  return PF_OK;
}

// Reads the device state of a slot into a host InstDev (FP64 layout for
// the scalars shared by both precisions).
template <typename R>
int read_inst(pf_batch *b, const pf_batch_slot &s, pf::InstDev<R> &I) {
  HIP_OK(hipMemcpyAsync(&I, (char *)b->inst_dev[s.prec] + (size_t)s.idx * sizeof(I), sizeof(I),
                        hipMemcpyDeviceToHost, b->stream));
  HIP_OK(hipStreamSynchronize(b->stream));
  return PF_OK;
}

int ensure_inst_cap(pf_batch *b, int prec, int need) {
  if (need <= b->inst_cap[prec]) return PF_OK;
  const int cap = std::max(need, std::max(8, 2 * b->inst_cap[prec]));
  void *nw = nullptr;
  HIP_OK(hipMalloc(&nw, (size_t)cap * inst_size(prec)));
  if (b->inst_dev[prec]) {
    HIP_OK(hipMemcpyAsync(nw, b->inst_dev[prec], (size_t)b->slot_of[prec].size() * inst_size(prec),
                          hipMemcpyDeviceToDevice, b->stream));
    HIP_OK(hipStreamSynchronize(b->stream));
    HIP_OK(hipFree(b->inst_dev[prec]));
  }
  b->inst_dev[prec] = nw;
  b->inst_cap[prec] = cap;
  return PF_OK;
}

}  // namespace

extern "C" {

int pf_batch_create(size_t hbm_budget, pf_batch **out) {
  if (!out) return fail_msg(PF_E_ARG, "null output");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail_msg(PF_E_NODEV, "no HIP device visible");
  pf_batch *b = new pf_batch();
  for (int c = 0; c < pf::kNClasses; ++c)  // 16x16 until autotuned
    for (int k = 0; k < pf::kNKinds; ++k) b->cls_shape[c][k] = 3;
  if (hipGetDevice(&b->device) != hipSuccess) {
    delete b;
    return fail_msg(PF_E_HIP, "hipGetDevice");
  }
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    delete b;
    return fail_msg(PF_E_HIP, "hipStreamCreate");
  }
  if (hipStreamCreateWithFlags(&b->stream2, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming) != hipSuccess) {
    pf_batch_destroy(b);
    return fail_msg(PF_E_HIP, "hipStreamCreate / hipEventCreate");
  }
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) fr = 0;  // then only an explicit budget admits
  b->budget = hbm_budget ? hbm_budget : (size_t)(0.9 * (double)fr);
  if (hipHostMalloc((void **)&b->pinned, 8 * sizeof(double)) != hipSuccess) b->pinned = nullptr;
  *out = b;
  return PF_OK;
}

int pf_batch_destroy(pf_batch *b) {
  if (!b) return PF_OK;
  // an error of work still in flight is reported here; the batch is freed
  // either way
  const hipError_t e = b->stream ? hipStreamSynchronize(b->stream) : hipSuccess;
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  for (auto &s : b->slots)
    for (void *p : s.buf)
      if (p) (void)hipFree(p);
  for (int prec = 0; prec < 2; ++prec)
    if (b->inst_dev[prec]) (void)hipFree(b->inst_dev[prec]);
  if (b->tiles_dev) (void)hipFree(b->tiles_dev);
  if (b->chunks_dev) (void)hipFree(b->chunks_dev);
  if (b->diag_part) (void)hipFree(b->diag_part);
  for (int prec = 0; prec < 2; ++prec)
    if (b->mass_part[prec]) (void)hipFree(b->mass_part[prec]);
  for (int prec = 0; prec < 2; ++prec)
    if (b->mass_due[prec]) (void)hipFree(b->mass_due[prec]);
  if (b->pinned) (void)hipHostFree(b->pinned);
  if (b->stream2) (void)hipStreamSynchronize(b->stream2);
  if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
  if (b->ev_join) (void)hipEventDestroy(b->ev_join);
  if (b->stream2) (void)hipStreamDestroy(b->stream2);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  if (e != hipSuccess) return fail_msg(PF_E_HIP, "pending work failed: %s", hipGetErrorString(e));
  return PF_OK;
}

int pf_batch_add(pf_batch *b, const pf_params *p, const pf_loop *L, int32_t nsteps,
                 const double *phi, const double *c, const double *T, int32_t *slot) {
  int rc = check_batch(b);
  if (rc != PF_OK) return rc;
  if (!p || !L || !phi || !c || !T || !slot) return fail_msg(PF_E_ARG, "null argument");
  *slot = -1;
  if (p->nx < 6 || p->ny < 6 || p->nz < 6)
    return fail_msg(PF_E_ARG, "batched grids need every extent >= 6 (got %dx%dx%d)", p->nx, p->ny,
                    p->nz);
  if ((long long)p->nx * p->ny * (p->precision == PF_FP64 ? 8 : 4) >= (1LL << 32))
    return fail_msg(PF_E_ARG, "plane too large (a plane of a field must be < 4 GiB)");
  if (p->precision != PF_FP64 && p->precision != PF_FP32_MIXED)
    return fail_msg(PF_E_ARG, "unknown precision %d", p->precision);
  if (nsteps < 0) return fail_msg(PF_E_ARG, "negative step count");
  const int nv = p->n_varreduras < 1 ? 1 : p->n_varreduras;
  bool any = false;
  for (auto &s : b->slots) any = any || s.used;
  if (any && nv != b->nv)
    return fail_msg(PF_E_ARG, "instances of one batch share n_varreduras (%d, got %d)", b->nv, nv);
  pf_batch_slot s;
  s.used = true;
  s.prec = p->precision;
  s.p = *p;
  s.L = *L;
  s.nx = p->nx; s.ny = p->ny; s.nz = p->nz;
  s.nv = nv;
  s.n = (size_t)p->nx * p->ny * p->nz;
  const size_t esz = s.prec == PF_FP64 ? 8 : 4;
  s.bytes = 6 * s.n * esz;
  if (b->used + s.bytes > b->budget)
    return fail_msg(PF_E_NOMEM, "instance needs %zu bytes, %zu of %zu in use", s.bytes, b->used,
                    b->budget);
  s.steps_total = s.steps_left = nsteps;
  // the solute-mass ledger as pf_run_steps keeps it by default
  s.mass_every = s.prec == PF_FP32_MIXED ? 100 : 0;
  s.host_mass = pf::host_interior_mass(c, s.nx, s.ny, s.nz);
  for (int f = 0; f < 6; ++f) {
    if (hipMalloc(&s.buf[f], s.n * esz) != hipSuccess) {
      for (int g = 0; g < f; ++g) (void)hipFree(s.buf[g]);
      return fail_msg(PF_E_NOMEM, "hipMalloc of %zu bytes failed", s.n * esz);
    }
  }
  auto cleanup = [&](int code) {
    for (void *q : s.buf)
      if (q) (void)hipFree(q);
    return code;
  };
  // phi_old, c_old, T_old into both buffers of each pair (the new buffers are
  // fully rewritten before they are read; T faces stay TOLD in both)
  if ((rc = upload_field(b, s.buf[0], s.n, s.prec, phi)) != PF_OK) return cleanup(rc);
  if (s.prec == PF_FP32_MIXED) s.c_off = c[s.n - 1];  // as pf_init: the top corner cell
  if ((rc = upload_field(b, s.buf[2], s.n, s.prec, c, s.c_off)) != PF_OK) return cleanup(rc);
  if ((rc = upload_field(b, s.buf[4], s.n, s.prec, T)) != PF_OK) return cleanup(rc);
  if ((rc = upload_field(b, s.buf[5], s.n, s.prec, T)) != PF_OK) return cleanup(rc);
  // temperatura of step 1 (principal.f90:98) with DPP = 0: c[1] is dead
  // until the first solute update, so it serves as the zero DPP field
  if (hipMemsetAsync(s.buf[3], 0, s.n * esz, b->stream) != hipSuccess)
    return cleanup(fail_msg(PF_E_HIP, "hipMemsetAsync"));
  const pf::KParams kt = pf::kparams_from(*p);
  hipError_t e;
  if (s.prec == PF_FP64)
    e = pf::launch_temperatura<double>(kt, (const double *)s.buf[4], (const double *)s.buf[0],
                                       (const double *)s.buf[3], (double *)s.buf[5], b->stream);
  else
    e = pf::launch_temperatura<float>(kt, (const float *)s.buf[4], (const float *)s.buf[0],
                                      (const float *)s.buf[3], (float *)s.buf[5], b->stream);
  if (e != hipSuccess) return cleanup(fail_msg(PF_E_HIP, "temperatura: %s", hipGetErrorString(e)));
  // device state
  const int prec = s.prec;
  if ((rc = ensure_inst_cap(b, prec, (int)b->slot_of[prec].size() + 1)) != PF_OK)
    return cleanup(rc);
  int id = -1;
  for (size_t q = 0; q < b->slots.size(); ++q)
    if (!b->slots[q].used) {
      id = (int)q;
      break;
    }
  if (id < 0) {
    id = (int)b->slots.size();
    b->slots.push_back(pf_batch_slot());
  }
  s.idx = (int)b->slot_of[prec].size();
  if (prec == PF_FP64) {
    pf::InstDev<double> I = make_inst<double>(s);
    set_geo(I, *b, s);
    e = hipMemcpyAsync((char *)b->inst_dev[prec] + (size_t)s.idx * sizeof I, &I, sizeof I,
                       hipMemcpyHostToDevice, b->stream);
  } else {
    pf::InstDev<float> I = make_inst<float>(s);
    set_geo(I, *b, s);
    e = hipMemcpyAsync((char *)b->inst_dev[prec] + (size_t)s.idx * sizeof I, &I, sizeof I,
                       hipMemcpyHostToDevice, b->stream);
  }
  if (e != hipSuccess || hipStreamSynchronize(b->stream) != hipSuccess)
    return cleanup(fail_msg(PF_E_HIP, "instance upload failed"));
  b->slot_of[prec].push_back(id);
  b->slots[id] = s;
  b->used += s.bytes;
  b->nv = nv;
  b->dirty = true;
  *slot = id;
  return PF_OK;
}

int pf_batch_remove(pf_batch *b, int32_t slot) {
  int rc = check_slot(b, slot);
  if (rc != PF_OK) return rc;
  HIP_OK(hipStreamSynchronize(b->stream));
  pf_batch_slot &s = b->slots[slot];
  const int prec = s.prec;
  auto &ids = b->slot_of[prec];
  // move the last device entry into the removed one's place
  const int last = (int)ids.size() - 1;
  if (s.idx != last) {
    const size_t sz = inst_size(prec);
    HIP_OK(hipMemcpyAsync((char *)b->inst_dev[prec] + (size_t)s.idx * sz,
                          (char *)b->inst_dev[prec] + (size_t)last * sz, sz,
                          hipMemcpyDeviceToDevice, b->stream));
    HIP_OK(hipStreamSynchronize(b->stream));
    ids[s.idx] = ids[last];
    b->slots[ids[s.idx]].idx = s.idx;
  }
  ids.pop_back();
  for (void *p : s.buf)
    if (p) HIP_OK(hipFree(p));
  b->used -= s.bytes;
  s = pf_batch_slot();
  b->dirty = true;
  return PF_OK;
}

int pf_batch_run(pf_batch *b, int32_t nsteps) {
  int rc = check_batch(b);
  if (rc != PF_OK) return rc;
  if (nsteps < 0) return fail_msg(PF_E_ARG, "negative step count");
  if (b->slot_of[0].empty() && b->slot_of[1].empty()) return PF_OK;
  if (b->dirty && (rc = rebuild(b)) != PF_OK) return rc;
  int left = nsteps;
  if (b->graph_steps > 0 && left >= b->graph_steps) {
    if (!b->graph_valid && (rc = capture_graph(b)) != PF_OK) return rc;
    while (left >= b->graph_steps) {
      HIP_OK(hipGraphLaunch(b->graph, b->stream));
      left -= b->graph_steps;
    }
  }
  if (left > 0 && (rc = enqueue_steps(b, left)) != PF_OK) return rc;
  for (auto &s : b->slots)
    if (s.used) s.steps_left = std::max(0, s.steps_left - nsteps);
  return PF_OK;
}

int pf_batch_set_graph_steps(pf_batch *b, int32_t k) {
  int rc = check_batch(b);
  if (rc != PF_OK) return rc;
  if (k < 0) return fail_msg(PF_E_ARG, "negative graph length");
  b->graph_steps = k;
  if (b->graph) {
    HIP_OK(hipStreamSynchronize(b->stream));
    (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
  }
  b->graph_valid = false;
  return PF_OK;
}

int pf_batch_synchronize(pf_batch *b) {
  int rc = check_batch(b);
  if (rc != PF_OK) return rc;
  HIP_OK(hipStreamSynchronize(b->stream));
  return PF_OK;
}

int pf_batch_status(pf_batch *b, int32_t slot, int32_t *steps_done, int32_t *steps_left) {
  int rc = check_slot(b, slot);
  if (rc != PF_OK) return rc;
  const pf_batch_slot &s = b->slots[slot];
  if (steps_done) *steps_done = s.steps_total - s.steps_left;
  if (steps_left) *steps_left = s.steps_left;
  return PF_OK;
}

int pf_batch_scalars(pf_batch *b, int32_t slot, pf_params *p, pf_loop *L) {
  int rc = check_slot(b, slot);
  if (rc != PF_OK) return rc;
  const pf_batch_slot &s = b->slots[slot];
  pf::KParams k;
  float time;
  if (s.prec == PF_FP64) {
    pf::InstDev<double> I;
    if ((rc = read_inst(b, s, I)) != PF_OK) return rc;
    k = I.k;
    time = I.time;
  } else {
    pf::InstDev<float> I;
    if ((rc = read_inst(b, s, I)) != PF_OK) return rc;
    k = I.k;
    time = I.time;
  }
  if (p) {
    *p = s.p;
    p->it = k.it;
    p->TP = k.TP;
    p->M0 = k.M0;
    p->Cs_eq = k.Cs_eq;
    p->Cl_eq = k.Cl_eq;
  }
  if (L) {
    *L = s.L;
    L->time = time;
  }
  return PF_OK;
}

int pf_batch_download(pf_batch *b, int32_t slot, double *phi, double *c, double *T) {
  int rc = check_slot(b, slot);
  if (rc != PF_OK) return rc;
  const pf_batch_slot &s = b->slots[slot];
  int po, co, to;
  if (s.prec == PF_FP64) {
    pf::InstDev<double> I;
    if ((rc = read_inst(b, s, I)) != PF_OK) return rc;
    po = I.po; co = I.co; to = I.to;
  } else {
    pf::InstDev<float> I;
    if ((rc = read_inst(b, s, I)) != PF_OK) return rc;
    po = I.po; co = I.co; to = I.to;
  }
  // after a step: phi[po], c[co] are the newest; T[to] holds the next step's
  // temperature, T[to ^ 1] the reference's TNEW of the last step
  if (s.prec == PF_FP64) {
    if ((rc = download_field<double>(b, s.buf[po], s.n, phi)) != PF_OK) return rc;
    if ((rc = download_field<double>(b, s.buf[2 + co], s.n, c)) != PF_OK) return rc;
    return download_field<double>(b, s.buf[4 + (to ^ 1)], s.n, T);
  }
  if ((rc = download_field<float>(b, s.buf[po], s.n, phi)) != PF_OK) return rc;
  if ((rc = download_field<float>(b, s.buf[2 + co], s.n, c, s.c_off)) != PF_OK) return rc;
  return download_field<float>(b, s.buf[4 + (to ^ 1)], s.n, T);
}

int pf_batch_diagnosticos(pf_batch *b, int32_t slot, double out[5]) {
  int rc = check_slot(b, slot);
  if (rc != PF_OK) return rc;
  if (!out) return fail_msg(PF_E_ARG, "null output");
  const pf_batch_slot &s = b->slots[slot];
  const int nblk = pf::diag_blocks(s.ny, s.nz);
  if (nblk + 1 > b->diag_cap) {
    if (b->diag_part) HIP_OK(hipFree(b->diag_part));
    b->diag_part = nullptr;
    b->diag_cap = 0;
    HIP_OK(hipMalloc((void **)&b->diag_part, sizeof(double) * 5 * (size_t)(nblk + 1)));
    b->diag_cap = nblk + 1;
  }
  pf_params p;
  if ((rc = pf_batch_scalars(b, slot, &p, nullptr)) != PF_OK) return rc;
  int po, co;
  if (s.prec == PF_FP64) {
    pf::InstDev<double> I;
    if ((rc = read_inst(b, s, I)) != PF_OK) return rc;
    po = I.po; co = I.co;
  } else {
    pf::InstDev<float> I;
    if ((rc = read_inst(b, s, I)) != PF_OK) return rc;
    po = I.po; co = I.co;
  }
  pf::KParams k = pf::kparams_from(p);
  k.c_off = s.c_off;
  double *outd = b->diag_part + 5 * nblk;
  hipError_t e;
  if (s.prec == PF_FP64)
    e = pf::launch_diag<double>(k, (const double *)s.buf[po], (const double *)s.buf[2 + co],
                                b->diag_part, nblk, outd, b->stream);
  else
    e = pf::launch_diag<float>(k, (const float *)s.buf[po], (const float *)s.buf[2 + co],
                               b->diag_part, nblk, outd, b->stream);
  if (e != hipSuccess) return fail_msg(PF_E_HIP, "diagnostics: %s", hipGetErrorString(e));
  HIP_OK(hipMemcpyAsync(out, outd, 5 * sizeof(double), hipMemcpyDeviceToHost, b->stream));
  HIP_OK(hipStreamSynchronize(b->stream));
  return PF_OK;
}

int pf_batch_bytes(const pf_batch *b, size_t *used, size_t *budget) {
  int rc = check_batch(b);
  if (rc != PF_OK) return rc;
  if (used) *used = b->used;
  if (budget) *budget = b->budget;
  return PF_OK;
}

int pf_batch_set_tiling(pf_batch *b, int32_t size_class, int32_t shape) {
  int rc = check_batch(b);
  if (rc != PF_OK) return rc;
  int sh[pf::kNKinds];
  if (shape >= 0 && shape < pf::kNShapes) {
    for (int k = 0; k < pf::kNKinds; ++k) sh[k] = shape;
  } else if ((shape & ~0xfff) == 0x1000) {  // PF_BATCH_TILES(a, b, s)
    for (int k = 0; k < pf::kNKinds; ++k) sh[k] = (shape >> (4 * k)) & 0xf;
  } else {
    sh[0] = -1;
  }
  bool ok = size_class >= 0 && size_class < pf::kNClasses;
  for (int k = 0; k < pf::kNKinds; ++k) ok = ok && sh[k] >= 0 && sh[k] < pf::kNShapes;
  if (!ok)
    return fail_msg(PF_E_ARG, "size class in [0,%d), shape in [0,%d) or PF_BATCH_TILES(a, b, s)",
                    pf::kNClasses, pf::kNShapes);
  HIP_OK(hipStreamSynchronize(b->stream));
  for (int k = 0; k < pf::kNKinds; ++k) b->cls_shape[size_class][k] = sh[k];
  b->dirty = true;
  return PF_OK;
}

int pf_batch_get_tiling(const pf_batch *b, int32_t size_class, int32_t *shape) {
  int rc = check_batch(b);
  if (rc != PF_OK) return rc;
  if (size_class < 0 || size_class >= pf::kNClasses) return fail_msg(PF_E_ARG, "bad size class");
  const int *sh = b->cls_shape[size_class];
  if (shape)
    *shape = sh[0] == sh[1] && sh[1] == sh[2] ? sh[0] : PF_BATCH_TILES(sh[0], sh[1], sh[2]);
  return PF_OK;
}

int pf_batch_set_group_target(pf_batch *b, int32_t target_groups) {
  int rc = check_batch(b);
  if (rc != PF_OK) return rc;
  if (target_groups < 1) return fail_msg(PF_E_ARG, "target_groups must be >= 1");
  HIP_OK(hipStreamSynchronize(b->stream));
  b->group_target = target_groups;
  b->dirty = true;
  return PF_OK;
}

int32_t pf_batch_size_class(int32_t precision, int32_t nx, int32_t ny) {
  return pf::size_class(precision == PF_FP32_MIXED ? 1 : 0, nx, ny);
}

// Times each tile shape for the size class (-1: every class present) and
// each kernel kind (sweep A, sweep B or the split form's phase + T kernel,
// the split form's solute kernel) as the sweeps of the whole batch take with
// that choice, launched in the order of a step (the other classes and kinds
// keep theirs: how a class's work-groups interleave with the rest, and what
// the kernel before leaves in the caches, matter more than its time alone),
// and keeps the fastest.  Every kind is timed with the roles of the first
// phase sweep (read the current phi, c, T; write the step's dead buffers),
// so no step is taken.
int pf_batch_autotune(pf_batch *b, int32_t size_class, int32_t reps) {
  int rc = check_batch(b);
  if (rc != PF_OK) return rc;
  if (reps < 1) return fail_msg(PF_E_ARG, "reps must be >= 1");
  if (size_class < -1 || size_class >= pf::kNClasses) return fail_msg(PF_E_ARG, "bad size class");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIP_OK(hipEventCreate(&e0));
  HIP_OK(hipEventCreate(&e1));
  const int nkinds = b->split_b ? pf::kNKinds : pf::kKindS;
  for (int cls = 0; cls < pf::kNClasses && rc == PF_OK; ++cls) {
    if (size_class >= 0 && cls != size_class) continue;
    bool present = false;
    for (const auto &s : b->slots)
      present = present || (s.used && pf::size_class(s.prec, s.nx, s.ny) == cls);
    if (!present) continue;
    for (int kind = 0; kind < nkinds && rc == PF_OK; ++kind) {
      float best = 1e30f;
      int bs = b->cls_shape[cls][kind];
      for (int shape = 0; shape < pf::kNShapes && rc == PF_OK; ++shape) {
        b->cls_shape[cls][kind] = shape;
        if ((rc = rebuild(b)) != PF_OK) break;
        hipError_t e = hipSuccess;
        for (int r = 0; r <= reps && e == hipSuccess; ++r) {
          if (r == 1) e = hipEventRecord(e0, b->stream);
          // timing launches: the ledgers are left alone
          for (int k = 0; k < nkinds && e == hipSuccess; ++k) {
            if (k == pf::kKindS) continue;
            for (int prec = 0; prec < 2; ++prec)
              for (int kk = k; kk <= (k == pf::kKindB ? nkinds - 1 : k); ++kk)
                for (const auto &g : b->groups)
                  if (e == hipSuccess && g.kind == kk && g.prec == prec)
                    e = launch_group(b, g, 1, 0, false, b->stream);
          }
        }
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventRecord(e1, b->stream);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) {
          rc = fail_msg(PF_E_HIP, "autotune: %s", hipGetErrorString(e));
          break;
        }
        if (ms < best) {
          best = ms;
          bs = shape;
        }
      }
      b->cls_shape[cls][kind] = bs;
    }
  }
  if (hipStreamSynchronize(b->stream) != hipSuccess && rc == PF_OK)
    rc = fail_msg(PF_E_HIP, "autotune: %s", hipGetErrorString(hipGetLastError()));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->dirty = true;
  return rc;
}

int pf_batch_set_sweep_split(pf_batch *b, int32_t on) {
  int rc = check_batch(b);
  if (rc != PF_OK) return rc;
  HIP_OK(hipStreamSynchronize(b->stream));
  b->split_b = on != 0;
  b->dirty = true;  // recapture the steps
  return PF_OK;
}

int pf_batch_set_mass_correction(pf_batch *b, int32_t slot, int32_t every) {
  int rc = check_slot(b, slot);
  if (rc != PF_OK) return rc;
  if (every < 0) return fail_msg(PF_E_ARG, "negative correction period");
  pf_batch_slot &s = b->slots[slot];
  if (s.steps_left != s.steps_total)
    return fail_msg(PF_E_STATE, "the correction period is set before the instance's first step");
  HIP_OK(hipStreamSynchronize(b->stream));
  const size_t off = s.prec == PF_FP64 ? offsetof(pf::InstDev<double>, mass_every)
                                       : offsetof(pf::InstDev<float>, mass_every);
  HIP_OK(hipMemcpyAsync((char *)b->inst_dev[s.prec] + (size_t)s.idx * inst_size(s.prec) + off,
                        &every, sizeof every, hipMemcpyHostToDevice, b->stream));
  HIP_OK(hipStreamSynchronize(b->stream));
  s.mass_every = every;
  b->dirty = true;  // the captured steps launch the correction kernels only where needed
  return PF_OK;
}

int pf_batch_mass_ledger(pf_batch *b, int32_t slot, double out[3]) {
  int rc = check_slot(b, slot);
  if (rc != PF_OK) return rc;
  if (!out) return fail_msg(PF_E_ARG, "null output");
  const pf_batch_slot &s = b->slots[slot];
  pf::MassLedger m;
  if (s.prec == PF_FP64) {
    pf::InstDev<double> I;
    if ((rc = read_inst(b, s, I)) != PF_OK) return rc;
    m = I.m;
  } else {
    pf::InstDev<float> I;
    if ((rc = read_inst(b, s, I)) != PF_OK) return rc;
    m = I.m;
  }
  out[0] = pf::ledger_value(m);
  out[1] = m.S;
  out[2] = m.f;
  return PF_OK;
}

}  // extern "C"

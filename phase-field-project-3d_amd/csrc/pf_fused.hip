// pf_fused.hip -- the fused time step of one instance on gfx950: launches of
// the z-march sweeps of pf_march.h (tile shapes, XCD-aware work-group order,
// autotuning of the tiling) and the FP32 form's solute-mass ledger.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "pf_common.h"
#include "pf_fused.h"
#include "pf_kernels.h"
#include "pf_march.h"
#include "pf_mass.h"
#include "pf_math.h"

namespace pf {
namespace {

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

template <typename R, int MODE, typename TL, int VAR>
__global__ void __launch_bounds__(TL::NT)
    __attribute__((amdgpu_waves_per_eu(march_waves<R, MODE, VAR>())))
    k_march(Geo g, KParams kp, SC<R> s, const R *__restrict__ po, const R *__restrict__ co,
            R *__restrict__ pn, R *__restrict__ cn, const R *__restrict__ Tcur,
            R *__restrict__ Tnext, R *__restrict__ dpp, int write_T,
            unsigned long long *__restrict__ macc, const R *__restrict__ cfo, R *__restrict__ cfn,
            int ncf) {
  const int nb = 8 * g.per_xcd;
  if ((int)blockIdx.x >= nb) {
    // the ncf work-groups after the tiles' copy the solute faces of the
    // previous step (rows only: the sweeps wrote the x faces of the interior
    // rows) into cfn from its interior and the snapshot cfo; this sweep reads
    // c only at interior cells, so the two are independent
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const long long cnt = shell_count(nx, ny, nz, 1, false);
    const int tid = threadIdx.y * TL::TX + threadIdx.x;
    for (long long t = (long long)((int)blockIdx.x - nb) * TL::NT + tid; t < cnt;
         t += (long long)ncf * TL::NT) {
      int i, j, k;
      shell_flat(t, nx, ny, nz, 1, i, j, k, false);
      solute_deco_cell(i, j, k, nx, ny, nz, cfo, cfn);
    }
    return;
  }
  int bx, by, sg;
  if (!tile_of(g, bx, by, sg)) return;
  march_tile<R, MODE, TL, VAR>(g, kp, s, po, co, pn, cn, Tcur, Tnext, dpp, write_T, macc, bx, by,
                                  sg);
}

// The solute update alone (split form of sweep B).  The ndeco work-groups
// after the tiles' decorate the phase sweep's phi (dpn, from the swept
// interior and the pre-sweep dpo) in the same launch: the solute reads only
// the pre-sweep phi and c, so the two are independent, and the decoration
// needs no launch of its own.  They run the shell rows only (the sweep wrote
// the x faces of the interior rows).
template <typename R, typename TL>
__global__ void __launch_bounds__(TL::NT)
    k_solute(Geo g, KParams kp, SC<R> s, const R *__restrict__ po, const R *__restrict__ co,
             R *__restrict__ cn, unsigned long long *__restrict__ macc, const R *__restrict__ dpo,
             R *__restrict__ dpn, int ndeco) {
  const int nb = 8 * g.per_xcd;
  if ((int)blockIdx.x >= nb) {  // decoration work-group
    const int nx = g.nx, ny = g.ny, nz = g.nz;
    const long long cnt = shell_count(nx, ny, nz, 2, false);
    const int tid = threadIdx.y * TL::TX + threadIdx.x;
    for (long long t = (long long)((int)blockIdx.x - nb) * TL::NT + tid; t < cnt;
         t += (long long)ndeco * TL::NT) {
      int i, j, k;
      shell_flat(t, nx, ny, nz, 2, i, j, k, false);
      phase_deco_cell(i, j, k, nx, ny, nz, dpo, dpn);
    }
    return;
  }
  int bx, by, sg;
  if (!tile_of(g, bx, by, sg)) return;
  solute_tile<R, TL>(g, kp, s, po, co, cn, macc, bx, by, sg);
}

using Tile64x8 = Tile<64, 8>;
using Tile32x8 = Tile<32, 8>;
using Tile32x16 = Tile<32, 16>;
using Tile16x16 = Tile<16, 16>;
constexpr int kTileX[PF_NVARIANTS] = {64, 32, 32, 16};
constexpr int kTileY[PF_NVARIANTS] = {8, 8, 16, 16};

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

template <typename R, typename TL, int VAR>
void launch_march_v(int mode, const Geo &g, const KParams &kp, const SC<R> &s, const R *po,
                    const R *co, R *pn, R *cn, const R *Tcur, R *Tnext, R *dpp, int write_T,
                    unsigned long long *mslot, hipStream_t st, const R *cfo, R *cfn) {
  long long ncf = 0;  // solute-face work-groups (sweep A only)
  if (cfn && mode == kSweepA) {
    ncf = (shell_count(g.nx, g.ny, g.nz, 1, false) + TL::NT - 1) / TL::NT;
    ncf = ncf < 1 ? 1 : (ncf > 2048 ? 2048 : ncf);
  }
  const dim3 blk(TL::TX, TL::TY, 1), grd(8 * g.per_xcd + (unsigned)ncf, 1, 1);
  const size_t lds = march_lds_bytes<R, TL>(mode);
  if (mode == kSweepB)
    hipLaunchKernelGGL((k_march<R, kSweepB, TL, VAR>), grd, blk, lds, st, g, kp, s, po, co,
                       pn, cn, Tcur, Tnext, dpp, write_T, mslot, nullptr, nullptr, 0);
  else if (mode == kSweepPT)
    hipLaunchKernelGGL((k_march<R, kSweepPT, TL, VAR>), grd, blk, lds, st, g, kp, s, po, co,
                       pn, cn, Tcur, Tnext, dpp, write_T, nullptr, nullptr, nullptr, 0);
  else
    hipLaunchKernelGGL((k_march<R, kSweepA, TL, VAR & kVarSmooth>), grd, blk, lds, st, g, kp, s,
                       po, co, pn, cn, Tcur, Tnext, dpp, write_T, nullptr, cfo, cfn, (int)ncf);
}

template <typename R, typename TL>
hipError_t launch_march_t(int mode, const Geo &g, const KParams &kp, const SC<R> &s,
                          const R *po, const R *co, R *pn, R *cn, const R *Tcur, R *Tnext, R *dpp,
                          int write_T, unsigned long long *mslot, hipStream_t st, const R *cfo,
                          R *cfn) {
  // the optional paths are compiled only into the kernels that need them
  const int var = (kp.AAA != 0.0 && kp.ruido_tipo == PF_RUIDO_SMOOTH ? kVarSmooth : 0) |
                  (kp.D != 0.0 ? kVarTDiff : 0);
#define PF_VAR(V) \
  launch_march_v<R, TL, V>(mode, g, kp, s, po, co, pn, cn, Tcur, Tnext, dpp, write_T, mslot, st, \
                           cfo, cfn)
  switch (var) {
    case 0: PF_VAR(0); break;
    case kVarSmooth: PF_VAR(kVarSmooth); break;
    case kVarTDiff: PF_VAR(kVarTDiff); break;
    default: PF_VAR(kVarAll); break;
  }
#undef PF_VAR
  return hipGetLastError();
}

// cfo/cfn (sweep A only): copy the previous step's solute faces into cfn in
// the same launch.
template <typename R>
hipError_t launch_march(int mode, int tile, const Geo &g, const KParams &kp, const R *po,
                        const R *co, R *pn, R *cn, const R *Tcur, R *Tnext, R *dpp, int write_T,
                        unsigned long long *mslot, hipStream_t st, const R *cfo = nullptr,
                        R *cfn = nullptr) {
  const SC<R> s = make_sc<R>(kp);
#define PF_MARCH(TL)                                                                           \
  return launch_march_t<R, TL>(mode, g, kp, s, po, co, pn, cn, Tcur, Tnext, dpp, write_T, mslot, \
                               st, cfo, cfn)
  switch (tile) {
    case 1: PF_MARCH(Tile32x8);
    case 2: PF_MARCH(Tile32x16);
    case 3: PF_MARCH(Tile16x16);
    default: PF_MARCH(Tile64x8);
  }
#undef PF_MARCH
}

template <typename R, typename TL>
hipError_t launch_solute_t(const Geo &g, const KParams &kp, const SC<R> &s, const R *po,
                           const R *co, R *cn, unsigned long long *mslot, const R *dpo, R *dpn,
                           hipStream_t st) {
  long long nd = 0;
  if (dpn) {
    nd = (shell_count(g.nx, g.ny, g.nz, 2, false) + TL::NT - 1) / TL::NT;
    nd = nd < 1 ? 1 : (nd > 2048 ? 2048 : nd);
  }
  const dim3 blk(TL::TX, TL::TY, 1), grd(8 * g.per_xcd + (unsigned)nd, 1, 1);
  const size_t lds = solute_lds_bytes<TL, R>();
  hipLaunchKernelGGL((k_solute<R, TL>), grd, blk, lds, st, g, kp, s, po, co, cn, mslot, dpo, dpn,
                     (int)nd);
  return hipGetLastError();
}

template <typename R>
hipError_t launch_solute_split(int tile, const Geo &g, const KParams &kp, const R *po, const R *co,
                               R *cn, unsigned long long *mslot, hipStream_t st,
                               const R *dpo = nullptr,
                               R *dpn = nullptr) {
  const SC<R> s = make_sc<R>(kp);
  switch (tile) {
    case 1: return launch_solute_t<R, Tile32x8>(g, kp, s, po, co, cn, mslot, dpo, dpn, st);
    case 2: return launch_solute_t<R, Tile32x16>(g, kp, s, po, co, cn, mslot, dpo, dpn, st);
    case 3: return launch_solute_t<R, Tile16x16>(g, kp, s, po, co, cn, mslot, dpo, dpn, st);
    default: return launch_solute_t<R, Tile64x8>(g, kp, s, po, co, cn, mslot, dpo, dpn, st);
  }
}

// The ledger is kept while it is valid (fused_mass_prepare starts it before
// a run); the solute launch adds its increments to the ledger's sum.
unsigned long long *mass_acc(const FusedState &S) {
  return S.mass_every > 0 && S.ledger_valid && S.mass ? &S.mass->acc : nullptr;
}

// Geometry of the launch that updates the solute: sweep B, or the solute
// kernel of the split form.
Geo solute_geo(const FusedState &S) {
  return S.split_b ? make_geo(S.nx, S.ny, S.nz, 2, S.nz - 1, S.seg_s, S.tile_s)
                   : make_geo(S.nx, S.ny, S.nz, 2, S.nz - 1, S.seg_b, S.tile_b);
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
      // with the previous step's solute faces, if still to copy
      tm.begin(PF_KT_SWEEP_A, st);
      e = launch_march<R>(kSweepA, S.tile_a, ga, k, po, (const R *)c[*c_old], pn, nullptr,
                          nullptr, nullptr, nullptr, 0, nullptr, st, (const R *)S.cf_src,
                          (R *)S.cf_dst);
      S.cf_src = nullptr;
      S.cf_dst = nullptr;
      tm.end(st);
      if (e != hipSuccess) return PF_E_HIP;
      tm.begin(PF_KT_BOUNDARY, st);
      e = launch_phase_boundary<R>(k, po, pn, st, false);  // x faces written by the sweep
      tm.end(st);
      if (e != hipSuccess) return PF_E_HIP;
    } else {
      const int cnw = NEWW(*c_old, *c_alias);
      // T of the next step goes into the buffer of this step's TOLD (dead now)
      const int t_cur = *T_alias ? *T_old : 1 - *T_old;
      const int t_nxt = 1 - t_cur;
      if (S.split_b) {
        // phase + T, then the solute alone (both read the same phi and c)
        tm.begin(PF_KT_SWEEP_B, st);
        e = launch_march<R>(kSweepPT, S.tile_b, gb, k, po, (const R *)c[*c_old], pn, nullptr,
                            (const R *)T[t_cur], (R *)T[t_nxt], (R *)dpp, fs.last ? 0 : 1,
                            nullptr, st);
        tm.end(st);
        if (e != hipSuccess) return PF_E_HIP;
        const Geo gs = solute_geo(S);
        // with the phase decoration of this sweep in the same launch
        tm.begin(PF_KT_SOLUTE, st);
        e = launch_solute_split<R>(S.tile_s, gs, k, po, (const R *)c[*c_old], (R *)c[cnw],
                                   mass_acc(S), st, po, pn);
        tm.end(st);
        if (e != hipSuccess) return PF_E_HIP;
      } else {
        tm.begin(PF_KT_SWEEP_B, st);
        e = launch_march<R>(kSweepB, S.tile_b, gb, k, po, (const R *)c[*c_old], pn, (R *)c[cnw],
                            (const R *)T[t_cur], (R *)T[t_nxt], (R *)dpp, fs.last ? 0 : 1,
                            mass_acc(S), st);
        tm.end(st);
        if (e != hipSuccess) return PF_E_HIP;
      }
      tm.begin(PF_KT_BOUNDARY, st);
      // x faces written by the sweeps; the split form's phase decoration ran
      // in the solute kernel's launch; the solute faces are copied now or by
      // the next sweep A
      e = S.split_b ? hipSuccess : launch_phase_boundary<R>(k, po, pn, st, false);
      if (e == hipSuccess && (fs.flush || fs.nv < 2)) {
        e = launch_solute_boundary<R>(k, (const R *)c[*c_old], (R *)c[cnw], st, false);
      } else if (e == hipSuccess) {
        S.cf_src = c[*c_old];
        S.cf_dst = c[cnw];
      }
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
  fused_set_split(s, true, 0);
  s.mass_every = prec == PF_FP32_MIXED ? 100 : 0;
  return PF_OK;
}

int fused_set_tile(FusedState &s, int tile_a, int tile_b, int target_groups) {
  if (tile_a < 0 || tile_a >= PF_NVARIANTS || tile_b < 0 || tile_b >= PF_NVARIANTS ||
      target_groups < 1)
    return PF_E_ARG;
  s.target = target_groups;
  fused_set_kind(s, 0, tile_a, target_groups);
  fused_set_kind(s, 1, tile_b, target_groups);
  fused_set_kind(s, 2, s.tile_s, target_groups);
  return PF_OK;
}

int fused_set_split(FusedState &s, bool on, int tile_s) {
  if (tile_s < 0 || tile_s >= PF_NVARIANTS) return PF_E_ARG;
  s.split_b = on;
  return fused_set_kind(s, 2, tile_s, s.tgt_s);
}

int fused_set_kind(FusedState &s, int kind, int tile, int target_groups) {
  if (kind < 0 || kind > 2 || tile < 0 || tile >= PF_NVARIANTS || target_groups < 1)
    return PF_E_ARG;
  const int seg = pick_seg(s.nx, s.ny, kind == 0 ? s.nz - 4 : s.nz - 2, tile, target_groups);
  if (kind == 0) { s.tile_a = tile; s.tgt_a = target_groups; s.seg_a = seg; }
  if (kind == 1) { s.tile_b = tile; s.tgt_b = target_groups; s.seg_b = seg; }
  if (kind == 2) { s.tile_s = tile; s.tgt_s = target_groups; s.seg_s = seg; }
  return PF_OK;
}

int fused_get_kind(const FusedState &s, int kind, int out[3]) {
  if (kind < 0 || kind > 2) return PF_E_ARG;
  out[0] = kind == 0 ? s.tile_a : kind == 1 ? s.tile_b : s.tile_s;
  out[1] = kind == 0 ? s.tgt_a : kind == 1 ? s.tgt_b : s.tgt_s;
  out[2] = kind == 0 ? s.seg_a : kind == 1 ? s.seg_b : s.seg_s;
  return PF_OK;
}

namespace {

template <typename R>
int autotune_impl(FusedState &S, const KParams &k, void *phi_old, void *c_old, void *T_old,
                  int reps, hipStream_t st, AutotuneResult *res) {
  const int nx = S.nx, ny = S.ny, nz = S.nz;
  const size_t n = (size_t)nx * ny * nz;
  // scratch outputs: phi, c, T and the second sweep's phi (the context's
  // fields are not touched)
  R *out = nullptr;
  if (hipMalloc((void **)&out, 4 * n * sizeof(R)) != hipSuccess) return PF_E_NOMEM;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = PF_OK;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) rc = PF_E_HIP;
  // work-group targets: powers of two and multiples of 768, the work-groups
  // that fill 256 CUs at three per CU (the FP64 sweeps' occupancy; the FP32
  // ones hold six): a count just above a whole number of such rounds leaves
  // most of the GPU idle for its last work-groups
  static const int targets[] = {1024, 1536, 2048, 2304, 3072, 3840, 4096, 4608, 6144, 8192};
  const R *po = (const R *)phi_old, *co = (const R *)c_old, *Tc = (const R *)T_old;
  // kinds: 0 sweep A, 1 fused sweep B, 2 phase + T of the split form,
  // 3 solute kernel of the split form
  float best[4] = {1e30f, 1e30f, 1e30f, 1e30f};
  int btile[4] = {S.tile_a, S.tile_b, S.tile_b, S.tile_s};
  int btgt[4] = {S.tgt_a, S.tgt_b, S.tgt_b, S.tgt_s};
  // average ms of `nrep` launches of one candidate after one warm-up launch
  auto time_cand = [&](int kind, int tile, int target, int nrep, float &ms) -> hipError_t {
    const int seg = pick_seg(nx, ny, kind ? nz - 2 : nz - 4, tile, target);
    const Geo g = kind ? make_geo(nx, ny, nz, 2, nz - 1, seg, tile)
                       : make_geo(nx, ny, nz, 3, nz - 2, seg, tile);
    hipError_t e = hipSuccess;
    for (int r = 0; r <= nrep && e == hipSuccess; ++r) {  // r = 0 is the warm-up
      if (r == 1 && (e = hipEventRecord(e0, st)) != hipSuccess) break;
      if (kind == 3)
        e = launch_solute_split<R>(tile, g, k, po, co, out + n, nullptr, st);
      else
        e = launch_march<R>(kind == 0 ? kSweepA : kind == 1 ? kSweepB : kSweepPT, tile, g, k,
                            po, co, out, out + n, Tc, out + 2 * n, nullptr, 1, nullptr, st);
    }
    ms = 0.f;
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    ms /= (float)nrep;
    return e;
  };
  struct Cand { int kind, tile, target; float ms; };
  std::vector<Cand> all;
  int cand = 0;
  for (int kind = 0; kind < 4 && rc == PF_OK; ++kind) {
    for (int tile = 0; tile < PF_NVARIANTS && rc == PF_OK; ++tile) {
      for (int target : targets) {
        float ms = 0.f;
        if (time_cand(kind, tile, target, reps, ms) != hipSuccess) { rc = PF_E_HIP; break; }
        if (res && cand < PF_AUTOTUNE_MAX) {
          res->cand[cand].sweep = kind;
          res->cand[cand].tile = tile;
          res->cand[cand].target = target;
          res->cand[cand].ms = ms;
        }
        ++cand;
        all.push_back({kind, tile, target, ms});
      }
    }
  }
  // the runners-up are often within timing noise of the first: time the
  // three fastest of each kind again, three times as long, and keep the two
  // fastest of those
  std::vector<Cand> top2[4];
  for (int kind = 0; kind < 4 && rc == PF_OK; ++kind) {
    std::vector<Cand> top;
    for (const Cand &c : all)
      if (c.kind == kind) top.push_back(c);
    std::sort(top.begin(), top.end(), [](const Cand &a, const Cand &b) { return a.ms < b.ms; });
    if (top.size() > 3) top.resize(3);
    for (Cand &c : top) {
      if (time_cand(kind, c.tile, c.target, 3 * reps, c.ms) != hipSuccess) { rc = PF_E_HIP; break; }
      if (c.ms < best[kind]) {
        best[kind] = c.ms;
        btile[kind] = c.tile;
        btgt[kind] = c.target;
      }
    }
    std::sort(top.begin(), top.end(), [](const Cand &a, const Cand &b) { return a.ms < b.ms; });
    if (top.size() > 2) top.resize(2);
    top2[kind] = top;
  }
  // Kernels timed one at a time do not see how they run one after another
  // in a step (the previous kernel's tail, what it left in the caches):
  // the finalists are compared as whole sequences -- sweep A, then either
  // the split pair (phase + T, solute) or the fused sweep B, reading and
  // writing the scratch fields as a step reads and writes its fields -- and
  // the fastest sequence is kept.
  auto geo_of = [&](int kind, int tile, int target) {
    const int seg = pick_seg(nx, ny, kind ? nz - 2 : nz - 4, tile, target);
    return kind ? make_geo(nx, ny, nz, 2, nz - 1, seg, tile) : make_geo(nx, ny, nz, 3, nz - 2, seg, tile);
  };
  auto time_seq = [&](const Cand &a, const Cand *pt, const Cand *sl, const Cand *fb, int nrep,
                      float &ms) -> hipError_t {
    const Geo ga = geo_of(0, a.tile, a.target);
    const Geo gb = pt ? geo_of(2, pt->tile, pt->target) : geo_of(1, fb->tile, fb->target);
    const Geo gs = sl ? geo_of(3, sl->tile, sl->target) : gb;
    R *phiH = out, *cN = out + n, *tN = out + 2 * n, *phiN = out + 3 * n;
    hipError_t e = hipSuccess;
    for (int r = 0; r <= nrep && e == hipSuccess; ++r) {  // r = 0 is the warm-up
      if (r == 1 && (e = hipEventRecord(e0, st)) != hipSuccess) break;
      e = launch_march<R>(kSweepA, a.tile, ga, k, po, co, phiH, nullptr, Tc, nullptr, nullptr, 1,
                          nullptr, st);
      if (e != hipSuccess) break;
      if (pt) {
        e = launch_march<R>(kSweepPT, pt->tile, gb, k, phiH, co, phiN, nullptr, Tc, tN, nullptr, 1,
                            nullptr, st);
        if (e == hipSuccess) e = launch_solute_split<R>(sl->tile, gs, k, phiH, co, cN, nullptr, st);
      } else {
        e = launch_march<R>(kSweepB, fb->tile, gb, k, phiH, co, phiN, cN, Tc, tN, nullptr, 1,
                            nullptr, st);
      }
    }
    ms = 0.f;
    if (e == hipSuccess) e = hipEventRecord(e1, st);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    ms /= (float)nrep;
    return e;
  };
  // Each sequence is timed in two rounds over all of them and keeps its
  // faster time, so that a drift of the clocks during the comparison does
  // not decide between sequences within a percent of each other.
  struct Seq {
    const Cand *a, *pt, *sl, *fb;
    float ms;
  };
  std::vector<Seq> seqs;
  for (const Cand &a : top2[0]) {
    for (const Cand &pt : top2[2])
      for (const Cand &sl : top2[3]) seqs.push_back({&a, &pt, &sl, nullptr, 1e30f});
    for (const Cand &fb : top2[1]) seqs.push_back({&a, nullptr, nullptr, &fb, 1e30f});
  }
  for (int round = 0; round < 2 && rc == PF_OK; ++round)
    for (Seq &q : seqs) {
      float ms = 0.f;
      if (time_seq(*q.a, q.pt, q.sl, q.fb, 3 * reps, ms) != hipSuccess) {
        rc = PF_E_HIP;
        break;
      }
      q.ms = std::min(q.ms, ms);
    }
  float best_seq = 1e30f;
  for (const Seq &q : seqs) {
    if (rc != PF_OK || !(q.ms < best_seq)) continue;
    best_seq = q.ms;
    S.split_b = q.pt != nullptr;
    fused_set_kind(S, 0, q.a->tile, q.a->target);
    if (q.pt) {
      fused_set_kind(S, 1, q.pt->tile, q.pt->target);
      fused_set_kind(S, 2, q.sl->tile, q.sl->target);
    } else {
      fused_set_kind(S, 1, q.fb->tile, q.fb->target);
    }
  }
  if (res) {
    res->n = cand < PF_AUTOTUNE_MAX ? cand : PF_AUTOTUNE_MAX;
    for (int q = 0; q < 4; ++q) res->best_ms[q] = best[q];
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (hipStreamSynchronize(st) != hipSuccess && rc == PF_OK) rc = PF_E_HIP;
  (void)hipFree(out);
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

// ---- solute-mass ledger (pf_mass.h) -------------------------------------------

namespace {

template <typename R>
__global__ void __launch_bounds__(256) k_mass_interior(const R *__restrict__ c, int nx, int ny,
                                                       int nz, double coff,
                                                       double *__restrict__ part) {
  __shared__ double sh[256];
  const double t = block_sum256(mass_rows(c, nx, ny, nz, (int)blockIdx.x, coff), sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// mode 0: the ledger starts from the interior sum of the partials (acc = 0);
// mode 1: S := interior sum, f := ledger / S.
__global__ void __launch_bounds__(256) k_mass_update(const double *__restrict__ part,
                                                     MassLedger *__restrict__ m, int mode) {
  __shared__ double sh[256];
  const double S = mass_fold(part, sh);
  if (threadIdx.x == 0) {
    if (mode == 0) {
      m->L0 = S;
      m->acc = 0ull;
    } else {
      m->S = S;
      m->f = ledger_value(*m) / S;
    }
  }
}

template <typename R>
__global__ void __launch_bounds__(256) k_mass_scale(R *__restrict__ c, int nx, int ny, int nz,
                                                    int it, uint64_t seed, double coff,
                                                    const MassLedger *__restrict__ m) {
  mass_scale_rows(c, nx, ny, nz, it, seed, m->f, (int)blockIdx.x, coff);
}

template <typename R>
hipError_t mass_interior(const FusedState &S, const void *c, double coff, hipStream_t st) {
  hipLaunchKernelGGL(k_mass_interior<R>, dim3(kMassBlocks), dim3(256), 0, st, (const R *)c, S.nx,
                     S.ny, S.nz, coff, S.mpart);
  return hipGetLastError();
}

}  // namespace

int fused_mass_prepare(FusedState &S, const void *c_old, double coff, hipStream_t st) {
  if (S.mass_every <= 0 || !S.enabled) return PF_OK;
  if (!S.mass) {
    if (hipMalloc((void **)&S.mass, sizeof(MassLedger)) != hipSuccess) return PF_E_NOMEM;
    if (hipMalloc((void **)&S.mpart, kMassBlocks * sizeof(double)) != hipSuccess)
      return PF_E_NOMEM;
    if (hipMemsetAsync(S.mass, 0, sizeof(MassLedger), st) != hipSuccess) return PF_E_HIP;
  }
  if (!S.ledger_valid) {
    if (S.host_mass_ok) {
      // the FP64 mass of the fields the context was created from
      MassLedger m{S.host_mass, 0ull, 0.0, 0.0};
      if (hipMemcpyAsync(S.mass, &m, sizeof m, hipMemcpyHostToDevice, st) != hipSuccess ||
          hipStreamSynchronize(st) != hipSuccess)
        return PF_E_HIP;
    } else {
      hipError_t e = S.prec == PF_FP64 ? mass_interior<double>(S, c_old, 0.0, st)
                                       : mass_interior<float>(S, c_old, coff, st);
      if (e != hipSuccess) return PF_E_HIP;
      hipLaunchKernelGGL(k_mass_update, dim3(1), dim3(256), 0, st, S.mpart, S.mass, 0);
      if (hipGetLastError() != hipSuccess) return PF_E_HIP;
    }
    S.ledger_valid = true;
  }
  return PF_OK;
}

int fused_mass_correct(FusedState &S, const KParams &k, void *c, hipStream_t st) {
  if (!S.ledger_valid || !S.mass) return PF_OK;
  const int it = k.it;
  const uint64_t seed = k.seed;
  hipError_t e = S.prec == PF_FP64 ? mass_interior<double>(S, c, 0.0, st)
                                   : mass_interior<float>(S, c, k.c_off, st);
  if (e != hipSuccess) return PF_E_HIP;
  hipLaunchKernelGGL(k_mass_update, dim3(1), dim3(256), 0, st, S.mpart, S.mass, 1);
  if (hipGetLastError() != hipSuccess) return PF_E_HIP;
  if (S.prec == PF_FP64) {
    hipLaunchKernelGGL(k_mass_scale<double>, dim3(kMassBlocks), dim3(256), 0, st, (double *)c,
                       S.nx, S.ny, S.nz, it, seed, 0.0, S.mass);
    e = hipGetLastError();
    if (e == hipSuccess) e = launch_solute_refresh<double>(k, (double *)c, st);
  } else {
    hipLaunchKernelGGL(k_mass_scale<float>, dim3(kMassBlocks), dim3(256), 0, st, (float *)c, S.nx,
                       S.ny, S.nz, it, seed, k.c_off, S.mass);
    e = hipGetLastError();
    if (e == hipSuccess) e = launch_solute_refresh<float>(k, (float *)c, st);
  }
  return e == hipSuccess ? PF_OK : PF_E_HIP;
}

int fused_mass_state(FusedState &S, MassLedger *out, hipStream_t st) {
  *out = MassLedger{0.0, 0ull, 0.0, 0.0};
  if (!S.mass) return PF_OK;
  if (hipMemcpyAsync(out, S.mass, sizeof *out, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return PF_E_HIP;
  return PF_OK;
}

int fused_mass_restore(FusedState &S, const MassLedger &m, hipStream_t st) {
  if (!S.mass) return PF_E_STATE;
  if (hipMemcpyAsync(S.mass, &m, sizeof m, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return PF_E_HIP;
  return PF_OK;
}

int fused_mass_read(FusedState &S, double out[3], hipStream_t st) {
  MassLedger m;
  out[0] = out[1] = out[2] = 0.0;
  if (fused_mass_state(S, &m, st) != PF_OK) return PF_E_HIP;
  if (!S.mass) return PF_OK;
  out[0] = ledger_value(m);
  out[1] = m.S;
  out[2] = m.f;
  return PF_OK;
}

void fused_free(FusedState &s) {
  s.enabled = false;
  if (s.mass) (void)hipFree(s.mass);
  if (s.mpart) (void)hipFree(s.mpart);
  s.mass = nullptr;
  s.mpart = nullptr;
  s.ledger_valid = false;
}
size_t fused_bytes(const FusedState &s) {
  return s.mass ? sizeof(MassLedger) + kMassBlocks * sizeof(double) : 0;
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

// pf_fused.h -- the fused time step (internal to libpfamd).
//
// One reference step with n_varreduras = nv becomes
//   [temperatura]            only on the first step of a pf_run_steps call
//   (nv-1) x sweep A         phase sweep: phi_old, c_old -> phi_new
//   1 x sweep B              last phase sweep + solute + the next step's
//                            temperature (or, on the last step of the call,
//                            DPP so that the reference state is complete)
//   boundary decorations     small kernels over the domain shell
// Buffer roles are swapped in place of the reference's copies.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "pf_common.h"
#include "pf_mass.h"

namespace pf {

// Tile shapes of the sweep kernels: 0 = 64x8, 1 = 32x8, 2 = 32x16, 3 = 16x16
// cells.  (Capping registers for a third FP64 wave per SIMD spills and runs
// slower.)
constexpr int PF_NVARIANTS = 4;

struct FusedState {
  int nx = 0, ny = 0, nz = 0, prec = PF_FP64;
  int tile_a = 0, tile_b = 0;  // tile shape of sweep A / sweep B
  int target = 2048;           // work-groups per sweep the z split aims for
  int seg_a = 0, seg_b = 0;    // z-planes per work-group segment
  int tgt_a = 2048, tgt_b = 2048, tgt_s = 2048;  // the targets the segments come from
  // split form of sweep B: phase + T, then the solute in its own kernel
  bool split_b = false;
  int tile_s = 0, seg_s = 0;
  bool enabled = false;
  // solute-mass ledger (pf_mass.h): period of the correction in steps
  // (0 = off), the device ledger, partial interior sums; the FP64 interior
  // mass of the host fields the context was created from (the ledger's
  // start while no step has changed them)
  int mass_every = 0;
  bool ledger_valid = false;
  MassLedger *mass = nullptr;
  double *mpart = nullptr;
  double host_mass = 0.0;
  bool host_mass_ok = false;
  // solute faces of the previous step still to copy (snapshot, new field)
  const void *cf_src = nullptr;
  void *cf_dst = nullptr;
};

// HIP-event pairs recorded around launches on one stream, by kernel kind.
// Only the kinds whose bit is set in `mask` are recorded: each record costs
// the stream a few microseconds, so a run that needs one kernel's duration
// records only that kind.
struct KTimer {
  bool on = false;
  uint32_t mask = 0xffffffffu;
  bool open = false;             // a record's first event is in, its second not
  std::vector<hipEvent_t> pool;  // 2 events per record
  std::vector<int> kind;         // kind of record r (events 2r, 2r+1)
  void begin(int k, hipStream_t s) {
    open = false;
    if (!on || !((mask >> k) & 1u)) return;
    const size_t r = kind.size();
    while (pool.size() < 2 * (r + 1)) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return;
      pool.push_back(e);
    }
    if (hipEventRecord(pool[2 * r], s) != hipSuccess) {
      on = false;  // timing only: stop timing
      return;
    }
    kind.push_back(k);
    open = true;
  }
  void end(hipStream_t s) {
    if (!on || !open) return;
    open = false;
    if (hipEventRecord(pool[2 * (kind.size() - 1) + 1], s) != hipSuccess) {
      on = false;
      kind.pop_back();
    }
  }
  void release() {
    for (hipEvent_t e : pool) (void)hipEventDestroy(e);
    pool.clear();
    kind.clear();
  }
};

struct FusedStep {
  bool first = false;  // run temperatura (T of this step) before the sweeps
  bool last = false;   // materialise DPP instead of the next step's T
  int nv = 1;          // phase sweeps per step
  // the solute faces must be complete when the step ends (last step of a
  // call, or a step the host or the mass correction reads c after);
  // otherwise they are copied by extra work-groups of the next sweep A
  bool flush = true;
};

int fused_init(FusedState &s, int nx, int ny, int nz, int prec);
// Tile shape per sweep and the work-group count the z split aims for.
int fused_set_tile(FusedState &s, int tile_a, int tile_b, int target_groups);
// Split form of sweep B on/off and the solute kernel's tile shape.
int fused_set_split(FusedState &s, bool on, int tile_s);
// Tile shape and work-group target of one kernel kind: 0 = sweep A, 1 =
// sweep B (fused, or the phase + T kernel of the split form), 2 = the
// solute kernel of the split form.
int fused_set_kind(FusedState &s, int kind, int tile, int target_groups);
int fused_get_kind(const FusedState &s, int kind, int out[3]);  // {tile, target, seg}
// Times every tile shape x z-split candidate of sweep A, the fused sweep B
// and the two kernels of its split form on the given fields (writing into
// scratch buffers of its own) and keeps the fastest; the split form is
// chosen when its two kernels together beat the fused sweep.
constexpr int PF_AUTOTUNE_MAX = 256;
struct AutotuneResult {
  int n = 0;
  struct { int sweep, tile, target; float ms; } cand[PF_AUTOTUNE_MAX];
  float best_ms[4] = {0, 0, 0, 0};  // sweep A, fused B, phase + T, solute
};
int fused_autotune(FusedState &s, const KParams &k, void *phi_old, void *c_old, void *T_old,
                   int reps, hipStream_t stream, AutotuneResult *res);
// Solute-mass ledger: prepare before a run (starts the ledger when it is not
// valid: from host_mass while that is valid, else from the stored field),
// correct the field (every mass_every steps), read {ledger, interior sum at
// the last correction, last factor}, read / restore the device record.
int fused_mass_prepare(FusedState &s, const void *c_old, double c_off, hipStream_t stream);
int fused_mass_correct(FusedState &s, const KParams &k, void *c, hipStream_t stream);
int fused_mass_read(FusedState &s, double out[3], hipStream_t stream);
int fused_mass_state(FusedState &s, MassLedger *out, hipStream_t stream);
int fused_mass_restore(FusedState &s, const MassLedger &m, hipStream_t stream);
void fused_free(FusedState &s);
size_t fused_bytes(const FusedState &s);
bool fused_supported(const FusedState &s, const pf_params &p);

// Advances one step.  The *_old / *_alias arguments are the context's
// old/new bookkeeping of each field pair (see pf_runtime.hip); they are
// updated to the reference's state after the end-of-step copy.  kt are the
// scalars temperatura sees (TP of the previous step), k those of the sweeps.
int fused_step(FusedState &s, const KParams &kt, const KParams &k, const FusedStep &fs,
               void *phi[2], int *phi_old, bool *phi_alias, void *c[2], int *c_old,
               bool *c_alias, void *T[2], int *T_old, bool *T_alias, void *dpp, double *probe,
               hipStream_t stream, KTimer &timer);

}  // namespace pf

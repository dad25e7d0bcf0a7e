/*
 * pf_capi.h -- C-ABI of the MI355X (gfx950) coupled phase-field time step.
 *
 * The reference (ileaof/phase-field-project-3D) runs its hot path as five
 * argument-less Fortran subroutines that talk through the module `variaveis`
 * (source/variaveis.f90:1-50) and keep the seven fields resident on the device
 * with an OpenMP data environment (source/principal.f90:60-64).  This header is
 * what replaces them: every entry point below names the reference subroutine or
 * loop it stands in for.  A Fortran caller binds it with ISO_C_BINDING (see the
 * module pf_capi in phase-field-project-3d_amd/fortran/pf_capi.f90); plain
 * pointers and sizes only, no C++ or torch types.
 *
 * Arrays are Fortran column-major (1:nx,1:ny,1:nz): the C offset of cell
 * (i,j,k) is (i-1) + (j-1)*nx + (k-1)*nx*ny.  Host arrays are always FP64;
 * the FP32-mixed form converts at upload/download.
 *
 * Conventions
 *   - Every call returns PF_OK (0) or a negative PF_E_* code; pf_last_error()
 *     gives a message for the calling thread's last failure.
 *   - One host thread per context; a context is not thread-safe.
 *   - All device work of a context goes to one HIP stream, in call order.
 *     Calls return without waiting, except the ones documented as synchronous
 *     (pf_diagnosticos, pf_download*, pf_probe, pf_finalize, pf_synchronize).
 */
#ifndef PF_CAPI_H
#define PF_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define PF_OK 0
#define PF_E_ARG (-1)   /* bad argument (null pointer, grid too small, ...)  */
#define PF_E_HIP (-2)   /* a HIP runtime call failed                          */
#define PF_E_NOMEM (-3) /* device allocation failed / over the HBM budget     */
#define PF_E_STATE (-4) /* call not valid in the context's current state      */
#define PF_E_NODEV (-5) /* no usable gfx950 device                            */

/* ---- precision forms --------------------------------------------------- */
#define PF_FP64 0       /* FP64 storage and arithmetic (the drop-in form)     */
#define PF_FP32_MIXED 1 /* FP32 fields + stencils; FP64 reductions/scalars    */

/* ---- noise forms (source/phase_body.inc:181-190) ------------------------ */
#define PF_RUIDO_CELL_POSITIVE 0 /* ran*AAA*16 phi^2 (1-phi)^2              */
#define PF_RUIDO_CELL_ZEROMEAN 1 /* (2 ran - 1)*AAA*16 phi^2 (1-phi)^2      */
#define PF_RUIDO_SMOOTH 2        /* ruido_suave (trilinear coarse lattice)   */

/*
 * Kernel-visible scalars: the subset of module `variaveis` that the field
 * kernels read (source/variaveis.f90:3-40).  POD, passed by pointer.
 * Ke is REAL(4) in the reference (variaveis.f90:28) and is kept as float.
 */
typedef struct pf_params {
  int32_t nx, ny, nz;   /* grid, each >= 5                                   */
  int32_t it;           /* current step, 1-based (principal.f90:93)          */
  int32_t n_varreduras; /* phase sweeps per step (entrada.f90:75-80)         */
  int32_t h_quintico;   /* 1 quintic h(phi), 0 cubic (phase_body.inc:159)    */
  int32_t ruido_tipo;   /* PF_RUIDO_*                                        */
  int32_t precision;    /* PF_FP64 or PF_FP32_MIXED                          */
  uint64_t seed;        /* noise seed; 0 = the reference's stream bit-exact  */
  double dx, dy, dz;    /* mesh spacing [m]                                  */
  double dt;            /* DELTA_T = dx^2/(7.5 D_liq) (entrada.f90:173)      */
  double D;             /* diffusivity in temperatura; the reference never
                           assigns it, so it is 0 there (variaveis.f90:3)     */
  double delta_H, cp_s; /* latent heat, solid heat capacity                  */
  double TP;            /* bath temperature seen by the kernels              */
  double Tm, Me, Vm, R; /* melting point, liquidus slope, molar vol., gas c. */
  double E0, W0, M0;    /* gradient coeff., barrier height, mobility         */
  double delta_eps;     /* DELTA_EPSLON, anisotropy strength                 */
  double AAA;           /* noise amplitude after the dt rescale              */
  double D_liq, D_sol;  /* solute diffusivities                              */
  double Cs_eq, Cl_eq;  /* informational: recomputed from TP by the phase call */
  float Ke;             /* partition coefficient, REAL(4)                    */
  int32_t reserved;     /* keep 8-byte size alignment; set to 0              */
} pf_params;

/*
 * Host loop state that principal.f90 keeps between kernel calls
 * (principal.f90:93-259): what pf_run_steps needs to advance TP, M0 and the
 * REAL(4) time accumulator on its own.
 */
typedef struct pf_loop {
  double T0;         /* initial temperature (entrada.dat line 5)             */
  double taxa_resfr; /* cooling rate [K/s]; 0 keeps TP fixed                 */
  float Sigma;       /* interface energy, REAL(4)                            */
  float Beta;        /* kinetic coefficient, REAL(4)                         */
  float ajuste;      /* mobility/E0 fit factor, REAL(4)                      */
  float time;        /* simulated time, REAL(4) accumulator (principal.f90:96) */
} pf_loop;

typedef struct pf_ctx pf_ctx; /* opaque: device fields, stream, scratch      */

/* ---- library ------------------------------------------------------------ */
const char *pf_version(void);
const char *pf_last_error(void);
/* Number of HIP devices visible (0 on a host without a GPU). */
int pf_device_count(int32_t *n);
/* The HIP device a context was created on (the current device at pf_init),
 * for the driver's timing block (principal.f90:380-383 prints the offload
 * device id). */
int pf_ctx_device(const pf_ctx *ctx, int32_t *dev);

/* ---- host scalar physics (bit-exact with the reference's host code) ------- */
/* source/variaveis.f90:73-92 (+ a seed term that is 0 for seed=0). */
double pf_pseudo_random(int32_t i, int32_t j, int32_t k, int32_t it, int32_t salt,
                        uint64_t seed);
/* source/variaveis.f90:98-111 */
double pf_ruido_suave(int32_t i, int32_t j, int32_t k, int32_t it, uint64_t seed);
/* source/W0_E0.f90:1-35: the REAL(4)-stepped search; returns the number of
 * loop increments taken (0-based index of the accepted try) or -1 when the
 * loop ran out without meeting the tolerance.  E0/W0 hold the last values. */
int32_t pf_w0_e0(double lambda, float ajuste, float sigma, double *W0, double *E0);
/* source/W0_E0.f90:19: the value the reference prints for every try,
 * |E0 - EPSLON_Try| (Inf for the first try, whose W0 is 0).  Writes the
 * first min(cap, tries) values to diff and returns the number of tries up to
 * and including the accepted one (pf_w0_e0's index + 1), or -1 as pf_w0_e0. */
int32_t pf_w0_e0_tries(double lambda, float ajuste, float sigma, double *diff, int32_t cap);
/* source/auxiliar.f90:1-49: thin-interface mobility M0(TP). */
double pf_auxiliar_m0(double TP, double Cl_eq, double Cs_eq, double R, double Vm,
                      double E0, double W0, float Sigma, float Beta, float Ke,
                      double Me, double D_liq, float ajuste);
/* source/equacao_campo_fase.f90:21-22 */
void pf_equilibrium(double TP, double Tm, float Ke, double Me, double *Cs_eq,
                    double *Cl_eq);
/* source/entrada.f90:173: DELTA_T = dx^2/(7.5 D_liq) (the deck's dt is unused). */
double pf_time_step(double dx, double D_liq);
/* source/entrada.f90:179-185: AAA rescaled to the 2-D code's dt, linearly for
 * ruido_tipo 0 and by sqrt otherwise. */
double pf_noise_scale(double AAA, double dt, double dx, double D_liq,
                      int32_t ruido_tipo);
/* source/entrada.f90:226-262: initial fields on the host.  tipo 0 planar
 * front j<=4, 1 sphere at (nx/2,ny/2,nz/2), 2 sphere at (3,3,3), 3 planar
 * base k<=4; new modes: 5 sphere at the bottom centre (nx/2,ny/2,1), 6
 * n_seeds spheres centred on k=1 at SplitMix64(seed_rng) positions in
 * [5,n-4].  Solid (dist^2 <= R^2): phi=1, c=Ke*C0; liquid: phi=0, c=C0;
 * phi_new=0, c_new=C0, T=T0, DPP=0. */
int pf_initial_fields(int32_t nx, int32_t ny, int32_t nz, int32_t tipo,
                      double R_semente, double T0, double C0, float Ke,
                      int32_t n_seeds, uint64_t seed_rng, double *phi_old,
                      double *phi_new, double *c_old, double *c_new,
                      double *T_old, double *T_new, double *dpp);

/* ---- context: source/principal.f90:60-64 (target enter data) ------------ */
/* Allocates the seven fields on the current device and uploads them.  Any
 * host pointer may be NULL, which leaves that device field zero-filled. */
int pf_init(const pf_params *p, const double *phi_old, const double *phi_new,
            const double *c_old, const double *c_new, const double *T_old,
            const double *T_new, const double *dpp, pf_ctx **out);
/* source/principal.f90:268-272 (target exit data). Synchronous: waits for
 * the context's queued work, then frees it either way; returns PF_E_HIP
 * (with the message) if that work failed. */
int pf_finalize(pf_ctx *ctx);

/* ---- the drop-in kernel calls (one per reference subroutine) ------------- */
/* source/temperatura.f90:1-80 + temp_body.inc. */
int pf_temperatura(pf_ctx *ctx, const pf_params *p);
/* source/equacao_campo_fase.f90:1-179 + phase_body.inc: one phase sweep
 * (interior update, clamp, DPP, the 12-plane boundary decoration).  If probe
 * is non-NULL the call is synchronous and fills probe[0..2] with the values
 * the reference prints (equacao_campo_fase.f90:174): phi_new at
 * (nx/2,10,max(nz/2,1)), TP, phi_new at (nx/2,min(100,ny-1),max(nz/2,1)). */
int pf_equacao_campo_fase(pf_ctx *ctx, const pf_params *p, double probe[3]);
/* source/principal.f90:127-150: phase_field_old <- phase_field_new between
 * sweeps (a buffer swap, no copy). */
int pf_phase_new_to_old(pf_ctx *ctx);
/* source/concentracao_soluto.f90:1-130 + conc_body.inc. */
int pf_concentracao_soluto(pf_ctx *ctx, const pf_params *p);
/* source/principal.f90:200-226: old <- new for phi, T and c (buffer swaps). */
int pf_end_step(pf_ctx *ctx);
/* source/diagnosticos.f90:1-58 + diag_body.inc over phase_field_new and
 * conc_soluto_new: out = {L2(phi), max phi, min phi, sum c, free energy}.
 * Deterministic fixed-order FP64 reduction.  Synchronous. */
int pf_diagnosticos(pf_ctx *ctx, const pf_params *p, double out[5]);
/* No counterpart in the reference: pf_diagnosticos for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving
 * loops).  The fields are the caller's device arrays of the stored type
 * (double for PF_FP64, float for PF_FP32_MIXED; p->precision says which),
 * out is 5 doubles of device memory ({L2, max, min, sum c, free energy} as
 * pf_diagnosticos), the work goes to `stream` (a hipStream_t; NULL = the
 * default stream) and the call returns without waiting.  No allocation:
 * the partial sums live in a fixed device array of the library (2048 x 5
 * doubles per device, held until the process ends), shared by every call
 * on that device -- calls that overlap on one device race.  Grids need
 * extents >= 3 and ny*nz < 2^31; same kernels, so the same bits as
 * pf_diagnosticos. */
int pf_diagnosticos_noalloc(const pf_params *p, const void *phi_new, const void *c_new,
                            double *out, void *stream);
/* source/principal.f90:182,248 (target update from).  Any pointer may be
 * NULL.  Synchronous. */
int pf_download(pf_ctx *ctx, double *phi_new, double *c_new, double *T_new);
/* All seven fields in the reference's naming (for checkpoints and tests). */
int pf_download_all(pf_ctx *ctx, double *phi_old, double *phi_new, double *c_old,
                    double *c_new, double *T_old, double *T_new, double *dpp);

/* ---- exact checkpoint ------------------------------------------------------
 * No counterpart in the reference (its only restart is the 8-digit Tecplot
 * snapshot read by tipo_inicial = 4, which the driver also supports).  The
 * file holds the seven fields in their stored precision, the scalars p and
 * L, the tiling and the solute-mass ledger, so that loading it and running
 * on gives the same bits as not having stopped.  Both synchronous. */
int pf_checkpoint_save(pf_ctx *ctx, const pf_params *p, const pf_loop *L, const char *path);
/* Creates a context on the current device from a checkpoint; fills *p, *L. */
int pf_checkpoint_load(const char *path, pf_params *p, pf_loop *L, pf_ctx **out);

/* ---- fused fast path ------------------------------------------------------ */
/* Runs nsteps iterations of the reference's time loop body
 * (principal.f90:96-226 without the output event): temperatura, the TP
 * update, n_varreduras phase sweeps, concentracao_soluto, auxiliar and the
 * end-of-step swap.  p->it, p->TP, p->M0, p->Cs_eq, p->Cl_eq and L->time are
 * advanced in place to the values the reference holds after the last step.
 * The sweeps are fused (phase sweep n_varreduras + solute + the next step's
 * temperature in one pass); the state left behind is the reference's.
 * If probes is non-NULL it receives 3 doubles per phase sweep (the values
 * pf_equacao_campo_fase reports), n_varreduras*nsteps*3 in total, and the
 * call is synchronous. */
int pf_run_steps(pf_ctx *ctx, pf_params *p, pf_loop *L, int32_t nsteps,
                 double *probes);
/* on = 0 makes pf_run_steps issue the drop-in kernels one reference call at
 * a time instead of the fused sweeps (default 1). */
int pf_set_fused(pf_ctx *ctx, int32_t on);
/* Work-group tiling of the fused sweeps (no counterpart in the reference,
 * whose loops are collapsed over the whole grid, equacao_campo_fase.f90:27).
 * Tile shapes: 0 = 64x8, 1 = 32x8, 2 = 32x16, 3 = 16x16 cells per work-group; the
 * z-march is split into segments so that a sweep launches about
 * target_groups work-groups.  The tiling never changes results: every cell
 * is computed by the same arithmetic in any tiling. */
int pf_set_tiling(pf_ctx *ctx, int32_t tile_a, int32_t tile_b, int32_t target_groups);
/* Sweep B as one fused kernel (on = 0) or split into phase + T and a
 * solute-only kernel (on = 1, the default until pf_autotune measures) whose
 * tile shape is tile_s.  Both forms give the same bits. */
int pf_set_sweep_split(pf_ctx *ctx, int32_t on, int32_t tile_s);
/* out = {split on, tile shape of the solute kernel, its z-planes per segment}. */
int pf_get_sweep_split(const pf_ctx *ctx, int32_t out[3]);
/* Tile shape and work-group target of one kernel kind (0 = sweep A, 1 =
 * sweep B or the phase + T kernel of its split form, 2 = the solute kernel
 * of the split form), as pf_autotune sets them per kind; get returns
 * {tile, target_groups, z-planes per segment}. */
int pf_set_sweep_tiling(pf_ctx *ctx, int32_t kind, int32_t tile, int32_t target_groups);
int pf_get_sweep_tiling(const pf_ctx *ctx, int32_t kind, int32_t out[3]);
/* out = {tile of sweep A, tile of sweep B, z-planes per segment A, B}. */
int pf_get_tiling(const pf_ctx *ctx, int32_t out[4]);
/* Times each tile shape x segment split of sweep A, of the fused sweep B and
 * of the two kernels of its split form (phase + T, then the solute alone) on
 * the context's current fields (reps launches each after one warm-up,
 * writing into scratch buffers it allocates and frees: the state is
 * untouched), keeps the fastest per kernel and uses the split form of sweep B
 * when its two kernels together are faster.  If ms is non-NULL it receives 4
 * doubles per candidate {kind (0 = A, 1 = fused B, 2 = phase + T, 3 =
 * solute), tile, target_groups, average ms}, at most 256 candidates, their
 * number in *ncand.  Synchronous. */
int pf_autotune(pf_ctx *ctx, const pf_params *p, int32_t reps, int32_t *ncand, double *ms);
/* Solute-mass correction of the fused path (no counterpart in the
 * reference, which is FP64 throughout).  The context keeps an FP64 ledger of
 * the interior ([2,n-1]^3) solute mass: the FP64 mass when it started plus
 * the sum of every increment the solute update computes, each rounded to
 * 2^-44 and summed as a 64-bit integer (exact, so the sum does not depend on
 * the tiling, on how a run is split into calls or on the batch it runs in).
 * It starts from the FP64 interior sum of the c_old array given to pf_init
 * while no step has run, else from the stored field.  Every `every` steps
 * (p->it a multiple of it) the FP64 interior sum of the stored field is
 * compared with the ledger and the interior is rescaled by ledger/sum (the
 * FP64 product rounded stochastically to FP32, unbiased, drawn from a
 * counter hash of cell and step), removing the drift that FP32 storage
 * rounding accumulates (and the FP32 rounding of the initial field).  The
 * ledger follows the scheme's own increments: the reference's solute update
 * is not conservative by itself (the face mobility takes the centre cell's
 * D), and that change is kept.  Default: 100 for PF_FP32_MIXED, 0 (off) for
 * PF_FP64.  Setting it restarts the ledger at the next pf_run_steps. */
int pf_set_mass_correction(pf_ctx *ctx, int32_t every);
/* Synchronous: out = {ledger, interior sum at the last correction, last
 * factor ledger/sum}; zeros before the first run with the correction on. */
int pf_mass_ledger(pf_ctx *ctx, double out[3]);
/* Wait for all queued work of the context. */
int pf_synchronize(pf_ctx *ctx);

/* ---- batched form ----------------------------------------------------------
 * Many independent instances (each a pf_init-style problem with its own grid,
 * precision, deck scalars and step count) advanced together: one launch per
 * sweep, precision and tile shape covers its instances (the FP64 and the
 * FP32-mixed instances' launches on two streams of the batch, side by side),
 * the loop scalars of principal.f90:96-157 (it, REAL(4) time, TP, Cs/Cl,
 * auxiliar M0) advance on the device per instance, and K steps are captured
 * once in a hipGraph and replayed.  No counterpart
 * in the reference, which runs one problem per process.  An instance's
 * result does not depend on the other instances or on the tiling: it equals
 * pf_run_steps on the same instance (tested), except that DPP is not kept --
 * including the solute-mass correction, which an instance keeps as
 * pf_run_steps does (default every 100 steps for PF_FP32_MIXED, off for
 * PF_FP64; the ledger starts from the FP64 c given to pf_batch_add).
 * Instances of one batch share n_varreduras; every extent must be >= 6. */
typedef struct pf_batch pf_batch;
/* hbm_budget: bytes of instance fields the batch may hold (0 = 90% of the
 * device's free memory now).  Instances use 6 fields (FP64: 48 B/cell,
 * FP32-mixed: 24 B/cell). */
int pf_batch_create(size_t hbm_budget, pf_batch **out);
/* Waits for the batch's queued work, then frees it either way; returns
 * PF_E_HIP (with the message) if that work failed. */
int pf_batch_destroy(pf_batch *b);
/* Admits an instance that will run nsteps steps from the given state
 * (phi_old, c_old, T_old as for pf_init; DPP = 0); runs its first
 * temperatura.  PF_E_NOMEM when it does not fit the budget.  *slot is the
 * instance's handle until pf_batch_remove. */
int pf_batch_add(pf_batch *b, const pf_params *p, const pf_loop *L, int32_t nsteps,
                 const double *phi_old, const double *c_old, const double *T_old,
                 int32_t *slot);
/* Frees an instance (synchronous). */
int pf_batch_remove(pf_batch *b, int32_t slot);
/* Enqueues nsteps steps of every instance that has steps left (each stops at
 * its own count) and returns without waiting. */
int pf_batch_run(pf_batch *b, int32_t nsteps);
/* Steps per captured graph (default 16; 0 = plain launches). */
int pf_batch_set_graph_steps(pf_batch *b, int32_t k);
int pf_batch_synchronize(pf_batch *b);
/* Host-side step accounting of an instance (no device wait). */
int pf_batch_status(pf_batch *b, int32_t slot, int32_t *steps_done, int32_t *steps_left);
/* Synchronous: the instance's scalars as pf_run_steps leaves them (p->it,
 * TP, M0, Cs_eq, Cl_eq; L->time). */
int pf_batch_scalars(pf_batch *b, int32_t slot, pf_params *p, pf_loop *L);
/* Synchronous: phi_new, c_new, T_new of the last completed step (any may be
 * NULL). */
int pf_batch_download(pf_batch *b, int32_t slot, double *phi, double *c, double *T);
/* Synchronous: diagnosticos of the instance (as pf_diagnosticos). */
int pf_batch_diagnosticos(pf_batch *b, int32_t slot, double out[5]);
int pf_batch_bytes(const pf_batch *b, size_t *used, size_t *budget);
/* Size classes: precision x plane-area bucket (<= 48^2, <= 96^2, <= 160^2,
 * larger); each class has a tile shape (0 = 64x8, 1 = 32x8, 2 = 32x16,
 * 3 = 16x16) per kernel kind: sweep A, sweep B (the split form's phase + T
 * kernel) and the split form's solute kernel (default 16x16 for each).
 * set_tiling takes one shape for every kind or PF_BATCH_TILES(a, b, s);
 * get_tiling returns the one shape when the three agree, else the packed
 * form. */
#define PF_BATCH_TILES(a, b, s) (0x1000 | (a) | ((b) << 4) | ((s) << 8))
int32_t pf_batch_size_class(int32_t precision, int32_t nx, int32_t ny);
int pf_batch_set_tiling(pf_batch *b, int32_t size_class, int32_t shape);
int pf_batch_get_tiling(const pf_batch *b, int32_t size_class, int32_t *shape);
/* Work-groups per sweep the z-march split aims for over the whole batch
 * (default 7168; segments of 8 to 16 planes): each instance gets its
 * share by cell count. */
int pf_batch_set_group_target(pf_batch *b, int32_t target_groups);
/* Times every tile shape for size_class (-1: each class present in turn)
 * and kernel kind, over the whole batch's launches of that kind, and keeps
 * the fastest per kind (no step is taken; the T_new reported for the
 * instances is not kept until their next step).  Synchronous. */
int pf_batch_autotune(pf_batch *b, int32_t size_class, int32_t reps);
/* Sweep B of the batched steps as one fused kernel (on = 0) or split into
 * a phase + T kernel and a solute kernel over the same tiles (on = 1, the
 * default), as pf_set_sweep_split; both forms give the same bits. */
int pf_batch_set_sweep_split(pf_batch *b, int32_t on);
/* The instance's solute-mass correction period (as pf_set_mass_correction;
 * 0 = off); only before its first step (PF_E_STATE after). */
int pf_batch_set_mass_correction(pf_batch *b, int32_t slot, int32_t every);
/* Synchronous: out = {ledger, interior sum at the last correction, last
 * factor ledger/sum} of the instance (as pf_mass_ledger). */
int pf_batch_mass_ledger(pf_batch *b, int32_t slot, double out[3]);

/* ---- measurement ----------------------------------------------------------
 * Per-kernel durations from HIP events recorded around each launch on the
 * context's stream (no counterpart in the reference, whose only timer is the
 * host wall clock at output events, principal.f90:308-408). */
#define PF_KT_SWEEP_A 0     /* fused phase sweep (all but the last sweep)   */
#define PF_KT_SWEEP_B 1     /* fused last sweep + (solute +) next temperature */
#define PF_KT_BOUNDARY 2    /* boundary decorations                         */
#define PF_KT_TEMPERATURA 3 /* standalone temperatura                       */
#define PF_KT_DROPIN 4      /* drop-in phase / solute kernels               */
#define PF_KT_MASS 5        /* solute-mass correction (FP32-mixed form)     */
#define PF_KT_SOLUTE 6      /* solute kernel of the split form of sweep B   */
#define PF_KT_COUNT 7
/* STREAM-style copy of `bytes` bytes (16-byte elements, coalesced, one read
 * and one write per element) repeated `iters` times on the current device,
 * in three forms (1, 4 and 8 loads in flight per thread); reports the best
 * copy bandwidth (read + write bytes / average kernel time).  Used as the
 * measured HBM peak and to calibrate HBM byte counters. */
int pf_stream_copy(size_t bytes, int32_t iters, double *gbs);
/* on != 0 starts recording (and clears earlier records); 0 stops. */
int pf_timing_enable(pf_ctx *ctx, int32_t on);
/* Records only launches of the kinds whose bit (1u << PF_KT_*) is set in
 * `kinds` (all kinds after pf_init).  Each record is a pair of HIP events
 * on the context's stream, which costs the stream a few microseconds per
 * launch: a run timed by the wall clock records only the kind whose
 * duration it needs.  No counterpart in the reference (a measurement aid). */
int pf_timing_kinds(pf_ctx *ctx, uint32_t kinds);
/* Synchronous: sums of durations [ms] and launch counts per kind since the
 * last enable; clears the records. */
int pf_timing_read(pf_ctx *ctx, double ms[PF_KT_COUNT], int64_t count[PF_KT_COUNT]);
/* Bytes of device memory the context holds. */
int pf_device_bytes(const pf_ctx *ctx, size_t *bytes);

#ifdef __cplusplus
}
#endif
#endif /* PF_CAPI_H */

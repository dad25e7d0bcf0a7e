// pf_runtime.hip -- the C-ABI context: device-resident fields, buffer swaps
// in place of the reference's whole-array copies, the drop-in calls and the
// fused time loop.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pf_capi.h"
#include "pf_common.h"
#include "pf_fused.h"
#include "pf_kernels.h"

static __device__ __align__(256) double scratch[2048 * 5];

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIP_TRY(x)                                                                  \
  do {                                                                              \
    hipError_t e_ = (x);                                                            \
    if (e_ != hipSuccess) return fail(PF_E_HIP, "%s: %s", #x, hipGetErrorString(e_)); \
  } while (0)

}  // namespace

int pf::fail_msg(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

// One field pair (old/new) with swap-instead-of-copy bookkeeping: `old_idx`
// names the buffer that holds the reference's *_old array; when `alias` is
// set the reference's *_new array equals *_old (right after an old<-new copy)
// and the other buffer holds stale data that the next kernel overwrites.
struct pf_pair {
  void *buf[2] = {nullptr, nullptr};
  int old_idx = 0;
  bool alias = false;
  void *old_() const { return buf[old_idx]; }
  void *new_() const { return alias ? buf[old_idx] : buf[1 - old_idx]; }
  void *new_w() { alias = false; return buf[1 - old_idx]; }  // for writing
  void swap() {  // old <- new
    if (!alias) old_idx = 1 - old_idx;
    alias = true;
  }
};

struct pf_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int nx = 0, ny = 0, nz = 0, prec = PF_FP64;
  size_t n = 0, esz = 8;
  pf_pair phi, c, T;
  void *dpp = nullptr;
  double *diag_part = nullptr, *diag_out = nullptr;
  int diag_nblk = 0;
  double *pinned = nullptr;  // 8 doubles of host-visible scratch
  double *probe_dev = nullptr;
  size_t probe_cap = 0;
  pf::FusedState fused;
  pf::KTimer timer;
  bool use_fused = true;
  size_t bytes = 0;
  // FP32-mixed: c is stored as c - c_off (the initial c of the top corner
  // cell, a liquid value), which keeps the small solute increments of the
  // near-uniform liquid that an FP32 c would round away; 0 for FP64
  double c_off = 0.0;
};

namespace {

int check_ctx(const pf_ctx *ctx, const pf_params *p) {
  if (!ctx) return fail(PF_E_ARG, "null context");
  if (p && (p->nx != ctx->nx || p->ny != ctx->ny || p->nz != ctx->nz))
    return fail(PF_E_ARG, "params grid %dx%dx%d differs from the context's %dx%dx%d", p->nx, p->ny,
                p->nz, ctx->nx, ctx->ny, ctx->nz);
  if (p && p->precision != ctx->prec)
    return fail(PF_E_ARG, "params precision %d differs from the context's %d", p->precision,
                ctx->prec);
  return PF_OK;
}

pf::KParams kparams(const pf_ctx *ctx, const pf_params &p) {
  pf::KParams k = pf::kparams_from(p);
  k.c_off = ctx->c_off;
  return k;
}

// Host FP64 <-> device field (FP64, or FP32 stored as value - off).
int upload(pf_ctx *ctx, void *dst, const double *src, double off = 0.0) {
  if (!src) {
    HIP_TRY(hipMemsetAsync(dst, 0, ctx->n * ctx->esz, ctx->stream));
    return PF_OK;
  }
  if (ctx->prec == PF_FP64) {
    HIP_TRY(hipMemcpyAsync(dst, src, ctx->n * 8, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PF_OK;
  }
  std::vector<float> tmp(ctx->n);
  for (size_t q = 0; q < ctx->n; ++q) tmp[q] = (float)(src[q] - off);
  HIP_TRY(hipMemcpyAsync(dst, tmp.data(), ctx->n * 4, hipMemcpyHostToDevice, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return PF_OK;
}

int download(pf_ctx *ctx, double *dst, const void *src, double off = 0.0) {
  if (!dst) return PF_OK;
  if (ctx->prec == PF_FP64) {
    HIP_TRY(hipMemcpyAsync(dst, src, ctx->n * 8, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PF_OK;
  }
  std::vector<float> tmp(ctx->n);
  HIP_TRY(hipMemcpyAsync(tmp.data(), src, ctx->n * 4, hipMemcpyDeviceToHost, ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  for (size_t q = 0; q < ctx->n; ++q) dst[q] = (double)tmp[q] + off;
  return PF_OK;
}

int launch_probe(pf_ctx *ctx, const void *field, double *dst) {
  size_t a, b;
  pf::probe_cells(ctx->nx, ctx->ny, ctx->nz, a, b);
  if (ctx->prec == PF_FP64)
    HIP_TRY(pf::launch_probe<double>((const double *)field, a, b, dst, ctx->stream));
  else
    HIP_TRY(pf::launch_probe<float>((const float *)field, a, b, dst, ctx->stream));
  return PF_OK;
}

int ensure_probe_cap(pf_ctx *ctx, size_t nd) {
  if (ctx->probe_cap >= nd) return PF_OK;
  if (ctx->probe_dev) HIP_TRY(hipFree(ctx->probe_dev));
  ctx->probe_dev = nullptr;
  HIP_TRY(hipMalloc(&ctx->probe_dev, nd * sizeof(double)));
  ctx->probe_cap = nd;
  return PF_OK;
}

// ---- one drop-in call each (used by the C-ABI and by the unfused loop) ----
int do_temperatura(pf_ctx *ctx, const pf::KParams &k) {
  hipError_t e;
  ctx->timer.begin(PF_KT_TEMPERATURA, ctx->stream);
  if (ctx->prec == PF_FP64)
    e = pf::launch_temperatura<double>(k, (const double *)ctx->T.old_(),
                                       (const double *)ctx->phi.old_(), (const double *)ctx->dpp,
                                       (double *)ctx->T.new_w(), ctx->stream);
  else
    e = pf::launch_temperatura<float>(k, (const float *)ctx->T.old_(),
                                      (const float *)ctx->phi.old_(), (const float *)ctx->dpp,
                                      (float *)ctx->T.new_w(), ctx->stream);
  ctx->timer.end(ctx->stream);
  if (e != hipSuccess) return fail(PF_E_HIP, "temperatura launch: %s", hipGetErrorString(e));
  return PF_OK;
}

int do_phase(pf_ctx *ctx, const pf::KParams &k) {
  hipError_t e;
  ctx->timer.begin(PF_KT_DROPIN, ctx->stream);
  if (ctx->prec == PF_FP64)
    e = pf::launch_phase<double>(k, (const double *)ctx->phi.old_(), (const double *)ctx->c.old_(),
                                 (double *)ctx->phi.new_w(), (double *)ctx->dpp, ctx->stream);
  else
    e = pf::launch_phase<float>(k, (const float *)ctx->phi.old_(), (const float *)ctx->c.old_(),
                                (float *)ctx->phi.new_w(), (float *)ctx->dpp, ctx->stream);
  ctx->timer.end(ctx->stream);
  if (e != hipSuccess) return fail(PF_E_HIP, "phase launch: %s", hipGetErrorString(e));
  return PF_OK;
}

int do_solute(pf_ctx *ctx, const pf::KParams &k) {
  ctx->fused.ledger_valid = false;  // c changes outside the ledger
  ctx->fused.host_mass_ok = false;
  hipError_t e;
  ctx->timer.begin(PF_KT_DROPIN, ctx->stream);
  if (ctx->prec == PF_FP64)
    e = pf::launch_solute<double>(k, (const double *)ctx->phi.old_(), (const double *)ctx->c.old_(),
                                  (double *)ctx->c.new_w(), ctx->stream);
  else
    e = pf::launch_solute<float>(k, (const float *)ctx->phi.old_(), (const float *)ctx->c.old_(),
                                 (float *)ctx->c.new_w(), ctx->stream);
  ctx->timer.end(ctx->stream);
  if (e != hipSuccess) return fail(PF_E_HIP, "solute launch: %s", hipGetErrorString(e));
  return PF_OK;
}

// STREAM-style copy: each thread moves U 16-byte vectors, all U loads issued
// before the stores, so that U loads per thread are in flight; for each u
// the 256 threads of a block touch 4 KiB contiguously.  One pass, no loop.
template <int U>
__global__ void __launch_bounds__(256) k_stream_copy(const double2 *__restrict__ a,
                                                     double2 *__restrict__ b, size_t n2) {
  const size_t base = (size_t)blockIdx.x * 256 * U + threadIdx.x;
  double2 v[U];
#pragma unroll
  for (int u = 0; u < U; ++u)
    if (base + (size_t)u * 256 < n2) v[u] = a[base + (size_t)u * 256];
#pragma unroll
  for (int u = 0; u < U; ++u)
    if (base + (size_t)u * 256 < n2) b[base + (size_t)u * 256] = v[u];
}

template <int U>
void launch_stream_copy(const double2 *a, double2 *b, size_t n2, hipStream_t s) {
  const size_t nb = (n2 + 256 * U - 1) / (256 * U);
  hipLaunchKernelGGL(k_stream_copy<U>, dim3((unsigned)nb), dim3(256), 0, s, a, b, n2);
}

// ---- diagnostics without allocation ------------------------------------------
// The partial sums of pf_diagnosticos_noalloc live in the fixed device array
// `scratch` (file scope, below the namespace): diag_blocks() blocks x 5
// doubles, at most kDiagScratchBlocks blocks, one copy per device (each
// device holds it until the process ends).  Host code never passes the
// array's name to a kernel: its device address on the current device comes
// from hipGetSymbolAddress, looked up once per device and kept in
// g_scratch_dev.
constexpr int kDiagScratchBlocks = 2048;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_dev[kMaxDevices] = {};

// The device address of `scratch` on device dev (looked up on first use, with
// dev current).
int diag_scratch(int dev, double **out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (dev < 0 || dev >= kMaxDevices) return fail(PF_E_ARG, "device %d out of range", dev);
  if (!g_scratch_dev[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(PF_E_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_scratch_dev[dev] = (double *)ptr;
  }
  *out = g_scratch_dev[dev];
  return PF_OK;
}

}  // namespace

extern "C" {

int pf_diagnosticos_noalloc(const pf_params *p, const void *phi_new, const void *c_new,
                            double *out, void *stream) {
  if (!p || !phi_new || !c_new || !out) return fail(PF_E_ARG, "null argument");
  if (p->nx < 3 || p->ny < 3 || p->nz < 3)
    return fail(PF_E_ARG, "grid %dx%dx%d: each extent must be >= 3", p->nx, p->ny, p->nz);
  if ((long long)p->ny * p->nz > 0x7fffffffLL)
    return fail(PF_E_ARG, "grid %dx%dx%d: ny*nz must be below 2^31", p->nx, p->ny, p->nz);
  if (p->precision != PF_FP64 && p->precision != PF_FP32_MIXED)
    return fail(PF_E_ARG, "unknown precision %d", p->precision);
  const int nblk = pf::diag_blocks(p->ny, p->nz);
  static_assert(sizeof(scratch) == sizeof(double) * 5 * kDiagScratchBlocks, "scratch size");
  if (nblk > kDiagScratchBlocks)
    return fail(PF_E_ARG, "%d partial blocks exceed the fixed scratch of %d", nblk,
                kDiagScratchBlocks);
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  double *part = nullptr;
  int rc = diag_scratch(dev, &part);
  if (rc != PF_OK) return rc;
  const pf::KParams k = pf::kparams_from(*p);  // the caller's fields: full c
  const hipStream_t st = (hipStream_t)stream;
  hipError_t e;
  if (p->precision == PF_FP64)
    e = pf::launch_diag<double>(k, (const double *)phi_new, (const double *)c_new, part, nblk, out,
                                st);
  else
    e = pf::launch_diag<float>(k, (const float *)phi_new, (const float *)c_new, part, nblk, out,
                               st);
  if (e != hipSuccess) return fail(PF_E_HIP, "diag launch: %s", hipGetErrorString(e));
  return PF_OK;
}

int pf_stream_copy(size_t bytes, int32_t iters, double *gbs) {
  if (!gbs || bytes < 1024 || iters < 1) return fail(PF_E_ARG, "bad stream-copy arguments");
  const size_t n2 = bytes / 16;
  double2 *a = nullptr, *b = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = PF_OK;
  if (hipMalloc(&a, n2 * 16) != hipSuccess || hipMalloc(&b, n2 * 16) != hipSuccess) {
    rc = fail(PF_E_NOMEM, "stream-copy buffers");
  } else if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess ||
             hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    rc = fail(PF_E_HIP, "stream-copy stream/events");
  } else {
    // the best of three copy forms: 1, 4 and 8 loads in flight per thread
    bool ok = hipMemsetAsync(a, 0, n2 * 16, s) == hipSuccess;
    double best = 0.0;
    for (int form = 0; form < 3 && ok; ++form) {
      auto launch = [&]() {
        if (form == 0) launch_stream_copy<1>(a, b, n2, s);
        else if (form == 1) launch_stream_copy<4>(a, b, n2, s);
        else launch_stream_copy<8>(a, b, n2, s);
      };
      launch();  // warm-up
      ok = ok && hipEventRecord(e0, s) == hipSuccess;
      for (int q = 0; q < iters; ++q) launch();
      ok = ok && hipGetLastError() == hipSuccess && hipEventRecord(e1, s) == hipSuccess;
      float ms = 0.f;
      ok = ok && hipEventSynchronize(e1) == hipSuccess &&
           hipEventElapsedTime(&ms, e0, e1) == hipSuccess && ms > 0.f;
      if (ok) best = std::max(best, 2.0 * (double)(n2 * 16) * iters / ((double)ms * 1e-3) / 1e9);
    }
    if (!ok) rc = fail(PF_E_HIP, "stream-copy timing");
    else *gbs = best;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return rc;
}

const char *pf_version(void) { return "pfamd 0.1 (gfx950)"; }
const char *pf_last_error(void) { return g_err.c_str(); }

int pf_device_count(int32_t *n) {
  if (!n) return fail(PF_E_ARG, "null output");
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
  *n = c;
  return PF_OK;
}

int pf_ctx_device(const pf_ctx *ctx, int32_t *dev) {
  if (!ctx || !dev) return fail(PF_E_ARG, "null context or output");
  *dev = ctx->device;
  return PF_OK;
}

int pf_init(const pf_params *p, const double *phi_old, const double *phi_new, const double *c_old,
            const double *c_new, const double *T_old, const double *T_new, const double *dpp,
            pf_ctx **out) {
  if (!p || !out) return fail(PF_E_ARG, "null argument");
  *out = nullptr;
  if (p->nx < 5 || p->ny < 5 || p->nz < 5)
    return fail(PF_E_ARG, "grid %dx%dx%d: each extent must be >= 5", p->nx, p->ny, p->nz);
  if (p->precision != PF_FP64 && p->precision != PF_FP32_MIXED)
    return fail(PF_E_ARG, "unknown precision %d", p->precision);
  // the sweeps address a plane of a field with 32-bit byte offsets
  if ((long long)p->nx * p->ny * (p->precision == PF_FP64 ? 8 : 4) >= (1LL << 32))
    return fail(PF_E_ARG, "plane %dx%d too large (a plane of a field must be < 4 GiB)", p->nx, p->ny);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(PF_E_NODEV, "no HIP device visible");
  pf_ctx *ctx = new pf_ctx();
  ctx->nx = p->nx; ctx->ny = p->ny; ctx->nz = p->nz;
  ctx->prec = p->precision;
  ctx->n = (size_t)p->nx * p->ny * p->nz;
  ctx->esz = ctx->prec == PF_FP64 ? 8 : 4;
  int rc = PF_OK;
  auto cleanup = [&](int code) {
    pf_finalize(ctx);
    return code;
  };
  if (hipGetDevice(&ctx->device) != hipSuccess) return cleanup(fail(PF_E_HIP, "hipGetDevice"));
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess)
    return cleanup(fail(PF_E_HIP, "hipStreamCreate"));
  void **bufs[7] = {&ctx->phi.buf[0], &ctx->phi.buf[1], &ctx->c.buf[0], &ctx->c.buf[1],
                    &ctx->T.buf[0], &ctx->T.buf[1], &ctx->dpp};
  const double *src[7] = {phi_old, phi_new, c_old, c_new, T_old, T_new, dpp};
  for (int f = 0; f < 7; ++f) {
    if (hipMalloc(bufs[f], ctx->n * ctx->esz) != hipSuccess) {
      *bufs[f] = nullptr;
      return cleanup(fail(PF_E_NOMEM, "hipMalloc of %zu bytes failed", ctx->n * ctx->esz));
    }
    ctx->bytes += ctx->n * ctx->esz;
  }
  ctx->diag_nblk = pf::diag_blocks(ctx->ny, ctx->nz);
  if (hipMalloc(&ctx->diag_part, sizeof(double) * 5 * (ctx->diag_nblk + 1)) != hipSuccess)
    return cleanup(fail(PF_E_NOMEM, "hipMalloc diag"));
  ctx->diag_out = ctx->diag_part + 5 * ctx->diag_nblk;
  if (hipHostMalloc(&ctx->pinned, 8 * sizeof(double)) != hipSuccess)
    return cleanup(fail(PF_E_NOMEM, "hipHostMalloc"));
  if (ctx->prec == PF_FP32_MIXED && c_old) ctx->c_off = c_old[ctx->n - 1];
  for (int f = 0; f < 7; ++f)
    if ((rc = upload(ctx, *bufs[f], src[f], (f == 2 || f == 3) ? ctx->c_off : 0.0)) != PF_OK)
      return cleanup(rc);
  if ((rc = pf::fused_init(ctx->fused, ctx->nx, ctx->ny, ctx->nz, ctx->prec)) != PF_OK)
    return cleanup(fail(rc, "fused path set-up failed"));
  // the solute-mass ledger starts from the FP64 mass of the caller's field,
  // not from its FP32 rounding
  ctx->fused.host_mass = c_old ? pf::host_interior_mass(c_old, ctx->nx, ctx->ny, ctx->nz) : 0.0;
  ctx->fused.host_mass_ok = true;
  *out = ctx;
  return PF_OK;
}

int pf_finalize(pf_ctx *ctx) {
  if (!ctx) return PF_OK;
  // an error of work still in flight is reported here; the context is freed
  // either way
  const hipError_t e = ctx->stream ? hipStreamSynchronize(ctx->stream) : hipSuccess;
  pf::fused_free(ctx->fused);
  ctx->timer.release();
  void *bufs[7] = {ctx->phi.buf[0], ctx->phi.buf[1], ctx->c.buf[0], ctx->c.buf[1],
                   ctx->T.buf[0], ctx->T.buf[1], ctx->dpp};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  if (ctx->diag_part) (void)hipFree(ctx->diag_part);
  if (ctx->probe_dev) (void)hipFree(ctx->probe_dev);
  if (ctx->pinned) (void)hipHostFree(ctx->pinned);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  if (e != hipSuccess) return fail(PF_E_HIP, "pending work failed: %s", hipGetErrorString(e));
  return PF_OK;
}

int pf_temperatura(pf_ctx *ctx, const pf_params *p) {
  int rc = check_ctx(ctx, p);
  if (rc == PF_OK && !p) rc = fail(PF_E_ARG, "null params");
  if (rc != PF_OK) return rc;
  return do_temperatura(ctx, kparams(ctx, *p));
}

int pf_equacao_campo_fase(pf_ctx *ctx, const pf_params *p, double probe[3]) {
  int rc = check_ctx(ctx, p);
  if (rc == PF_OK && !p) rc = fail(PF_E_ARG, "null params");
  if (rc != PF_OK) return rc;
  if ((rc = do_phase(ctx, kparams(ctx, *p))) != PF_OK) return rc;
  if (probe) {
    if ((rc = ensure_probe_cap(ctx, 2)) != PF_OK) return rc;
    if ((rc = launch_probe(ctx, ctx->phi.new_(), ctx->probe_dev)) != PF_OK) return rc;
    HIP_TRY(hipMemcpyAsync(ctx->pinned, ctx->probe_dev, 2 * sizeof(double), hipMemcpyDeviceToHost,
                           ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    probe[0] = ctx->pinned[0];
    probe[1] = p->TP;
    probe[2] = ctx->pinned[1];
  }
  return PF_OK;
}

int pf_phase_new_to_old(pf_ctx *ctx) {
  int rc = check_ctx(ctx, nullptr);
  if (rc != PF_OK) return rc;
  ctx->phi.swap();
  return PF_OK;
}

int pf_concentracao_soluto(pf_ctx *ctx, const pf_params *p) {
  int rc = check_ctx(ctx, p);
  if (rc == PF_OK && !p) rc = fail(PF_E_ARG, "null params");
  if (rc != PF_OK) return rc;
  return do_solute(ctx, kparams(ctx, *p));
}

int pf_end_step(pf_ctx *ctx) {
  int rc = check_ctx(ctx, nullptr);
  if (rc != PF_OK) return rc;
  ctx->phi.swap();
  ctx->T.swap();
  ctx->c.swap();
  return PF_OK;
}

int pf_diagnosticos(pf_ctx *ctx, const pf_params *p, double out[5]) {
  int rc = check_ctx(ctx, p);
  if (rc == PF_OK && (!p || !out)) rc = fail(PF_E_ARG, "null argument");
  if (rc != PF_OK) return rc;
  const pf::KParams k = kparams(ctx, *p);
  hipError_t e;
  if (ctx->prec == PF_FP64)
    e = pf::launch_diag<double>(k, (const double *)ctx->phi.new_(), (const double *)ctx->c.new_(),
                                ctx->diag_part, ctx->diag_nblk, ctx->diag_out, ctx->stream);
  else
    e = pf::launch_diag<float>(k, (const float *)ctx->phi.new_(), (const float *)ctx->c.new_(),
                               ctx->diag_part, ctx->diag_nblk, ctx->diag_out, ctx->stream);
  if (e != hipSuccess) return fail(PF_E_HIP, "diag launch: %s", hipGetErrorString(e));
  HIP_TRY(hipMemcpyAsync(ctx->pinned, ctx->diag_out, 5 * sizeof(double), hipMemcpyDeviceToHost,
                         ctx->stream));
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  memcpy(out, ctx->pinned, 5 * sizeof(double));
  return PF_OK;
}

int pf_download(pf_ctx *ctx, double *phi_new, double *c_new, double *T_new) {
  int rc = check_ctx(ctx, nullptr);
  if (rc != PF_OK) return rc;
  if ((rc = download(ctx, phi_new, ctx->phi.new_())) != PF_OK) return rc;
  if ((rc = download(ctx, c_new, ctx->c.new_(), ctx->c_off)) != PF_OK) return rc;
  return download(ctx, T_new, ctx->T.new_());
}

int pf_download_all(pf_ctx *ctx, double *phi_old, double *phi_new, double *c_old, double *c_new,
                    double *T_old, double *T_new, double *dpp) {
  int rc = check_ctx(ctx, nullptr);
  if (rc != PF_OK) return rc;
  const void *src[7] = {ctx->phi.old_(), ctx->phi.new_(), ctx->c.old_(), ctx->c.new_(),
                        ctx->T.old_(),   ctx->T.new_(),   ctx->dpp};
  double *dst[7] = {phi_old, phi_new, c_old, c_new, T_old, T_new, dpp};
  for (int f = 0; f < 7; ++f)
    if ((rc = download(ctx, dst[f], src[f], (f == 2 || f == 3) ? ctx->c_off : 0.0)) != PF_OK)
      return rc;
  return PF_OK;
}

// ---- binary checkpoint ------------------------------------------------------
// Layout: CkptHeader, pf_params, pf_loop, then the seven fields in the
// reference's naming (phi_old, phi_new, c_old, c_new, T_old, T_new, DPP), each
// n elements of the context's stored type (raw device bytes: exact for FP32).
namespace {
struct CkptHeader {
  char magic[8];  // "PFCKPT01"
  int32_t version, prec, nx, ny, nz;
  int32_t params_size, loop_size;
  int32_t use_fused, tile_a, tile_b, target;
  int32_t mass_every, ledger_valid, host_mass_ok, pad;
  pf::MassLedger ledger;
  double host_mass;
  double c_off;  // the stored c is c - c_off (FP32-mixed)
};
const char kCkptMagic[8] = {'P', 'F', 'C', 'K', 'P', 'T', '0', '3'};
}  // namespace

int pf_checkpoint_save(pf_ctx *ctx, const pf_params *p, const pf_loop *L, const char *path) {
  int rc = check_ctx(ctx, p);
  if (rc == PF_OK && (!p || !L || !path)) rc = fail(PF_E_ARG, "null argument");
  if (rc != PF_OK) return rc;
  CkptHeader h;
  memset(&h, 0, sizeof h);
  memcpy(h.magic, kCkptMagic, 8);
  h.version = 3;
  h.prec = ctx->prec;
  h.nx = ctx->nx; h.ny = ctx->ny; h.nz = ctx->nz;
  h.params_size = (int32_t)sizeof(pf_params);
  h.loop_size = (int32_t)sizeof(pf_loop);
  h.use_fused = ctx->use_fused ? 1 : 0;
  h.tile_a = ctx->fused.tile_a; h.tile_b = ctx->fused.tile_b; h.target = ctx->fused.target;
  h.mass_every = ctx->fused.mass_every;
  h.ledger_valid = ctx->fused.ledger_valid ? 1 : 0;
  h.host_mass_ok = ctx->fused.host_mass_ok ? 1 : 0;
  h.host_mass = ctx->fused.host_mass;
  h.c_off = ctx->c_off;
  if (h.ledger_valid && pf::fused_mass_state(ctx->fused, &h.ledger, ctx->stream) != PF_OK)
    return fail(PF_E_HIP, "reading the ledger failed");
  const void *src[7] = {ctx->phi.old_(), ctx->phi.new_(), ctx->c.old_(), ctx->c.new_(),
                        ctx->T.old_(),   ctx->T.new_(),   ctx->dpp};
  std::vector<char> buf(ctx->n * ctx->esz);
  FILE *f = fopen(path, "wb");
  if (!f) return fail(PF_E_ARG, "cannot open %s for writing", path);
  bool ok = fwrite(&h, sizeof h, 1, f) == 1 && fwrite(p, sizeof *p, 1, f) == 1 &&
            fwrite(L, sizeof *L, 1, f) == 1;
  for (int q = 0; q < 7 && ok; ++q) {
    if (hipMemcpyAsync(buf.data(), src[q], buf.size(), hipMemcpyDeviceToHost, ctx->stream) !=
            hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess) {
      fclose(f);
      return fail(PF_E_HIP, "checkpoint download failed");
    }
    ok = fwrite(buf.data(), 1, buf.size(), f) == buf.size();
  }
  ok = (fclose(f) == 0) && ok;
  if (!ok) return fail(PF_E_ARG, "writing %s failed", path);
  return PF_OK;
}

int pf_checkpoint_load(const char *path, pf_params *p, pf_loop *L, pf_ctx **out) {
  if (!path || !p || !L || !out) return fail(PF_E_ARG, "null argument");
  *out = nullptr;
  FILE *f = fopen(path, "rb");
  if (!f) return fail(PF_E_ARG, "cannot open %s", path);
  CkptHeader h;
  pf_params pp;
  pf_loop ll;
  bool ok = fread(&h, sizeof h, 1, f) == 1;
  if (ok && (memcmp(h.magic, kCkptMagic, 8) != 0 || h.version != 3 ||
             h.params_size != (int32_t)sizeof(pf_params) || h.loop_size != (int32_t)sizeof(pf_loop))) {
    fclose(f);
    return fail(PF_E_ARG, "%s is not a checkpoint of this library version", path);
  }
  ok = ok && fread(&pp, sizeof pp, 1, f) == 1 && fread(&ll, sizeof ll, 1, f) == 1;
  if (!ok || pp.nx != h.nx || pp.ny != h.ny || pp.nz != h.nz || pp.precision != h.prec) {
    fclose(f);
    return fail(PF_E_ARG, "%s: truncated or inconsistent header", path);
  }
  pf_ctx *ctx = nullptr;
  int rc = pf_init(&pp, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, &ctx);
  if (rc != PF_OK) {
    fclose(f);
    return rc;
  }
  // pf_init leaves buffer 0 = old, 1 = new of each pair
  void *dst[7] = {ctx->phi.buf[0], ctx->phi.buf[1], ctx->c.buf[0], ctx->c.buf[1],
                  ctx->T.buf[0],   ctx->T.buf[1],   ctx->dpp};
  std::vector<char> buf(ctx->n * ctx->esz);
  for (int q = 0; q < 7 && rc == PF_OK; ++q) {
    if (fread(buf.data(), 1, buf.size(), f) != buf.size())
      rc = fail(PF_E_ARG, "%s: truncated field data", path);
    else if (hipMemcpyAsync(dst[q], buf.data(), buf.size(), hipMemcpyHostToDevice, ctx->stream) !=
                 hipSuccess ||
             hipStreamSynchronize(ctx->stream) != hipSuccess)
      rc = fail(PF_E_HIP, "checkpoint upload failed");
  }
  fclose(f);
  if (rc == PF_OK) {
    ctx->use_fused = h.use_fused != 0;
    if (pf::fused_set_tile(ctx->fused, h.tile_a, h.tile_b, h.target) != PF_OK)
      rc = fail(PF_E_ARG, "%s: bad tiling", path);
  }
  if (rc == PF_OK) {
    ctx->fused.mass_every = h.mass_every;
    ctx->fused.ledger_valid = false;
    ctx->fused.host_mass = h.host_mass;
    ctx->fused.host_mass_ok = h.host_mass_ok != 0;
    ctx->c_off = h.c_off;
    if (h.ledger_valid && h.mass_every > 0) {
      // allocate the ledger, then put the saved record back
      if (pf::fused_mass_prepare(ctx->fused, ctx->c.old_(), ctx->c_off, ctx->stream) != PF_OK ||
          pf::fused_mass_restore(ctx->fused, h.ledger, ctx->stream) != PF_OK)
        rc = fail(PF_E_HIP, "restoring the solute-mass ledger failed");
    }
  }
  if (rc != PF_OK) {
    pf_finalize(ctx);
    return rc;
  }
  *p = pp;
  *L = ll;
  *out = ctx;
  return PF_OK;
}

int pf_synchronize(pf_ctx *ctx) {
  int rc = check_ctx(ctx, nullptr);
  if (rc != PF_OK) return rc;
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  return PF_OK;
}

int pf_device_bytes(const pf_ctx *ctx, size_t *bytes) {
  int rc = check_ctx(ctx, nullptr);
  if (rc != PF_OK) return rc;
  if (!bytes) return fail(PF_E_ARG, "null output");
  *bytes = ctx->bytes + pf::fused_bytes(ctx->fused);
  return PF_OK;
}

int pf_set_fused(pf_ctx *ctx, int32_t on) {
  int rc = check_ctx(ctx, nullptr);
  if (rc != PF_OK) return rc;
  ctx->use_fused = on != 0;
  return PF_OK;
}

int pf_set_tiling(pf_ctx *ctx, int32_t tile_a, int32_t tile_b, int32_t target_groups) {
  int rc = check_ctx(ctx, nullptr);
  if (rc != PF_OK) return rc;
  if (pf::fused_set_tile(ctx->fused, tile_a, tile_b, target_groups) != PF_OK)
    return fail(PF_E_ARG, "tile shapes must be in [0,%d) and target_groups >= 1",
                pf::PF_NVARIANTS);
  return PF_OK;
}

int pf_set_sweep_split(pf_ctx *ctx, int32_t on, int32_t tile_s) {
  int rc = check_ctx(ctx, nullptr);
  if (rc != PF_OK) return rc;
  if (pf::fused_set_split(ctx->fused, on != 0, tile_s) != PF_OK)
    return fail(PF_E_ARG, "tile shape must be in [0,%d)", pf::PF_NVARIANTS);
  return PF_OK;
}

int pf_get_tiling(const pf_ctx *ctx, int32_t out[4]) {
  int rc = check_ctx(ctx, nullptr);
  if (rc != PF_OK) return rc;
  if (!out) return fail(PF_E_ARG, "null output");
  out[0] = ctx->fused.tile_a;
  out[1] = ctx->fused.tile_b;
  out[2] = ctx->fused.seg_a;
  out[3] = ctx->fused.seg_b;
  return PF_OK;
}

int pf_set_sweep_tiling(pf_ctx *ctx, int32_t kind, int32_t tile, int32_t target_groups) {
  int rc = check_ctx(ctx, nullptr);
  if (rc != PF_OK) return rc;
  if (pf::fused_set_kind(ctx->fused, kind, tile, target_groups) != PF_OK)
    return fail(PF_E_ARG, "kind must be 0..2, tile shape in [0,%d), target_groups >= 1",
                pf::PF_NVARIANTS);
  return PF_OK;
}

int pf_get_sweep_tiling(const pf_ctx *ctx, int32_t kind, int32_t out[3]) {
  int rc = check_ctx(ctx, nullptr);
  if (rc != PF_OK) return rc;
  if (!out) return fail(PF_E_ARG, "null output");
  int v[3];
  if (pf::fused_get_kind(ctx->fused, kind, v) != PF_OK) return fail(PF_E_ARG, "kind must be 0..2");
  for (int q = 0; q < 3; ++q) out[q] = v[q];
  return PF_OK;
}

int pf_get_sweep_split(const pf_ctx *ctx, int32_t out[3]) {
  int rc = check_ctx(ctx, nullptr);
  if (rc != PF_OK) return rc;
  if (!out) return fail(PF_E_ARG, "null output");
  out[0] = ctx->fused.split_b ? 1 : 0;
  out[1] = ctx->fused.tile_s;
  out[2] = ctx->fused.seg_s;
  return PF_OK;
}

int pf_autotune(pf_ctx *ctx, const pf_params *p, int32_t reps, int32_t *ncand, double *ms) {
  int rc = check_ctx(ctx, p);
  if (rc == PF_OK && !p) rc = fail(PF_E_ARG, "null params");
  if (rc == PF_OK && reps < 1) rc = fail(PF_E_ARG, "reps must be >= 1");
  if (rc != PF_OK) return rc;
  if (!ctx->fused.enabled) return fail(PF_E_STATE, "grid too small for the fused sweeps");
  HIP_TRY(hipSetDevice(ctx->device));
  pf::AutotuneResult res;
  rc = pf::fused_autotune(ctx->fused, kparams(ctx, *p), ctx->phi.old_(), ctx->c.old_(),
                          ctx->T.old_(), reps, ctx->stream, &res);
  if (rc == PF_E_NOMEM) return fail(rc, "no device memory for the autotune scratch fields");
  if (rc != PF_OK) return fail(rc, "autotune launch failed: %s",
                               hipGetErrorString(hipGetLastError()));
  if (ncand) *ncand = res.n;
  if (ms)
    for (int q = 0; q < res.n; ++q) {
      ms[4 * q + 0] = res.cand[q].sweep;
      ms[4 * q + 1] = res.cand[q].tile;
      ms[4 * q + 2] = res.cand[q].target;
      ms[4 * q + 3] = res.cand[q].ms;
    }
  return PF_OK;
}

int pf_set_mass_correction(pf_ctx *ctx, int32_t every) {
  int rc = check_ctx(ctx, nullptr);
  if (rc != PF_OK) return rc;
  if (every < 0) return fail(PF_E_ARG, "negative correction period");
  ctx->fused.mass_every = every;
  ctx->fused.ledger_valid = false;  // restarts from the field at the next run
  return PF_OK;
}

int pf_mass_ledger(pf_ctx *ctx, double out[3]) {
  int rc = check_ctx(ctx, nullptr);
  if (rc != PF_OK) return rc;
  if (!out) return fail(PF_E_ARG, "null output");
  if (pf::fused_mass_read(ctx->fused, out, ctx->stream) != PF_OK)
    return fail(PF_E_HIP, "reading the ledger failed");
  return PF_OK;
}

int pf_timing_enable(pf_ctx *ctx, int32_t on) {
  int rc = check_ctx(ctx, nullptr);
  if (rc != PF_OK) return rc;
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  ctx->timer.kind.clear();
  ctx->timer.on = on != 0;
  return PF_OK;
}

int pf_timing_kinds(pf_ctx *ctx, uint32_t kinds) {
  int rc = check_ctx(ctx, nullptr);
  if (rc != PF_OK) return rc;
  ctx->timer.mask = kinds;
  return PF_OK;
}

int pf_timing_read(pf_ctx *ctx, double ms[PF_KT_COUNT], int64_t count[PF_KT_COUNT]) {
  int rc = check_ctx(ctx, nullptr);
  if (rc == PF_OK && (!ms || !count)) rc = fail(PF_E_ARG, "null output");
  if (rc != PF_OK) return rc;
  HIP_TRY(hipStreamSynchronize(ctx->stream));
  for (int q = 0; q < PF_KT_COUNT; ++q) {
    ms[q] = 0.0;
    count[q] = 0;
  }
  pf::KTimer &t = ctx->timer;
  for (size_t r = 0; r < t.kind.size(); ++r) {
    float e = 0.f;
    HIP_TRY(hipEventElapsedTime(&e, t.pool[2 * r], t.pool[2 * r + 1]));
    ms[t.kind[r]] += (double)e;
    count[t.kind[r]] += 1;
  }
  t.kind.clear();
  return PF_OK;
}

// principal.f90:96-226 without the output event.  The host scalars (time,
// TP, Cs_eq/Cl_eq, M0) advance exactly as in the reference; the field work
// goes to the fused kernels when the context supports them.
int pf_run_steps(pf_ctx *ctx, pf_params *p, pf_loop *L, int32_t nsteps, double *probes) {
  int rc = check_ctx(ctx, p);
  if (rc == PF_OK && (!p || !L)) rc = fail(PF_E_ARG, "null argument");
  if (rc == PF_OK && nsteps < 0) rc = fail(PF_E_ARG, "negative step count");
  if (rc != PF_OK) return rc;
  if (nsteps == 0) return PF_OK;
  const int nv = p->n_varreduras < 1 ? 1 : p->n_varreduras;
  if (probes && (rc = ensure_probe_cap(ctx, (size_t)2 * nv * nsteps)) != PF_OK) return rc;
  std::vector<double> tp_seen;
  if (probes) tp_seen.reserve((size_t)nsteps);
  const bool fused = ctx->use_fused && pf::fused_supported(ctx->fused, *p);
  pf::FusedState &fst = ctx->fused;
  if (fused) {
    if ((rc = pf::fused_mass_prepare(fst, ctx->c.old_(), ctx->c_off, ctx->stream)) != PF_OK)
      return fail(rc, "solute-mass ledger set-up failed");
  } else {
    fst.ledger_valid = false;  // the unfused kernels do not keep the ledger
  }
  fst.host_mass_ok = false;  // the fields move on from the caller's
  for (int s = 0; s < nsteps; ++s) {
    const bool first = (s == 0), last = (s == nsteps - 1);
    p->it += 1;
    const int it = p->it;
    L->time = (float)(p->dt + (double)L->time);  // principal.f90:96
    const pf::KParams k_temp = kparams(ctx, *p);      // temperatura sees TP of step it-1
    if (L->taxa_resfr > 0.0) p->TP = L->T0 - L->taxa_resfr * p->dt * (double)it;  // :104
    pf_equilibrium(p->TP, p->Tm, p->Ke, p->Me, &p->Cs_eq, &p->Cl_eq);
    const pf::KParams k = kparams(ctx, *p);
    if (probes) tp_seen.push_back(p->TP);
    double *pr = probes ? ctx->probe_dev + (size_t)2 * nv * s : nullptr;
    if (fused) {
      pf::FusedStep fs;
      fs.first = first;
      fs.last = last;
      fs.nv = nv;
      fs.flush = last || (fst.mass_every > 0 && it % fst.mass_every == 0);
      rc = pf::fused_step(ctx->fused, k_temp, k, fs, ctx->phi.buf, &ctx->phi.old_idx,
                          &ctx->phi.alias, ctx->c.buf, &ctx->c.old_idx, &ctx->c.alias, ctx->T.buf,
                          &ctx->T.old_idx, &ctx->T.alias, ctx->dpp, pr, ctx->stream,
                          ctx->timer);
      if (rc != PF_OK) return rc;
      if (fst.mass_every > 0 && it % fst.mass_every == 0) {  // c_old holds this step's c
        ctx->timer.begin(PF_KT_MASS, ctx->stream);
        rc = pf::fused_mass_correct(fst, k, ctx->c.buf[ctx->c.old_idx], ctx->stream);
        ctx->timer.end(ctx->stream);
        if (rc != PF_OK) return fail(rc, "solute-mass correction failed");
      }
    } else {
      if ((rc = do_temperatura(ctx, k_temp)) != PF_OK) return rc;
      for (int sw = 1; sw <= nv; ++sw) {
        if (sw > 1) ctx->phi.swap();  // principal.f90:127-150
        if ((rc = do_phase(ctx, k)) != PF_OK) return rc;
        if (pr && (rc = launch_probe(ctx, ctx->phi.new_(), pr + 2 * (sw - 1))) != PF_OK) return rc;
      }
      if ((rc = do_solute(ctx, k)) != PF_OK) return rc;
      ctx->phi.swap();  // principal.f90:200-226
      ctx->T.swap();
      ctx->c.swap();
    }
    // auxiliar.f90 (principal.f90:157): M0 for the next step
    p->M0 = pf_auxiliar_m0(p->TP, p->Cl_eq, p->Cs_eq, p->R, p->Vm, p->E0, p->W0, L->Sigma, L->Beta,
                           p->Ke, p->Me, p->D_liq, L->ajuste);
  }
  if (probes) {
    std::vector<double> tmp((size_t)2 * nv * nsteps);
    HIP_TRY(hipMemcpyAsync(tmp.data(), ctx->probe_dev, tmp.size() * sizeof(double),
                           hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (int s = 0; s < nsteps; ++s)
      for (int sw = 0; sw < nv; ++sw) {
        const size_t q = (size_t)s * nv + sw;
        probes[3 * q + 0] = tmp[2 * q];
        probes[3 * q + 1] = tp_seen[s];
        probes[3 * q + 2] = tmp[2 * q + 1];
      }
  }
  return PF_OK;
}

}  // extern "C"

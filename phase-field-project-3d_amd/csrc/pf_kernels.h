// pf_kernels.h -- launchers of the gfx950 kernels (internal to libpfamd).
#pragma once

#include <hip/hip_runtime.h>

#include "pf_common.h"

namespace pf {

// Drop-in kernels: one per reference subroutine (pf_kernels_dropin.hip).
template <typename R>
hipError_t launch_temperatura(const KParams &p, const R *Told, const R *phi, const R *dpp, R *Tnew,
                              hipStream_t s);
template <typename R>
hipError_t launch_phase(const KParams &p, const R *po, const R *co, R *pn, R *dpp, hipStream_t s);
template <typename R>
hipError_t launch_solute(const KParams &p, const R *po, const R *co, R *cn, hipStream_t s);
template <typename R>
hipError_t launch_diag(const KParams &p, const R *ph, const R *cc, double *part, int nblk,
                       double *out, hipStream_t s);

// The boundary decorations alone (also used by the fused sweeps).  xf =
// false leaves out the x-face cells of the interior rows, which the fused
// sweeps write themselves.
template <typename R>
hipError_t launch_phase_boundary(const KParams &p, const R *po, R *pn, hipStream_t s,
                                 bool xf = true);
template <typename R>
hipError_t launch_solute_boundary(const KParams &p, const R *co, R *cn, hipStream_t s,
                                  bool xf = true);
// Face cells take their interior source cell's value again (in place).
template <typename R>
hipError_t launch_solute_refresh(const KParams &p, R *c, hipStream_t s);

// The two cells equacao_campo_fase.f90:174 prints (0-based offsets):
// (nx/2, 10, max(nz/2,1)) and (nx/2, min(100,ny-1), max(nz/2,1)).
inline void probe_cells(int nx, int ny, int nz, size_t &a, size_t &b) {
  const int kz = nz / 2 > 1 ? nz / 2 : 1;
  const int j2 = ny - 1 < 100 ? ny - 1 : 100;
  const int i = nx / 2;
  const int j1 = ny >= 10 ? 10 : ny;  // the reference reads out of bounds below 10
  a = (size_t)(i - 1) + (size_t)nx * ((size_t)(j1 - 1) + (size_t)ny * (size_t)(kz - 1));
  b = (size_t)(i - 1) + (size_t)nx * ((size_t)(j2 - 1) + (size_t)ny * (size_t)(kz - 1));
}
// Copies f[a], f[b] (as double) to out[0..1] on the stream.
template <typename R>
hipError_t launch_probe(const R *f, size_t a, size_t b, double *out, hipStream_t s);

// Number of partial blocks the diagnostics reduction uses for a grid.
inline int diag_blocks(int ny, int nz) {
  const int rows = ny * nz;
  return rows < 2048 ? rows : 2048;
}

}  // namespace pf

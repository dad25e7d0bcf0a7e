// pf_mass.h -- the solute-mass ledger of the fused forms, shared by the
// single-instance loop (pf_fused.hip) and the batched form (pf_batch.hip).
//
// The ledger is the FP64 interior ([2,n-1]^3) solute mass the scheme's own
// increments imply: the FP64 mass when it started plus the sum of every
// increment the solute update computed (before the increment is rounded into
// the stored field).  Every `every` steps the FP64 interior sum of the stored
// field is compared with it and the interior is rescaled by ledger/sum,
// which removes the drift that FP32 storage rounding accumulates.
//
// Increments are summed as 64-bit integers in units of 2^-44 (each cell's
// increment rounded to the nearest unit, ties to even): integer addition is
// associative, so the sum does not depend on the order in which work-groups
// add their share, on the tiling, on how a run is split into calls, or on
// what else runs in a batch -- a corrected run gives the same bits in the
// single-instance and the batched forms.  The unit keeps the rounding of a
// cell's increment below 3e-14 (an FP32 increment of a liquid cell is
// rounded to ~1e-15 absolute already); the range, 2^19 in units of c, is
// far above the interior mass change of any run.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pf_common.h"

namespace pf {

constexpr double kMassUnit = 0x1p-44;   // one ledger unit, in units of c
constexpr double kMassPerUnit = 0x1p44;

// Device record of one ledger.
struct MassLedger {
  double L0;                // FP64 interior mass when the ledger started
  unsigned long long acc;   // increments since then (two's complement, ledger units)
  double S;                 // FP64 interior sum of the stored field at the last correction
  double f;                 // ledger / S at the last correction
};

PF_HD double ledger_value(const MassLedger &m) {
#pragma clang fp contract(off)
  return m.L0 + (double)(long long)m.acc * kMassUnit;
}

// One increment in ledger units (exact conversion of an integer-valued
// double below 2^63).
template <typename R>
__device__ __forceinline__ long long mass_units(R v) {
  return (long long)rint((double)v * kMassPerUnit);
}

// Work-group sum of the per-thread ledger units, added to *dst with one
// integer atomic (exact, so the order in which work-groups add does not
// matter).  Every thread of the work-group calls it; nt threads, a multiple
// of 64.
template <int NT>
__device__ __forceinline__ void mass_flush(long long v, int tid, unsigned long long *dst) {
  __shared__ long long wsum[NT / 64];
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  if ((tid & 63) == 0) wsum[tid >> 6] = v;
  __syncthreads();
  if (tid == 0) {
    long long t = 0;
    for (int w = 0; w < NT / 64; ++w) t += wsum[w];
    atomicAdd(dst, (unsigned long long)t);
  }
}

// Work-groups of the interior-sum kernels and their fixed decomposition:
// block b of kMassBlocks sums the rows r = b, b + kMassBlocks, ... of the
// interior ((j,k) = (2 + r % (ny-2), 2 + r / (ny-2))), 256 threads striding
// over i; the block sums, then the kMassBlocks partials, are folded by
// fixed trees.  Single instances and batched instances use the same
// decomposition, so their sums are equal bit for bit.
constexpr int kMassBlocks = 1024;

__device__ __forceinline__ double block_sum256(double v, double *sh) {
  const int t = threadIdx.x;
  __syncthreads();  // sh may still be read from a previous call
  sh[t] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) sh[t] += sh[t + w];
    __syncthreads();
  }
  return sh[0];
}

template <typename R>
__device__ __forceinline__ double mass_rows(const R *__restrict__ c, int nx, int ny, int nz,
                                            int b) {
  const int nrow = (ny - 2) * (nz - 2);
  const size_t sz = (size_t)nx * ny;
  double acc = 0.0;
  for (int r = b; r < nrow; r += kMassBlocks) {
    const int j = 2 + r % (ny - 2), k = 2 + r / (ny - 2);
    const size_t row = (size_t)(j - 1) * nx + (size_t)(k - 1) * sz;
    for (int i = 2 + (int)threadIdx.x; i <= nx - 1; i += 256) acc += (double)c[row + i - 1];
  }
  return acc;
}

// The interior sum from the kMassBlocks partials (one 256-thread block).
__device__ __forceinline__ double mass_fold(const double *__restrict__ part, double *sh) {
  double a = 0.0;
  for (int q = threadIdx.x; q < kMassBlocks; q += 256) a += part[q];
  return block_sum256(a, sh);
}

// Rescales the interior rows of block b by m.f.  FP32: the FP64 product is
// rounded stochastically to one of its two neighbouring floats with the
// probability that makes the expectation exact (a counter hash of cell and
// step, salt 4, draws the choice), because the factor is typically
// 1 -+ 1e-9, far below one FP32 ulp, and round-to-nearest would leave every
// cell as it was.
template <typename R>
__device__ __forceinline__ void mass_scale_rows(R *__restrict__ c, int nx, int ny, int nz, int it,
                                                uint64_t seed, double f, int b) {
  const int nrow = (ny - 2) * (nz - 2);
  const size_t sz = (size_t)nx * ny;
  for (int r = b; r < nrow; r += kMassBlocks) {
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

// FP64 interior mass of a host field (the ledger's start when the stored
// field is FP32: its rounding would otherwise enter the ledger).  Defined in
// pf_host.cpp.
double host_interior_mass(const double *c, int nx, int ny, int nz);

}  // namespace pf

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

// c is the stored field: c itself (FP64) or c - coff (FP32-mixed).
template <typename R>
__device__ __forceinline__ double mass_rows(const R *__restrict__ c, int nx, int ny, int nz,
                                            int b, double coff) {
  const int nrow = (ny - 2) * (nz - 2);
  const size_t sz = (size_t)nx * ny;
  double acc = 0.0;
  for (int r = b; r < nrow; r += kMassBlocks) {
    const int j = 2 + r % (ny - 2), k = 2 + r / (ny - 2);
    const size_t row = (size_t)(j - 1) * nx + (size_t)(k - 1) * sz;
    for (int i = 2 + (int)threadIdx.x; i <= nx - 1; i += 256) {
      if constexpr (sizeof(R) == 8) acc += (double)c[row + i - 1];
      else acc += (double)c[row + i - 1] + coff;
    }
  }
  return acc;
}

// The interior sum from the kMassBlocks partials (one 256-thread block).
__device__ __forceinline__ double mass_fold(const double *__restrict__ part, double *sh) {
  double a = 0.0;
  for (int q = threadIdx.x; q < kMassBlocks; q += 256) a += part[q];
  return block_sum256(a, sh);
}

// Rescales the interior rows of block b by f.  FP32: the FP64 product is
// rounded stochastically to one of its two neighbouring floats with the
// probability that makes the expectation exact (a counter hash of cell and
// step, salt 4, draws the choice), because the factor is typically
// 1 -+ 1e-9, far below one FP32 ulp, and round-to-nearest would leave every
// cell as it was.
template <typename R>
__device__ __forceinline__ void mass_scale_rows(R *__restrict__ c, int nx, int ny, int nz, int it,
                                                uint64_t seed, double f, int b, double coff) {
  const int nrow = (ny - 2) * (nz - 2);
  const size_t sz = (size_t)nx * ny;
  for (int r = b; r < nrow; r += kMassBlocks) {
    const int j = 2 + r % (ny - 2), k = 2 + r / (ny - 2);
    const size_t row = (size_t)(j - 1) * nx + (size_t)(k - 1) * sz;
    for (int i = 2 + (int)threadIdx.x; i <= nx - 1; i += 256) {
      const size_t t = row + i - 1;
      if constexpr (sizeof(R) == 8) {
        c[t] = (R)((double)c[t] * f);
      } else {
        // the stored offset of the rescaled c
        const double x = ((double)c[t] + coff) * f - coff;
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

// 32-bit counter hash of a cell and step for csum: the "lowbias32"
// finaliser (C. Wellons, hash-prospector) over an odd-multiplier mix of the
// coordinates.  col = cell_hash_col(i, j, it) once per column, then
// cell_hash(col, k) per plane.
__device__ __forceinline__ uint32_t cell_hash_col(int i, int j, int it) {
  return (uint32_t)i * 0x9E3779B1u ^ (uint32_t)j * 0x85EBCA77u ^ (uint32_t)it * 0x27D4EB2Fu;
}
__device__ __forceinline__ uint32_t cell_hash(uint32_t col, int k) {
  uint32_t h = col ^ ((uint32_t)k * 0xC2B2AE3Du);
  h ^= h >> 16;
  h *= 0x7FEB352Du;
  h ^= h >> 15;
  h *= 0x846CA68Bu;
  h ^= h >> 16;
  return h;
}

// A stored value after an increment, cst + inc (c after the solute update,
// phi after a phase sweep).  FP64: rounded to nearest.  FP32: rounded
// stochastically to one of the two floats around the exact sum, with the
// probability that makes the expected stored value exact (TwoSum gives the
// rounding error e of the nearest sum s; the neighbour of s on e's side is
// taken with probability |e| / gap), the draw from a 32-bit counter hash of
// the cell and the step.  Increments far below an FP32 ulp are then kept
// on average instead of all being rounded away: those of c in the solid,
// whose diffusivity is 1e-4 of the liquid's, and those of phi as it
// approaches 1 in the solid, where round-to-nearest stalled phi below its
// FP64 trajectory and biased the solute flux across the walls (the FP32
// form's interior mass drifted 4.3e-7 from the FP64 run's over 5,000 steps
// of a growth deck with c alone rounded so, 5.2e-8 with phi too).
template <typename R>
__device__ __forceinline__ R csum(R cst, R inc, uint32_t col, int k);
template <>
__device__ __forceinline__ double csum<double>(double cst, double inc, uint32_t, int) {
  return cst + inc;
}
// (col, k) name the draw, cell_hash(col, k) below.  It is hashed only in
// waves where some lane's sum is inexact (e != 0): a lane with e == 0 keeps
// s whatever the draw, so the result depends on the lane's own values
// alone, and waves of exact sums -- zero increments in uniform regions --
// skip the hash.
template <>
__device__ __forceinline__ float csum<float>(float cst, float inc, uint32_t col, int k) {
#pragma clang fp contract(off)
  const float s = cst + inc;
  const float bb = s - cst;
  const float e = (cst - (s - bb)) + (inc - bb);  // cst + inc = s + e exactly
  if (__builtin_amdgcn_ballot_w64(e != 0.f) == 0) return s;
  // the float next to s on e's side (s == 0 implies e == 0: both operands are
  // multiples of the smallest denormal, so is their sum)
  const int step = (e > 0.f) == (s > 0.f) ? 1 : -1;
  const float nb = __int_as_float(__float_as_int(s) + step);
  const float u = (float)(cell_hash(col, k) >> 8) * 0x1p-24f;  // [0, 1)
  return u * fabsf(nb - s) < fabsf(e) ? nb : s;  // false for e == 0 and for NaN
}

// Salts that keep the draws of one step's FP32 stores apart: the solute
// store draws from cell_hash(col, k), the phase stores from
// cell_hash(col ^ salt, k) -- kSaltPhiFirst for the sweeps before the last
// one of a step, kSaltPhiLast for the last (the drop-in kernel, called once
// per sweep, uses kSaltPhiFirst for every sweep).
constexpr uint32_t kSaltPhiFirst = 0x1B56C4E9u, kSaltPhiLast = 0x68E31DA4u;

// FP64 interior mass of a host field (the ledger's start when the stored
// field is FP32: its rounding would otherwise enter the ledger).  Defined in
// pf_host.cpp.
double host_interior_mass(const double *c, int nx, int ny, int nz);

}  // namespace pf

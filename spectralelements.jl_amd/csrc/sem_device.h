// Device-side helpers shared by the fused apply and the PCG kernels.
#pragma once

#include "sem_common.h"

namespace sem {

__device__ __forceinline__ double nanmax(double a, double b) {
  return (b > a || b != b) ? b : a;
}

// Fixed-order block sum of v over nthreads threads (red has nthreads entries,
// nthreads need not be a power of two). Every block computes bit-identical
// results for identical inputs.
__device__ inline double block_sum(double v, double* red, int nthreads) {
  const int tid = threadIdx.x;
  red[tid] = v;
  __syncthreads();
  int n = nthreads;
  while (n > 1) {
    const int h = (n + 1) / 2;
    if (tid < n / 2) red[tid] = red[tid] + red[tid + h];
    __syncthreads();
    n = h;
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

__device__ inline double block_max(double v, double* red, int nthreads) {
  const int tid = threadIdx.x;
  red[tid] = v;
  __syncthreads();
  int n = nthreads;
  while (n > 1) {
    const int h = (n + 1) / 2;
    if (tid < n / 2) red[tid] = nanmax(red[tid], red[tid + h]);
    __syncthreads();
    n = h;
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

// Sum of part[0..n) and max of part[n..2n), identical in every block.
__device__ inline void reduce_partials(const double* part, int n, double* red, int nthreads,
                                       double* sum, double* mx) {
  double s = 0.0, m = 0.0;
  for (int q = threadIdx.x; q < n; q += nthreads) {
    s += part[q];
    m = nanmax(m, part[n + q]);
  }
  *sum = block_sum(s, red, nthreads);
  *mx = block_max(m, red, nthreads);
}

__device__ inline double reduce_sum_array(const double* part, int n, double* red, int nthreads) {
  double s = 0.0;
  for (int q = threadIdx.x; q < n; q += nthreads) s += part[q];
  return block_sum(s, red, nthreads);
}

// Reads state->done once per block and broadcasts it, so the whole block
// takes the same branch.
__device__ inline bool block_done(const PcgState* st, double* red) {
  if (threadIdx.x == 0) red[0] = (double)st->done;
  __syncthreads();
  const bool d = red[0] != 0.0;
  __syncthreads();
  return d;
}

// 1/(multiplicity along one direction) of local row i (P points per element)
__device__ __forceinline__ double mult1(int i, int P, int L, int per) {
  const int a = i % P;
  if (a == 0 && (i > 0 || per)) return 0.5;
  if (a == P - 1 && (i < L - 1 || per)) return 0.5;
  return 1.0;
}

__device__ __forceinline__ bool mask_on(int i, int j, int L, int NC, int bc) {
  if ((bc & SEM_BC_XMIN) && i == 0) return false;
  if ((bc & SEM_BC_XMAX) && i == L - 1) return false;
  if ((bc & SEM_BC_YMIN) && j == 0) return false;
  if ((bc & SEM_BC_YMAX) && j == NC - 1) return false;
  return true;
}

template <typename T>
__device__ __forceinline__ T mval(int i, int j, int L, int NC, int bc) {
  return mask_on(i, j, L, NC, bc) ? T(1) : T(0);
}

}  // namespace sem

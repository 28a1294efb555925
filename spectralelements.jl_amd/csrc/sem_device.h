// Device-side helpers shared by the fused apply and the PCG kernels.
#pragma once

#include "sem_common.h"

namespace sem {

__device__ __forceinline__ double nanmax(double a, double b) {
  return (b > a || b != b) ? b : a;
}

// Fixed-order reductions within a 64-lane wavefront: a butterfly over the
// lanes; lane 0's result has a fixed operand order and is the one used.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = nanmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Fixed-order block sum of v over nthreads threads (red has nthreads entries,
// nthreads need not be a power of two). Every block computes bit-identical
// results for identical inputs. (A tree over LDS: it adds no registers to
// the register-heavy apply kernels that call it.)
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

// The same reductions for blocks of NT threads, NT a multiple of 64: lane
// shuffles within each wavefront, then one LDS step across the waves. Fixed
// order, bit-identical across blocks.
//
// Block sum returned to every thread with two barriers: red[0] holds the
// result until the caller's next barrier, so red must not be written before
// one.
template <int NT>
__device__ inline double block_sum_w_nb(double v, double* red) {
  static_assert(NT % 64 == 0 && NT <= 4096, "whole wavefronts");
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  v = wave_sum(v);
  if (lane == 0) red[wv] = v;
  __syncthreads();
  if (wv == 0) {
    double t = lane < NT / 64 ? red[lane] : 0.0;
    t = wave_sum(t);
    if (lane == 0) red[0] = t;   // wave 0 has read red[0..NT/64) already
  }
  __syncthreads();
  return red[0];
}

// Block sum of s and block max of m in one pass, valid in thread 0 only
// (one barrier). red must hold 2 NT / 64 values and be free.
template <int NT>
__device__ inline void block_sum_max_to0(double& s, double& m, double* red) {
  static_assert(NT % 64 == 0 && NT <= 4096, "whole wavefronts");
  constexpr int NW = NT / 64;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  s = wave_sum(s);
  m = wave_max(m);
  if (lane == 0) {
    red[wv] = s;
    red[NW + wv] = m;
  }
  __syncthreads();
  if (wv == 0) {
    double ts = lane < NW ? red[lane] : 0.0;
    double tm = lane < NW ? red[NW + lane] : 0.0;
    s = wave_sum(ts);
    m = wave_max(tm);
  }
}

// Sum of part[0..n) and max of part[n..2n), identical in every block, fused
// with the stop-flag broadcast: the sum and the max climb the same LDS trees
// as block_sum / block_max side by side, and state->done, read by thread 0,
// is published by the first barrier (one tree's barriers instead of two
// trees' plus two).
__device__ inline void reduce_partials_flag(const double* part, int n, double* red, double* red2,
                                            int* flag, int nthreads, const PcgState* st,
                                            double* sum, double* mx, bool* stopped) {
  const int tid = threadIdx.x;
  double s = 0.0, m = 0.0;
  for (int q = tid; q < n; q += nthreads) {
    s += part[q];
    m = nanmax(m, part[n + q]);
  }
  if (tid == 0) *flag = st->done;
  red[tid] = s;
  red2[tid] = m;
  __syncthreads();
  int c = nthreads;
  while (c > 1) {
    const int h = (c + 1) / 2;
    if (tid < c / 2) {
      red[tid] = red[tid] + red[tid + h];
      red2[tid] = nanmax(red2[tid], red2[tid + h]);
    }
    __syncthreads();
    c = h;
  }
  *sum = red[0];
  *mx = red2[0];
  *stopped = *flag != 0;
  __syncthreads();
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

// QQ' of one global node g through the sorted map (perm: local copies by
// global id, seg: offsets): the copies' sum in a fixed order (pairs of
// consecutive copies first), masked by the first copy's mask when bc >= 0,
// written to every copy. u may alias y (all copies are read first).
template <typename T>
__device__ __forceinline__ void gsmap_node(const int32_t* __restrict__ perm,
                                           const int32_t* __restrict__ seg, const T* u, T* y,
                                           long g, int bc, int L, int NC) {
  const int s0 = seg[g], s1 = seg[g + 1];
  const int n = s1 - s0;
  T s;
  if (n == 1) {
    s = u[perm[s0]];
  } else if (n == 2) {
    s = u[perm[s0]] + u[perm[s0 + 1]];
  } else if (n == 4) {
    s = (u[perm[s0]] + u[perm[s0 + 1]]) + (u[perm[s0 + 2]] + u[perm[s0 + 3]]);
  } else {
    // general segment: pairwise sums of consecutive pairs, left to right
    s = T(0);
    int k = s0;
    bool first = true;
    for (; k + 1 < s1; k += 2) {
      const T p = u[perm[k]] + u[perm[k + 1]];
      s = first ? p : s + p;
      first = false;
    }
    if (k < s1) s = first ? u[perm[k]] : s + u[perm[k]];
  }
  if (bc >= 0) {
    const int l = perm[s0];
    s = s * (mask_on(l % L, l / L, L, NC, bc) ? T(1) : T(0));
  }
  for (int k = s0; k < s1; ++k) y[perm[k]] = s;
}

// Chebyshev-Jacobi direction update d' = c1 d + c2 dinv .* r: the smoother's
// vector step and the fused apply epilogue evaluate this one expression with
// an explicit fused multiply-add (left to the compiler, the contraction it
// picks may differ between the two call sites), so the fused and unfused
// smoothers give identical bits.
template <typename T>
__device__ __forceinline__ T cheb_dir(T c1, T d, T c2, T dinv, T r) {
  const T t = c2 * (dinv * r);
  if constexpr (sizeof(T) == 8) return __builtin_fma(c1, d, t);
  else return __builtin_fmaf(c1, d, t);
}

// Fused apply epilogue (the kinds of ApplyEpi) for node copy t: its
// operands are loaded first, together with the apply's other loads, and the
// update is stored once the copy's assembled, masked value w is final: r -= w,
// and for KIND 1 the direction update d' (stored in y; dold = the copy's
// input value) and x += d'.
template <typename T>
struct EpiIn {
  T r, dinv, x;
};
template <typename T, int KIND>
__device__ __forceinline__ EpiIn<T> epi_load(long t, const T* r, const T* x, const T* dinv) {
  EpiIn<T> v{};
  v.r = r[t];
  if constexpr (KIND == 1 || KIND == 3) {
    v.dinv = dinv[t];
    v.x = x[t];
  }
  return v;
}
template <typename T, int KIND>
__device__ __forceinline__ void epi_store(long t, T w, T dold, const EpiIn<T>& v, T* r, T* x,
                                          T c1, T c2, T* y, T* d, T theta) {
  const T rv = v.r - w;
  r[t] = rv;
  if constexpr (KIND == 1) {
    const T dv = cheb_dir(c1, dold, c2, v.dinv, rv);
    y[t] = dv;
    x[t] = v.x + dv;
  } else if constexpr (KIND == 3) {
    const T dv = v.dinv * rv / theta;   // the Chebyshev start's expression
    d[t] = dv;
    x[t] = v.x + dv;
  }
}

// Seam geometry of a fused apply launch (see AxPlan): strips of CXP local
// rows, segments of SP local columns.
struct SeamGeom {
  int P, L, NC, CXP, nstrips, SP, nsegs, per_x, per_y, bc;
};

// Final value of node copy (i, j) of y = M QQ'(H) when the strip kernel left
// the seam copies unsummed (the k_seams rules, evaluated per copy): returns
// the stored value for copies the strip kernel finished. Every copy of a node
// evaluates the same expression in the same order, so all copies get
// identical bits.
template <typename T>
__device__ __forceinline__ T seam_final(const T* __restrict__ y, int i, int j, const SeamGeom& g) {
  const long L = g.L;
  const int r = i % g.CXP;
  int iL = -1, iR = -1;
  if (r == 0 && i > 0) {
    iL = i - 1;
    iR = i;
  } else if (r == g.CXP - 1 && i < g.L - 1) {
    iL = i;
    iR = i + 1;
  } else if (g.per_x && (i == 0 || i == g.L - 1)) {
    iL = g.L - 1;
    iR = 0;
  }
  if (iL >= 0) {
    // x seam: raw values on both sides, on every y-interface too
    const int b = j % g.P;
    int jl = -1, ju = -1;
    if (b == 0 && (j > 0 || g.per_y)) {
      jl = j > 0 ? j - 1 : g.NC - 1;
      ju = j;
    } else if (b == g.P - 1 && (j < g.NC - 1 || g.per_y)) {
      jl = j;
      ju = j + 1 < g.NC ? j + 1 : 0;
    }
    T s;
    if (jl >= 0) {
      const T sl = y[iL + L * jl] + y[iR + L * jl];
      const T su = y[iL + L * ju] + y[iR + L * ju];
      s = sl + su;
    } else {
      s = y[iL + L * j] + y[iR + L * j];
    }
    return s * mval<T>(i, j, g.L, g.NC, g.bc);
  }
  // y seam: x-summed values on both sides
  const int q = j % g.SP;
  int jL = -1, jR = -1;
  if (q == 0 && j > 0) {
    jL = j - 1;
    jR = j;
  } else if (q == g.SP - 1 && j < g.NC - 1) {
    jL = j;
    jR = j + 1;
  } else if (g.per_y && (j == 0 || j == g.NC - 1)) {
    jL = g.NC - 1;
    jR = 0;
  }
  if (jL >= 0) {
    const T s = y[i + L * jL] + y[i + L * jR];
    return s * mval<T>(i, j, g.L, g.NC, g.bc);
  }
  return y[i + L * j];
}

}  // namespace sem

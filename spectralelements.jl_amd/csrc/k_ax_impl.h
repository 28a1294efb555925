// Fused SEM operator apply  y = M .* QQ'(nu A u + k B u)  for gfx950
// (kernel templates; instantiated per P by k_ax_inst.hip, dispatched by
// k_ax.hip).
//
// Reference: laplace (lapl.jl:70-81), hlmz (hlmz.jl:12-19), gatherScatter
// (gatherScatter.jl:8-16, dense QQtx*u*QQty'), mask (mask.jl:10-18), composed as
// opLHS (diffusion.jl:36-45).
//
// Work decomposition (fields stay in the reference's column-major local layout,
// L = P*Ex rows, NC = P*Ey columns):
//  * A workgroup owns a strip of CX consecutive elements in x (CX*P local rows,
//    one thread per row) and sweeps a segment of S element rows upward.
//  * "Row owner": thread (e, a) holds row a of element e for all P columns b,
//    so every global load/store is lane-contiguous along x (fully coalesced).
//    The s-direction contractions (u Ds', ws Ds) run in registers.
//  * "Column owner": the same thread, renamed (e, b'=a), reads column b' of
//    element e from LDS and runs the r-direction contractions (Dr u, Dr' wr)
//    in registers. Four LDS transposes per element row; Dr/Ds are uniform
//    (scalar loads).
//  * Gather-scatter inside the workgroup: x-pairs are summed through LDS, then
//    y-pairs through a one-value register carry between consecutive element
//    rows, so every node gets fl(fl(u1+u2)+fl(u3+u4)) with x-pairs first --
//    the value the reference's two 0/1 GEMMs produce.
//  * Nodes whose partner copies live in another workgroup (strip seams in x,
//    segment seams in y, periodic wraps) are written unsummed and finished by
//    k_seams in the same order. Every node copy is finalised exactly once.
//  * The assembled FMA apply at P < 16 splits the workgroup into RY = 2 or 4
//    groups of CX/RY elements on consecutive element rows side by side: the
//    y-pairs between the rows are summed through LDS, each row writing its
//    own copies, so a segment is RY rows without a longer sweep.
//  * No atomics; the result is deterministic.
#pragma once
#include <algorithm>

#include "sem_device.h"
#include "_gen/gll_dtables.h"

// Tuning constants: threads per strip workgroup (one per local row, so a
// strip is SEM_AX_THREADS / P elements wide) and the largest P * sizeof(T)
// for which the geometric factors are staged one phase early.
// Stores of the kernel's outputs as non-temporal (streaming) writes: the
// apply's y (MODE 0 and 1; plain applies stream from HBM 0.7 us faster at
// 256x256, N = 7, MALL-resident ones 1 us slower) and the PCG iteration's p,
// x and w (MODE 2; off: its outputs are read again at once).
#ifndef SEM_AX_NT_STORE
#define SEM_AX_NT_STORE 1
#endif
#ifndef SEM_AX_NT_PCG
#define SEM_AX_NT_PCG 0
#endif
// the seam pass's loads and stores of y as non-temporal ones (plain apply;
// off: 40.0 vs 38.7 us at 256x256, N = 7)
#ifndef SEM_SEAMS_NT
#define SEM_SEAMS_NT 0
#endif
// unsummed seam copies (re-read by the seam pass) with ordinary stores: with
// the factors' non-temporal loads 38.65 vs 38.7 us at 256x256, N = 7 (42.7 vs
// 42.3 us before them)
#ifndef SEM_AX_SEAM_CACHED
#define SEM_AX_SEAM_CACHED 1
#endif
// the geometric factors and mass read with non-temporal loads (MODE 0 and 1;
// MODE 2), and u (MODE 0 and 1). At 256x256, N = 7 from HBM: factors in
// the apply 42.3 -> 38.9 us (MALL-resident: 36.7 -> 39.3 us); factors in the
// PCG iteration 88.6 -> 89.4 us; u as well 40.7 us. Only the first is on.
#ifndef SEM_AX_NT_LOAD_G
#define SEM_AX_NT_LOAD_G 1
#endif
#ifndef SEM_AX_NT_LOAD_G_PCG
#define SEM_AX_NT_LOAD_G_PCG 0
#endif
#ifndef SEM_AX_NT_LOAD_U
#define SEM_AX_NT_LOAD_U 0
#endif
#ifndef SEM_AX_THREADS
#define SEM_AX_THREADS 256
#endif
#ifndef SEM_AX_EARLYG_BYTES
#define SEM_AX_EARLYG_BYTES 80
#endif
#ifndef SEM_AX_INPLACE
#define SEM_AX_INPLACE 1
#endif

namespace sem {
namespace ax {

template <typename T>
struct AxP {
  const T* __restrict__ u;
  T* __restrict__ y;
  const T* __restrict__ G11;
  const T* __restrict__ G12;
  const T* __restrict__ G22;
  // The single-pass FP16/BF16 operators store G11, G12, G22 and B as FP16,
  // each scaled by a power of two (gs = the inverse scales, exact)
  const _Float16* __restrict__ h11;
  const _Float16* __restrict__ h12;
  const _Float16* __restrict__ h22;
  const _Float16* __restrict__ hB;
  float gs[4];
  const T* __restrict__ B;
  const T* __restrict__ Dr;   // P x P column-major: Dr(i + P*j) = Dr_ij
  const T* __restrict__ Ds;
  const double* __restrict__ zr;   // GLL nodes and weights (deflated MFMA forms)
  const double* __restrict__ wr;
  const T* __restrict__ nu_f;
  const T* __restrict__ k_f;
  T nu, k;
  int has_mass;
  int Ex, Ey, S, nstrips;
  int per_x, per_y, bc;
  // MODE 2 (fused PCG iteration): u is formed on load as z + beta p_old
  // (z = d .* r or r) and written back to p; <u, y> partials go to pcg.Q
  const T* __restrict__ r;
  const T* __restrict__ d;
  T* __restrict__ p;
  // x += alpha_{k-1} p_{k-1} while p_{k-1} is read (the update kernel
  // leaves x alone)
  T* __restrict__ x;
  PcgDev pcg;
  // fused epilogue (EPI 1 to 3; see ApplyEpi): er, ex, ed and y are written, u is
  // only read, and none of them alias
  T* er;
  T* ex;
  const T* edinv;
  T ec1, ec2;
  T* ed;       // EPI 3: the direction written by the smoother's start
  T etheta;
};

// ---- MFMA contractions for 16x16 element blocks (MF != 0) ----------------
// One v_mfma_f32_16x16x16{f16,bf16} per element contraction (K = 16); lane l
// holds A[l%16][4(l/16)+q], B[4(l/16)+q][l%16] and D[4(l/16)+i][l%16].
// MF = 1: FP16 inputs; 2: BF16 inputs; 3 / 4: FP16 / BF16 inputs split as
// hi + lo with three products (lo*hi + hi*lo + hi*hi); accumulation is FP32.
// MF = 5 / 6: FP32 / FP64 inputs on v_mfma_{f32,f64}_16x16x4 (four K-blocks;
// lane l holds A[l%16][4kb + l/16] and B[4kb + l/16][l%16]), exact in the
// field precision: the 16x16 contractions without per-lane D constants. The
// FP64 form returns D interleaved: register r of lane l is D[l/16 + 4r][l%16].
typedef _Float16 mf_h4 __attribute__((ext_vector_type(4)));
typedef short mf_s4 __attribute__((ext_vector_type(4)));
typedef float mf_f4 __attribute__((ext_vector_type(4)));

template <int MF>
struct MfFrag {
  using type = mf_h4;
};
template <>
struct MfFrag<2> {
  using type = mf_s4;
};
template <>
struct MfFrag<4> {
  using type = mf_s4;
};

__device__ __forceinline__ unsigned short mf_bf16(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
  return (unsigned short)(u >> 16);
}

template <int MF>
__device__ __forceinline__ void mf_pack(const float v[4], typename MfFrag<MF>::type& hi,
                                        typename MfFrag<MF>::type& lo) {
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    if constexpr (MF == 1 || MF == 3) {
      const _Float16 h = (_Float16)v[q];
      hi[q] = h;
      if constexpr (MF == 3) lo[q] = (_Float16)(v[q] - (float)h);
    } else {
      const unsigned short b = mf_bf16(v[q]);
      hi[q] = (short)b;
      if constexpr (MF == 4) lo[q] = (short)mf_bf16(v[q] - __uint_as_float((unsigned)b << 16));
    }
  }
}

template <int MF>
__device__ __forceinline__ mf_f4 mf_mma(typename MfFrag<MF>::type a, typename MfFrag<MF>::type b,
                                        mf_f4 c) {
  if constexpr (MF == 1 || MF == 3) return __builtin_amdgcn_mfma_f32_16x16x16f16(a, b, c, 0, 0, 0);
  else return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

typedef double mf_d4 __attribute__((ext_vector_type(4)));

// 16x16x16 product from four 16x16x4 steps: a[kb] = A[m][4kb + kg],
// b[kb] = B[4kb + kg][n]
__device__ __forceinline__ mf_f4 mf_contract_k4(const float a[4], const float b[4]) {
  mf_f4 c = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[kb], b[kb], c, 0, 0, 0);
  return c;
}
__device__ __forceinline__ mf_d4 mf_contract_k4(const double a[4], const double b[4]) {
  mf_d4 c = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int kb = 0; kb < 4; ++kb) c = __builtin_amdgcn_mfma_f64_16x16x4f64(a[kb], b[kb], c, 0, 0, 0);
  return c;
}

template <int MF>
__device__ __forceinline__ mf_f4 mf_contract(typename MfFrag<MF>::type ah,
                                             typename MfFrag<MF>::type al,
                                             typename MfFrag<MF>::type bh,
                                             typename MfFrag<MF>::type bl) {
  mf_f4 c = {0.f, 0.f, 0.f, 0.f};
  if constexpr (MF >= 3) {
    c = mf_mma<MF>(al, bh, c);
    c = mf_mma<MF>(ah, bl, c);
  }
  return mf_mma<MF>(ah, bh, c);
}

// MODE 0: element-local H u (no assembly); MODE 1: assembled + masked;
// MODE 2: one fused PCG iteration body (see AxP). MF selects the MFMA
// contraction path (T = float, P = 16 only).
// RY = 2 or 4 (assembled apply or PCG iteration on the FMA path; 4 in the
// apply only): the workgroup's threads form
// RY groups of CX / RY elements each, on RY consecutive element rows processed
// side by side; the y-pairs between them are summed through LDS, so a segment
// is RY element rows and 1 / RY of the segment boundaries of one-row
// workgroups remain for k_seams, while every thread still issues one element
// row's loads.
template <typename T, int P, int CX, int MODE, int MF = 0, int RY = 1, int EPI = 0>
__global__ void __launch_bounds__(CX* P)
    __attribute__((amdgpu_waves_per_eu((MF == 1 || MF == 2) ? 3 : 1))) k_ax(const AxP<T> prm) {
  constexpr int NR = CX * P;
  constexpr int RS = NR + 1;
  constexpr int CXr = CX / RY;   // elements per strip row
  constexpr int NRr = CXr * P;   // threads per element row
  static_assert(RY == 1 || (MF == 0 && CX % RY == 0 &&
                            ((RY == 2 && MODE != 0) || (RY == 4 && MODE == 1))),
                "rows side by side: assembled FMA apply, strip width a multiple of RY");
  static_assert(MF == 0 || (P == 16 && NR == 256), "MFMA path: 16x16 blocks");
  static_assert(EPI == 0 || (MODE == 1 && MF == 0 && RY == 2), "fused epilogue: two-row apply");
  static_assert(MF == 0 || MF == 6 || sizeof(T) == 4, "MFMA modes 1..5 take FP32 fields");
  static_assert(MF != 6 || sizeof(T) == 8, "MFMA mode 6 takes FP64 fields");
  // The FMA path (MF == 0) runs C1, R2 and C2 in place: in each of them a
  // thread reads a set of s0 slots and then rewrites exactly that set (C1/C2:
  // column a of element e; R2: row tid), so s1 aliases s0 there and the LDS
  // per workgroup drops by P * RS values. SEM_AX_INPLACE=0 keeps two buffers.
  constexpr bool kInPlace = MF == 0 && SEM_AX_INPLACE;
  __shared__ T s0[P * RS];
  __shared__ T s1_[kInPlace ? 1 : P * RS];
  T* const s1 = kInPlace ? s0 : s1_;
  // x-pair exchange buffer (R3 reads s1, which is s0 on the in-place path)
  __shared__ T xbuf[(MF || kInPlace) ? 2 * NR : 1];
  __shared__ double red[NR];
  // PCG mode: the max tree of the partials and the stop flag
  __shared__ double red2[MODE == 2 ? NR : 1];
  __shared__ int flag;
  // RY > 1: each row's top line, then each row's bottom line
  __shared__ T ybuf[RY > 1 ? 2 * NR : 1];
  // MFMA A-operands: D and D' (Dr = Ds); hi and lo parts (MF 1..4) or the
  // four K-blocks (MF 5, 6)
  using Frag = typename MfFrag<(MF >= 1 && MF <= 4) ? MF : 1>::type;
  Frag aD, aDl, aT, aTl;
  T aDk[4], aTk[4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  // Single-pass FP16/BF16 forms (MF 1, 2) deflate every contracted 1-D
  // vector v by its linear interpolant l(z) = c0 + c1 z through the two end
  // nodes before rounding it: A v = A (v - l) + c0 (A 1) + c1 (A z), the last
  // two terms exact in FP32. For smooth fields v - l is O(h^2 v'') instead of
  // O(v), which removes the rounding error of the constant and linear parts
  // (the modes a low-precision operator corrupts most, relative to A v).
  // Derivatives: A = D, D 1 = 0, D z = 1. Weak-form transposes: D' acts on
  // fluxes that carry the GLL weight of the contracted index, so the
  // operand is A = D' diag(w / wm) on f = flux .* (wm / w) (wm = max w keeps
  // f in FP16 range), with A 1 = (e_N - e_0) / wm and
  // A z = (e_N + e_0 - w) / wm (GLL quadrature is exact for both integrals).
  constexpr bool kDefl = MF == 1 || MF == 2;
  // deflation tables (nodes z_k, wm / w_k, (A z)_k, (A 1)_k), in LDS rather
  // than registers: the PCG-mode kernel sits at the register limit of three
  // waves per SIMD
  __shared__ float4 sdefl[kDefl ? 16 : 1];
  if constexpr (MF >= 1 && MF <= 4) {
    const int m = lane & 15, kg = lane >> 4;
    float dv[4], tv[4];
    double wm = 0.0;
    if constexpr (kDefl)
      for (int q = 0; q < 16; ++q) wm = fmax(wm, prm.wr[q]);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int k = 4 * kg + q;
      dv[q] = (float)prm.Dr[m + 16 * k];
      tv[q] = kDefl ? (float)(prm.Dr[k + 16 * m] * (prm.wr[k] / wm)) : (float)prm.Dr[k + 16 * m];
    }
    if (kDefl && threadIdx.x < 16) {
      const int k = threadIdx.x;
      const double eN = k == 15 ? 1.0 : 0.0, e0 = k == 0 ? 1.0 : 0.0;
      sdefl[k] = make_float4((float)prm.zr[k], (float)(wm / prm.wr[k]),
                             (float)((eN + e0 - prm.wr[k]) / wm), (float)((eN - e0) / wm));
    }
    mf_pack<MF>(dv, aD, aDl);
    mf_pack<MF>(tv, aT, aTl);
  } else if constexpr (MF >= 5) {
    const int m = lane & 15, kg = lane >> 4;
#pragma unroll
    for (int kb = 0; kb < 4; ++kb) {
      aDk[kb] = prm.Dr[m + 16 * (4 * kb + kg)];   // D[m][4kb + kg]
      aTk[kb] = prm.Dr[(4 * kb + kg) + 16 * m];   // D'[m][4kb + kg]
    }
  }

  // geometric factors and mass in the operator's storage precision
  constexpr bool kHalfG = MF == 1 || MF == 2;
  auto ld_g11 = [&](int idx) -> T {
    if constexpr (kHalfG) return T((float)prm.h11[idx] * prm.gs[0]);
    else if constexpr (MODE == 2 ? SEM_AX_NT_LOAD_G_PCG : SEM_AX_NT_LOAD_G) return __builtin_nontemporal_load(&prm.G11[idx]);
    else return prm.G11[idx];
  };
  auto ld_g12 = [&](int idx) -> T {
    if constexpr (kHalfG) return T((float)prm.h12[idx] * prm.gs[1]);
    else if constexpr (MODE == 2 ? SEM_AX_NT_LOAD_G_PCG : SEM_AX_NT_LOAD_G) return __builtin_nontemporal_load(&prm.G12[idx]);
    else return prm.G12[idx];
  };
  auto ld_g22 = [&](int idx) -> T {
    if constexpr (kHalfG) return T((float)prm.h22[idx] * prm.gs[2]);
    else if constexpr (MODE == 2 ? SEM_AX_NT_LOAD_G_PCG : SEM_AX_NT_LOAD_G) return __builtin_nontemporal_load(&prm.G22[idx]);
    else return prm.G22[idx];
  };
  auto ld_b = [&](int idx) -> T {
    if constexpr (kHalfG) return T((float)prm.hB[idx] * prm.gs[3]);
    else if constexpr (MODE == 2 ? SEM_AX_NT_LOAD_G_PCG : SEM_AX_NT_LOAD_G) return __builtin_nontemporal_load(&prm.B[idx]);
    else return prm.B[idx];
  };

  const int tid = threadIdx.x;
  constexpr bool kNT = (MODE == 2) ? SEM_AX_NT_PCG : SEM_AX_NT_STORE;
  auto st_out = [&](T* dst, T v) {
    if constexpr (kNT) __builtin_nontemporal_store(v, dst);
    else *dst = v;
  };
  auto st_seam = [&](T* dst, T v) {
    if constexpr (kNT && !SEM_AX_SEAM_CACHED) __builtin_nontemporal_store(v, dst);
    else *dst = v;
  };
  // PCG: the iterate is advanced here, one iteration late
  constexpr bool kXLate = MODE == 2;
  T beta = T(0), alpha_prev = T(0);
  bool first = false;
  if (MODE == 2) {
    double t, res;
    bool stopped;
    reduce_partials_flag(prm.pcg.part_prev, prm.pcg.n3, red, red2, &flag, NR, prm.pcg.state, &t,
                         &res, &stopped);
    if (stopped) return;
    const int k = prm.pcg.k;
    // tol < 0: relative stop test norm(r, Inf) <= |tol| norm(b, Inf), where
    // norm(b, Inf) is iteration 1's res (x0 = 0, r = b); iteration 1 keeps
    // the absolute value in the state for the later launches
    double tol = prm.pcg.tol;
    if (tol < 0.0) {
      tol = (k == 1) ? -tol * res : prm.pcg.state->tol_eff;
      if (k == 1 && blockIdx.x == 0 && tid == 0) prm.pcg.state->tol_eff = tol;
    }
    const bool done = !(res > tol) || (k - 1 == prm.pcg.maxiter);
    if (done) {
      if (blockIdx.x == 0 && tid == 0) {
        prm.pcg.state->iters = k - 1;
        prm.pcg.state->converged = res <= tol ? 1 : 0;
        prm.pcg.state->res = res;
        prm.pcg.state->done = 1;
      }
      return;
    }
    first = (k == 1);
    if (!first) beta = T(t / prm.pcg.state->t[(k - 1) & 1]);
    if (kXLate && !first) alpha_prev = T(prm.pcg.state->alpha[(k - 1) & 1]);
    if (blockIdx.x == 0 && tid == 0) prm.pcg.state->t[k & 1] = t;
  }
  double dot = 0.0;
  T carry_u = T(0);
  const int strip = blockIdx.x % prm.nstrips;
  const int seg = blockIdx.x / prm.nstrips;
  const int Ex = prm.Ex, Ey = prm.Ey;
  const int L = P * Ex;
  const int NC = P * Ey;
  const int ex0 = strip * CXr;
  const int ncx = min(CXr, Ex - ex0);
  // row group hh of the workgroup (RY > 1), its thread tr = (e, a), and eo =
  // the element's slot in the workgroup's LDS rows
  const int hh = RY > 1 ? tid / NRr : 0;
  const int tr = tid - hh * NRr;
  const int e = tr / P;
  const int a = tr - e * P;
  const int eo = hh * CXr + e;
  const int ey0 = seg * prm.S;
  const int ey1 = min(Ey, ey0 + prm.S);
  // this thread's element rows: the whole segment, or row ey0 + hh
  const int eyA = RY > 1 ? ey0 + hh : ey0;
  const int eyB = RY > 1 ? eyA + 1 : ey1;
  const bool rowact = RY == 1 || eyA < Ey;
  const bool act = e < ncx && rowact;
  const int i = ex0 * P + tr;
  // Dr = Ds = derivMat(gausslobatto(P)) (mesh.jl:70-74, nr == ns) is a
  // compile-time table generated by the library's own host routine (the
  // mesh checks bit-equality at creation), so every coefficient is an
  // immediate the compiler rematerialises in SGPRs: no loads, no spills.
#define Dr(ix) T(gll_d<P>(ix))
#define Ds(ix) T(gll_d<P>(ix))

  // x-edge classification of this thread's row
  bool xseam = false, xpair = false;
  int xside = -1;
  if (MODE != 0 && act) {
    const int ex = ex0 + e;
    if (a == 0) {
      xside = 0;
      if (ex > 0 || prm.per_x) {
        if (e > 0) xpair = true; else xseam = true;
      }
    } else if (a == P - 1) {
      xside = 1;
      if (ex < Ex - 1 || prm.per_x) {
        if (e < ncx - 1) xpair = true; else xseam = true;
      }
    }
  }

  // Inactive threads (past the last element of a partial strip) load from a
  // valid row so that every load is unconditional and the compiler can issue
  // a phase's loads back to back; their results are never stored.
  const int ic = act ? i : L - 1;
  // G is staged in registers one phase early when it fits, so its latency
  // hides under the C1 contraction.
  constexpr bool kEarlyG = P * sizeof(T) <= SEM_AX_EARLYG_BYTES;
  const bool plain = prm.nu_f == nullptr && prm.k_f == nullptr;

  T carry = T(0);
  T carry_hi = T(0);   // RY > 1: the copy of the top line from the row above
  for (int ey = eyA; ey < eyB; ++ey) {
    const int jb = ey * P;
    const int base = i + L * jb;
    // an idle row group (past Ey) loads the last element row
    const int bc_ = ic + L * (rowact ? jb : (Ey - 1) * P);
    T uv[P], us[P];
    // the row's geometric factors, issued right after its other loads so
    // that all of them are in flight together (their use is two phases later)
    T g11[kEarlyG ? P : 1], g12[kEarlyG ? P : 1], g22[kEarlyG ? P : 1];
    auto load_g = [&]() {
      if (kEarlyG) {
#pragma unroll
        for (int b = 0; b < (kEarlyG ? P : 1); ++b) {
          g11[b] = ld_g11(bc_ + L * b);
          g12[b] = ld_g12(bc_ + L * b);
          g22[b] = ld_g22(bc_ + L * b);
        }
      }
    };
    // R1: load row a of element e, s-contraction us(a,b) = sum_k u(a,k) Ds_{b,k}
    if (MODE == 2) {
      T rv[P], zv[P];
#pragma unroll
      for (int b = 0; b < P; ++b) rv[b] = prm.r[bc_ + L * b];
      if (prm.d) {
#pragma unroll
        for (int b = 0; b < P; ++b) zv[b] = prm.d[bc_ + L * b] * rv[b];
      } else {
#pragma unroll
        for (int b = 0; b < P; ++b) zv[b] = rv[b];
      }
      if (first) {
#pragma unroll
        for (int b = 0; b < P; ++b) uv[b] = zv[b];
      } else if constexpr (kXLate) {
        T po[P], xo[P];
#pragma unroll
        for (int b = 0; b < P; ++b) po[b] = prm.p[bc_ + L * b];
#pragma unroll
        for (int b = 0; b < P; ++b) xo[b] = prm.x[bc_ + L * b];
#pragma unroll
        for (int b = 0; b < P; ++b) uv[b] = zv[b] + beta * po[b];
        if (act) {
#pragma unroll
          for (int b = 0; b < P; ++b) st_out(&prm.x[base + L * b], xo[b] + alpha_prev * po[b]);
        }
      } else {
#pragma unroll
        for (int b = 0; b < P; ++b) uv[b] = zv[b] + beta * prm.p[bc_ + L * b];
      }
      if (act) {
#pragma unroll
        for (int b = 0; b < P; ++b) st_out(&prm.p[base + L * b], uv[b]);
      }
    } else {
#pragma unroll
      for (int b = 0; b < P; ++b) {
        if constexpr (SEM_AX_NT_LOAD_U) uv[b] = __builtin_nontemporal_load(&prm.u[bc_ + L * b]);
        else uv[b] = prm.u[bc_ + L * b];
      }
      load_g();
    }
    if constexpr (MF == 0) {
#pragma unroll
      for (int b = 0; b < P; ++b) {
        T acc = T(0);
#pragma unroll
        for (int k = 0; k < P; ++k) acc += uv[k] * Ds(b + P * k);
        us[b] = acc;
      }
    }
#pragma unroll
    for (int b = 0; b < P; ++b) s0[b * RS + tid] = uv[b];
    // PCG mode: after the branches of R1 (a load issued before a branch
    // would be waited for where its paths join)
    if (MODE == 2) load_g();
    __syncthreads();
    if constexpr (MF != 0) {
      // M1: per element (wave wv takes elements wv, wv+4, ...):
      // ur = D U -> s1 [b][r];  us' = D U' -> s0 [b][r] (over u, after it is read)
      const int n = lane & 15, kg = lane >> 4;
      for (int e2 = wv; e2 < ncx; e2 += 4) {
        if constexpr (MF <= 4) {
          float bu[4], bt[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            bu[q] = (float)s0[n * RS + e2 * 16 + 4 * kg + q];
            bt[q] = (float)s0[(4 * kg + q) * RS + e2 * 16 + n];
          }
          float c1u = 0.f, c1t = 0.f;
          if constexpr (kDefl) {
            float dz[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) dz[q] = sdefl[4 * kg + q].x;
            // end nodes of column n (r-contraction) and of row n (s-contraction)
            const float u0 = (float)s0[n * RS + e2 * 16], u1 = (float)s0[n * RS + e2 * 16 + 15];
            const float t0 = (float)s0[e2 * 16 + n], t1 = (float)s0[15 * RS + e2 * 16 + n];
            const float c0u = 0.5f * (u0 + u1), c0t = 0.5f * (t0 + t1);
            c1u = 0.5f * (u1 - u0);
            c1t = 0.5f * (t1 - t0);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              bu[q] = bu[q] - c0u - c1u * dz[q];
              bt[q] = bt[q] - c0t - c1t * dz[q];
            }
          }
          Frag bh, bl, th, tl;
          mf_pack<MF>(bu, bh, bl);
          mf_pack<MF>(bt, th, tl);
          const mf_f4 ur4 = mf_contract<MF>(aD, aDl, bh, bl);
          const mf_f4 us4 = mf_contract<MF>(aD, aDl, th, tl);
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            s1[n * RS + e2 * 16 + 4 * kg + q] = kDefl ? ur4[q] + c1u : ur4[q];
            s0[(4 * kg + q) * RS + e2 * 16 + n] = kDefl ? us4[q] + c1t : us4[q];
          }
        } else {
          T bu[4], bt[4];
#pragma unroll
          for (int kb = 0; kb < 4; ++kb) {
            bu[kb] = s0[n * RS + e2 * 16 + 4 * kb + kg];
            bt[kb] = s0[(4 * kb + kg) * RS + e2 * 16 + n];
          }
          const auto ur4 = mf_contract_k4(aDk, bu);
          const auto us4 = mf_contract_k4(aDk, bt);
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int row = MF == 6 ? kg + 4 * q : 4 * kg + q;
            s1[n * RS + e2 * 16 + row] = ur4[q];
            s0[row * RS + e2 * 16 + n] = us4[q];
          }
        }
      }
    } else {
    // C1: column owner (e, b'=a): ur(:, b') = Dr u(:, b')
      T v[P];
#pragma unroll
      for (int k = 0; k < P; ++k) v[k] = s0[a * RS + eo * P + k];
#pragma unroll
      for (int a2 = 0; a2 < P; ++a2) {
        T acc = T(0);
#pragma unroll
        for (int k = 0; k < P; ++k) acc += Dr(a2 + P * k) * v[k];
        s1[a * RS + eo * P + a2] = acc;
      }
    }
    __syncthreads();
    // R2: wr = G11 ur + G12 us, ws = G12 ur + G22 us; part2(a,b) = sum_k ws(a,k) Ds_{k,b}
    T p2[P];
    {
      T wsv[P];
#pragma unroll
      for (int b = 0; b < P; ++b) {
        const T ur = s1[b * RS + tid];
        T x11, x12, x22;
        if (kEarlyG) {
          x11 = g11[kEarlyG ? b : 0];
          x12 = g12[kEarlyG ? b : 0];
          x22 = g22[kEarlyG ? b : 0];
        } else {
          x11 = ld_g11(bc_ + L * b);
          x12 = ld_g12(bc_ + L * b);
          x22 = ld_g22(bc_ + L * b);
        }
        const T usb = MF ? s0[b * RS + tid] : us[b];
        const T wr = x11 * ur + x12 * usb;
        wsv[b] = x12 * ur + x22 * usb;
        s0[b * RS + tid] = wr;
        if constexpr (MF != 0) s1[b * RS + tid] = wsv[b];
      }
      if constexpr (MF == 0) {
#pragma unroll
        for (int b = 0; b < P; ++b) {
          T acc = T(0);
#pragma unroll
          for (int k = 0; k < P; ++k) acc += wsv[k] * Ds(k + P * b);
          p2[b] = acc;
        }
      }
    }
    // B is loaded unconditionally (G22 stands in without a mass term: those
    // lines were just read, so they come from cache): a load under a runtime
    // branch would make the compiler wait for every outstanding load at the
    // join, before the barrier, instead of letting B's latency hide under C2
    T bv[P];
    if constexpr (kHalfG) {
      const _Float16* bsrc = (prm.has_mass && plain) ? prm.hB : prm.h22;
      const float sc = (prm.has_mass && plain) ? prm.gs[3] : prm.gs[2];
#pragma unroll
      for (int b = 0; b < P; ++b) bv[b] = T((float)bsrc[bc_ + L * b] * sc);
    } else {
      const T* bsrc = (prm.has_mass && plain) ? prm.B : prm.G22;
#pragma unroll
      for (int b = 0; b < P; ++b) {
        if constexpr (MODE == 2 ? SEM_AX_NT_LOAD_G_PCG : SEM_AX_NT_LOAD_G) bv[b] = __builtin_nontemporal_load(&bsrc[bc_ + L * b]);
        else bv[b] = bsrc[bc_ + L * b];
      }
    }
    __syncthreads();
    if constexpr (MF != 0) {
      // M2: part1 = D' wr -> s1 [b][r]; part2' = D' ws' -> s0 [b][r] (after the reads)
      const int n = lane & 15, kg = lane >> 4;
      for (int e2 = wv; e2 < ncx; e2 += 4) {
        if constexpr (MF <= 4) {
          float bw[4], bs[4];
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            bw[q] = (float)s0[n * RS + e2 * 16 + 4 * kg + q];
            bs[q] = (float)s1[(4 * kg + q) * RS + e2 * 16 + n];
          }
          float c0w = 0.f, c1w = 0.f, c0s = 0.f, c1s = 0.f;
          float4 dq[4];
          if constexpr (kDefl) {
#pragma unroll
            for (int q = 0; q < 4; ++q) dq[q] = sdefl[4 * kg + q];
            const float dwi0 = sdefl[0].y, dwi15 = sdefl[15].y;
            const float w0 = (float)s0[n * RS + e2 * 16] * dwi0;
            const float w1 = (float)s0[n * RS + e2 * 16 + 15] * dwi15;
            const float v0 = (float)s1[e2 * 16 + n] * dwi0;
            const float v1 = (float)s1[15 * RS + e2 * 16 + n] * dwi15;
            c0w = 0.5f * (w0 + w1);
            c1w = 0.5f * (w1 - w0);
            c0s = 0.5f * (v0 + v1);
            c1s = 0.5f * (v1 - v0);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              bw[q] = bw[q] * dq[q].y - c0w - c1w * dq[q].x;
              bs[q] = bs[q] * dq[q].y - c0s - c1s * dq[q].x;
            }
          }
          Frag wh, wl, sh, sl;
          mf_pack<MF>(bw, wh, wl);
          mf_pack<MF>(bs, sh, sl);
          const mf_f4 p1 = mf_contract<MF>(aT, aTl, wh, wl);
          const mf_f4 p2t = mf_contract<MF>(aT, aTl, sh, sl);
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            s1[n * RS + e2 * 16 + 4 * kg + q] =
                kDefl ? p1[q] + (c0w * dq[q].w + c1w * dq[q].z) : p1[q];
            s0[(4 * kg + q) * RS + e2 * 16 + n] =
                kDefl ? p2t[q] + (c0s * dq[q].w + c1s * dq[q].z) : p2t[q];
          }
        } else {
          T bw[4], bs[4];
#pragma unroll
          for (int kb = 0; kb < 4; ++kb) {
            bw[kb] = s0[n * RS + e2 * 16 + 4 * kb + kg];
            bs[kb] = s1[(4 * kb + kg) * RS + e2 * 16 + n];
          }
          const auto p1 = mf_contract_k4(aTk, bw);
          const auto p2t = mf_contract_k4(aTk, bs);
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            const int row = MF == 6 ? kg + 4 * q : 4 * kg + q;
            s1[n * RS + e2 * 16 + row] = p1[q];
            s0[row * RS + e2 * 16 + n] = p2t[q];
          }
        }
      }
    } else {
    // C2: column owner: part1(:, b') = Dr' wr(:, b')
      T v[P];
#pragma unroll
      for (int k = 0; k < P; ++k) v[k] = s0[a * RS + eo * P + k];
#pragma unroll
      for (int a2 = 0; a2 < P; ++a2) {
        T acc = T(0);
#pragma unroll
        for (int k = 0; k < P; ++k) acc += Dr(k + P * a2) * v[k];
        s1[a * RS + eo * P + a2] = acc;
      }
    }
    __syncthreads();
    // fused epilogue operands of this thread's copies, in flight while H
    // and the pair sums are formed (clamped rows for inactive threads)
    EpiIn<T> ein[EPI ? P : 1];
    if constexpr (EPI != 0) {
#pragma unroll
      for (int b = 0; b < P; ++b) ein[b] = epi_load<T, EPI>(bc_ + L * b, prm.er, prm.ex, prm.edinv);
    }
    // R3: H = nu (part1 + part2) + k (B u)
    T h[P];
    if (plain) {
      if (prm.has_mass) {
#pragma unroll
        for (int b = 0; b < P; ++b)
          h[b] = prm.nu * (s1[b * RS + tid] + (MF ? s0[b * RS + tid] : p2[b])) +
                 prm.k * (bv[b] * uv[b]);
      } else {
#pragma unroll
        for (int b = 0; b < P; ++b)
          h[b] = prm.nu * (s1[b * RS + tid] + (MF ? s0[b * RS + tid] : p2[b]));
      }
    } else {
#pragma unroll
      for (int b = 0; b < P; ++b) {
        const int idx = bc_ + L * b;
        const T au = s1[b * RS + tid] + (MF ? s0[b * RS + tid] : p2[b]);
        const T nuv = prm.nu_f ? prm.nu_f[idx] : prm.nu;
        T hv = nuv * au;
        if (prm.has_mass) {
          const T kv = prm.k_f ? prm.k_f[idx] : prm.k;
          hv = hv + kv * (ld_b(idx) * uv[b]);
        }
        h[b] = hv;
      }
    }
    if (MODE == 2 && act) {
      // <p, M QQ'(H)>_mult = sum over all local copies of p .* H for a
      // continuous, masked p, so the partial needs no assembled values
#pragma unroll
      for (int b = 0; b < P; ++b) dot += (double)uv[b] * (double)h[b];
    }
    if (MODE == 0) {
      if (act) {
#pragma unroll
        for (int b = 0; b < P; ++b) st_out(&prm.y[base + L * b], h[b]);
      }
      if constexpr (MF != 0) __syncthreads();  // s0, s1 were read in R3
      continue;
    }
    // x-pairs inside the strip through LDS: s0 on the two-buffer FMA path
    // (s0 was last read in C2), else the exchange buffer (R3 reads s1, which
    // is s0 in place, and both buffers on the MFMA path)
    T* xb = (MF || kInPlace) ? xbuf : s0;
    if (xside >= 0) {
#pragma unroll
      for (int b = 0; b < P; ++b) xb[(xside * CX + eo) * P + b] = h[b];
    }
    __syncthreads();
    if (xpair) {
      if (xside == 1) {
#pragma unroll
        for (int b = 0; b < P; ++b) h[b] = h[b] + xb[(eo + 1) * P + b];
      } else {
#pragma unroll
        for (int b = 0; b < P; ++b) h[b] = xb[(CX + eo - 1) * P + b] + h[b];
      }
    }
    if constexpr (RY > 1) {
      // y-pairs of the rows: every row publishes its top and bottom lines
      // (x-pairs already summed) and takes the partner copies: carry from the
      // row below, carry_hi from the row above; each row writes only its own
      // copies
      ybuf[hh * NRr + tr] = h[P - 1];
      ybuf[NR + hh * NRr + tr] = h[0];
      __syncthreads();
      if (hh > 0) carry = ybuf[(hh - 1) * NRr + tr];
      if (hh < RY - 1) carry_hi = ybuf[NR + (hh + 1) * NRr + tr];
    }
    if (act) {
      const int bc = prm.bc;
      // final (masked) write of a node copy; with a fused epilogue the
      // final value feeds it instead (uin: the copy's input value; eb: its
      // slot among this thread's preloaded epilogue operands)
      auto fin = [&](int idx, int j, T val, T uin, int eb) {
        const T wv = val * mval<T>(i, j, L, NC, bc);
        if constexpr (EPI == 0) st_out(&prm.y[idx], wv);
        else epi_store<T, EPI>(idx, wv, uin, ein[eb], prm.er, prm.ex, prm.ec1, prm.ec2, prm.y,
                               prm.ed, prm.etheta);
      };
#pragma unroll
      for (int b = 0; b < P; ++b) {
        const int j = jb + b;
        const int idx = base + L * b;
        if (b == 0) {
          const bool ysh = ey > 0 || prm.per_y;
          if (ysh) {
            if (ey > ey0) {
              if (xseam) {
                st_seam(&prm.y[idx], h[0]);
                if (RY == 1) st_seam(&prm.y[idx - L], carry);
              } else {
                const T s = carry + h[0];
                fin(idx, j, s, uv[0], 0);
                if (RY == 1) fin(idx - L, j - 1, s, carry_u, 0);
              }
            } else {
              st_seam(&prm.y[idx], h[0]);  // y seam: finished by k_seams
            }
          } else {
            if (xseam) st_seam(&prm.y[idx], h[0]);
            else fin(idx, j, h[0], uv[0], 0);
          }
        } else if (b == P - 1) {
          const bool ysh = ey < Ey - 1 || prm.per_y;
          if (ysh) {
            if (ey < ey1 - 1) {
              if constexpr (RY > 1) {
                if (xseam) st_seam(&prm.y[idx], h[P - 1]);
                else fin(idx, j, h[P - 1] + carry_hi, uv[P - 1], P - 1);
              } else {
                carry = h[P - 1];
                carry_u = uv[P - 1];
              }
            } else {
              st_seam(&prm.y[idx], h[P - 1]);  // y seam
            }
          } else {
            if (xseam) st_seam(&prm.y[idx], h[P - 1]);
            else fin(idx, j, h[P - 1], uv[P - 1], P - 1);
          }
        } else {
          if (xseam) st_seam(&prm.y[idx], h[b]);
          else fin(idx, j, h[b], uv[b], b);
        }
      }
    }
    __syncthreads();  // s0 is rewritten by the next element row
  }
  if (MODE == 2) {
    const double tot = block_sum(dot, red, NR);
    if (tid == 0) prm.pcg.Q[blockIdx.x] = tot;
  }
#undef Dr
#undef Ds
}

// Finishes the node copies the strips could not sum: y seams (segment
// boundaries and the periodic y wrap), then x seams (strip boundaries and the
// periodic x wrap). Grid row blockIdx.y is one seam line: y seam g < nys
// (threads over the L rows), else x seam s = blockIdx.y - nys (threads over
// the NC columns); the line's operands are uniform per block and there is
// no per-thread integer division. Items touch disjoint copies, so there are
// no races. The kernel is latency-bound (a few MB in thousands of short
// blocks): 128-thread blocks measured best (1 item per thread beat 4).
constexpr int kSeamBlock = 128;
// EPI: the fused epilogue of the apply (see ApplyEpi) on the copies finished
// here; u is the apply's input (the old direction for EPI 1).
template <typename T, int EPI = 0>
__global__ void __launch_bounds__(kSeamBlock)
k_seams(T* y, int L, int NC, int P, int CXP, int nstrips, int SP, int nsegs, int per_x, int per_y,
        int bc, int nys, const T* u, T* er, T* ex, const T* edinv, T ec1, T ec2, T* ed, T etheta) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  // a copy's epilogue operands (and input value) are loaded together with
  // the seam copies, then the update is stored from the final value
  struct Copy {
    long t;
    EpiIn<T> e;
    T u;
  };
  auto pre = [&](int i, int j) {
    Copy c{};
    c.t = i + (long)L * j;
    if constexpr (EPI != 0) {
      c.e = epi_load<T, EPI>(c.t, er, ex, edinv);
      if constexpr (EPI == 1) c.u = u[c.t];
    }
    return c;
  };
  auto ly = [&](long t) -> T {
    if constexpr (SEM_SEAMS_NT && EPI == 0) return __builtin_nontemporal_load(&y[t]);
    else return y[t];
  };
  auto fin = [&](int i, int j, const Copy& c, T s) {
    const T wv = s * mval<T>(i, j, L, NC, bc);
    if constexpr (EPI == 0) {
      if constexpr (SEM_SEAMS_NT) __builtin_nontemporal_store(wv, &y[c.t]);
      else y[c.t] = wv;
    }
    else epi_store<T, EPI>(c.t, wv, c.u, c.e, er, ex, ec1, ec2, y, ed, etheta);
  };
  if ((int)blockIdx.y < nys) {
    const int g = blockIdx.y;
    int jL, jR;
    if (g < nsegs - 1) {
      jR = (g + 1) * SP;
      jL = jR - 1;
    } else {
      jL = NC - 1;
      jR = 0;
    }
    const int i = q;
    if (i >= L) return;
    const int r = i % CXP;
    const bool xs_row = (r == 0 && i > 0) || (r == CXP - 1 && i < L - 1) ||
                        (per_x && (i == 0 || i == L - 1));
    if (xs_row) return;   // summed with the four-copy rule by the x-seam items
    const Copy cL = pre(i, jL), cR = pre(i, jR);
    const T sm = ly(cL.t) + ly(cR.t);
    fin(i, jL, cL, sm);
    fin(i, jR, cR, sm);
  } else {
    const int s = blockIdx.y - nys;
    int iL, iR;
    if (s < nstrips - 1) {
      iR = (s + 1) * CXP;
      iL = iR - 1;
    } else {
      iL = L - 1;
      iR = 0;
    }
    const int j = q;
    if (j >= NC) return;
    const int b = j % P;
    if (b == 0 && (j > 0 || per_y)) return;  // summed by the item of the row below
    if (b == P - 1 && (j < NC - 1 || per_y)) {
      const int jU = (j + 1 == NC) ? 0 : j + 1;
      const Copy c0 = pre(iL, j), c1 = pre(iR, j), c2 = pre(iL, jU), c3 = pre(iR, jU);
      const T sl = ly(c0.t) + ly(c1.t);
      const T su = ly(c2.t) + ly(c3.t);
      const T sm = sl + su;
      fin(iL, j, c0, sm);
      fin(iR, j, c1, sm);
      fin(iL, jU, c2, sm);
      fin(iR, jU, c3, sm);
    } else {
      const Copy c0 = pre(iL, j), c1 = pre(iR, j);
      const T sm = ly(c0.t) + ly(c1.t);
      fin(iL, j, c0, sm);
      fin(iR, j, c1, sm);
    }
  }
}

// Element rows swept per workgroup; forced > 0 overrides the heuristic (the
// SEM_AX_ROWS hook, read when an operator is created, lets the tests exercise
// every seam configuration).
inline int rows_per_segment(const Layout& g, int cx, bool pcg, int forced) {
  if (forced > 0) return forced;
  const int nstrips = (g.Ex + cx - 1) / cx;
  // Plain apply: one element row per workgroup until the grid exceeds 4096
  // workgroups (seams are finished by a separate light kernel). PCG: the
  // update kernel finishes seams inline, so two rows per workgroup while at
  // least 1024 workgroups remain.
  const long target = pcg ? 1024 : 4096;
  int S = 1;
  if (pcg) {
    while (S < g.Ey && (long)nstrips * ((g.Ey + 2 * S - 1) / (2 * S)) >= target) S *= 2;
  } else {
    while (S < g.Ey && (long)nstrips * ((g.Ey + S - 1) / S) > target) S *= 2;
  }
  return S;
}

template <typename T, int P>
AxP<T> make_params(const Layout& g, const ApplyArgs& a, const AxPlan& pl) {
  AxP<T> prm{};
  prm.u = static_cast<const T*>(a.u);
  prm.y = static_cast<T*>(a.y);
  prm.G11 = static_cast<const T*>(a.G11);
  prm.G12 = static_cast<const T*>(a.G12);
  prm.G22 = static_cast<const T*>(a.G22);
  prm.B = static_cast<const T*>(a.B);
  prm.h11 = static_cast<const _Float16*>(a.G11);
  prm.h12 = static_cast<const _Float16*>(a.G12);
  prm.h22 = static_cast<const _Float16*>(a.G22);
  prm.hB = static_cast<const _Float16*>(a.B);
  for (int q = 0; q < 4; ++q) prm.gs[q] = a.gscale[q];
  prm.Dr = static_cast<const T*>(a.Dr);
  prm.zr = a.zr;
  prm.wr = a.wr;
  prm.Ds = static_cast<const T*>(a.Ds);
  prm.nu_f = static_cast<const T*>(a.nu_f);
  prm.k_f = static_cast<const T*>(a.k_f);
  prm.nu = T(a.nu);
  prm.k = T(a.k);
  prm.has_mass = a.has_mass;
  prm.Ex = g.Ex;
  prm.Ey = g.Ey;
  prm.nstrips = pl.nstrips;
  prm.S = pl.S;
  prm.per_x = g.per_x;
  prm.per_y = g.per_y;
  prm.bc = a.bc;
  return prm;
}

template <typename T, int P, int CX, int MF>
void launch_k1(hipStream_t st, dim3 grid, dim3 block, const AxP<T>& prm, int mode) {
  if (mode == 2) hipLaunchKernelGGL((k_ax<T, P, CX, 2, MF>), grid, block, 0, st, prm);
  else if (mode == 1) hipLaunchKernelGGL((k_ax<T, P, CX, 1, MF>), grid, block, 0, st, prm);
  else hipLaunchKernelGGL((k_ax<T, P, CX, 0, MF>), grid, block, 0, st, prm);
}
// the two-row apply with a fused epilogue (kind 1 to 3): P = 2, 4, 8, the
// multigrid levels of N = 7 and 15
template <typename T, int P, int CX>
void launch_k1_epi(hipStream_t st, dim3 grid, dim3 block, const AxP<T>& prm, int kind) {
  if constexpr (CX % 2 == 0 && (P == 2 || P == 4 || P == 8)) {
    if (kind == 1) hipLaunchKernelGGL((k_ax<T, P, CX, 1, 0, 2, 1>), grid, block, 0, st, prm);
    if (kind == 2) hipLaunchKernelGGL((k_ax<T, P, CX, 1, 0, 2, 2>), grid, block, 0, st, prm);
    if (kind == 3) hipLaunchKernelGGL((k_ax<T, P, CX, 1, 0, 2, 3>), grid, block, 0, st, prm);
  }
}
template <typename T, int P, int CX>
void launch_k1_rows(hipStream_t st, dim3 grid, dim3 block, const AxP<T>& prm, int ry, int mode) {
  if constexpr (P < 16) {
    if constexpr (CX % 4 == 0)
      if (ry == 4 && mode == 1) hipLaunchKernelGGL((k_ax<T, P, CX, 1, 0, 4>), grid, block, 0, st, prm);
    if constexpr (CX % 2 == 0) {
      if (ry == 2 && mode == 1) hipLaunchKernelGGL((k_ax<T, P, CX, 1, 0, 2>), grid, block, 0, st, prm);
      if (ry == 2 && mode == 2) hipLaunchKernelGGL((k_ax<T, P, CX, 2, 0, 2>), grid, block, 0, st, prm);
    }
  }
}

template <typename T, int P>
int launch_apply_p(hipStream_t st, const Layout& g, const ApplyArgs& a, const PcgLaunch* pc) {
  constexpr int CX = SEM_AX_THREADS / P;
  const AxPlan pl = make_plan(g, a.assemble != 0, pc != nullptr, a.rows);
  AxP<T> prm = make_params<T, P>(g, a, pl);
  const dim3 grid(pl.nk1), block(CX * P);
  int mode = a.assemble ? 1 : 0;
  if (pc) {
    if (!pc->x) {
      set_error("PCG apply needs the iterate x");
      return SEM_ERR_ARG;
    }
    prm.r = static_cast<const T*>(pc->r);
    prm.d = static_cast<const T*>(pc->d);
    prm.p = static_cast<T*>(pc->p);
    prm.x = static_cast<T*>(pc->x);
    prm.u = nullptr;
    prm.pcg = pc->dev;
    mode = 2;
  }
  // 16x16 blocks run their contractions on matrix cores in every precision
  // (FP64 / FP32: exact MFMA forms); a.fma16 (the SEM_AX_FMA16=1 hook, read
  // when the operator is created) selects the FMA path.
  int mf = a.mfma;
  if (mf == 0 && P == 16 && CX * P == 256 && !a.fma16) mf = sizeof(T) == 8 ? 6 : 5;
  const int epi = a.epi.kind;
  if (epi != 0) {
    if (!(epi >= 1 && epi <= 3) || mf != 0 || mode != 1 || pl.RY != 2 ||
        !(P == 2 || P == 4 || P == 8) || !a.epi.r ||
        (epi != 2 && (!a.epi.x || !a.epi.dinv)) || (epi == 3 && !a.epi.d)) {
      set_error("fused apply epilogue %d: unsupported launch", epi);
      return SEM_ERR_UNSUPPORTED;
    }
    prm.er = static_cast<T*>(a.epi.r);
    prm.ex = static_cast<T*>(a.epi.x);
    prm.edinv = static_cast<const T*>(a.epi.dinv);
    prm.ec1 = T(a.epi.c1);
    prm.ec2 = T(a.epi.c2);
    prm.ed = static_cast<T*>(a.epi.d);
    prm.etheta = T(a.epi.theta);
    launch_k1_epi<T, P, CX>(st, grid, block, prm, epi);
  } else if (mf == 0 && pl.RY > 1) {
    launch_k1_rows<T, P, CX>(st, grid, block, prm, pl.RY, mode);
  } else if (mf == 0) {
    launch_k1<T, P, CX, 0>(st, grid, block, prm, mode);
  } else {
    if constexpr (P == 16 && CX * P == 256) {
      if constexpr (sizeof(T) == 8) {
        if (mf != 6) {
          set_error("FP64 fields take the FP64 contraction (mode %d requested)", mf);
          return SEM_ERR_UNSUPPORTED;
        }
        launch_k1<T, P, CX, 6>(st, grid, block, prm, mode);
      } else {
        switch (mf) {
          case 1: launch_k1<T, P, CX, 1>(st, grid, block, prm, mode); break;
          case 2: launch_k1<T, P, CX, 2>(st, grid, block, prm, mode); break;
          case 3: launch_k1<T, P, CX, 3>(st, grid, block, prm, mode); break;
          case 4: launch_k1<T, P, CX, 4>(st, grid, block, prm, mode); break;
          case 5: launch_k1<T, P, CX, 5>(st, grid, block, prm, mode); break;
          default:
            set_error("unknown MFMA contraction mode %d", mf);
            return SEM_ERR_ARG;
        }
      }
    } else {
      set_error("MFMA contractions need nr == ns == 16");
      return SEM_ERR_UNSUPPORTED;
    }
  }
  // In PCG mode the seam copies are finished by the update kernel instead.
  if (a.assemble && pl.nk2 > 0 && !pc) {
    const int nys = pl.nsegs - 1 + (g.per_y ? 1 : 0);
    const int nxs = pl.nstrips - 1 + (g.per_x ? 1 : 0);
    const dim3 sgrid((std::max(g.L, g.NC) + kSeamBlock - 1) / kSeamBlock, nys + nxs);
    if (epi == 1)
      hipLaunchKernelGGL((k_seams<T, 1>), sgrid, dim3(kSeamBlock), 0, st, prm.y, g.L, g.NC, P,
                         pl.CX * P, pl.nstrips, pl.S * P, pl.nsegs, g.per_x, g.per_y, a.bc, nys,
                         prm.u, prm.er, prm.ex, prm.edinv, prm.ec1, prm.ec2, prm.ed, prm.etheta);
    else if (epi == 2)
      hipLaunchKernelGGL((k_seams<T, 2>), sgrid, dim3(kSeamBlock), 0, st, prm.y, g.L, g.NC, P,
                         pl.CX * P, pl.nstrips, pl.S * P, pl.nsegs, g.per_x, g.per_y, a.bc, nys,
                         prm.u, prm.er, prm.ex, prm.edinv, prm.ec1, prm.ec2, prm.ed, prm.etheta);
    else if (epi == 3)
      hipLaunchKernelGGL((k_seams<T, 3>), sgrid, dim3(kSeamBlock), 0, st, prm.y, g.L, g.NC, P,
                         pl.CX * P, pl.nstrips, pl.S * P, pl.nsegs, g.per_x, g.per_y, a.bc, nys,
                         prm.u, prm.er, prm.ex, prm.edinv, prm.ec1, prm.ec2, prm.ed, prm.etheta);
    else
      hipLaunchKernelGGL((k_seams<T, 0>), sgrid, dim3(kSeamBlock), 0, st, prm.y, g.L, g.NC, P,
                         pl.CX * P, pl.nstrips, pl.S * P, pl.nsegs, g.per_x, g.per_y, a.bc, nys,
                         nullptr, nullptr, nullptr, nullptr, T(0), T(0), nullptr, T(1));
  }
  SEM_HIP(hipGetLastError());
  return SEM_OK;
}

}  // namespace ax
}  // namespace sem

// Fused Jacobi-PCG for the SEM operator, restating pcg.jl:16-60 on the GPU.
//
// Per iteration k (all scalars stay on the device):
//   apply kernel (k_ax MODE 2): checks the stop test on the partials of the
//       previous update (norm(r, Inf) <= tol, or k-1 == maxiter), forms
//       t = <r, h>, beta = t / t_prev, u = h + beta u (h = d .* r or r)
//       and, while the old u is in registers, x += alpha_{k-1} u_old;
//       w = M .* QQ'(A u) except at workgroup seams, and the partials of
//       <u, w>_mult, which equal sum(u .* H) over the unassembled element
//       outputs H because u is continuous and masked;
//   update kernel: finishes the seam copies of w on the fly (same order as
//       the seam kernel of the plain apply), alpha = t / <u, w> (kept in the
//       state for the next apply), r -= alpha w, and the partials of <r, h>
//       and norm(r, Inf) for the next iteration.
// Two launches per iteration; after the stop test one pass adds the last
// alpha u to x. x follows the same sequence as x = x + a * u in pcg.jl:53.
// Partials are summed in a fixed order by every block that needs them, so
// every block sees bit-identical scalars and the solve is deterministic.
// The host launches iterations in growing batches and reads the done flag
// between batches; iterations launched after convergence exit at once.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <mutex>

#include "sem_device.h"

namespace sem {
namespace {

// init and partials passes: one block per partial, 1024 threads, 32-bit
// node indices (meshes hold fewer than 2^31 local nodes)
constexpr int kInitThreads = 1024;

template <typename T>
__global__ void __launch_bounds__(kInitThreads)
k_pcg_init(const T* __restrict__ b, T* __restrict__ x, T* __restrict__ r, T* __restrict__ p,
           const T* __restrict__ d, int P, int L, int NC, int per_x, int per_y, long n,
           double* __restrict__ part, int n3) {
  __shared__ double red[kInitThreads];
  double rho = 0.0, mx = 0.0;
  const int nn = (int)n;
  for (int t = blockIdx.x * kInitThreads + threadIdx.x; t < nn; t += gridDim.x * kInitThreads) {
    const int i = t % L;
    const int j = t / L;
    const T bv = b[t];
    x[t] = T(0);
    r[t] = bv;
    p[t] = T(0);   // the first apply reads p (times beta = 0)
    const T z = d ? d[t] * bv : bv;
    rho += (double)bv * (double)z * (mult1(i, P, L, per_x) * mult1(j, P, NC, per_y));
    mx = nanmax(mx, fabs((double)bv));
  }
  rho = block_sum(rho, red, kInitThreads);
  mx = block_max(mx, red, kInitThreads);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = rho;
    part[n3 + blockIdx.x] = mx;
  }
}

// x-direction facts of local row i, packed: bit 0 = half multiplicity in x
// (mult1), bits 1-2 = seam kind of the strip kernel's unsummed copies
// (0 none, 1 pair (i-1, i), 2 pair (i, i+1), 3 periodic wrap (L-1, 0)),
// bit 3 = masked by an x Dirichlet side.
__device__ inline int row_info(int i, const SeamGeom& g) {
  int v = 0;
  const int a = i % g.P;
  if ((a == 0 && (i > 0 || g.per_x)) || (a == g.P - 1 && (i < g.L - 1 || g.per_x))) v |= 1;
  const int r = i % g.CXP;
  int kind = 0;
  if (r == 0 && i > 0) kind = 1;
  else if (r == g.CXP - 1 && i < g.L - 1) kind = 2;
  else if (g.per_x && (i == 0 || i == g.L - 1)) kind = 3;
  v |= kind << 1;
  if (((g.bc & SEM_BC_XMIN) && i == 0) || ((g.bc & SEM_BC_XMAX) && i == g.L - 1)) v |= 8;
  return v;
}

// y-direction facts of local column j: x = bit 0 half multiplicity in y,
// bit 1 masked by a y Dirichlet side, bit 2 a segment y-seam column (the
// strip kernel left its copies and those of column jR unsummed), bit 3 the
// seam pair is the periodic wrap (NC-1, 0); y = jL (jR = jL + 1 unless wrap)
__device__ inline int2 col_info(int j, const SeamGeom& g) {
  int f = 0, jl = 0;
  const int b = j % g.P;
  if ((b == 0 && (j > 0 || g.per_y)) || (b == g.P - 1 && (j < g.NC - 1 || g.per_y))) f |= 1;
  if (((g.bc & SEM_BC_YMIN) && j == 0) || ((g.bc & SEM_BC_YMAX) && j == g.NC - 1)) f |= 2;
  const int qy = j % g.SP;
  if (qy == 0 && j > 0) {
    f |= 4;
    jl = j - 1;
  } else if (qy == g.SP - 1 && j < g.NC - 1) {
    f |= 4;
    jl = j;
  } else if (g.per_y && (j == 0 || j == g.NC - 1)) {
    f |= 4 | 8;
    jl = g.NC - 1;
  }
  return make_int2(f, jl);
}

// Row and column fact tables of one solve (written once, before the
// iterations): the update kernel reads them instead of recomputing integer
// remainders per node (which made it scalar-instruction bound).
__global__ void k_pcg_tables(int* __restrict__ rinfo, int2* __restrict__ cinfo, SeamGeom g) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < g.L) rinfo[t] = row_info(t, g);
  if (t < g.NC) cinfo[t] = col_info(t, g);
}

// rows whose fact table is copied into LDS (56 KiB, so that with the 8 KiB
// reduction array the block stays within 64 KiB); longer rows read it from
// global memory
constexpr int kRowTable = 14336;

// One node of the PCG update: r -= alpha wf, partials of <r, z>_mult (m =
// the node's mult) and norm(r, Inf). With keep == false the node is left as
// it is (its old r is stored back) and adds nothing to the partials: a
// select, not a branch, so the bulk pass has no divergent control flow. The
// max ignores NaN here; the block restores NaN from the <r, z> partial, which
// is NaN whenever some r is.
// The iterate is not touched here: the next iteration's apply adds alpha p
// while it reads p anyway, and k_pcg_xlast adds the last one.
// Non-temporal hints of the one-node-per-thread update: w loads, d loads,
// r stores (compile-time switches, all off: at 256x256, N = 7 the FP64
// iteration took 89.3 us with streaming w loads, 89.3 us with streaming r
// stores as well and 90.5 us with streaming d loads, against 88.4 us)
#ifndef SEM_PCG_NT_W
#define SEM_PCG_NT_W 0
#endif
#ifndef SEM_PCG_NT_D
#define SEM_PCG_NT_D 0
#endif
#ifndef SEM_PCG_NT_R
#define SEM_PCG_NT_R 0
#endif
template <typename T, bool NT>
__device__ __forceinline__ T ld_hint(const T* p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}
template <typename T, bool HAS_D>
__device__ __forceinline__ void upd_node(T* __restrict__ r, const T* __restrict__ d, long t,
                                         T wf, T alpha, double m, bool keep, double& rho,
                                         double& mx) {
  const T ro = r[t];
  const T rv = ro - alpha * wf;
  if constexpr (SEM_PCG_NT_R) __builtin_nontemporal_store(keep ? rv : ro, &r[t]);
  else r[t] = keep ? rv : ro;
  T z = rv;
  if constexpr (HAS_D) z = ld_hint<T, SEM_PCG_NT_D>(&d[t]) * rv;
  rho += (double)rv * (double)z * (keep ? m : 0.0);
  mx = fmax(mx, keep ? fabs((double)rv) : 0.0);
}

// r -= alpha w (w's seam copies finished here with the k_seams
// rules: same operands, same order, so every copy gets identical bits);
// partials of <r, z>_mult and norm(r, Inf) for the next iteration. A block
// walks whole columns j (its share of NC). Bulk pass: each thread takes rows
// tid, tid + kUpdThreads, ... except the strip-seam rows; the per-row facts
// come from the row table, the per-column facts (segment y-seam, mask, mult)
// from the column table in one uniform load, so the bulk has no integer
// division and no divergent seam branches. Seam pass: the 2 (nstrips - 1)
// strip-seam rows (+ 2 periodic wrap rows) of the block's columns, with the
// four-copy sums at element y-interfaces.
constexpr int kUpdThreads = 1024;   // 16 waves per block: latency hiding for FP32
// most update blocks (and block partials of <r, z>, norm(r, Inf)) per solve
#ifndef SEM_PCG_NPART
#define SEM_PCG_NPART 256
#endif

// HAS_D: a diagonal preconditioner is present. VEC = 2: two adjacent rows per
// thread with 8-byte accesses (FP32, L even). FULL: L is a multiple of
// VEC * kUpdThreads, so the row loop needs no bound test.
template <typename T, bool HAS_D, bool FULL, int VEC = 1>
__global__ void __launch_bounds__(kUpdThreads)
k_pcg_update(T* __restrict__ r, const T* __restrict__ w, const T* __restrict__ d,
             const double* __restrict__ Q, int nQ, PcgState* __restrict__ st, int k,
             double* __restrict__ part, int n3, SeamGeom sg, const int* __restrict__ rtab,
             const int2* __restrict__ ctab, long n) {
  const int NC = sg.NC;
  const int Lr = sg.L;
  const long L = sg.L;
  __shared__ double red[kUpdThreads];
  // independent loads first, so their latencies overlap: this thread's share
  // of the <p, Ap> partials and t = <r, z> of this iteration (both valid
  // memory even when the solve has already stopped)
  double qs = 0.0;
  for (int q = threadIdx.x; q < nQ; q += kUpdThreads) qs += Q[q];
  const double tk = st->t[k & 1];
  // done is written only by the apply launches, so it cannot change while
  // this kernel runs: every thread reads the same value, no broadcast needed
  if (st->done) return;
  // one-row-per-thread passes read the row table from LDS when it fits (an
  // LDS read per node is cheaper than a global one there; the paired-row FP32
  // pass reads one 8-byte global entry per pair, which measured faster);
  // either way the table is the solve's, so no integer remainders here
  extern __shared__ int rlds[];
  const bool in_lds = VEC == 1 && Lr <= kRowTable;   // uniform
  if (in_lds)
    for (int i = threadIdx.x; i < Lr; i += kUpdThreads) rlds[i] = rtab[i];
  const int* __restrict__ rinfo = in_lds ? rlds : rtab;
  // synchronises the block (also for the table); red is next written by the
  // final reduction, after the barrier that precedes the seam pass
  const double pw = block_sum_w_nb<kUpdThreads>(qs, red);
  const T alpha = T(tk / pw);
  if (blockIdx.x == 0 && threadIdx.x == 0) st->alpha[k & 1] = (double)alpha;
  const int cpb = (NC + (int)gridDim.x - 1) / (int)gridDim.x;
  const int j0 = blockIdx.x * cpb;
  const int j1 = min(NC, j0 + cpb);
  double rho = 0.0, mx = 0.0;
  // bulk: every row but the strip-seam rows
  for (int j = j0; j < j1; ++j) {
    const int2 ci = ctab[j];   // uniform
    const double my = (ci.x & 1) ? 0.5 : 1.0;
    const bool masky = (ci.x & 2) != 0;
    const bool yseam = (ci.x & 4) != 0;
    const int jL = ci.y, jR = (ci.x & 8) ? 0 : ci.y + 1;
    const long cj = L * j;
    if constexpr (VEC == 2) {
      static_assert(sizeof(T) == 4, "paired rows: FP32");
      using T2 = float2;
      const long cL = L * (yseam ? jL : 0), cR = L * (yseam ? jR : 0);
#pragma unroll 2
      for (int i0 = 0; i0 < Lr; i0 += 2 * kUpdThreads) {
        const int i = i0 + 2 * (int)threadIdx.x;
        if (!FULL && i >= Lr) break;
        const long t = i + cj;
        const int2 info = *reinterpret_cast<const int2*>(rinfo + i);
        T2 wf;
        if (yseam) {
          const T2 wl = *reinterpret_cast<const T2*>(w + i + cL);
          const T2 wr2 = *reinterpret_cast<const T2*>(w + i + cR);
          const T m0 = ((info.x & 8) || masky) ? T(0) : T(1);
          const T m1 = ((info.y & 8) || masky) ? T(0) : T(1);
          wf.x = (wl.x + wr2.x) * m0;
          wf.y = (wl.y + wr2.y) * m1;
        } else {
          wf = *reinterpret_cast<const T2*>(w + t);
        }
        const T2 ro = *reinterpret_cast<const T2*>(r + t);
        const bool k0 = !(info.x & 6), k1 = !(info.y & 6);
        T2 rn;
        rn.x = ro.x - alpha * wf.x;
        rn.y = ro.y - alpha * wf.y;
        T2 rs;
        rs.x = k0 ? rn.x : ro.x;
        rs.y = k1 ? rn.y : ro.y;
        *reinterpret_cast<T2*>(r + t) = rs;
        T2 z = rn;
        if constexpr (HAS_D) {
          const T2 dv = *reinterpret_cast<const T2*>(d + t);
          z.x = dv.x * rn.x;
          z.y = dv.y * rn.y;
        }
        rho += (double)rn.x * (double)z.x * (k0 ? (info.x & 1 ? 0.5 : 1.0) * my : 0.0);
        rho += (double)rn.y * (double)z.y * (k1 ? (info.y & 1 ? 0.5 : 1.0) * my : 0.0);
        mx = fmax(mx, k0 ? fabs((double)rn.x) : 0.0);
        mx = fmax(mx, k1 ? fabs((double)rn.y) : 0.0);
      }
    } else if (yseam) {   // uniform: a segment y-seam column
      const long cL = L * jL, cR = L * jR;
#pragma unroll 2
      for (int i0 = 0; i0 < Lr; i0 += kUpdThreads) {
        const int i = i0 + (int)threadIdx.x;
        if (!FULL && i >= Lr) break;
        const int info = rinfo[i];
        const T mv = ((info & 8) || masky) ? T(0) : T(1);
        const T wf = (ld_hint<T, SEM_PCG_NT_W>(&w[i + cL]) + ld_hint<T, SEM_PCG_NT_W>(&w[i + cR])) * mv;
        upd_node<T, HAS_D>(r, d, i + cj, wf, alpha, (info & 1 ? 0.5 : 1.0) * my, !(info & 6),
                           rho, mx);
      }
    } else {
#pragma unroll 2
      for (int i0 = 0; i0 < Lr; i0 += kUpdThreads) {
        const int i = i0 + (int)threadIdx.x;
        if (!FULL && i >= Lr) break;
        const int info = rinfo[i];
        const long t = i + cj;
        upd_node<T, HAS_D>(r, d, t, ld_hint<T, SEM_PCG_NT_W>(&w[t]), alpha, (info & 1 ? 0.5 : 1.0) * my, !(info & 6), rho,
                           mx);
      }
    }
  }
  __syncthreads();   // the bulk stored the seam rows' old r; the seam pass follows
  // strip-seam rows: s*CXP - 1 and s*CXP for s = 1 .. nstrips-1, then the
  // periodic wrap rows 0 and L-1
  const int nsr = 2 * (sg.nstrips - 1);
  const int nxr = nsr + (sg.per_x ? 2 : 0);
  const int ncol = max(0, j1 - j0);
  for (int item = threadIdx.x; item < nxr * ncol; item += kUpdThreads) {
    const int q = item % nxr;
    const int j = j0 + item / nxr;
    int i, iL, iR;
    if (q < nsr) {
      const int s = q / 2 + 1;
      i = (q & 1) ? s * sg.CXP : s * sg.CXP - 1;
      iL = s * sg.CXP - 1;
      iR = s * sg.CXP;
    } else {
      i = (q == nsr) ? 0 : Lr - 1;
      iL = Lr - 1;
      iR = 0;
    }
    const int info = rinfo[i];
    const int P = sg.P, per_y = sg.per_y;
    const int b = j % P;
    const bool yhalf = (b == 0 && (j > 0 || per_y)) || (b == P - 1 && (j < NC - 1 || per_y));
    int jl = -1, ju = -1;   // element y-interface: four copies
    if (b == 0 && (j > 0 || per_y)) {
      jl = j > 0 ? j - 1 : NC - 1;
      ju = j;
    } else if (b == P - 1 && (j < NC - 1 || per_y)) {
      jl = j;
      ju = j + 1 < NC ? j + 1 : 0;
    }
    const bool masky = ((sg.bc & SEM_BC_YMIN) && j == 0) || ((sg.bc & SEM_BC_YMAX) && j == NC - 1);
    const T mv = ((info & 8) || masky) ? T(0) : T(1);
    T sv;
    if (jl >= 0) {
      const T sl = w[iL + L * jl] + w[iR + L * jl];
      const T su = w[iL + L * ju] + w[iR + L * ju];
      sv = sl + su;
    } else {
      sv = w[iL + L * j] + w[iR + L * j];
    }
    upd_node<T, HAS_D>(r, d, i + L * j, sv * mv, alpha,
                       (info & 1 ? 0.5 : 1.0) * (yhalf ? 0.5 : 1.0), true, rho, mx);
  }
  block_sum_max_to0<kUpdThreads>(rho, mx, red);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = rho;
    part[n3 + blockIdx.x] = rho != rho ? rho : mx;   // a NaN residual makes rho NaN
  }
}

// x += alpha p of the last iteration (the apply that would have added it
// found the stop test fired); nothing when no iteration ran.
template <typename T>
__global__ void k_pcg_xlast(T* __restrict__ x, const T* __restrict__ p,
                            const PcgState* __restrict__ st, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int it = st->iters;
  if (it < 1) return;
  x[t] = x[t] + T(st->alpha[it & 1]) * p[t];
}

// After k_pcg_xlast: that alpha p is in x, so a solve continued from the
// next iteration (kept search direction) must not add it again.
__global__ void k_pcg_alpha_clear(PcgState* st) {
  const int it = st->iters;
  if (it >= 1) st->alpha[it & 1] = 0.0;
}

// Element diagonal of nu A + k B (nu, k scalars or fields):
// nu [sum_i Dr_ia^2 G11_ib + 2 Dr_aa Ds_bb G12_ab + sum_j Ds_jb^2 G22_aj] + k B_ab
__global__ void k_jacobi_local(const double* __restrict__ G11, const double* __restrict__ G12,
                               const double* __restrict__ G22, const double* __restrict__ B,
                               const double* __restrict__ Dr, const double* __restrict__ Ds,
                               double nu, const double* __restrict__ nu_f, double kc,
                               const double* __restrict__ k_f, int has_mass, int nr, int ns,
                               int L, long n, double* __restrict__ dl) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = (int)(t % L);
  const int j = (int)(t / L);
  const int a = i % nr, i0 = i - a;
  const int b = j % ns, j0 = j - b;
  double t1 = 0, t3 = 0;
  for (int q = 0; q < nr; ++q) {
    const double dr = Dr[q + nr * a];
    t1 += dr * dr * G11[(i0 + q) + (long)L * j];
  }
  for (int q = 0; q < ns; ++q) {
    const double ds = Ds[q + ns * b];
    t3 += ds * ds * G22[i + (long)L * (j0 + q)];
  }
  const double t2 = 2.0 * Dr[a + nr * a] * Ds[b + ns * b] * G12[t];
  const double nuv = nu_f ? nu_f[t] : nu;
  double v = nuv * (t1 + t2 + t3);
  if (has_mass) v = v + (k_f ? k_f[t] : kc) * B[t];
  dl[t] = v;
}

__global__ void k_inv_masked(const double* __restrict__ d, double* __restrict__ dinv, int L,
                             int NC, int bc, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = (int)(t % L);
  const int j = (int)(t / L);
  dinv[t] = mask_on(i, j, L, NC, bc) ? 1.0 / d[t] : 0.0;
}

inline int nblk(long n, int bs = 256) { return (int)((n + bs - 1) / bs); }

// Partials of <r, h> and norm(r, Inf) of the current residual (after a
// residual replacement).
template <typename T>
__global__ void __launch_bounds__(kInitThreads)
k_pcg_partials(const T* __restrict__ r, const T* __restrict__ d, int P, int L, int NC, int per_x,
               int per_y, long n, double* __restrict__ part, int n3) {
  __shared__ double red[kInitThreads];
  double rho = 0.0, mx = 0.0;
  const int nn = (int)n;
  for (int t = blockIdx.x * kInitThreads + threadIdx.x; t < nn; t += gridDim.x * kInitThreads) {
    const int i = t % L;
    const int j = t / L;
    const T rv = r[t];
    const T z = d ? d[t] * rv : rv;
    rho += (double)rv * (double)z * (mult1(i, P, L, per_x) * mult1(j, P, NC, per_y));
    mx = nanmax(mx, fabs((double)rv));
  }
  rho = block_sum(rho, red, kInitThreads);
  mx = block_max(mx, red, kInitThreads);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = rho;
    part[n3 + blockIdx.x] = mx;
  }
}

// A resumable fused PCG: init from b, run iterations until the stop test
// fires, optionally replace the residual and continue with the same search
// direction and iteration count (used by the mixed-precision solver).
template <typename T>
struct PcgRun {
  sem_ctx* ctx = nullptr;
  const sem_operator* op = nullptr;
  const T* d = nullptr;
  Layout g{};
  long n = 0;
  AxPlan pl{};
  int n3 = 0, nQ = 0;
  SeamGeom sg{};
  ApplyArgs aa{};
  T *r = nullptr, *p = nullptr, *w = nullptr;
  double *part = nullptr, *Q = nullptr;
  int* rtab = nullptr;
  int2* ctab = nullptr;
  PcgState* state = nullptr;
  PcgState* hs = nullptr;

  // launch geometry; returns the workspace bytes the solve needs
  size_t plan(sem_ctx* c, const sem_operator* o, const T* dd) {
    ctx = c;
    op = o;
    d = dd;
    const sem_mesh* m = op->mesh;
    g = m->layout();
    n = (long)m->nl;
    pl = make_plan(g, true, true, op->ax_rows);
    n3 = (int)std::min<long>(SEM_PCG_NPART, std::max<long>(1, (n + 2047) / 2048));
    nQ = pl.nk1;
    sg = SeamGeom{g.P, g.L, g.NC, pl.CX * g.P, pl.nstrips, pl.S * g.P, pl.nsegs, g.per_x,
                  g.per_y, op->desc.bc_bits};
    return 3 * vec_bytes() + part_bytes() + q_bytes() + tab_bytes() + 256;
  }
  size_t tab_bytes() const {
    return ((sizeof(int) * g.L + 255) / 256) * 256 + ((sizeof(int2) * g.NC + 255) / 256) * 256;
  }
  size_t vec_bytes() const { return ((sizeof(T) * n + 255) / 256) * 256; }
  size_t part_bytes() const { return sizeof(double) * 4 * n3; }
  size_t q_bytes() const { return ((sizeof(double) * nQ + 255) / 256) * 256; }

  // binds the workspace (r, p, w | part[2][2*n3] | Q[nQ] | row table |
  // column table | state) and the host copy of the state, and writes the
  // row and column fact tables of the update kernel
  int bind(void* ws, PcgState* host_state) {
    const size_t vb = vec_bytes(), pb = part_bytes(), qb = q_bytes(), tb = tab_bytes();
    char* w0 = static_cast<char*>(ws);
    r = reinterpret_cast<T*>(w0);
    p = reinterpret_cast<T*>(w0 + vb);
    w = reinterpret_cast<T*>(w0 + 2 * vb);
    part = reinterpret_cast<double*>(w0 + 3 * vb);
    Q = reinterpret_cast<double*>(w0 + 3 * vb + pb);
    rtab = reinterpret_cast<int*>(w0 + 3 * vb + pb + qb);
    ctab = reinterpret_cast<int2*>(w0 + 3 * vb + pb + qb + ((sizeof(int) * g.L + 255) / 256) * 256);
    state = reinterpret_cast<PcgState*>(w0 + 3 * vb + pb + qb + tb);
    hs = host_state;
    aa = args_of(op, nullptr, w);   // factors, precision, recomputed geometry
    aa.bc = op->desc.bc_bits;
    aa.assemble = 1;
    hipLaunchKernelGGL(k_pcg_tables, dim3(nblk(std::max(g.L, g.NC))), dim3(256), 0, ctx->stream,
                       rtab, ctab, sg);
    SEM_HIP(hipGetLastError());
    return SEM_OK;
  }

  // workspace from the context (grown on demand), state read back through
  // the context's pinned buffer
  int setup(sem_ctx* c, const sem_operator* o, const T* dd) {
    const size_t bytes = plan(c, o, dd);
    SEM_TRY(ensure_ws(ctx, bytes));
    SEM_TRY(ensure_pinned(ctx, 128));  // PcgState, then a scalar slot at byte 64
    return bind(ctx->ws, static_cast<PcgState*>(ctx->pinned));
  }

  // x = 0, r = b, partials of r for iteration 1, state cleared
  int init(const T* b, T* x) {
    hipStream_t st = ctx->stream;
    SEM_HIP(hipMemsetAsync(state, 0, sizeof(PcgState), st));
    hipLaunchKernelGGL((k_pcg_init<T>), dim3(n3), dim3(kInitThreads), 0, st, b, x, r, p, d, g.P,
                       g.L,
                       g.NC, g.per_x, g.per_y, n, part, n3);
    SEM_HIP(hipGetLastError());
    return SEM_OK;
  }

  // after r was replaced: partials of r for iteration k, done flag cleared
  int reprime(int k) {
    hipStream_t st = ctx->stream;
    SEM_HIP(hipMemsetAsync(&state->done, 0, sizeof(int), st));
    hipLaunchKernelGGL((k_pcg_partials<T>), dim3(n3), dim3(kInitThreads), 0, st, r, d, g.P, g.L,
                       g.NC,
                       g.per_x, g.per_y, n, part + 2 * n3 * ((k - 1) & 1), n3);
    SEM_HIP(hipGetLastError());
    return SEM_OK;
  }

  void launch_update(hipStream_t st, int k) {
    // FP32 with rows in pairs: L even (8-byte aligned pairs, since every
    // column then starts at an even offset)
    const bool pairs = sizeof(T) == 4 && g.L % 2 == 0;
    const bool full = g.L % ((pairs ? 2 : 1) * kUpdThreads) == 0;
    double* pk = part + 2 * n3 * (k & 1);
    const size_t lds = (!pairs && g.L <= kRowTable) ? sizeof(int) * g.L : 0;
    auto go = [&](auto kern) {
      hipLaunchKernelGGL(kern, dim3(n3), dim3(kUpdThreads), lds, st, r, w, d, Q, nQ, state, k,
                         pk, n3, sg, rtab, ctab, n);
    };
    if constexpr (sizeof(T) == 4) {
      if (pairs) {
        if (d) full ? go(k_pcg_update<T, true, true, 2>) : go(k_pcg_update<T, true, false, 2>);
        else full ? go(k_pcg_update<T, false, true, 2>) : go(k_pcg_update<T, false, false, 2>);
        return;
      }
    }
    if (d) full ? go(k_pcg_update<T, true, true>) : go(k_pcg_update<T, true, false>);
    else full ? go(k_pcg_update<T, false, true>) : go(k_pcg_update<T, false, false>);
  }

  // launches iterations k_start.. until the stop test fires (norm(r, Inf) <=
  // tol or k - 1 == maxiter); on return hs holds the final state
  // hint: a fresh solve starts with a batch the size of this operator's last
  // converged solve (repeated solves: time steps, adjoints, FD) and records it.
  // ev_end (optional) is recorded before each state read, so it has completed
  // when this returns; the last update of x is left queued on the stream.
  int run(T* x, double tol, int maxiter, int k_start, bool hint = false,
          hipEvent_t ev_end = nullptr) {
    hipStream_t st = ctx->stream;
    int k = k_start,
        batch = (hint && op->last_iters > 0) ? std::min(op->last_iters + 2, 1024) : 4;
    const long kmax = (long)maxiter + 1;
    for (;;) {
      for (int q = 0; q < batch && k <= kmax; ++q, ++k) {
        PcgLaunch pc{};
        pc.r = r;
        pc.d = d;
        pc.p = p;
        pc.x = x;
        pc.dev.part_prev = part + 2 * n3 * ((k - 1) & 1);
        pc.dev.n3 = n3;
        pc.dev.Q = Q;
        pc.dev.state = state;
        pc.dev.tol = tol;
        pc.dev.maxiter = maxiter;
        pc.dev.k = k;
        SEM_TRY(launch_apply_pcg<T>(st, g, aa, pc));
        launch_update(st, k);
      }
      SEM_HIP(hipGetLastError());
      if (ev_end) SEM_HIP(hipEventRecord(ev_end, st));
      SEM_HIP(hipMemcpyAsync(hs, state, sizeof(PcgState), hipMemcpyDeviceToHost, st));
      SEM_HIP(hipStreamSynchronize(st));
      if (hs->done) {
        if (hint && hs->converged) op->last_iters = hs->iters;
        hipLaunchKernelGGL((k_pcg_xlast<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                           st, x, p, state, n);
        hipLaunchKernelGGL(k_pcg_alpha_clear, dim3(1), dim3(1), 0, st, state);
        SEM_HIP(hipGetLastError());
        return SEM_OK;
      }
      if (k > kmax) {
        set_error("pcg: iteration limit passed without the stop test firing");
        return SEM_ERR_BREAKDOWN;
      }
      batch = std::min(batch * 2, 64);
    }
  }
};

template <typename T>
int pcg_solve(PcgRun<T>& R, sem_ctx* ctx, const sem_operator* op, const T* b, T* x, double tol,
              int maxiter, sem_pcg_info* info);

template <typename T>
int pcg_run_t(sem_ctx* ctx, const sem_operator* op, const T* d, const T* b, T* x, double tol,
              int maxiter, sem_pcg_info* info) {
  PcgRun<T> R;
  SEM_TRY(R.setup(ctx, op, d));
  return pcg_solve(R, ctx, op, b, x, tol, maxiter, info);
}

template <typename T>
int pcg_solve(PcgRun<T>& R, sem_ctx* ctx, const sem_operator* op, const T* b, T* x, double tol,
              int maxiter, sem_pcg_info* info) {
  hipStream_t st = ctx->stream;
  SEM_HIP(hipEventRecord(ctx->ev0, st));
  SEM_TRY(R.init(b, x));
  // ev1 completes inside run (before its last state read); the final x
  // update stays queued on the stream, so no second wait here
  SEM_TRY(R.run(x, tol, maxiter, 1, true, ctx->ev1));
  float ms = 0.f;
  SEM_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  const PcgState* hs = R.hs;
  if (info) {
    info->iters = hs->iters;
    info->converged = hs->converged;
    info->res = hs->res;
    info->elapsed_ms = ms;
    info->outer = 0;
    info->inner_total = hs->iters;
  }
  if (hs->res != hs->res) {
    set_error("pcg: residual is not finite after %d iterations", hs->iters);
    return SEM_ERR_NONFINITE;
  }
  return hs->converged ? SEM_OK : SEM_WARN_MAXITER;
}

__global__ void k_resid(const double* __restrict__ b, double* __restrict__ r, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) r[t] = b[t] - r[t];
}

__global__ void k_to_f32(const double* __restrict__ r, float* __restrict__ r32, double s,
                         long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) r32[t] = (float)(r[t] * s);
}

__global__ void k_accum_zero(double* __restrict__ x, float* __restrict__ e, double s, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) {
    x[t] = x[t] + s * (double)e[t];
    e[t] = 0.f;
  }
}

// p *= f and the stored <r, z> of iteration k-1 *= f^2 (f a power of two, so
// both are exact): the inner recurrence continues on a rescaled residual.
__global__ void k_rescale(float* __restrict__ p, float f, PcgState* state, int k, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = p[t] * f;
  if (t == 0) state->t[(k - 1) & 1] = state->t[(k - 1) & 1] * ((double)f * (double)f);
}

// power of two nearest below v (v > 0, finite)
double pow2_near(double v) { return std::ldexp(1.0, std::ilogb(v)); }

// Mixed precision with FP64 residual replacement ("reliable updates"): one
// FP32 fused PCG whose iterate is accumulated into x in FP64 and whose
// recursive residual is replaced by the true residual b - A64 x (fused FP64
// operator) whenever it has dropped by the factor delta since the last
// replacement, or (when the direction is kept) after kChunk inner iterations. The search direction and the
// iteration count carry over, so the Krylov space is not restarted. At every
// replacement the inner vectors are rescaled by a power of two so that the
// residual is O(1) (FP16 contraction inputs stay in their exponent range).
// A cycle whose true residual is not finite or grew tenfold is rolled back
// and the Krylov space restarted from the last accepted x. Operators with
// single-pass FP16/BF16 contraction inputs restart the Krylov space at every
// replacement (classical defect correction): their operator error is too
// large for the carried search direction to stay useful. The others keep it
// until a replacement shows no progress, then restart from there on. Stops when the
// true FP64 residual meets tol; gives up after max_cycles replacements, four
// without a 10% improvement, or max_inner inner iterations in total.
constexpr int kChunk = 1000;

int mixed_run(sem_ctx* ctx, const sem_operator* op64, const sem_operator* op_lo,
              const double* d64, const double* b, double* x, double tol, double delta,
              int max_cycles, int max_inner, sem_pcg_info* info) {
  const sem_mesh* m = op64->mesh;
  const long n = (long)m->nl;
  hipStream_t st = ctx->stream;
  // r64, xb (FP64), x32, d32 (FP32), reduction partials: the context's second
  // workspace (the first holds the inner PCG's vectors)
  double *r64 = nullptr, *xb = nullptr, *nrm = nullptr;
  float *x32 = nullptr, *d32 = nullptr;
  auto cleanup = [&]() { (void)hipStreamSynchronize(st); };
  {
    const size_t b8 = ((sizeof(double) * n + 255) / 256) * 256;
    const size_t b4 = ((sizeof(float) * n + 255) / 256) * 256;
    SEM_TRY(ensure_ws2(ctx, 2 * b8 + 2 * b4 + sizeof(double) * 260));
    char* w0 = static_cast<char*>(ctx->ws2);
    r64 = reinterpret_cast<double*>(w0);
    xb = reinterpret_cast<double*>(w0 + b8);
    x32 = reinterpret_cast<float*>(w0 + 2 * b8);
    d32 = d64 ? reinterpret_cast<float*>(w0 + 2 * b8 + b4) : nullptr;
    nrm = reinterpret_cast<double*>(w0 + 2 * b8 + 2 * b4);
  }
  int s = SEM_OK;
  if (d64) s = launch_convert_f64_f32(st, n, d64, d32);
  PcgRun<float> R;
  if (s == SEM_OK) s = R.setup(ctx, op_lo, d32);
  if (s != SEM_OK) {
    cleanup();
    return s;
  }
  const bool debug = std::getenv("SEM_DEBUG_MIXED") != nullptr;
  // The FP32, FP16 and split (x3) operators keep the search direction across
  // replacements; the BF16 operator restarts it (measured at 128x128
  // elements, N = 15: kept, BF16 needs 28,000-30,000 inner iterations,
  // restarted 21,000).
  // SEM_MIXED_RELIABLE=1 / 0 forces one policy.
  const char* menv = std::getenv("SEM_MIXED_RELIABLE");
  bool reliable = menv ? std::atoi(menv) != 0
                       : (op_lo->precision == SEM_F32 || op_lo->precision == SEM_F16_MFMA ||
                          op_lo->precision == SEM_F16X3_MFMA ||
                          op_lo->precision == SEM_BF16X3_MFMA);
  hipEvent_t e0 = ctx->ev0, e1 = ctx->ev1;   // the context's timing events
  if (hipEventRecord(e0, st) != hipSuccess ||
      hipMemsetAsync(x, 0, sizeof(double) * n, st) != hipSuccess ||
      hipMemsetAsync(x32, 0, sizeof(float) * n, st) != hipSuccess ||
      hipMemcpyAsync(r64, b, sizeof(double) * n, hipMemcpyDeviceToDevice, st) != hipSuccess) {
    set_error("sem_pcg_mixed: HIP call failed during set-up");
    cleanup();
    return SEM_ERR_HIP;
  }
  double* hres = reinterpret_cast<double*>(static_cast<char*>(ctx->pinned) + 64);
  auto true_norm = [&](double* out) -> int {
    SEM_TRY(launch_norm_inf(st, n, r64, nrm, 256, nrm + 256));
    SEM_HIP(hipMemcpyAsync(hres, nrm + 256, sizeof(double), hipMemcpyDeviceToHost, st));
    SEM_HIP(hipStreamSynchronize(st));
    *out = *hres;
    return SEM_OK;
  };
  auto true_residual = [&](double* out) -> int {
    SEM_TRY(sem_operator_apply(ctx, op64, x, r64));
    hipLaunchKernelGGL(k_resid, dim3(nblk(n)), dim3(256), 0, st, b, r64, n);
    return true_norm(out);
  };
  double res = 0.0;
  if ((s = true_norm(&res)) != SEM_OK) {
    cleanup();
    return s;
  }
  int cycles = 0, stall = 0, k = 1, total = 0, restarts = 0, status = SEM_WARN_MAXITER;
  double best = res, scale = 1.0;
  if (!(res > tol)) status = (res <= tol) ? SEM_OK : SEM_ERR_NONFINITE;
  // (re)starts the inner solve from the residual in r64 (norm res)
  auto start_inner = [&]() -> int {
    scale = pow2_near(res);
    hipLaunchKernelGGL(k_to_f32, dim3(nblk(n)), dim3(256), 0, st, r64, R.r, 1.0 / scale, n);
    SEM_HIP(hipMemsetAsync(R.state, 0, sizeof(PcgState), st));
    // the first apply of a restarted cycle reads p times beta = 0: p must be
    // finite (the old direction may not be, after a rolled-back cycle)
    SEM_HIP(hipMemsetAsync(R.p, 0, sizeof(float) * n, st));
    k = 1;
    return R.reprime(1);
  };
  if (status == SEM_WARN_MAXITER) {
    s = start_inner();
    for (;;) {
      if (s != SEM_OK) break;
      const double tol_c = std::max(delta * res, 0.5 * tol) / scale;
      const int k0 = k;
      // reliable updates replace the residual at least every kChunk iterations;
      // a restarted cycle runs its inner solve to tol_c (a cap would restart
      // the Krylov space mid-solve)
      const int cap = reliable ? kChunk : max_inner;
      s = R.run(x32, tol_c, k0 - 1 + std::min(cap, std::max(1, max_inner - total)), k0);
      if (s != SEM_OK) break;
      k = R.hs->iters + 1;
      total += k - k0;
      const double inner_res = R.hs->res;
      if (hipMemcpyAsync(xb, x, sizeof(double) * n, hipMemcpyDeviceToDevice, st) != hipSuccess) {
        set_error("sem_pcg_mixed: copy failed");
        s = SEM_ERR_HIP;
        break;
      }
      hipLaunchKernelGGL(k_accum_zero, dim3(nblk(n)), dim3(256), 0, st, x, x32, scale, n);
      const double res_prev = res;
      if ((s = true_residual(&res)) != SEM_OK) break;
      ++cycles;
      if (debug)
        std::fprintf(stderr, "[mixed] cycle %d: inner %d its (total %d), inner res %.3e, "
                     "true res %.3e\n", cycles, k - k0, total, inner_res * scale, res);
      if (res <= tol) {
        status = SEM_OK;
        break;
      }
      const bool bad = !(res < 10.0 * res_prev) || !std::isfinite(inner_res);
      if (bad) {
        // roll back to the last accepted iterate, restart the Krylov space
        if (hipMemcpyAsync(x, xb, sizeof(double) * n, hipMemcpyDeviceToDevice, st) != hipSuccess) {
          set_error("sem_pcg_mixed: copy failed");
          s = SEM_ERR_HIP;
          break;
        }
        if ((s = true_residual(&res)) != SEM_OK) break;
        ++restarts;
      }
      stall = (res < 0.9 * best) ? 0 : stall + 1;
      best = std::min(best, res);
      // a replacement without progress: the kept direction has gone stale,
      // restart the Krylov space from here on
      if (stall > 0 && reliable && !menv) reliable = false;
      if (cycles >= max_cycles || total >= max_inner) break;
      if (stall >= 4) {
        status = SEM_WARN_STALLED;
        break;
      }
      if (bad || !reliable) {
        s = start_inner();
      } else {
        // replace the recursive residual; rescale p and <r, z> to match
        const double ns = pow2_near(res);
        hipLaunchKernelGGL(k_to_f32, dim3(nblk(n)), dim3(256), 0, st, r64, R.r, 1.0 / ns, n);
        hipLaunchKernelGGL(k_rescale, dim3(nblk(n)), dim3(256), 0, st, R.p,
                           (float)(scale / ns), R.state, k, n);
        if (hipGetLastError() != hipSuccess) {
          set_error("sem_pcg_mixed: kernel launch failed");
          s = SEM_ERR_HIP;
          break;
        }
        scale = ns;
        s = R.reprime(k);
      }
    }
  }
  float ms = 0;
  if (s == SEM_OK && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess)) {
    set_error("sem_pcg_mixed: timing event failed");
    s = SEM_ERR_HIP;
  }
  if (s == SEM_OK) (void)hipEventElapsedTime(&ms, e0, e1);
  cleanup();
  if (s != SEM_OK) return s;
  if (!std::isfinite(res)) status = SEM_ERR_NONFINITE;
  if (debug)
    std::fprintf(stderr, "[mixed] done: %d cycles, %d restarts, %d inner its, res %.3e\n",
                 cycles, restarts, total, res);
  if (info) {
    info->iters = total;
    info->inner_total = total;
    info->outer = cycles;
    info->res = res;
    info->converged = status == SEM_OK;
    info->elapsed_ms = ms;
  }
  if (status == SEM_ERR_NONFINITE) set_error("sem_pcg_mixed: residual is not finite");
  return status;
}

// ---- allocation-free PCG: the workspace of sem_pcg (r, p, w, the block
// partials, the <p, Ap> partials and the device state) in a fixed device
// array instead of the context's grown-on-demand buffer. One array per GPU,
// shared by every call on that GPU.
constexpr size_t kScratchBytes = (size_t(24) << 20) + (size_t(128) << 10);
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
void* g_scratch_ptr[kMaxDevices] = {};
std::mutex g_scratch_mu;

}  // namespace

// Device address of this GPU's scratch array, looked up once per device with
// that device current (sem_ctx_create calls this after hipSetDevice).
int init_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) {
    set_error("device %d outside the %d devices the scratch table holds", device, kMaxDevices);
    return SEM_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return SEM_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    set_error("hipGetSymbolAddress(scratch) failed on device %d: %s", device,
              hipGetErrorString(e));
    return SEM_ERR_HIP;
  }
  g_scratch_ptr[device] = p;
  return SEM_OK;
}

namespace {

template <typename T>
int pcg_noalloc_t(sem_ctx* ctx, const sem_operator* op, const T* d, const T* b, T* x, double tol,
                  int maxiter, sem_pcg_info* info) {
  void* ws = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    ws = (ctx->device >= 0 && ctx->device < kMaxDevices) ? g_scratch_ptr[ctx->device] : nullptr;
  }
  if (!ws) {
    set_error("sem_pcg_noalloc: the scratch array of device %d is not set up", ctx->device);
    return SEM_ERR_HIP;
  }
  PcgRun<T> R;
  const size_t need = R.plan(ctx, op, d);
  if (need > kScratchBytes) {
    set_error("sem_pcg_noalloc: %lld local nodes need %zu bytes of scratch, more than its %zu "
              "(at most about %zu local nodes in %s)",
              (long long)op->mesh->nl, need, kScratchBytes,
              (kScratchBytes - (size_t(128) << 10)) / (3 * sizeof(T)),
              sizeof(T) == 8 ? "FP64" : "FP32");
    return SEM_ERR_UNSUPPORTED;
  }
  PcgState hstate{};   // read back by a plain copy: no pinned buffer
  SEM_TRY(R.bind(ws, &hstate));
  return pcg_solve(R, ctx, op, b, x, tol, maxiter, info);
}

}  // namespace

int launch_jacobi_local(hipStream_t st, const sem_operator* op, double* dl) {
  const sem_mesh* m = op->mesh;
  const sem_op_desc& ds = op->desc;
  const double* G11 = ds.G11 ? ds.G11 : m->f[SEM_G11];
  const double* G12 = ds.G12 ? ds.G12 : m->f[SEM_G12];
  const double* G22 = ds.G22 ? ds.G22 : m->f[SEM_G22];
  const double* B = ds.B ? ds.B : m->f[SEM_B];
  hipLaunchKernelGGL(k_jacobi_local, dim3(nblk(m->nl)), dim3(256), 0, st, G11, G12, G22, B,
                     m->Dr, m->Ds, ds.nu, ds.nu_f, ds.k, ds.k_f, op->has_mass ? 1 : 0, m->nr,
                     m->ns, m->L, (long)m->nl, dl);
  SEM_HIP(hipGetLastError());
  return SEM_OK;
}

int launch_inv_masked(hipStream_t st, const Layout& g, int bits, const double* d, double* dinv) {
  const long n = (long)g.L * g.NC;
  hipLaunchKernelGGL(k_inv_masked, dim3(nblk(n)), dim3(256), 0, st, d, dinv, g.L, g.NC, bits, n);
  SEM_HIP(hipGetLastError());
  return SEM_OK;
}

int pcg_run(sem_ctx* ctx, const sem_operator* op, int precond, const void* d, const void* b,
            void* x, double tol, int maxiter, sem_pcg_info* info) {
  if (precond == SEM_PRECOND_IDENTITY) d = nullptr;
  if (op->generic || precond == SEM_PRECOND_DIAG_LOCAL)
    return pcg_generic(ctx, op, static_cast<const double*>(d), static_cast<const double*>(b),
                       static_cast<double*>(x), tol, maxiter, info);
  if (op->precision == SEM_F64)
    return pcg_run_t<double>(ctx, op, static_cast<const double*>(d),
                             static_cast<const double*>(b), static_cast<double*>(x), tol, maxiter,
                             info);
  return pcg_run_t<float>(ctx, op, static_cast<const float*>(d), static_cast<const float*>(b),
                          static_cast<float*>(x), tol, maxiter, info);
}

}  // namespace sem

using namespace sem;

extern "C" int sem_pcg_noalloc(sem_ctx* ctx, const sem_operator* op, int precond, const void* d,
                               const void* b, void* x, double tol, int maxiter,
                               sem_pcg_info* info) {
  SEM_ARG(ctx && op && b && x, "NULL argument");
  SEM_ARG(op->desc.assemble, "sem_pcg_noalloc needs an assembled operator (desc.assemble = 1)");
  SEM_ARG(precond == SEM_PRECOND_IDENTITY || (precond == SEM_PRECOND_DIAG && d),
          "sem_pcg_noalloc: precond must be SEM_PRECOND_IDENTITY, or SEM_PRECOND_DIAG with d");
  SEM_ARG(x != b, "sem_pcg_noalloc: x must not alias b");
  SEM_ARG(maxiter >= 0, "sem_pcg_noalloc: maxiter < 0");
  SEM_ARG(tol == tol, "sem_pcg_noalloc: tol is NaN");
  if (precond == SEM_PRECOND_IDENTITY) d = nullptr;
  if (op->generic) {
    set_error("sem_pcg_noalloc: meshes with nr != ns or nr > %d take sem_pcg", kMaxP);
    return SEM_ERR_UNSUPPORTED;
  }
  if (op->precision == SEM_F64)
    return pcg_noalloc_t<double>(ctx, op, static_cast<const double*>(d),
                                 static_cast<const double*>(b), static_cast<double*>(x), tol,
                                 maxiter, info);
  return pcg_noalloc_t<float>(ctx, op, static_cast<const float*>(d), static_cast<const float*>(b),
                              static_cast<float*>(x), tol, maxiter, info);
}

extern "C" int sem_pcg_mixed(sem_ctx* ctx, const sem_operator* op64, const sem_operator* op_lo,
                             const double* d64, const double* b, double* x, double tol,
                             double inner_tol, int max_outer, int max_inner,
                             sem_pcg_info* info) {
  SEM_ARG(ctx && op64 && op_lo && b && x, "NULL argument");
  SEM_ARG(op64->precision == SEM_F64, "sem_pcg_mixed: op64 must be SEM_F64");
  SEM_ARG(op_lo->precision != SEM_F64, "sem_pcg_mixed: op_lo must be a low-precision operator");
  SEM_ARG(op64->mesh == op_lo->mesh, "sem_pcg_mixed: operators on different meshes");
  SEM_ARG(op64->desc.assemble && op_lo->desc.assemble, "sem_pcg_mixed: operators must assemble");
  if (op64->generic || op_lo->generic) {
    set_error("sem_pcg_mixed: needs nr == ns <= %d", kMaxP);
    return SEM_ERR_UNSUPPORTED;
  }
  SEM_ARG(max_outer >= 1 && max_inner >= 1 && inner_tol > 0 && inner_tol < 1 && tol >= 0,
          "sem_pcg_mixed: bad limits");
  return mixed_run(ctx, op64, op_lo, d64, b, x, tol, inner_tol, max_outer, max_inner, info);
}

namespace sem {

}  // namespace sem

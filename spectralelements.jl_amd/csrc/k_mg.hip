// Multigrid preconditioner for the fused SEM operator and PCG with the
// V-cycle as opM (pcg.jl:16-60 with a stronger preconditioner than the
// identity / Jacobi of diffusion.jl:47-49; the reference's own next step is
// the commented element fast-diagonalisation sketch of lapl.jl:105-118).
//
// Hierarchy: p-multigrid on the element grid (N -> N/2 -> ... -> 1), then
// h-multigrid on bilinear elements (2 x 2 elements -> 1). Every level is a
// library mesh + fused operator (k_ax + k_seams), rediscretised with the same
// scalar nu, k and Dirichlet bits on nodes that sample the fine geometry.
// Smoother: Chebyshev-Jacobi (Saad, Iterative Methods for Sparse Linear
// Systems, Alg. 12.1) with host-side coefficients; transfers: one thread per
// output node, the element's interpolation rows from a small table.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "_gen/gll_dtables.h"
#include "sem_device.h"
#include "sem_l0.h"

namespace sem {
namespace {

inline int nblk(long n, int bs = 256) { return (int)((n + bs - 1) / bs); }

// The V-cycle's vector kernels are typed by the cycle's precision T (FP64,
// or FP32 for the FP32 cycle); the coefficients come in FP64 from the host.

// Chebyshev start: d = dinv .* r / theta; x = d (x0 = 0, and r = b copied
// from b) or x += d.
template <typename T>
__global__ void k_cheb_start(const T* __restrict__ b, T* __restrict__ r,
                             const T* __restrict__ dinv, T* __restrict__ d, T* __restrict__ x,
                             double theta, int from_zero, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  T rv;
  if (from_zero) {
    rv = b[t];
    r[t] = rv;
  } else {
    rv = r[t];
  }
  const T dv = dinv[t] * rv / T(theta);
  d[t] = dv;
  x[t] = from_zero ? dv : x[t] + dv;
}

// One Chebyshev step after w = A d: r -= w; d = c1 d + c2 dinv .* r; x += d
template <typename T>
__global__ void k_cheb_step(const T* __restrict__ w, T* __restrict__ r,
                            const T* __restrict__ dinv, T* __restrict__ d, T* __restrict__ x,
                            double c1, double c2, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const T rv = r[t] - w[t];
  r[t] = rv;
  const T dv = cheb_dir(T(c1), d[t], T(c2), dinv[t], rv);
  d[t] = dv;
  x[t] = x[t] + dv;
}

template <typename T>
__global__ void k_sub(T* __restrict__ r, const T* __restrict__ w, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) r[t] = r[t] - w[t];
}

// Transfer geometry: fine element ef = (ec << h) + sub; J holds nsub
// row-major (pf x pc) blocks for x, then for y.
struct Xfer {
  int pf, pc, hx, hy;
  int Lf, NCf, Lc, NCc;
  int per_x, per_y;
};

// e = P xc (per fine element: Jx xc Jy'), x += e. Row owner: thread i of
// grid row eyf holds fine row i of fine element row eyf; it forms the
// x-contraction s(bb) = sum_aa Jx(a, aa) xc(aa, bb) once and then every
// column b of the element, so each coarse value is loaded pc / pf times per
// fine node and there is one integer division per thread. Same operation
// order as the one-node-per-thread form: v = sum_bb Jy(b, bb) s(bb).
constexpr int kXferMaxP = 16;
template <typename T>
__global__ void __launch_bounds__(256) k_prolong(const T* __restrict__ xc, const T* __restrict__ J,
                                                  T* __restrict__ e, T* __restrict__ x, Xfer g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= g.Lf) return;
  const int eyf = blockIdx.y;
  const int exf = i / g.pf, a = i - exf * g.pf;
  const int sx = exf & g.hx, sy = eyf & g.hy;
  const int exc = exf >> g.hx, eyc = eyf >> g.hy;
  const int blk = g.pf * g.pc;
  const T* jx = J + sx * blk + a * g.pc;
  const T* jy = J + (g.hx ? 2 : 1) * blk + sy * blk;
  const long c0 = (long)exc * g.pc + (long)g.Lc * ((long)eyc * g.pc);
  T sv[kXferMaxP];
#pragma unroll
  for (int bb = 0; bb < kXferMaxP; ++bb) {
    if (bb < g.pc) {
      T sacc = T(0);
      for (int aa = 0; aa < g.pc; ++aa) sacc += jx[aa] * xc[c0 + aa + (long)g.Lc * bb];
      sv[bb] = sacc;
    }
  }
  const long f0 = i + (long)g.Lf * ((long)eyf * g.pf);
  for (int b = 0; b < g.pf; ++b) {
    T v = T(0);
#pragma unroll
    for (int bb = 0; bb < kXferMaxP; ++bb)
      if (bb < g.pc) v += jy[b * g.pc + bb] * sv[bb];
    const long t = f0 + (long)g.Lf * b;
    e[t] = v;
    x[t] = x[t] + v;
  }
}

// rc (element-local) = Jx' (mult_f .* rf) Jy summed over the fine nodes of
// the coarse element; assembled and masked afterwards by the coarse QQ'.
// A workgroup takes one coarse element row (grid row) and the fine rows of
// ce whole coarse elements: it stages those rows of the element row's fine
// columns in LDS with the x multiplicity applied (coalesced loads; the
// multiplicities are 1 or 1/2, so the product is exact), then forms the
// x-contractions s(a, c) = sum_(sx, af) Jx(af, a) rf(af, c) once per coarse
// row and fine column, and each coarse node v = sum_c wy(c, b) s(a, c) with
// wy = Jy my: the operation order of the one-node-per-thread form, so the
// bits are the same.
template <typename T>
__global__ void __launch_bounds__(256) k_restrict(const T* __restrict__ rf, const T* __restrict__ J,
                                                   T* __restrict__ rc, Xfer g) {
  extern __shared__ unsigned char xfer_lds[];
  const int nsx = g.hx ? 2 : 1, nsy = g.hy ? 2 : 1;
  const int fw = nsx * g.pf;          // fine rows per coarse element
  const int ce = 256 / fw;            // coarse elements per workgroup
  const int FR = ce * fw;             // fine rows staged
  const int nfc = nsy * g.pf;         // fine columns of the coarse element row
  const int rows = ce * g.pc;         // coarse rows of the workgroup
  const int blk = g.pf * g.pc;
  T* tile = reinterpret_cast<T*>(xfer_lds);   // [nfc][FR]
  T* sx_ = tile + FR * nfc;                   // [nfc][rows]
  T* jx = sx_ + rows * nfc;                   // [nsx * blk]: J's x blocks
  T* wy = jx + nsx * blk;                     // [nfc][pc]: Jy times the y multiplicity
  const int eyc = blockIdx.y;
  const int i0 = blockIdx.x * FR;
  const int jf0 = eyc * nfc;
  for (int q = threadIdx.x; q < FR * nfc; q += blockDim.x) {
    const int c = q / FR, r = q - c * FR;
    const int ifn = i0 + r;
    T v = T(0);
    if (ifn < g.Lf) v = T(mult1(ifn, g.pf, g.Lf, g.per_x)) * rf[ifn + (long)g.Lf * (jf0 + c)];
    tile[c * FR + r] = v;
  }
  for (int q = threadIdx.x; q < nsx * blk; q += blockDim.x) jx[q] = J[q];
  for (int q = threadIdx.x; q < nfc * g.pc; q += blockDim.x) {
    const int c = q / g.pc, b = q - c * g.pc;
    const int sy = c / g.pf, bf = c - sy * g.pf;
    const int jf = ((eyc << g.hy) + sy) * g.pf + bf;
    wy[q] = J[nsx * blk + sy * blk + bf * g.pc + b] * T(mult1(jf, g.pf, g.NCf, g.per_y));
  }
  __syncthreads();
  for (int q = threadIdx.x; q < rows * nfc; q += blockDim.x) {
    const int c = q / rows, rl = q - c * rows;
    const int exl = rl / g.pc, a = rl - exl * g.pc;
    const T* trow = tile + c * FR + exl * fw;
    T sacc = T(0);
    for (int sx = 0; sx < nsx; ++sx)
      for (int af = 0; af < g.pf; ++af) sacc += jx[sx * blk + af * g.pc + a] * trow[sx * g.pf + af];
    sx_[c * rows + rl] = sacc;
  }
  __syncthreads();
  const int exc0 = i0 / fw;
  for (int o = threadIdx.x; o < rows * g.pc; o += blockDim.x) {
    const int b = o / rows, rl = o - b * rows;
    const int ic = exc0 * g.pc + rl;
    if (ic >= g.Lc) continue;
    T v = T(0);
    for (int c = 0; c < nfc; ++c) v += wy[c * g.pc + b] * sx_[c * rows + rl];
    rc[ic + (long)g.Lc * (eyc * g.pc + b)] = v;
  }
}

// coarse coordinates: the fine field interpolated at the coarse nodes; Jc is
// (pc x pf) row-major, node a of a coarse element lies in fine sub-element
// (subbits >> a) & 1 (h-merge only)
__global__ void k_coarse_coords(const double* __restrict__ xf, const double* __restrict__ Jc,
                                int subbits, double* __restrict__ xc, Xfer g, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = (int)(t % g.Lc), j = (int)(t / g.Lc);
  const int exc = i / g.pc, a = i - exc * g.pc;
  const int eyc = j / g.pc, b = j - eyc * g.pc;
  const int exf = (exc << g.hx) + (g.hx ? (subbits >> a) & 1 : 0);
  const int eyf = (eyc << g.hy) + (g.hy ? (subbits >> b) & 1 : 0);
  const long f0 = (long)exf * g.pf + (long)g.Lf * ((long)eyf * g.pf);
  double v = 0.0;
  for (int bf = 0; bf < g.pf; ++bf) {
    double s = 0.0;
    for (int af = 0; af < g.pf; ++af) s += Jc[a * g.pf + af] * xf[f0 + af + (long)g.Lf * bf];
    v += Jc[b * g.pf + bf] * s;
  }
  xc[t] = v;
}

// ---- the bilinear coarse levels in one workgroup ---------------------------
// Levels of 2 x 2-node elements with at most kBlockNodes local nodes are
// cycled by one 1024-thread workgroup, phase after phase with a barrier in
// between, instead of one launch per vector pass (each of those is bound by
// dispatch latency at these sizes). The element operator follows the fused
// apply's P = 2 arithmetic; QQ' sums x-pairs, then y-pairs, as the fused
// apply and its seam pass do; restriction, the coarse QQ' and prolongation
// are the one-node-per-thread forms of the separate kernels.
constexpr int kBlockThreads = 1024;
constexpr long kBlockNodes = 1024;
constexpr int kBlockMaxLevels = 12;
constexpr int kBlockMaxDeg = 16;

template <typename T>
struct BlockLevel {
  int Ex, Ey, L, NC, per_x, per_y, bc, has_mass;
  int n;                         // local nodes
  long ng;                       // global nodes (its QQ' map)
  const int32_t* perm;
  const int32_t* seg;
  const T *G11, *G12, *G22, *B;
  T nu, k;
  T *x, *b, *r, *d, *w, *e;
  const T* dinv;
  const T* J;                    // transfer rows from the next level onto this one
  Xfer xf;
  int deg;                       // Chebyshev degree on this level
  double theta;
  double c1[kBlockMaxDeg], c2[kBlockMaxDeg];   // step i = 1 .. deg - 1
};
template <typename T>
struct BlockCycle {
  int nlev;
  BlockLevel<T> lv[kBlockMaxLevels];
};

// element-local H u at every copy (the fused apply's P = 2 operations)
template <typename T>
__device__ void blk_local(const BlockLevel<T>& V, const T* u, T* h) {
#define D2(ix) T(gll_d<2>(ix))
  for (int t = threadIdx.x; t < V.n; t += blockDim.x) {
    const int i = t % V.L, j = t / V.L;
    const int a = i & 1, b = j & 1;
    const int t0 = (i - a) + V.L * (j - b);
    auto at = [&](int al, int be) { return t0 + al + V.L * be; };
    T uu[2][2], g11[2][2], g12[2][2], g22[2][2];
#pragma unroll
    for (int al = 0; al < 2; ++al)
#pragma unroll
      for (int be = 0; be < 2; ++be) {
        uu[al][be] = u[at(al, be)];
        g11[al][be] = V.G11[at(al, be)];
        g12[al][be] = V.G12[at(al, be)];
        g22[al][be] = V.G22[at(al, be)];
      }
    // us(al, be) = sum_k u(al, k) Ds(be + 2k); ur(al, be) = sum_k Dr(al + 2k) u(k, be)
    T us[2][2], ur[2][2];
#pragma unroll
    for (int al = 0; al < 2; ++al)
#pragma unroll
      for (int be = 0; be < 2; ++be) {
        T acc = T(0);
#pragma unroll
        for (int k = 0; k < 2; ++k) acc += uu[al][k] * D2(be + 2 * k);
        us[al][be] = acc;
        T acc2 = T(0);
#pragma unroll
        for (int k = 0; k < 2; ++k) acc2 += D2(al + 2 * k) * uu[k][be];
        ur[al][be] = acc2;
      }
    T wr[2][2], ws[2][2];
#pragma unroll
    for (int al = 0; al < 2; ++al)
#pragma unroll
      for (int be = 0; be < 2; ++be) {
        wr[al][be] = g11[al][be] * ur[al][be] + g12[al][be] * us[al][be];
        ws[al][be] = g12[al][be] * ur[al][be] + g22[al][be] * us[al][be];
      }
    // this copy's row a and column b, selected (no indexing of register
    // arrays by runtime values)
    const T wsa[2] = {a ? ws[1][0] : ws[0][0], a ? ws[1][1] : ws[0][1]};
    const T wrb[2] = {b ? wr[0][1] : wr[0][0], b ? wr[1][1] : wr[1][0]};
    const T db[2] = {b ? D2(2) : D2(0), b ? D2(3) : D2(1)};   // Ds(k + 2b)
    const T da[2] = {a ? D2(2) : D2(0), a ? D2(3) : D2(1)};   // Dr(k + 2a)
    const T uab = a ? (b ? uu[1][1] : uu[1][0]) : (b ? uu[0][1] : uu[0][0]);
    T p2 = T(0), p1 = T(0);
#pragma unroll
    for (int k = 0; k < 2; ++k) p2 += wsa[k] * db[k];
#pragma unroll
    for (int k = 0; k < 2; ++k) p1 += da[k] * wrb[k];
    T hv;
    if (V.has_mass) hv = V.nu * (p1 + p2) + V.k * (V.B[t] * uab);
    else hv = V.nu * (p1 + p2);
    h[t] = hv;
  }
#undef D2
}

// the assembled, masked value of every copy from the element-local values h,
// fed to the epilogue (0: y = w; 1: Chebyshev step on r, d, x in place;
// 2: r -= w)
template <typename T, int KIND>
__device__ void blk_final(const BlockLevel<T>& V, const T* h, T* y, double c1, double c2) {
  for (int t = threadIdx.x; t < V.n; t += blockDim.x) {
    const int i = t % V.L, j = t / V.L;
    const int ex = i >> 1, a = i & 1, ey = j >> 1, b = j & 1;
    const bool hx = a == 0 ? (ex > 0 || V.per_x) : (ex < V.Ex - 1 || V.per_x);
    const bool hy = b == 0 ? (ey > 0 || V.per_y) : (ey < V.Ey - 1 || V.per_y);
    const int ip = a == 0 ? (ex > 0 ? i - 1 : V.L - 1) : (ex < V.Ex - 1 ? i + 1 : 0);
    const int jp = b == 0 ? (ey > 0 ? j - 1 : V.NC - 1) : (ey < V.Ey - 1 ? j + 1 : 0);
    const int il = a == 0 ? ip : i, ir = a == 0 ? i : ip;   // left, right copies
    const int jl = b == 0 ? jp : j, ju = b == 0 ? j : jp;   // lower, upper copies
    auto xs = [&](int jj) {
      const long cj = (long)V.L * jj;
      return hx ? h[il + cj] + h[ir + cj] : h[i + cj];
    };
    const T s = hy ? xs(jl) + xs(ju) : xs(j);
    const T wv = s * mval<T>(i, j, V.L, V.NC, V.bc);
    if constexpr (KIND == 0) {
      y[t] = wv;
    } else {
      const T rv = V.r[t] - wv;
      V.r[t] = rv;
      if constexpr (KIND == 1) {
        const T dv = cheb_dir(T(c1), V.d[t], T(c2), V.dinv[t], rv);
        V.d[t] = dv;
        V.x[t] = V.x[t] + dv;
      }
    }
  }
}

// Chebyshev smoothing of degree deg on level V (the separate kernels' steps;
// from zero: x = 0, r = b), with the residual of the final x when want_r
template <typename T>
__device__ void blk_cheby(const BlockLevel<T>& V, int from_zero, int deg, bool want_r) {
  for (int t = threadIdx.x; t < V.n; t += blockDim.x) {
    T rv;
    if (from_zero) {
      rv = V.b[t];
      V.r[t] = rv;
    } else {
      rv = V.r[t];
    }
    const T dv = V.dinv[t] * rv / T(V.theta);
    V.d[t] = dv;
    V.x[t] = from_zero ? dv : V.x[t] + dv;
  }
  __syncthreads();
  for (int it = 1; it < deg; ++it) {
    blk_local<T>(V, V.d, V.w);
    __syncthreads();
    blk_final<T, 1>(V, V.w, nullptr, V.c1[it], V.c2[it]);
    __syncthreads();
  }
  if (want_r) {
    blk_local<T>(V, V.d, V.w);
    __syncthreads();
    blk_final<T, 2>(V, V.w, nullptr, 0.0, 0.0);
    __syncthreads();
  }
}

// the V-cycle on levels lv[0] (the finest taken here) .. lv[nlev - 1]; lv[0].b
// holds the right-hand side and lv[0].x receives the result
template <typename T>
__global__ void __launch_bounds__(kBlockThreads) k_block_cycle(const BlockCycle<T>* __restrict__ cc) {
  const int nlev = cc->nlev;
  for (int l = 0; l < nlev - 1; ++l) {
    const BlockLevel<T>& F = cc->lv[l];
    const BlockLevel<T>& C = cc->lv[l + 1];
    blk_cheby<T>(F, 1, F.deg, true);
    // restriction of r onto C.b (per coarse copy), then the coarse QQ', masked
    const Xfer& g = F.xf;
    const int nsx = g.hx ? 2 : 1, nsy = g.hy ? 2 : 1, blk = g.pf * g.pc;
    for (int t = threadIdx.x; t < C.n; t += blockDim.x) {
      const int i = t % g.Lc, j = t / g.Lc;
      const int exc = i / g.pc, a = i - exc * g.pc;
      const int eyc = j / g.pc, b = j - eyc * g.pc;
      const T* Jy0 = F.J + nsx * blk;
      T v = T(0);
      for (int sy = 0; sy < nsy; ++sy) {
        const int eyf = (eyc << g.hy) + sy;
        for (int bf = 0; bf < g.pf; ++bf) {
          const int jf = eyf * g.pf + bf;
          const T wy = Jy0[sy * blk + bf * g.pc + b] * T(mult1(jf, g.pf, g.NCf, g.per_y));
          T sacc = T(0);
          for (int sx = 0; sx < nsx; ++sx) {
            const int exf = (exc << g.hx) + sx;
            for (int af = 0; af < g.pf; ++af) {
              const int ifn = exf * g.pf + af;
              sacc += F.J[sx * blk + af * g.pc + a] *
                      (T(mult1(ifn, g.pf, g.Lf, g.per_x)) * F.r[ifn + (long)g.Lf * jf]);
            }
          }
          v += wy * sacc;
        }
      }
      C.b[t] = v;
    }
    __syncthreads();
    for (long q = threadIdx.x; q < C.ng; q += blockDim.x)
      gsmap_node(C.perm, C.seg, C.b, C.b, q, C.bc, C.L, C.NC);
    __syncthreads();
  }
  {
    const BlockLevel<T>& V = cc->lv[nlev - 1];
    blk_cheby<T>(V, 1, V.deg, false);
  }
  for (int l = nlev - 2; l >= 0; --l) {
    const BlockLevel<T>& F = cc->lv[l];
    const BlockLevel<T>& C = cc->lv[l + 1];
    const Xfer& g = F.xf;
    const int blk = g.pf * g.pc;
    for (int t = threadIdx.x; t < F.n; t += blockDim.x) {
      const int i = t % g.Lf, j = t / g.Lf;
      const int exf = i / g.pf, a = i - exf * g.pf;
      const int eyf = j / g.pf, b = j - eyf * g.pf;
      const int sx = exf & g.hx, sy = eyf & g.hy;
      const int exc = exf >> g.hx, eyc = eyf >> g.hy;
      const T* jx = F.J + sx * blk + a * g.pc;
      const T* jy = F.J + (g.hx ? 2 : 1) * blk + sy * blk + b * g.pc;
      const long c0 = (long)exc * g.pc + (long)g.Lc * ((long)eyc * g.pc);
      T v = T(0);
      for (int bb = 0; bb < g.pc; ++bb) {
        T sacc = T(0);
        for (int aa = 0; aa < g.pc; ++aa) sacc += jx[aa] * C.x[c0 + aa + (long)g.Lc * bb];
        v += jy[bb] * sacc;
      }
      F.e[t] = v;
      F.x[t] = F.x[t] + v;
    }
    __syncthreads();
    blk_local<T>(F, F.e, F.w);
    __syncthreads();
    blk_final<T, 2>(F, F.w, nullptr, 0.0, 0.0);
    __syncthreads();
    blk_cheby<T>(F, 0, F.deg, false);
  }
}

// hashed start vector of the power iteration: M .* (h(g) / 2^32 - 1/2),
// h(g) = (g * 2654435761 + 12345) mod 2^32 of the global node id
__global__ void k_hash_start(const int32_t* __restrict__ l2g, double* __restrict__ v, int L,
                             int NC, int bc, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const uint32_t h = (uint32_t)l2g[t] * 2654435761u + 12345u;
  const int i = (int)(t % L), j = (int)(t / L);
  v[t] = mask_on(i, j, L, NC, bc) ? (double)h / 4294967296.0 - 0.5 : 0.0;
}

__global__ void k_mul2(const double* __restrict__ a, const double* __restrict__ b,
                       double* __restrict__ y, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) y[t] = a[t] * b[t];
}

__global__ void k_div_scalar(const double* __restrict__ a, const double* __restrict__ s,
                             double* __restrict__ y, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) y[t] = a[t] / s[0];
}

// v .* D (D = 1 ./ dinv where dinv != 0, else 0): the diagonal weight of the
// Rayleigh quotient's denominator
__global__ void k_times_diag(const double* __restrict__ v, const double* __restrict__ dinv,
                             double* __restrict__ y, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) {
    const double di = dinv[t];
    y[t] = di != 0.0 ? v[t] * (1.0 / di) : 0.0;
  }
}

// x += alpha p, r -= alpha w with alpha = <z, r> / <p, A p> read from the
// device (pcg.jl:50-52)
__global__ void k_xr(double* __restrict__ x, double* __restrict__ r, const double* __restrict__ p,
                     const double* __restrict__ w, const double* __restrict__ tz,
                     const double* __restrict__ tpw, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const double alpha = *tz / *tpw;
  if (t < n) {
    x[t] = x[t] + alpha * p[t];
    r[t] = r[t] - alpha * w[t];
  }
}

// p = a z + beta p with beta = t / t_prev, or (t - s) / t_prev when s is
// given (the flexible form of the FP32 cycle), read from the device
// (pcg.jl:49)
__global__ void k_pdir(double a, const double* __restrict__ z, double* p,
                       const double* __restrict__ tz, const double* __restrict__ tprev,
                       const double* __restrict__ s, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const double beta = (s ? *tz - *s : *tz) / *tprev;
  if (t < n) p[t] = a * z[t] + beta * p[t];
}

// Interpolation matrix rows with exact unit rows where an output point equals
// an input point (so copies of a shared node interpolate to identical bits).
std::vector<double> interp_rows(int no, const double* xo, int ni, const double* xi) {
  std::vector<double> Jc((size_t)no * ni), J((size_t)no * ni);
  interpmat(no, xo, ni, xi, Jc.data());   // column-major
  for (int i = 0; i < no; ++i)
    for (int j = 0; j < ni; ++j) J[(size_t)i * ni + j] = Jc[i + (size_t)no * j];
  for (int i = 0; i < no; ++i)
    for (int j = 0; j < ni; ++j)
      if (xo[i] == xi[j]) {
        for (int q = 0; q < ni; ++q) J[(size_t)i * ni + q] = q == j ? 1.0 : 0.0;
        break;
      }
  return J;
}

}  // namespace
}  // namespace sem

using namespace sem;

struct sem_mg {
  struct Level {
    sem_mesh* mesh = nullptr;
    bool own_mesh = false;
    const sem_operator* op = nullptr;
    bool own_op = false;
    double* dinv = nullptr;
    double lmax = 0.0;
    int h = 0;            // the next level merges this level's elements 2 x 2
    double* J = nullptr;  // transfer rows from the next level onto this one
    Xfer xf{};
    long nJ = 0;
    double *x = nullptr, *b = nullptr, *r = nullptr, *d = nullptr, *w = nullptr, *e = nullptr;
    // the FP32 cycle: the level operator in FP32 and FP32 copies of the above
    const sem_operator* op32 = nullptr;
    float *fx = nullptr, *fb = nullptr, *fr = nullptr, *fd = nullptr, *fw = nullptr,
          *fe = nullptr, *fdinv = nullptr, *fJ = nullptr;
    // the level's applies carry the smoother's vector updates as a fused
    // epilogue (FP64 / FP32 cycle)
    bool fuse64 = false, fuse32 = false;
    // the device blocks the level vectors above live in (one allocation per
    // precision)
    double* blk64 = nullptr;
    float* blk32 = nullptr;
  };
  sem_mg_params prm{};
  int device = 0;
  const sem_operator* op = nullptr;
  bool f32 = false;
  // smoother updates fused into the level applies (SEM_MG_FUSE=0 at
  // creation turns it off: a test hook)
  bool fuse = true;
  // the bilinear coarse levels lc .. last cycled by one workgroup (lc < 0:
  // none; SEM_MG_BLOCK=0 at creation turns it off), with their level
  // descriptors in device memory per cycle precision
  int lc = -1;
  bool block = true;
  void* bdesc64 = nullptr;
  void* bdesc32 = nullptr;
  std::vector<Level> lv;
  // PCG vectors on the fine level (pq: the previous residual of the flexible
  // recurrence), reduction partials and the scalar slot
  double *pr = nullptr, *pp = nullptr, *pw = nullptr, *pz = nullptr, *pq = nullptr,
         *part = nullptr;
};

namespace {

constexpr int kParts = 256;
// device scalars after the partials: [0] reduction result, [1], [2] <z, r> of
// the even and odd PCG iterations, [3] <p, A p>, [4] the flexible beta's s
constexpr int kScalars = 8;

void mg_free(sem_mg* mg) {
  if (!mg) return;
  (void)hipDeviceSynchronize();
  for (auto& L : mg->lv) {
    double* bufs[] = {L.blk64, L.J};
    for (double* p : bufs)
      if (p) (void)hipFree(p);
    float* fbufs[] = {L.blk32, L.fJ};
    for (float* p : fbufs)
      if (p) (void)hipFree(p);
    if (L.op32) sem_operator_destroy(const_cast<sem_operator*>(L.op32));
    if (L.own_op) sem_operator_destroy(const_cast<sem_operator*>(L.op));
    if (L.own_mesh) sem_mesh_destroy(L.mesh);
  }
  double* bufs[] = {mg->pr, mg->pp, mg->pw, mg->pz, mg->pq, mg->part};
  for (double* p : bufs)
    if (p) (void)hipFree(p);
  if (mg->bdesc64) (void)hipFree(mg->bdesc64);
  if (mg->bdesc32) (void)hipFree(mg->bdesc32);
  delete mg;
}

template <typename T>
int dalloc(T** p, long n) {
  SEM_HIP(hipMalloc(p, sizeof(T) * std::max<long>(n, 1)));
  return SEM_OK;
}

// a level's buffers and operator in the cycle's precision
template <typename T>
struct LvBuf {
  T *x, *b, *r, *d, *w, *e, *dinv, *J;
  const sem_operator* op;
  bool fuse;
};
template <typename T>
LvBuf<T> bufs(sem_mg::Level& L);
template <>
LvBuf<double> bufs<double>(sem_mg::Level& L) {
  return {L.x, L.b, L.r, L.d, L.w, L.e, L.dinv, L.J, L.op, L.fuse64};
}
template <>
LvBuf<float> bufs<float>(sem_mg::Level& L) {
  return {L.fx, L.fb, L.fr, L.fd, L.fw, L.fe, L.fdinv, L.fJ, L.op32, L.fuse32};
}
// the fused Chebyshev step leaves the new direction in w: the two buffers
// trade roles
template <typename T>
void swap_dw(sem_mg::Level& L);
template <>
void swap_dw<double>(sem_mg::Level& L) {
  std::swap(L.d, L.w);
}
template <>
void swap_dw<float>(sem_mg::Level& L) {
  std::swap(L.fd, L.fw);
}

// host value of a device reduction result
int fetch(sem_ctx* ctx, const double* dev, double* out) {
  SEM_TRY(ensure_pinned(ctx, 128));
  double* h = reinterpret_cast<double*>(static_cast<char*>(ctx->pinned) + 64);
  SEM_HIP(hipMemcpyAsync(h, dev, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  SEM_HIP(hipStreamSynchronize(ctx->stream));
  *out = *h;
  return SEM_OK;
}

int wdot_host(sem_ctx* ctx, sem_mg* mg, const Layout& g, const double* a, const double* b,
              double* out) {
  SEM_TRY(launch_wdot(ctx->stream, g, a, b, mg->part, kParts, mg->part + kParts));
  return fetch(ctx, mg->part + kParts, out);
}

int norm_host(sem_ctx* ctx, sem_mg* mg, long n, const double* a, double* out) {
  SEM_TRY(launch_norm_inf(ctx->stream, n, a, mg->part, kParts, mg->part + kParts));
  return fetch(ctx, mg->part + kParts, out);
}

int apply(sem_ctx* ctx, const sem_operator* op, const void* u, void* y) {
  return sem_operator_apply(ctx, op, u, y);
}

// r -= A v (fused into the apply's final writes where the level allows;
// w is scratch either way); with start_theta the fused form also starts the
// Chebyshev smoother from the new r (d = dinv .* r / theta; x += d) and
// returns *started = true
template <typename T>
int residual_update(sem_ctx* ctx, const LvBuf<T>& L, const T* v, const double* start_theta = nullptr,
                    bool* started = nullptr) {
  if (started) *started = false;
  if (L.fuse) {
    ApplyEpi e;
    e.kind = start_theta ? 3 : 2;
    e.r = L.r;
    if (start_theta) {
      e.x = L.x;
      e.dinv = L.dinv;
      e.d = L.d;
      e.theta = *start_theta;
    }
    SEM_TRY(operator_apply_epi(ctx, L.op, v, L.w, e));
    if (started) *started = start_theta != nullptr;
    return SEM_OK;
  }
  SEM_TRY(apply(ctx, L.op, v, L.w));
  const long n = (long)L.op->mesh->nl;
  hipLaunchKernelGGL((k_sub<T>), dim3(nblk(n)), dim3(256), 0, ctx->stream, L.r, L.w, n);
  SEM_HIP(hipGetLastError());
  return SEM_OK;
}

// Chebyshev iteration of degree deg for D^-1 A on [lmax / ratio, lmax]; from
// zero (x = 0, r = b) or from (x, r). With want_r the residual of the final
// x is formed in r (one more apply).
template <typename T>
int cheby(sem_ctx* ctx, sem_mg* mg, sem_mg::Level& Lv, const T* b, int from_zero, int deg,
          bool want_r, bool started = false) {
  hipStream_t st = ctx->stream;
  LvBuf<T> L = bufs<T>(Lv);
  const long n = (long)Lv.mesh->nl;
  const double lmax = Lv.lmax, lmin = Lv.lmax / mg->prm.ratio;
  const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
  const double sigma = theta / delta;
  double rho = 1.0 / sigma;
  if (!started) {   // else the residual update formed d and advanced x
    hipLaunchKernelGGL((k_cheb_start<T>), dim3(nblk(n)), dim3(256), 0, st, b, L.r, L.dinv, L.d,
                       L.x, theta, from_zero, n);
    SEM_HIP(hipGetLastError());
  }
  for (int i = 1; i < deg; ++i) {
    const double rho_n = 1.0 / (2.0 * sigma - rho);
    if (L.fuse) {
      // w = A d with r -= w; d' = c1 d + c2 dinv .* r; x += d' fused into
      // its final writes (d' lands in w, which becomes the direction)
      ApplyEpi e;
      e.kind = 1;
      e.r = L.r;
      e.x = L.x;
      e.dinv = L.dinv;
      e.c1 = rho_n * rho;
      e.c2 = 2.0 * rho_n / delta;
      SEM_TRY(operator_apply_epi(ctx, L.op, L.d, L.w, e));
      swap_dw<T>(Lv);
      L = bufs<T>(Lv);
    } else {
      SEM_TRY(apply(ctx, L.op, L.d, L.w));
      hipLaunchKernelGGL((k_cheb_step<T>), dim3(nblk(n)), dim3(256), 0, st, L.w, L.r, L.dinv,
                         L.d, L.x, rho_n * rho, 2.0 * rho_n / delta, n);
      SEM_HIP(hipGetLastError());
    }
    rho = rho_n;
  }
  if (want_r) SEM_TRY(residual_update<T>(ctx, L, L.d));
  return SEM_OK;
}

// transfer launches; the geometry is checked against the kernels' limits
template <typename T>
int launch_prolong(hipStream_t st, const T* xc, const T* J, T* e, T* x, const Xfer& g) {
  if (g.pc > kXferMaxP || g.pf > kXferMaxP || g.NCf % g.pf != 0) {
    set_error("multigrid prolongation: unsupported transfer (%d -> %d points)", g.pc, g.pf);
    return SEM_ERR_UNSUPPORTED;
  }
  hipLaunchKernelGGL((k_prolong<T>), dim3(nblk(g.Lf), g.NCf / g.pf), dim3(256), 0, st, xc, J, e,
                     x, g);
  SEM_HIP(hipGetLastError());
  return SEM_OK;
}
template <typename T>
int launch_restrict(hipStream_t st, const T* rf, const T* J, T* rc, const Xfer& g) {
  const int fw = (g.hx ? 2 : 1) * g.pf, nfc = (g.hy ? 2 : 1) * g.pf;
  if (fw > 256 || g.pc > kXferMaxP || g.NCc % g.pc != 0) {
    set_error("multigrid restriction: unsupported transfer (%d -> %d points)", g.pf, g.pc);
    return SEM_ERR_UNSUPPORTED;
  }
  const int FR = (256 / fw) * fw, rows = (256 / fw) * g.pc;
  const int nsx = g.hx ? 2 : 1;
  const size_t lds =
      sizeof(T) * ((size_t)FR * nfc + (size_t)rows * nfc + (size_t)nsx * g.pf * g.pc + (size_t)nfc * g.pc);
  if (lds > 65536) {
    set_error("multigrid restriction: %zu bytes of LDS for %d -> %d points", lds, g.pf, g.pc);
    return SEM_ERR_UNSUPPORTED;
  }
  hipLaunchKernelGGL((k_restrict<T>), dim3((g.Lf + FR - 1) / FR, g.NCc / g.pc), dim3(256), lds, st,
                     rf, J, rc, g);
  SEM_HIP(hipGetLastError());
  return SEM_OK;
}

// V-cycle on level l: x = V(b) in the level buffers of precision T
template <typename T>
int vcycle(sem_ctx* ctx, sem_mg* mg, int l) {
  hipStream_t st = ctx->stream;
  if (l == mg->lc) {
    const void* desc = sizeof(T) == 8 ? mg->bdesc64 : mg->bdesc32;
    hipLaunchKernelGGL((k_block_cycle<T>), dim3(1), dim3(kBlockThreads), 0, st,
                       static_cast<const BlockCycle<T>*>(desc));
    SEM_HIP(hipGetLastError());
    return SEM_OK;
  }
  sem_mg::Level& Lv = mg->lv[l];
  LvBuf<T> L = bufs<T>(Lv);
  const int last = (int)mg->lv.size() - 1;
  if (l == last) return cheby<T>(ctx, mg, Lv, L.b, 1, mg->prm.degree_coarse, false);
  SEM_TRY(cheby<T>(ctx, mg, Lv, L.b, 1, mg->prm.degree, true));
  L = bufs<T>(Lv);   // the smoother may have swapped d and w
  sem_mg::Level& Cv = mg->lv[l + 1];
  const LvBuf<T> C = bufs<T>(Cv);
  SEM_TRY(launch_restrict<T>(st, L.r, L.J, C.b, Lv.xf));
  SEM_HIP(hipGetLastError());
  SEM_TRY(launch_gsmap_masked(st, Cv.mesh->map, Cv.mesh->layout(), Cv.op->desc.bc_bits, C.b,
                              C.b));
  SEM_TRY(vcycle<T>(ctx, mg, l + 1));
  SEM_TRY(launch_prolong<T>(st, C.x, L.J, L.e, L.x, Lv.xf));
  // the post-smoother's start (from r) rides on the residual update
  const double theta = 0.5 * (Lv.lmax + Lv.lmax / mg->prm.ratio);
  bool started = false;
  SEM_TRY(residual_update<T>(ctx, L, L.e, &theta, &started));
  return cheby<T>(ctx, mg, Lv, L.b, 0, mg->prm.degree, false, started);
}

// lmax = safety x Rayleigh quotient after npow power steps of D^-1 A
int estimate_lmax(sem_ctx* ctx, sem_mg* mg, sem_mg::Level& L) {
  hipStream_t st = ctx->stream;
  const sem_mesh* m = L.mesh;
  const long n = (long)m->nl;
  const Layout g = m->layout();
  double* v = L.d;
  double* w = L.w;
  double* e = L.e;
  hipLaunchKernelGGL(k_hash_start, dim3(nblk(n)), dim3(256), 0, st, m->l2g, v, m->L, m->NC,
                     L.op->desc.bc_bits, n);
  SEM_HIP(hipGetLastError());
  // v <- D^-1 A v; the Rayleigh quotient is scale-free, so the iterate is
  // rescaled by its max norm only every kRenorm steps (|lambda| < 4 keeps the
  // growth in between far from overflow), which saves the reduction and the
  // division launches of the other steps
  constexpr int kRenorm = 8;
  for (int it = 0; it < mg->prm.npow; ++it) {
    SEM_TRY(apply(ctx, L.op, v, w));
    hipLaunchKernelGGL(k_mul2, dim3(nblk(n)), dim3(256), 0, st, L.dinv, w, e, n);
    if ((it + 1) % kRenorm == 0) {
      SEM_TRY(launch_norm_inf(st, n, e, mg->part, kParts, mg->part + kParts));
      hipLaunchKernelGGL(k_div_scalar, dim3(nblk(n)), dim3(256), 0, st, e, mg->part + kParts, v,
                         n);
    } else {
      std::swap(v, e);
    }
    SEM_HIP(hipGetLastError());
  }
  double num = 0.0, den = 0.0;
  SEM_TRY(apply(ctx, L.op, v, w));
  SEM_TRY(wdot_host(ctx, mg, g, v, w, &num));
  hipLaunchKernelGGL(k_times_diag, dim3(nblk(n)), dim3(256), 0, st, v, L.dinv, e, n);
  SEM_TRY(wdot_host(ctx, mg, g, v, e, &den));
  if (!(den > 0.0) || !(num > 0.0) || !std::isfinite(num / den)) {
    set_error("sem_mg_create: eigenvalue estimate failed on a level with p = %d, %d x %d "
              "elements (Rayleigh quotient %g / %g)", m->nr, m->Ex, m->Ey, num, den);
    return SEM_ERR_BREAKDOWN;
  }
  L.lmax = mg->prm.safety * num / den;
  return SEM_OK;
}

// the level's vectors (x and b only below the finest level) in one block,
// each starting on a 512-byte boundary
int level_buffers(sem_mg::Level& L, bool own_xb) {
  const long n = (long)L.mesh->nl;
  const long sz = (n + 63) / 64 * 64;
  SEM_TRY(dalloc(&L.blk64, sz * (own_xb ? 7 : 5)));
  double* p = L.blk64;
  L.dinv = p;
  L.r = p + sz;
  L.d = p + 2 * sz;
  L.w = p + 3 * sz;
  L.e = p + 4 * sz;
  if (own_xb) {
    L.x = p + 5 * sz;
    L.b = p + 6 * sz;
  }
  return SEM_OK;
}

// the first level of the block cycle: the smallest l >= 1 from which every
// level is bilinear with at most kBlockNodes local nodes and every transfer
// a 2 x 2 element merge (-1: none)
int block_start(const sem_mg* mg) {
  const int nl = (int)mg->lv.size();
  if (!mg->block || nl > kBlockMaxLevels + 1) return -1;
  if (mg->prm.degree > kBlockMaxDeg || mg->prm.degree_coarse > kBlockMaxDeg) return -1;
  int lc = -1;
  for (int l = nl - 1; l >= 1; --l) {
    const sem_mg::Level& L = mg->lv[l];
    const sem_mesh* m = L.mesh;
    bool ok = m->nr == 2 && m->ns == 2 && m->nl <= kBlockNodes;
    if (l < nl - 1) ok = ok && L.h == 1 && L.xf.pf == 2 && L.xf.pc == 2;
    if (!ok) break;
    lc = l;
  }
  if (lc >= 0 && nl - lc > kBlockMaxLevels) lc = -1;
  return lc;
}

template <typename T>
int block_desc(sem_mg* mg, void** out) {
  const int nl = (int)mg->lv.size();
  const int lc = mg->lc;
  BlockCycle<T>* h = new (std::nothrow) BlockCycle<T>();
  if (!h) {
    set_error("multigrid: out of host memory");
    return SEM_ERR_HIP;
  }
  h->nlev = nl - lc;
  for (int q = 0; q < h->nlev; ++q) {
    sem_mg::Level& L = mg->lv[lc + q];
    const LvBuf<T> B = bufs<T>(L);
    const sem_operator* op = B.op;
    const sem_mesh* m = L.mesh;
    BlockLevel<T>& V = h->lv[q];
    V.Ex = m->Ex;
    V.Ey = m->Ey;
    V.L = m->L;
    V.NC = m->NC;
    V.per_x = m->per_x;
    V.per_y = m->per_y;
    V.bc = op->desc.bc_bits;
    V.has_mass = op->has_mass ? 1 : 0;
    V.n = (int)m->nl;
    V.ng = (long)m->map->n_global;
    V.perm = m->map->perm;
    V.seg = m->map->seg;
    V.G11 = static_cast<const T*>(op->G11);
    V.G12 = static_cast<const T*>(op->G12);
    V.G22 = static_cast<const T*>(op->G22);
    V.B = static_cast<const T*>(op->B);
    V.nu = T(op->desc.nu);
    V.k = T(op->desc.k);
    V.x = B.x;
    V.b = B.b;
    V.r = B.r;
    V.d = B.d;
    V.w = B.w;
    V.e = B.e;
    V.dinv = B.dinv;
    V.J = B.J;
    V.xf = L.xf;
    const bool last = lc + q == nl - 1;
    V.deg = last ? mg->prm.degree_coarse : mg->prm.degree;
    // the coefficients cheby() forms on the host, step by step
    const double lmax = L.lmax, lmin = L.lmax / mg->prm.ratio;
    const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin);
    const double sigma = theta / delta;
    double rho = 1.0 / sigma;
    V.theta = theta;
    for (int i = 1; i < V.deg; ++i) {
      const double rho_n = 1.0 / (2.0 * sigma - rho);
      V.c1[i] = rho_n * rho;
      V.c2[i] = 2.0 * rho_n / delta;
      rho = rho_n;
    }
  }
  int s = dalloc(reinterpret_cast<unsigned char**>(out), (long)sizeof(BlockCycle<T>));
  if (s == SEM_OK && hipMemcpy(*out, h, sizeof(BlockCycle<T>), hipMemcpyHostToDevice) != hipSuccess) {
    set_error("multigrid: copying the block-cycle levels failed");
    s = SEM_ERR_HIP;
  }
  delete h;
  return s;
}

int build(sem_ctx* ctx, sem_mg* mg) {
  hipStream_t st = ctx->stream;
  const sem_operator* op0 = mg->op;
  {
    sem_mg::Level L0;
    L0.mesh = const_cast<sem_mesh*>(op0->mesh);
    L0.op = op0;
    mg->lv.push_back(L0);
  }
  while ((int)mg->lv.size() < mg->prm.max_levels) {
    sem_mg::Level& F = mg->lv.back();
    const sem_mesh* mf = F.mesh;
    const int N = mf->nr - 1;
    int pc, h;
    if (N > 1) {
      pc = N / 2 + 1;
      h = 0;
    } else if (mf->Ex % 2 == 0 && mf->Ey % 2 == 0 && mf->Ex >= 4 && mf->Ey >= 4) {
      pc = 2;
      h = 1;
    } else {
      break;
    }
    const int pf = mf->nr;
    std::vector<double> zc(pc), wc(pc);
    gll(pc, zc.data(), wc.data());
    sem_mesh* mc = nullptr;
    SEM_TRY(sem_mesh_create(ctx, pc, pc, mf->Ex >> h, mf->Ey >> h, mf->per_x, mf->per_y,
                            SEM_DEFORM_NONE, 0.0, &mc));
    sem_mg::Level C;
    C.mesh = mc;
    C.own_mesh = true;
    mg->lv.push_back(C);
    sem_mg::Level& Fl = mg->lv[mg->lv.size() - 2];
    Fl.h = h;
    Xfer xf{};
    xf.pf = pf;
    xf.pc = pc;
    xf.hx = h;
    xf.hy = h;
    xf.Lf = mf->L;
    xf.NCf = mf->NC;
    xf.Lc = mc->L;
    xf.NCc = mc->NC;
    xf.per_x = mf->per_x;
    xf.per_y = mf->per_y;
    Fl.xf = xf;
    // transfer rows: fine GLL nodes of each sub-element in the coarse
    // reference coordinate, for x then y (square elements: identical)
    const std::vector<double>& zf = mf->hzr;
    std::vector<double> Jall;
    for (int dir = 0; dir < 2; ++dir)
      for (int s = 0; s < (h ? 2 : 1); ++s) {
        std::vector<double> xo(pf);
        for (int i = 0; i < pf; ++i) xo[i] = h ? (zf[i] + 2.0 * s - 1.0) / 2.0 : zf[i];
        std::vector<double> J = interp_rows(pf, xo.data(), pc, zc.data());
        Jall.insert(Jall.end(), J.begin(), J.end());
      }
    Fl.nJ = (long)Jall.size();
    SEM_TRY(dalloc(&Fl.J, (long)Jall.size()));
    SEM_HIP(hipMemcpy(Fl.J, Jall.data(), sizeof(double) * Jall.size(), hipMemcpyHostToDevice));
    // coarse coordinates: the fine geometry at the coarse nodes
    int subbits = 0;
    std::vector<double> xo(pc);
    for (int a = 0; a < pc; ++a) {
      const int sub = (h && zc[a] > 0.0) ? 1 : 0;
      subbits |= sub << a;
      xo[a] = h ? 2.0 * zc[a] + 1.0 - 2.0 * sub : zc[a];
    }
    std::vector<double> Jc = interp_rows(pc, xo.data(), pf, zf.data());
    double* dJc = nullptr;
    SEM_TRY(dalloc(&dJc, (long)Jc.size()));
    SEM_HIP(hipMemcpy(dJc, Jc.data(), sizeof(double) * Jc.size(), hipMemcpyHostToDevice));
    const long nc = (long)mc->nl;
    hipLaunchKernelGGL(k_coarse_coords, dim3(nblk(nc)), dim3(256), 0, st, mf->f[SEM_X], dJc,
                       subbits, mc->f[SEM_X], xf, nc);
    hipLaunchKernelGGL(k_coarse_coords, dim3(nblk(nc)), dim3(256), 0, st, mf->f[SEM_Y], dJc,
                       subbits, mc->f[SEM_Y], xf, nc);
    const hipError_t e = hipGetLastError();
    (void)hipStreamSynchronize(st);
    (void)hipFree(dJc);
    SEM_HIP(e);
    SEM_TRY(sem_mesh_set_coords(ctx, mc, mc->f[SEM_X], mc->f[SEM_Y]));
  }
  // operators, diagonals, buffers, eigenvalue estimates
  for (size_t l = 0; l < mg->lv.size(); ++l) {
    sem_mg::Level& L = mg->lv[l];
    if (l > 0) {
      sem_op_desc d{};
      d.nu = op0->desc.nu;
      d.k = op0->desc.k;
      d.bc_bits = op0->desc.bc_bits;
      d.assemble = 1;
      sem_operator* oc = nullptr;
      SEM_TRY(sem_operator_create(ctx, L.mesh, &d, SEM_F64, &oc));
      L.op = oc;
      L.own_op = true;
    }
    SEM_TRY(level_buffers(L, l > 0));
    SEM_TRY(sem_operator_jacobi(ctx, L.op, L.dinv));
    SEM_TRY(estimate_lmax(ctx, mg, L));   // uses the level's d, w, e only
    L.fuse64 = mg->fuse && operator_epi_ok(L.op);
  }
  // the FP32 cycle: every level's operator rebuilt in FP32 and FP32 copies
  // of the diagonals and transfer rows (the eigenvalue bounds are the FP64
  // ones)
  if (mg->f32) {
    for (size_t l = 0; l < mg->lv.size(); ++l) {
      sem_mg::Level& L = mg->lv[l];
      const long n = (long)L.mesh->nl;
      sem_op_desc d{};
      d.nu = op0->desc.nu;
      d.k = op0->desc.k;
      d.bc_bits = op0->desc.bc_bits;
      d.assemble = 1;
      sem_operator* o32 = nullptr;
      SEM_TRY(sem_operator_create(ctx, L.mesh, &d, SEM_F32, &o32));
      L.op32 = o32;
      L.fuse32 = mg->fuse && operator_epi_ok(o32);
      const long sz = (n + 127) / 128 * 128;
      float** fb[] = {&L.fx, &L.fb, &L.fr, &L.fd, &L.fw, &L.fe, &L.fdinv};
      SEM_TRY(dalloc(&L.blk32, sz * 7));
      for (int q = 0; q < 7; ++q) *fb[q] = L.blk32 + q * sz;
      SEM_TRY(launch_convert_f64_f32(st, n, L.dinv, L.fdinv));
      if (L.nJ > 0) {
        SEM_TRY(dalloc(&L.fJ, L.nJ));
        SEM_TRY(launch_convert_f64_f32(st, L.nJ, L.J, L.fJ));
      }
    }
  }
  const long n0 = (long)op0->mesh->nl;
  SEM_TRY(dalloc(&mg->pr, n0));
  SEM_TRY(dalloc(&mg->pp, n0));
  SEM_TRY(dalloc(&mg->pw, n0));
  SEM_TRY(dalloc(&mg->pz, n0));
  if (mg->f32) SEM_TRY(dalloc(&mg->pq, n0));
  mg->lc = block_start(mg);
  if (mg->lc >= 1) {
    SEM_TRY(block_desc<double>(mg, &mg->bdesc64));
    if (mg->f32) SEM_TRY(block_desc<float>(mg, &mg->bdesc32));
  }
  SEM_HIP(hipStreamSynchronize(st));
  return SEM_OK;
}

int check_op(const sem_operator* op) {
  SEM_ARG(op->precision == SEM_F64, "multigrid: the operator must be SEM_F64");
  if (op->generic) {
    set_error("multigrid: needs nr == ns <= %d", kMaxP);
    return SEM_ERR_UNSUPPORTED;
  }
  SEM_ARG(op->desc.assemble, "multigrid: the operator must assemble (desc.assemble = 1)");
  if (op->desc.nu_f || op->desc.k_f || op->desc.G11 || op->desc.G12 || op->desc.G22 ||
      op->desc.B) {
    set_error("multigrid: field nu/k and factor overrides are not supported (scalar nu, k only)");
    return SEM_ERR_UNSUPPORTED;
  }
  return SEM_OK;
}

}  // namespace

extern "C" {

int sem_mg_default_params(sem_mg_params* p) {
  SEM_ARG(p, "NULL argument");
  p->degree = 3;
  p->ratio = 20.0;
  p->degree_coarse = 8;
  p->npow = 20;
  p->safety = 1.1;
  p->max_levels = 32;
  p->precision = SEM_F64;
  return SEM_OK;
}

int sem_mg_create(sem_ctx* ctx, const sem_operator* op, const sem_mg_params* params,
                  sem_mg** out) {
  SEM_ARG(ctx && op && out, "NULL argument");
  SEM_TRY(check_op(op));
  sem_mg_params p;
  sem_mg_default_params(&p);
  if (params) p = *params;
  SEM_ARG(p.degree >= 1 && p.degree_coarse >= 1 && p.ratio > 1.0 && p.npow >= 1 &&
              p.safety >= 1.0 && p.max_levels >= 1,
          "sem_mg_create: bad parameters");
  SEM_ARG(p.precision == SEM_F64 || p.precision == SEM_F32,
          "sem_mg_create: precision must be SEM_F64 or SEM_F32");
  sem_mg* mg = new (std::nothrow) sem_mg();
  SEM_ARG(mg, "out of host memory");
  mg->prm = p;
  mg->device = ctx->device;
  mg->op = op;
  mg->f32 = p.precision == SEM_F32;
  const char* fz = std::getenv("SEM_MG_FUSE");
  mg->fuse = !(fz && std::atoi(fz) == 0);
  const char* fb = std::getenv("SEM_MG_BLOCK");
  mg->block = !(fb && std::atoi(fb) == 0);
  if (dalloc(&mg->part, kParts + kScalars) != SEM_OK) {
    mg_free(mg);
    return SEM_ERR_HIP;
  }
  const int s = build(ctx, mg);
  if (s != SEM_OK) {
    mg_free(mg);
    return s;
  }
  *out = mg;
  return SEM_OK;
}

int sem_mg_destroy(sem_mg* mg) {
  mg_free(mg);
  return SEM_OK;
}

int sem_mg_num_levels(const sem_mg* mg, int* nlevels) {
  SEM_ARG(mg && nlevels, "NULL argument");
  *nlevels = (int)mg->lv.size();
  return SEM_OK;
}

int sem_mg_level_info(const sem_mg* mg, int level, int* p, int* Ex, int* Ey, double* lmax) {
  SEM_ARG(mg, "NULL argument");
  SEM_ARG(level >= 0 && level < (int)mg->lv.size(), "sem_mg_level_info: no level %d", level);
  const sem_mg::Level& L = mg->lv[level];
  if (p) *p = L.mesh->nr;
  if (Ex) *Ex = L.mesh->Ex;
  if (Ey) *Ey = L.mesh->Ey;
  if (lmax) *lmax = L.lmax;
  return SEM_OK;
}

int sem_mg_set_lmax(sem_mg* mg, int level, double lmax) {
  SEM_ARG(mg, "NULL argument");
  SEM_ARG(level >= 0 && level < (int)mg->lv.size(), "sem_mg_set_lmax: no level %d", level);
  SEM_ARG(lmax > 0.0 && std::isfinite(lmax), "sem_mg_set_lmax: lmax must be positive");
  mg->lv[level].lmax = lmax;
  return SEM_OK;
}

int sem_mg_vcycle(sem_ctx* ctx, sem_mg* mg, const double* r, double* z) {
  SEM_ARG(ctx && mg && r && z, "NULL argument");
  SEM_ARG(r != z, "sem_mg_vcycle: r and z must not alias");
  sem_mg::Level& L0 = mg->lv[0];
  if (mg->f32) {
    const long n = (long)L0.mesh->nl;
    SEM_TRY(launch_convert_f64_f32(ctx->stream, n, r, L0.fb));
    SEM_TRY(vcycle<float>(ctx, mg, 0));
    return launch_convert_f32_f64(ctx->stream, n, L0.fx, z);
  }
  L0.b = const_cast<double*>(r);
  L0.x = z;
  const int s = vcycle<double>(ctx, mg, 0);
  L0.b = nullptr;
  L0.x = nullptr;
  return s;
}

int sem_pcg_mg(sem_ctx* ctx, const sem_operator* op, sem_mg* mg, const double* b, double* x,
               double tol, int maxiter, sem_pcg_info* info) {
  SEM_ARG(ctx && op && mg && b && x, "NULL argument");
  SEM_ARG(op == mg->op, "sem_pcg_mg: mg was built for another operator");
  SEM_ARG(b != x, "sem_pcg_mg: b and x must not alias");
  SEM_ARG(maxiter >= 0 && tol >= 0.0, "sem_pcg_mg: bad limits");
  hipStream_t st = ctx->stream;
  const sem_mesh* m = op->mesh;
  const Layout g = m->layout();
  const long n = (long)m->nl;
  SEM_HIP(hipEventRecord(ctx->ev0, st));
  SEM_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, st));
  SEM_HIP(hipMemcpyAsync(mg->pr, b, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
  double res = 0.0;
  double* sc = mg->part + kParts;
  SEM_TRY(norm_host(ctx, mg, n, mg->pr, &res));
  int k = 0;
  bool conv = true;
  while (res > tol) {
    if (k == maxiter) {
      conv = false;
      break;
    }
    ++k;
    SEM_TRY(sem_mg_vcycle(ctx, mg, mg->pr, mg->pz));
    // the scalars stay on the device: the loop waits once per iteration, for
    // the residual norm of its convergence test
    double* tz = sc + 1 + (k & 1);
    double* tprev = sc + 1 + ((k - 1) & 1);
    SEM_TRY(launch_wdot(st, g, mg->pr, mg->pz, mg->part, kParts, tz));
    if (k == 1) {
      SEM_HIP(hipMemcpyAsync(mg->pp, mg->pz, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    } else {
      // FP32 cycle: the flexible (Polak-Ribiere) beta = <z_k, r_k - r_{k-1}>
      // / t_prev, which tolerates the cycle's rounding; FP64: pcg.jl:49
      if (mg->f32) SEM_TRY(launch_wdot(st, g, mg->pq, mg->pz, mg->part, kParts, sc + 4));
      hipLaunchKernelGGL(k_pdir, dim3(nblk(n)), dim3(256), 0, st, 1.0, mg->pz, mg->pp, tz, tprev,
                         mg->f32 ? sc + 4 : nullptr, n);
      SEM_HIP(hipGetLastError());
    }
    if (mg->f32)
      SEM_HIP(hipMemcpyAsync(mg->pq, mg->pr, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    SEM_TRY(apply(ctx, op, mg->pp, mg->pw));
    SEM_TRY(launch_wdot(st, g, mg->pp, mg->pw, mg->part, kParts, sc + 3));
    hipLaunchKernelGGL(k_xr, dim3(nblk(n)), dim3(256), 0, st, x, mg->pr, mg->pp, mg->pw, tz,
                       sc + 3, n);
    SEM_HIP(hipGetLastError());
    SEM_TRY(norm_host(ctx, mg, n, mg->pr, &res));
    if (!std::isfinite(res)) break;
  }
  SEM_HIP(hipEventRecord(ctx->ev1, st));
  SEM_HIP(hipEventSynchronize(ctx->ev1));
  float ms = 0.f;
  SEM_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  if (info) {
    info->iters = k;
    info->converged = conv && res <= tol;
    info->res = res;
    info->elapsed_ms = ms;
    info->outer = 0;
    info->inner_total = k;
  }
  if (!std::isfinite(res)) {
    set_error("sem_pcg_mg: residual is not finite after %d iterations", k);
    return SEM_ERR_NONFINITE;
  }
  return (conv && res <= tol) ? SEM_OK : SEM_WARN_MAXITER;
}

}  // extern "C"

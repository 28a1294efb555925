// Setup and utility kernels: coordinates and FP64 geometry (mesh.jl:98-123,
// jac.jl:24-40), mult, masks, the map-based gather-scatter, ABu, reductions,
// grad / grad' and the adjoint dJ/dnu kernel, dealiased advection.
#include "sem_common.h"
#include "sem_device.h"

namespace sem {
namespace {

constexpr double kPi = 3.14159265358979323846;

__global__ void k_coords(double* __restrict__ x, double* __restrict__ y,
                         const double* __restrict__ xe, const double* __restrict__ ye, int L,
                         long n, int deform, double amp) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = (int)(t % L);
  const int j = (int)(t / L);
  double X = xe[i], Y = ye[j];
  if (deform == SEM_DEFORM_SINSIN) {
    const double s = amp * sin(kPi * X) * sin(kPi * Y);
    x[t] = X + s;
    y[t] = Y + s;
  } else if (deform == SEM_DEFORM_AFFINE01) {
    x[t] = 0.5 * (X + 1.0);
    y[t] = 0.5 * (Y + 1.0);
  } else {
    x[t] = X;
    y[t] = Y;
  }
}

// One thread per local node: xr = Dr x (within the element column), xs = x Ds',
// then J, Ji, rx, ry, sx, sy, B = J (wr ws), Bi, G11, G12, G22 as mesh.jl:114-123.
__global__ void k_geometry(const double* __restrict__ x, const double* __restrict__ y,
                           const double* __restrict__ Dr, const double* __restrict__ Ds,
                           const double* __restrict__ wr, const double* __restrict__ ws, int nr,
                           int ns, int L, long n, double* Jac, double* Jaci, double* rx,
                           double* ry, double* sx, double* sy, double* B, double* Bi,
                           double* G11, double* G12, double* G22) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = (int)(t % L);
  const int j = (int)(t / L);
  const int a = i % nr, i0 = i - a;
  const int b = j % ns, j0 = j - b;
  double xr = 0, yr = 0, xs = 0, ys = 0;
  for (int k = 0; k < nr; ++k) {
    const double d = Dr[a + nr * k];
    xr += d * x[(i0 + k) + (long)L * j];
    yr += d * y[(i0 + k) + (long)L * j];
  }
  for (int k = 0; k < ns; ++k) {
    const double d = Ds[b + ns * k];
    xs += x[i + (long)L * (j0 + k)] * d;
    ys += y[i + (long)L * (j0 + k)] * d;
  }
  const double J = xr * ys - xs * yr;
  const double Ji = 1.0 / J;
  const double vrx = Ji * ys, vry = -Ji * xs, vsx = -Ji * yr, vsy = Ji * xr;
  const double w = wr[a] * ws[b];
  const double vB = J * w;
  Jac[t] = J;
  Jaci[t] = Ji;
  rx[t] = vrx;
  ry[t] = vry;
  sx[t] = vsx;
  sy[t] = vsy;
  B[t] = vB;
  Bi[t] = 1.0 / vB;
  G11[t] = vB * (vrx * vrx + vry * vry);
  G12[t] = vB * (vrx * vsx + vry * vsy);
  G22[t] = vB * (vsx * vsx + vsy * vsy);
}

__global__ void k_mult(double* __restrict__ mult, int nr, int ns, int L, int NC, int per_x,
                       int per_y, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = (int)(t % L);
  const int j = (int)(t / L);
  // n == 1 point per element cannot occur (nr, ns >= 2)
  mult[t] = mult1(i, nr, L, per_x) * mult1(j, ns, NC, per_y);
}


__global__ void k_mask_fill(double* __restrict__ M, int L, int NC, int bc, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = (int)(t % L);
  const int j = (int)(t / L);
  M[t] = mask_on(i, j, L, NC, bc) ? 1.0 : 0.0;
}

__global__ void k_mask_bc(const double* __restrict__ u, double* __restrict__ y, int L, int NC,
                          int bc, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = (int)(t % L);
  const int j = (int)(t / L);
  y[t] = (mask_on(i, j, L, NC, bc) ? 1.0 : 0.0) * u[t];
}

__global__ void k_mul(const double* __restrict__ a, const double* __restrict__ u,
                      double* __restrict__ y, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  y[t] = a ? a[t] * u[t] : u[t];
}

// Map-based QQ': one thread per global node, copies summed pairwise in map
// order ((v0+v1)+(v2+v3)), the sum written back to every copy. u and y may be
// the same array (every copy of a node is read before any is written, and
// nodes are disjoint). bc >= 0 also applies the Dirichlet mask of the
// structured layout (L rows, NC columns), a property of the global node.
template <typename T>
__global__ void k_gsmap(const int32_t* __restrict__ perm, const int32_t* __restrict__ seg,
                        const T* u, T* y, long ng, int bc, int L, int NC) {
  const long g = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ng) return;
  gsmap_node(perm, seg, u, y, g, bc, L, NC);
}

__global__ void k_f64_to_f32(const double* __restrict__ x, float* __restrict__ y, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) y[t] = (float)x[t];
}
__global__ void k_f64_to_f16(const double* __restrict__ x, double scale,
                             _Float16* __restrict__ y, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) y[t] = (_Float16)(float)(x[t] * scale);
}
__global__ void k_f32_to_f64(const float* __restrict__ x, double* __restrict__ y, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) y[t] = (double)x[t];
}

// ABu, ABu.jl:9-37. Stage 1 (Br on groups of nb rows): tmp (m1 x n);
// stage 2 (As on blocks of na columns): y (m1 x n2). trans: operands transposed.
__global__ void k_abu_r(const double* __restrict__ u, const double* __restrict__ Br, int m,
                        int n, int mb, int nb, int trans, double* __restrict__ out) {
  const int m1 = (m / nb) * mb;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)m1 * n) return;
  const int r = (int)(t % m1);
  const int c = (int)(t / m1);
  const int grp = r / mb, rr = r - grp * mb;
  double acc = 0;
  for (int k = 0; k < nb; ++k) {
    // Br is mb x nb column-major; transposed operand is stored nb x mb
    const double bv = trans ? Br[k + nb * rr] : Br[rr + mb * k];
    acc += bv * u[(grp * nb + k) + (long)m * c];
  }
  out[t] = acc;
}

__global__ void k_abu_s(const double* __restrict__ u, const double* __restrict__ As, int m,
                        int n, int ma, int na, int trans, double* __restrict__ out) {
  const int n2 = (n / na) * ma;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)m * n2) return;
  const int r = (int)(t % m);
  const int c = (int)(t / m);
  const int blk = c / ma, cc = c - blk * ma;
  double acc = 0;
  for (int l = 0; l < na; ++l) {
    const double av = trans ? As[l + na * cc] : As[cc + ma * l];
    acc += u[r + (long)m * (blk * na + l)] * av;
  }
  out[t] = acc;
}

// ---- reductions (deterministic two-stage: fixed per-block partials, then one
// block sums them in index order)
__device__ double block_reduce_sum(double v, double* sh) {
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (tid < s) sh[tid] = sh[tid] + sh[tid + s];
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}
__device__ double block_reduce_max(double v, double* sh) {
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (tid < s) {
      const double o = sh[tid + s];
      // NaN propagates (the reference's norm(NaN) is NaN)
      sh[tid] = (o > sh[tid] || o != o) ? o : sh[tid];
    }
    __syncthreads();
  }
  const double r = sh[0];
  __syncthreads();
  return r;
}

// Block partials of sum(a .* b .* mult): block q takes columns
// [q cpb, (q + 1) cpb) whole, one contiguous run of nodes; its 1,024 threads
// step through the run four nodes each per pass with their loads issued
// together (one block per CU: sixteen waves with eight loads each in
// flight). Node f of the run lies in column j0 + f / L, row f % L (32-bit
// arithmetic: a run holds fewer than 2^31 nodes).
constexpr int kWdotThreads = 1024;
__global__ void __launch_bounds__(kWdotThreads)
k_wdot_part(const double* __restrict__ a, const double* __restrict__ b, int P, int L, int NC,
            int per_x, int per_y, double* part) {
  __shared__ double sh[kWdotThreads];
  const int cpb = (NC + (int)gridDim.x - 1) / (int)gridDim.x;
  const int j0 = blockIdx.x * cpb, j1 = min(NC, j0 + cpb);
  const int cnt = j1 > j0 ? (j1 - j0) * L : 0;
  const double* __restrict__ ab = a + (long)L * j0;
  const double* __restrict__ bb = b + (long)L * j0;
  const int bs = blockDim.x;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int f0 = threadIdx.x; f0 < cnt; f0 += 4 * bs) {
    double av[4], bv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int f = min(f0 + u * bs, cnt - 1);
      av[u] = ab[f];
      bv[u] = bb[f];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int f = f0 + u * bs;
      if (f < cnt) {
        const int jj = f / L;
        const int i = f - jj * L;
        acc[u] += av[u] * bv[u] * (mult1(i, P, L, per_x) * mult1(j0 + jj, P, NC, per_y));
      }
    }
  }
  const double s = block_reduce_sum((acc[0] + acc[1]) + (acc[2] + acc[3]), sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// max that keeps a NaN once seen (from either side)
__device__ __forceinline__ double nan_max(double m, double v) {
  return (v > m || v != v) ? v : m;
}

// four independent loads in flight per step (the max is exact and
// order-free, so the result does not depend on the grouping)
__global__ void __launch_bounds__(256)
k_absmax_part(const double* __restrict__ a, long n, double* part) {
  __shared__ double sh[256];
  const long stride = (long)gridDim.x * blockDim.x;
  double m0 = 0, m1 = 0, m2 = 0, m3 = 0;
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  for (; t + 3 * stride < n; t += 4 * stride) {
    const double v0 = a[t], v1 = a[t + stride], v2 = a[t + 2 * stride], v3 = a[t + 3 * stride];
    m0 = nan_max(m0, fabs(v0));
    m1 = nan_max(m1, fabs(v1));
    m2 = nan_max(m2, fabs(v2));
    m3 = nan_max(m3, fabs(v3));
  }
  for (; t < n; t += stride) m0 = nan_max(m0, fabs(a[t]));
  const double m = nan_max(nan_max(m0, m1), nan_max(m2, m3));
  const double r = block_reduce_max(m, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

// op 0: out = sum, 1: out = max (NaN-propagating), 2: out += scale * sum
__global__ void __launch_bounds__(256) k_final(const double* part, int n, int op, double* out,
                                               double scale = 1.0) {
  __shared__ double sh[256];
  double acc = 0;
  for (int t = threadIdx.x; t < n; t += blockDim.x) {
    const double v = part[t];
    if (op != 1) acc += v;
    else acc = (v > acc || v != v) ? v : acc;
  }
  const double r = op != 1 ? block_reduce_sum(acc, sh) : block_reduce_max(acc, sh);
  if (threadIdx.x == 0) out[0] = op == 2 ? out[0] + scale * r : r;
}

// ---- right-hand-side composition and weighted reductions
// y = mask(sum_j c_j (B if weighted_j) x_j): the pointwise half of makeRHS!
// (convectionDiffusion.jl:93-110, diffusion.jl:51-65); y may alias any x_j
// (each thread reads its node of every term before writing it).
__global__ void k_combine(CombineTerms terms, const double* __restrict__ B, double* y, int L,
                          int NC, int bc, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double bt = B[t];
  double acc = 0.0;
  for (int j = 0; j < terms.n; ++j) {
    const double v = terms.x[j][t];
    acc += terms.c[j] * (terms.w[j] ? bt * v : v);
  }
  if (bc >= 0) {
    const int i = (int)(t % L);
    const int jj = (int)(t / L);
    acc *= mask_on(i, jj, L, NC, bc) ? 1.0 : 0.0;
  }
  y[t] = acc;
}

// four strided nodes per step, their loads issued together, each summed
// into its own accumulator; HB, HW: b, w given (a branch-free loop body)
template <bool HB, bool HW>
__global__ void __launch_bounds__(256)
k_dot_part(const double* __restrict__ a, const double* __restrict__ b,
           const double* __restrict__ w, long n, double* part) {
  __shared__ double sh[256];
  auto term = [&](long i) {
    if constexpr (HB && HW) return a[i] * b[i] * w[i];
    else if constexpr (HB) return a[i] * b[i];
    else if constexpr (HW) return a[i] * w[i];
    else return a[i];
  };
  const long stride = (long)gridDim.x * blockDim.x;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  for (; t + 3 * stride < n; t += 4 * stride) {
    double v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = term(t + u * stride);
#pragma unroll
    for (int u = 0; u < 4; ++u) acc[u] += v[u];
  }
  for (; t < n; t += stride) acc[0] += term(t);
  const double s = block_reduce_sum((acc[0] + acc[1]) + (acc[2] + acc[3]), sh);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// wr = G11 a + G12 b, ws = G12 a + G22 b (lapl.jl:75-76, 95-96)
__global__ void k_gprod(const double* __restrict__ G11, const double* __restrict__ G12,
                        const double* __restrict__ G22, const double* __restrict__ a,
                        const double* __restrict__ b, double* __restrict__ wr,
                        double* __restrict__ ws, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double av = a[t], bv = b[t];
  wr[t] = G11[t] * av + G12[t] * bv;
  ws[t] = G12[t] * av + G22[t] * bv;
}

// elementwise y = a .* b (op 0), a ./ b (op 1), alpha ./ a (op 2); y may
// alias a or b
__global__ void k_pointwise(int op, double alpha, const double* a, const double* b, double* y,
                            long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double av = a[t];
  double v;
  switch (op) {
    case 0: v = av * b[t]; break;
    case 1: v = av / b[t]; break;
    case 2: v = alpha / av; break;
    case 3: v = av / alpha; break;
    default: v = alpha * av; break;
  }
  y[t] = v;
}

// out = alpha x + beta y (out may alias x or y)
__global__ void k_axpby(double alpha, const double* x, double beta, const double* y, double* out,
                        long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = alpha * x[t] + beta * y[t];
}

// nu folded into the geometric factors: G_ij <- nu .* G_ij (semPS.jl:161-163)
__global__ void k_vardiff(const double* __restrict__ nu, const double* __restrict__ G11,
                          const double* __restrict__ G12, const double* __restrict__ G22,
                          double* __restrict__ o11, double* __restrict__ o12,
                          double* __restrict__ o22, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double v = nu[t];
  o11[t] = v * G11[t];
  o12[t] = v * G12[t];
  o22[t] = v * G22[t];
}

// ---- gradient (grad.jl:15-34), its transpose, and the dJ/dnu kernel
__global__ void k_grad(const double* __restrict__ u, const double* __restrict__ Dr,
                       const double* __restrict__ Ds, const double* __restrict__ rx,
                       const double* __restrict__ ry, const double* __restrict__ sx,
                       const double* __restrict__ sy, int nr, int ns, int L, long n,
                       double* __restrict__ ux, double* __restrict__ uy) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = (int)(t % L);
  const int j = (int)(t / L);
  const int a = i % nr, i0 = i - a;
  const int b = j % ns, j0 = j - b;
  double ur = 0, us = 0;
  for (int k = 0; k < nr; ++k) ur += Dr[a + nr * k] * u[(i0 + k) + (long)L * j];
  for (int k = 0; k < ns; ++k) us += u[i + (long)L * (j0 + k)] * Ds[b + ns * k];
  ux[t] = rx[t] * ur + sx[t] * us;
  uy[t] = ry[t] * ur + sy[t] * us;
}

__global__ void k_grad_t(const double* __restrict__ vx, const double* __restrict__ vy,
                         const double* __restrict__ Dr, const double* __restrict__ Ds,
                         const double* __restrict__ rx, const double* __restrict__ ry,
                         const double* __restrict__ sx, const double* __restrict__ sy, int nr,
                         int ns, int L, long n, double* __restrict__ y) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = (int)(t % L);
  const int j = (int)(t / L);
  const int a = i % nr, i0 = i - a;
  const int b = j % ns, j0 = j - b;
  double acc = 0;
  // Dr'(rx vx + ry vy): sum over rows k of the element column
  for (int k = 0; k < nr; ++k) {
    const long q = (i0 + k) + (long)L * j;
    acc += Dr[k + nr * a] * (rx[q] * vx[q] + ry[q] * vy[q]);
  }
  double acc2 = 0;
  for (int k = 0; k < ns; ++k) {
    const long q = i + (long)L * (j0 + k);
    acc2 += (sx[q] * vx[q] + sy[q] * vy[q]) * Ds[k + ns * b];
  }
  y[t] = acc + acc2;
}

// g = B .* (grad(lambda) . grad(u)): with G_ij <- nu G_ij the element form is
// lambda' A(nu) u = sum_nodes nu B (grad lambda . grad u), so d/dnu_local is
// the integrand B (grad lambda . grad u).
__global__ void k_adjoint_dnu(const double* __restrict__ u, const double* __restrict__ lam,
                              const double* __restrict__ Dr, const double* __restrict__ Ds,
                              const double* __restrict__ rx, const double* __restrict__ ry,
                              const double* __restrict__ sx, const double* __restrict__ sy,
                              const double* __restrict__ B, int nr, int ns, int L, long n,
                              double* __restrict__ g) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = (int)(t % L);
  const int j = (int)(t / L);
  const int a = i % nr, i0 = i - a;
  const int b = j % ns, j0 = j - b;
  double ur = 0, us = 0, lr = 0, ls = 0;
  for (int k = 0; k < nr; ++k) {
    const long q = (i0 + k) + (long)L * j;
    ur += Dr[a + nr * k] * u[q];
    lr += Dr[a + nr * k] * lam[q];
  }
  for (int k = 0; k < ns; ++k) {
    const long q = i + (long)L * (j0 + k);
    us += u[q] * Ds[b + ns * k];
    ls += lam[q] * Ds[b + ns * k];
  }
  const double ux = rx[t] * ur + sx[t] * us, uy = ry[t] * ur + sy[t] * us;
  const double lx = rx[t] * lr + sx[t] * ls, ly = ry[t] * lr + sy[t] * ls;
  g[t] = B[t] * (lx * ux + ly * uy);
}

// Dealiased advection, advect.jl:45-64, one workgroup per element, with
// J = Jr (x) Js the element interpolation to the dealiasing grid:
//   forward:   y = J' [ B_D ((J vx)(J Tx) + (J vy)(J Ty)) ]
//   transpose: the same linear map in T transposed factor by factor,
//              y = grad'( J'[(J vx) B_D (J w)], J'[(J vy) B_D (J w)] )
// with grad'(qx, qy) = Dr'(rx qx + ry qy) + (sx qx + sy qy) Ds.
// The element's fields and the four contraction matrices are staged in LDS
// once; each phase maps threads to the output points of one contraction and
// computes every field of that contraction together, so barriers separate
// phases (5 forward, 6 transpose), not fields. The per-node factors (rx, ry,
// sx, sy; B_D) are loaded into registers before the first barrier.
constexpr int kMaxPD = 24;
constexpr int kAdvThreads = 256;
static_assert(kMaxP * kMaxP <= kAdvThreads, "one velocity-grid node per thread");

__host__ __device__ constexpr int advect_lds_doubles(int p, int pd) {
  return 8 * pd * p + 7 * p * p + 2 * pd * pd;
}

__global__ void __launch_bounds__(kAdvThreads)
k_advect(const double* __restrict__ T, const double* __restrict__ vx,
         const double* __restrict__ vy, const double* __restrict__ Dr,
         const double* __restrict__ Ds, const double* __restrict__ rx,
         const double* __restrict__ ry, const double* __restrict__ sx,
         const double* __restrict__ sy, const double* __restrict__ BD,
         const double* __restrict__ Jr, const double* __restrict__ Js, int p, int pd, int Ex,
         int Ey, double* __restrict__ y, int transpose) {
  extern __shared__ double lds[];
  const int pp = p * p, dp = pd * p, dd = pd * pd;
  double* sJr = lds;        // pd x p, column-major
  double* sJs = sJr + dp;
  double* sDr = sJs + dp;   // p x p
  double* sDs = sDr + pp;
  double* sT = sDs + pp;    // the element's T (forward) or w (transpose)
  double* sF = sT + pp;     // four velocity-grid fields
  double* sA = sF + 4 * pp; // four pd x p row interpolants
  double* sC = sA + 4 * dp; // two dealiased fields
  double* sE = sC + 2 * dd; // two p x pd back-interpolants
  const int tid = threadIdx.x;
  const int ex = blockIdx.x % Ex, ey = blockIdx.x / Ex;
  const long L = (long)p * Ex, LD = (long)pd * Ex;
  const long g0 = (long)ex * p + L * ((long)ey * p);
  const long d0 = (long)ex * pd + LD * ((long)ey * pd);
  // this thread's velocity-grid node (tid < pp) and dealiased nodes
  const int a0 = tid % p, b0 = tid / p;
  const long gn = g0 + a0 + L * b0;
  const bool own = tid < pp;
  double frx = 0, fry = 0, fsx = 0, fsy = 0;
  double bd[3] = {0, 0, 0};
  if (own) {
    frx = rx[gn];
    fry = ry[gn];
    fsx = sx[gn];
    fsy = sy[gn];
    sT[tid] = T[gn];
    sF[pp + tid] = vx[gn];
    sF[2 * pp + tid] = vy[gn];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const int q = tid + i * kAdvThreads;
    if (q < dd) bd[i] = BD[d0 + q % pd + LD * (q / pd)];
  }
  for (int q = tid; q < dp; q += kAdvThreads) {
    sJr[q] = Jr[q];
    sJs[q] = Js[q];
  }
  for (int q = tid; q < pp; q += kAdvThreads) {
    sDr[q] = Dr[q];
    sDs[q] = Ds[q];
  }
  __syncthreads();
  if (!transpose) {
    // F = (Tx, Ty, vx, vy): the gradient of T on the velocity grid
    if (own) {
      double ur = 0, us = 0;
      for (int k = 0; k < p; ++k) ur += sDr[a0 + p * k] * sT[k + p * b0];
      for (int k = 0; k < p; ++k) us += sT[a0 + p * k] * sDs[b0 + p * k];
      sF[tid] = frx * ur + fsx * us;
      sF[3 * pp + tid] = sF[2 * pp + tid];
      sF[2 * pp + tid] = sF[pp + tid];
      sF[pp + tid] = fry * ur + fsy * us;
    }
    __syncthreads();
    // rows: A_f = Jr F_f (pd x p)
    for (int q = tid; q < dp; q += kAdvThreads) {
      const int ad = q % pd, b = q / pd;
      double c0 = 0, c1 = 0, c2 = 0, c3 = 0;
      for (int k = 0; k < p; ++k) {
        const double j = sJr[ad + pd * k];
        const int f = k + p * b;
        c0 += j * sF[f];
        c1 += j * sF[pp + f];
        c2 += j * sF[2 * pp + f];
        c3 += j * sF[3 * pp + f];
      }
      sA[q] = c0;
      sA[dp + q] = c1;
      sA[2 * dp + q] = c2;
      sA[3 * dp + q] = c3;
    }
    __syncthreads();
    // columns: X_f = A_f Js', then c = B_D (X_vx X_Tx + X_vy X_Ty)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const int q = tid + i * kAdvThreads;
      if (q < dd) {
        const int ad = q % pd, bdd = q / pd;
        double x0 = 0, x1 = 0, x2 = 0, x3 = 0;
        for (int k = 0; k < p; ++k) {
          const double j = sJs[bdd + pd * k];
          const int f = ad + pd * k;
          x0 += sA[f] * j;
          x1 += sA[dp + f] * j;
          x2 += sA[2 * dp + f] * j;
          x3 += sA[3 * dp + f] * j;
        }
        sC[q] = (x2 * x0 + x3 * x1) * bd[i];
      }
    }
    __syncthreads();
    // back, rows: E = Jr' C (p x pd)
    for (int q = tid; q < dp; q += kAdvThreads) {
      const int a = q % p, bdd = q / p;
      double e = 0;
      for (int k = 0; k < pd; ++k) e += sJr[k + pd * a] * sC[k + pd * bdd];
      sE[q] = e;
    }
    __syncthreads();
    // back, columns: y = E Js
    if (own) {
      double acc = 0;
      for (int k = 0; k < pd; ++k) acc += sE[a0 + p * k] * sJs[k + pd * b0];
      y[gn] = acc;
    }
  } else {
    // F = (w, vx, vy) with w in sT
    if (own) sF[tid] = sT[tid];
    __syncthreads();
    for (int q = tid; q < dp; q += kAdvThreads) {
      const int ad = q % pd, b = q / pd;
      double c0 = 0, c1 = 0, c2 = 0;
      for (int k = 0; k < p; ++k) {
        const double j = sJr[ad + pd * k];
        const int f = k + p * b;
        c0 += j * sF[f];
        c1 += j * sF[pp + f];
        c2 += j * sF[2 * pp + f];
      }
      sA[q] = c0;
      sA[dp + q] = c1;
      sA[2 * dp + q] = c2;
    }
    __syncthreads();
    // c = B_D (J w); C_0 = (J vx) c, C_1 = (J vy) c
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const int q = tid + i * kAdvThreads;
      if (q < dd) {
        const int ad = q % pd, bdd = q / pd;
        double x0 = 0, x1 = 0, x2 = 0;
        for (int k = 0; k < p; ++k) {
          const double j = sJs[bdd + pd * k];
          const int f = ad + pd * k;
          x0 += sA[f] * j;
          x1 += sA[dp + f] * j;
          x2 += sA[2 * dp + f] * j;
        }
        const double c = x0 * bd[i];
        sC[q] = x1 * c;
        sC[dd + q] = x2 * c;
      }
    }
    __syncthreads();
    for (int q = tid; q < dp; q += kAdvThreads) {
      const int a = q % p, bdd = q / p;
      double e0 = 0, e1 = 0;
      for (int k = 0; k < pd; ++k) {
        const double j = sJr[k + pd * a];
        e0 += j * sC[k + pd * bdd];
        e1 += j * sC[dd + k + pd * bdd];
      }
      sE[q] = e0;
      sE[dp + q] = e1;
    }
    __syncthreads();
    // S_f = E_f Js on this thread's node, then the grad' operands
    // P1 = rx S0 + ry S1, P2 = sx S0 + sy S1 (reusing sF)
    if (own) {
      double s0 = 0, s1 = 0;
      for (int k = 0; k < pd; ++k) {
        const double j = sJs[k + pd * b0];
        s0 += sE[a0 + p * k] * j;
        s1 += sE[dp + a0 + p * k] * j;
      }
      sF[tid] = frx * s0 + fry * s1;
      sF[pp + tid] = fsx * s0 + fsy * s1;
    }
    __syncthreads();
    if (own) {
      double acc = 0, acc2 = 0;
      for (int k = 0; k < p; ++k) acc += sDr[k + p * a0] * sF[k + p * b0];
      for (int k = 0; k < p; ++k) acc2 += sF[pp + a0 + p * k] * sDs[k + p * b0];
      y[gn] = acc + acc2;
    }
  }
}

inline int nblk(long n, int bs = 256) { return (int)((n + bs - 1) / bs); }

}  // namespace

int launch_coords(hipStream_t st, const sem_mesh* m, const double* xe, const double* ye,
                  int deform, double amp) {
  hipLaunchKernelGGL(k_coords, dim3(nblk(m->nl)), dim3(256), 0, st, m->f[SEM_X], m->f[SEM_Y], xe,
                     ye, m->L, (long)m->nl, deform, amp);
  SEM_HIP(hipGetLastError());
  return SEM_OK;
}

int launch_geometry(hipStream_t st, sem_mesh* m) {
  double** f = m->f;
  hipLaunchKernelGGL(k_geometry, dim3(nblk(m->nl)), dim3(256), 0, st, f[SEM_X], f[SEM_Y], m->Dr,
                     m->Ds, m->wr, m->ws, m->nr, m->ns, m->L, (long)m->nl, f[SEM_JAC],
                     f[SEM_JACI], f[SEM_RX], f[SEM_RY], f[SEM_SX], f[SEM_SY], f[SEM_B],
                     f[SEM_BI], f[SEM_G11], f[SEM_G12], f[SEM_G22]);
  SEM_HIP(hipGetLastError());
  return SEM_OK;
}

// ---- structured numbering on the device (semq.jl:6-25) -------------------
// Per direction (small host tables): the global index of each local row
// (gof), and for each global index its copies (c0, and c1 or -1 for a node
// shared by two elements or the periodic wrap) and the exclusive prefix of the
// copy counts (pre). Global node g = gx + nxg gy then owns the local copies
// {ic + L jc} in ascending order, starting at seg[g] = prey[gy] L + cy[gy]
// prex[gx]: the map the host builds by a counting sort of l2g, without the
// host pass over every local node.
struct AxisMap {
  const int32_t* gof;
  const int32_t* c0;
  const int32_t* c1;
  const int32_t* pre;
};

__global__ void k_l2g(int32_t* __restrict__ l2g, AxisMap ax, AxisMap ay, int L, int NC, long nxg) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= L) return;
  for (int j = blockIdx.y; j < NC; j += gridDim.y)
    l2g[i + (long)L * j] = (int32_t)(ax.gof[i] + nxg * ay.gof[j]);
}

__global__ void k_gsmap_build(int32_t* __restrict__ perm, int32_t* __restrict__ seg, AxisMap ax,
                              AxisMap ay, int L, int nxg, int nyg, long ng, long nl) {
  const int gx = blockIdx.x * blockDim.x + threadIdx.x;
  if (gx >= nxg) return;
  for (int gy = blockIdx.y; gy < nyg; gy += gridDim.y) {
  const long g = gx + (long)nxg * gy;
  const int cx = ax.c1[gx] >= 0 ? 2 : 1, cy = ay.c1[gy] >= 0 ? 2 : 1;
  const long s0 = (long)ay.pre[gy] * L + (long)cy * ax.pre[gx];
  seg[g] = (int32_t)s0;
  if (g == ng - 1) seg[ng] = (int32_t)nl;
  long k = s0;
  for (int q = 0; q < cy; ++q) {
    const long jc = q == 0 ? ay.c0[gy] : ay.c1[gy];
    for (int p = 0; p < cx; ++p) {
      const int ic = p == 0 ? ax.c0[gx] : ax.c1[gx];
      perm[k++] = (int32_t)(ic + L * jc);
    }
  }
  }
}

int launch_mult(hipStream_t st, const sem_mesh* m) {
  hipLaunchKernelGGL(k_mult, dim3(nblk(m->nl)), dim3(256), 0, st, m->f[SEM_MULT], m->nr, m->ns,
                     m->L, m->NC, m->per_x, m->per_y, (long)m->nl);
  SEM_HIP(hipGetLastError());
  return SEM_OK;
}

int launch_numbering(hipStream_t st, const sem_mesh* m, const int32_t* const tabs[8], int nxg,
                     int nyg, int32_t* l2g, int32_t* perm, int32_t* seg) {
  const AxisMap ax{tabs[0], tabs[1], tabs[2], tabs[3]};
  const AxisMap ay{tabs[4], tabs[5], tabs[6], tabs[7]};
  hipLaunchKernelGGL(k_l2g, dim3(nblk(m->L), std::min(m->NC, 32768)), dim3(256), 0, st, l2g, ax,
                     ay, m->L, m->NC, (long)nxg);
  hipLaunchKernelGGL(k_gsmap_build, dim3(nblk(nxg), std::min(nyg, 32768)), dim3(256), 0, st, perm,
                     seg, ax, ay, m->L, nxg, nyg, (long)nxg * nyg, (long)m->nl);
  SEM_HIP(hipGetLastError());
  return SEM_OK;
}

int launch_mask_fill(hipStream_t st, const Layout& g, int bits, double* M) {
  const long n = (long)g.L * g.NC;
  hipLaunchKernelGGL(k_mask_fill, dim3(nblk(n)), dim3(256), 0, st, M, g.L, g.NC, bits, n);
  SEM_HIP(hipGetLastError());
  return SEM_OK;
}

int launch_mask_bc(hipStream_t st, const Layout& g, int bits, const double* u, double* y) {
  const long n = (long)g.L * g.NC;
  hipLaunchKernelGGL(k_mask_bc, dim3(nblk(n)), dim3(256), 0, st, u, y, g.L, g.NC, bits, n);
  SEM_HIP(hipGetLastError());
  return SEM_OK;
}

int launch_mask(hipStream_t st, int64_t n, const double* M, const double* u, double* y) {
  hipLaunchKernelGGL(k_mul, dim3(nblk(n)), dim3(256), 0, st, M, u, y, (long)n);
  SEM_HIP(hipGetLastError());
  return SEM_OK;
}

int launch_mul(hipStream_t st, int64_t n, const double* a, const double* u, double* y) {
  return launch_mask(st, n, a, u, y);
}

int launch_gsmap(hipStream_t st, const sem_gsmap* map, const double* u, double* y) {
  hipLaunchKernelGGL((k_gsmap<double>), dim3(nblk(map->n_global)), dim3(256), 0, st, map->perm,
                     map->seg, u, y, (long)map->n_global, -1, 1, 1);
  SEM_HIP(hipGetLastError());
  return SEM_OK;
}

int launch_gsmap_masked(hipStream_t st, const sem_gsmap* map, const Layout& g, int bits,
                        const double* u, double* y) {
  hipLaunchKernelGGL((k_gsmap<double>), dim3(nblk(map->n_global)), dim3(256), 0, st, map->perm,
                     map->seg, u, y, (long)map->n_global, bits, g.L, g.NC);
  SEM_HIP(hipGetLastError());
  return SEM_OK;
}

int launch_gsmap_masked(hipStream_t st, const sem_gsmap* map, const Layout& g, int bits,
                        const float* u, float* y) {
  hipLaunchKernelGGL((k_gsmap<float>), dim3(nblk(map->n_global)), dim3(256), 0, st, map->perm,
                     map->seg, u, y, (long)map->n_global, bits, g.L, g.NC);
  SEM_HIP(hipGetLastError());
  return SEM_OK;
}

int launch_gsmap_f32(hipStream_t st, const sem_gsmap* map, const float* u, float* y) {
  hipLaunchKernelGGL((k_gsmap<float>), dim3(nblk(map->n_global)), dim3(256), 0, st, map->perm,
                     map->seg, u, y, (long)map->n_global, -1, 1, 1);
  SEM_HIP(hipGetLastError());
  return SEM_OK;
}

int launch_convert_f64_f32(hipStream_t st, int64_t n, const double* x, float* y) {
  hipLaunchKernelGGL(k_f64_to_f32, dim3(nblk(n)), dim3(256), 0, st, x, y, (long)n);
  SEM_HIP(hipGetLastError());
  return SEM_OK;
}
int launch_convert_f64_f16(hipStream_t st, int64_t n, const double* x, double scale, void* y) {
  hipLaunchKernelGGL(k_f64_to_f16, dim3(nblk(n)), dim3(256), 0, st, x, scale,
                     static_cast<_Float16*>(y), (long)n);
  SEM_HIP(hipGetLastError());
  return SEM_OK;
}
int launch_convert_f32_f64(hipStream_t st, int64_t n, const float* x, double* y) {
  hipLaunchKernelGGL(k_f32_to_f64, dim3(nblk(n)), dim3(256), 0, st, x, y, (long)n);
  SEM_HIP(hipGetLastError());
  return SEM_OK;
}

int launch_abu(hipStream_t st, int m, int n, const double* u, const double* As, int ma, int na,
               const double* Br, int mb, int nb, int trans, double* tmp, double* y) {
  // with trans the operands are read as their transposes: Br' is (nb x mb)
  const int rmb = trans ? nb : mb, rnb = trans ? mb : nb;
  const int rma = trans ? na : ma, rna = trans ? ma : na;
  const double* cur = u;
  int m1 = m;
  if (Br) {
    m1 = (m / rnb) * rmb;
    double* dst = As ? tmp : y;
    hipLaunchKernelGGL(k_abu_r, dim3(nblk((long)m1 * n)), dim3(256), 0, st, u, Br, m, n, rmb, rnb,
                       trans, dst);
    cur = dst;
  }
  if (As) {
    const int n2 = (n / rna) * rma;
    hipLaunchKernelGGL(k_abu_s, dim3(nblk((long)m1 * n2)), dim3(256), 0, st, cur, As, m1, n, rma,
                       rna, trans, y);
  }
  if (!Br && !As) SEM_HIP(hipMemcpyAsync(y, u, sizeof(double) * (size_t)m * n,
                                         hipMemcpyDeviceToDevice, st));
  SEM_HIP(hipGetLastError());
  return SEM_OK;
}

int launch_wdot(hipStream_t st, const Layout& g, const double* a, const double* b, double* part,
                int nparts, double* out) {
  if ((int64_t)((g.NC + nparts - 1) / nparts) * g.L > INT32_MAX) {
    set_error("weighted dot: %d columns of %d rows in %d parts exceed 2^31 nodes per part", g.NC,
              g.L, nparts);
    return SEM_ERR_UNSUPPORTED;
  }
  hipLaunchKernelGGL(k_wdot_part, dim3(nparts), dim3(kWdotThreads), 0, st, a, b, g.P, g.L, g.NC,
                     g.per_x, g.per_y, part);
  hipLaunchKernelGGL(k_final, dim3(1), dim3(256), 0, st, part, nparts, 0, out);
  SEM_HIP(hipGetLastError());
  return SEM_OK;
}

int launch_combine(hipStream_t st, const sem_mesh* m, const CombineTerms& terms, int bits,
                   double* y) {
  const long n = (long)m->nl;
  hipLaunchKernelGGL(k_combine, dim3(nblk(n)), dim3(256), 0, st, terms, m->f[SEM_B], y, m->L,
                     m->NC, bits, n);
  SEM_HIP(hipGetLastError());
  return SEM_OK;
}

static int launch_dot_part(hipStream_t st, int64_t n, const double* a, const double* b,
                           const double* w, double* part, int nparts) {
  if (b && w)
    hipLaunchKernelGGL((k_dot_part<true, true>), dim3(nparts), dim3(256), 0, st, a, b, w, (long)n,
                       part);
  else if (b)
    hipLaunchKernelGGL((k_dot_part<true, false>), dim3(nparts), dim3(256), 0, st, a, b, w,
                       (long)n, part);
  else if (w)
    hipLaunchKernelGGL((k_dot_part<false, true>), dim3(nparts), dim3(256), 0, st, a, b, w,
                       (long)n, part);
  else
    hipLaunchKernelGGL((k_dot_part<false, false>), dim3(nparts), dim3(256), 0, st, a, b, w,
                       (long)n, part);
  SEM_HIP(hipGetLastError());
  return SEM_OK;
}

int launch_dot(hipStream_t st, int64_t n, const double* a, const double* b, const double* w,
               double* part, int nparts, double* out) {
  SEM_TRY(launch_dot_part(st, n, a, b, w, part, nparts));
  hipLaunchKernelGGL(k_final, dim3(1), dim3(256), 0, st, part, nparts, 0, out);
  SEM_HIP(hipGetLastError());
  return SEM_OK;
}

int launch_dot_acc(hipStream_t st, int64_t n, const double* a, const double* b, const double* w,
                   double scale, double* part, int nparts, double* acc) {
  SEM_TRY(launch_dot_part(st, n, a, b, w, part, nparts));
  hipLaunchKernelGGL(k_final, dim3(1), dim3(256), 0, st, part, nparts, 2, acc, scale);
  SEM_HIP(hipGetLastError());
  return SEM_OK;
}

int launch_gprod(hipStream_t st, int64_t n, const double* G11, const double* G12,
                 const double* G22, const double* a, const double* b, double* wr, double* ws) {
  hipLaunchKernelGGL(k_gprod, dim3(nblk(n)), dim3(256), 0, st, G11, G12, G22, a, b, wr, ws,
                     (long)n);
  SEM_HIP(hipGetLastError());
  return SEM_OK;
}

int launch_axpby(hipStream_t st, int64_t n, double alpha, const double* x, double beta,
                 const double* y, double* out) {
  hipLaunchKernelGGL(k_axpby, dim3(nblk(n)), dim3(256), 0, st, alpha, x, beta, y, out, (long)n);
  SEM_HIP(hipGetLastError());
  return SEM_OK;
}

int launch_pointwise(hipStream_t st, int op, int64_t n, double alpha, const double* a,
                     const double* b, double* y) {
  hipLaunchKernelGGL(k_pointwise, dim3(nblk(n)), dim3(256), 0, st, op, alpha, a, b, y, (long)n);
  SEM_HIP(hipGetLastError());
  return SEM_OK;
}

int launch_vardiff(hipStream_t st, const sem_mesh* m, const double* nu, double* G11, double* G12,
                   double* G22) {
  const long n = (long)m->nl;
  hipLaunchKernelGGL(k_vardiff, dim3(nblk(n)), dim3(256), 0, st, nu, m->f[SEM_G11],
                     m->f[SEM_G12], m->f[SEM_G22], G11, G12, G22, n);
  SEM_HIP(hipGetLastError());
  return SEM_OK;
}

int launch_norm_inf(hipStream_t st, int64_t n, const double* a, double* part, int nparts,
                    double* out) {
  hipLaunchKernelGGL(k_absmax_part, dim3(nparts), dim3(256), 0, st, a, (long)n, part);
  hipLaunchKernelGGL(k_final, dim3(1), dim3(256), 0, st, part, nparts, 1, out);
  SEM_HIP(hipGetLastError());
  return SEM_OK;
}

int launch_grad(hipStream_t st, const sem_mesh* m, const double* u, double* ux, double* uy) {
  double* const* f = m->f;
  hipLaunchKernelGGL(k_grad, dim3(nblk(m->nl)), dim3(256), 0, st, u, m->Dr, m->Ds, f[SEM_RX],
                     f[SEM_RY], f[SEM_SX], f[SEM_SY], m->nr, m->ns, m->L, (long)m->nl, ux, uy);
  SEM_HIP(hipGetLastError());
  return SEM_OK;
}

int launch_grad_t(hipStream_t st, const sem_mesh* m, const double* vx, const double* vy,
                  double* y) {
  double* const* f = m->f;
  hipLaunchKernelGGL(k_grad_t, dim3(nblk(m->nl)), dim3(256), 0, st, vx, vy, m->Dr, m->Ds,
                     f[SEM_RX], f[SEM_RY], f[SEM_SX], f[SEM_SY], m->nr, m->ns, m->L, (long)m->nl,
                     y);
  SEM_HIP(hipGetLastError());
  return SEM_OK;
}

int launch_adjoint_dnu(hipStream_t st, const sem_mesh* m, const double* u, const double* lam,
                       double* g) {
  double* const* f = m->f;
  hipLaunchKernelGGL(k_adjoint_dnu, dim3(nblk(m->nl)), dim3(256), 0, st, u, lam, m->Dr, m->Ds,
                     f[SEM_RX], f[SEM_RY], f[SEM_SX], f[SEM_SY], f[SEM_B], m->nr, m->ns, m->L,
                     (long)m->nl, g);
  SEM_HIP(hipGetLastError());
  return SEM_OK;
}

int launch_advect(hipStream_t st, const sem_mesh* mV, const sem_mesh* mD, const double* J,
                  const double* T, const double* vx, const double* vy, double* y,
                  int transpose) {
  const int p = mV->nr, pd = mD->nr;
  const int E = mV->Ex * mV->Ey;
  double* const* f = mV->f;
  // J holds Jr (pd x p) followed by Js (pd x p), column-major
  const size_t lds = sizeof(double) * advect_lds_doubles(p, pd);
  hipLaunchKernelGGL(k_advect, dim3(E), dim3(kAdvThreads), lds, st, T, vx, vy, mV->Dr, mV->Ds,
                     f[SEM_RX],
                     f[SEM_RY], f[SEM_SX], f[SEM_SY], mD->f[SEM_B], J, J + pd * p, p, pd, mV->Ex,
                     mV->Ey, y, transpose);
  SEM_HIP(hipGetLastError());
  return SEM_OK;
}

}  // namespace sem

// Mesh generation and point evaluation on the GPU: transfinite (Gordon-Hall)
// interpolation of four boundary curves (geom.jl:8-31) and evaluation of a
// nodal field at physical points by Newton inversion of the element map
// (semInterp, interp.jl:47-102).
#include <algorithm>
#include <cmath>
#include <vector>

#include "sem_device.h"
#include "sem_l0.h"

namespace sem {
namespace {

inline int nblk(long n, int bs = 256) { return (int)((n + bs - 1) / bs); }

// x(r_i, s_j) = (1-r)/2 xrm(s_j) + (1+r)/2 xrp(s_j) + (1-s)/2 xsm(r_i)
//             + (1+s)/2 xsp(r_i) - bilinear interpolant of the corners
// (corners taken from the r-edges: xrm/xrp at s = -1 and s = +1).
__global__ void k_gordon_hall(int nr, int ns, const double* __restrict__ r,
                              const double* __restrict__ s, const double* __restrict__ xrm,
                              const double* __restrict__ xrp, const double* __restrict__ xsm,
                              const double* __restrict__ xsp, double* __restrict__ x) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)nr * ns) return;
  const int i = (int)(t % nr), j = (int)(t / nr);
  const double rm = 0.5 * (1.0 - r[i]), rp = 0.5 * (1.0 + r[i]);
  const double sm = 0.5 * (1.0 - s[j]), sp = 0.5 * (1.0 + s[j]);
  const double c00 = xrm[0], c10 = xrp[0], c01 = xrm[ns - 1], c11 = xrp[ns - 1];
  const double er = rm * xrm[j] + rp * xrp[j];
  const double es = sm * xsm[i] + sp * xsp[i];
  const double cb = sm * (rm * c00 + rp * c10) + sp * (rm * c01 + rp * c11);
  x[t] = er + es - cb;
}

// Lagrange basis at x on nodes z (barycentric weights a): l_j(x), the
// product form of interpMat (interp.jl:10-35)
__device__ inline void lagrange(int n, const double* z, const double* a, double x, double* l) {
  double s[kMaxP], t[kMaxP];
  s[0] = 1.0;
  for (int j = 1; j < n; ++j) s[j] = s[j - 1] * (x - z[j - 1]);
  t[n - 1] = 1.0;
  for (int j = n - 2; j >= 0; --j) t[j] = t[j + 1] * (x - z[j + 1]);
  for (int j = 0; j < n; ++j) l[j] = a[j] * s[j] * t[j];
}

__device__ inline double tensor_eval(const double* __restrict__ f, long base, int L, int nr, int ns,
                                     const double* lr, const double* ls) {
  double v = 0.0;
  for (int b = 0; b < ns; ++b) {
    double acc = 0.0;
    for (int a = 0; a < nr; ++a) acc += lr[a] * f[base + a + (long)L * b];
    v += ls[b] * acc;
  }
  return v;
}

struct InterpArgs {
  const double* u;
  const double *x, *y, *rx, *ry, *sx, *sy;
  const double *zr, *zs, *ar, *as;   // GLL nodes and barycentric weights
  int nr, ns, Ex, Ey, L;
  const double *xp, *yp;
  int npts;
  double tol;
  int maxiter;
  const double* box;   // per element: xmin, xmax, ymin, ymax (4 Ex Ey)
  double* out;
  int* elem;           // element found (-1: none)
};

// One thread per point: scan the elements whose bounding box (widened by
// tol) holds the point; in each, Newton on the element map from the closest
// GLL node, r += rx dx + ry dy, s += sx dx + sy dy with the inverse Jacobian
// interpolated at (r, s), until |(dx, dy)| < tol; the first element where it
// converges inside [-1, 1]^2 gives the value sum l_a(r) l_b(s) u_ab.
__global__ void k_interp_points(InterpArgs g) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.npts) return;
  const double X = g.xp[p], Y = g.yp[p];
  double lr[kMaxP], ls[kMaxP];
  double result = NAN;
  int found = -1;
  const int ne = g.Ex * g.Ey;
  for (int e = 0; e < ne && found < 0; ++e) {
    const double* bx = g.box + 4 * (long)e;
    const double pad = g.tol + 1e-12 * (fabs(bx[1] - bx[0]) + fabs(bx[3] - bx[2]));
    if (X < bx[0] - pad || X > bx[1] + pad || Y < bx[2] - pad || Y > bx[3] + pad) continue;
    const int ex = e % g.Ex, ey = e / g.Ex;
    const long base = (long)ex * g.nr + (long)g.L * ((long)ey * g.ns);
    // start from the closest GLL node of the element
    double best = INFINITY;
    int ia = 0, ib = 0;
    for (int b = 0; b < g.ns; ++b)
      for (int a = 0; a < g.nr; ++a) {
        const long q = base + a + (long)g.L * b;
        const double d2 = (g.x[q] - X) * (g.x[q] - X) + (g.y[q] - Y) * (g.y[q] - Y);
        if (d2 < best) {
          best = d2;
          ia = a;
          ib = b;
        }
      }
    double r = g.zr[ia], s = g.zs[ib];
    bool ok = false;
    for (int it = 0;; ++it) {
      lagrange(g.nr, g.zr, g.ar, r, lr);
      lagrange(g.ns, g.zs, g.as, s, ls);
      const double dx = X - tensor_eval(g.x, base, g.L, g.nr, g.ns, lr, ls);
      const double dy = Y - tensor_eval(g.y, base, g.L, g.nr, g.ns, lr, ls);
      if (sqrt(dx * dx + dy * dy) < g.tol) {
        ok = true;
        break;
      }
      if (it == g.maxiter) break;
      const double rx = tensor_eval(g.rx, base, g.L, g.nr, g.ns, lr, ls);
      const double ry = tensor_eval(g.ry, base, g.L, g.nr, g.ns, lr, ls);
      const double sx = tensor_eval(g.sx, base, g.L, g.nr, g.ns, lr, ls);
      const double sy = tensor_eval(g.sy, base, g.L, g.nr, g.ns, lr, ls);
      const double dr = rx * dx + ry * dy, ds = sx * dx + sy * dy;
      if (!(fabs(r + dr) <= 1.0 + 1e-9 && fabs(s + ds) <= 1.0 + 1e-9)) break;  // left the element
      r += dr;
      s += ds;
    }
    if (ok) {
      found = e;
      result = tensor_eval(g.u, base, g.L, g.nr, g.ns, lr, ls);
    }
  }
  g.out[p] = result;
  if (g.elem) g.elem[p] = found;
}

// bounding box of every element
__global__ void k_elem_box(const double* __restrict__ x, const double* __restrict__ y, int nr,
                           int ns, int Ex, int Ey, int L, double* __restrict__ box) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= Ex * Ey) return;
  const int ex = e % Ex, ey = e / Ex;
  const long base = (long)ex * nr + (long)L * ((long)ey * ns);
  double x0 = INFINITY, x1 = -INFINITY, y0 = INFINITY, y1 = -INFINITY;
  for (int b = 0; b < ns; ++b)
    for (int a = 0; a < nr; ++a) {
      const long q = base + a + (long)L * b;
      x0 = fmin(x0, x[q]);
      x1 = fmax(x1, x[q]);
      y0 = fmin(y0, y[q]);
      y1 = fmax(y1, y[q]);
    }
  box[4 * (long)e] = x0;
  box[4 * (long)e + 1] = x1;
  box[4 * (long)e + 2] = y0;
  box[4 * (long)e + 3] = y1;
}

}  // namespace
}  // namespace sem

using namespace sem;

extern "C" {

int sem_gordon_hall(sem_ctx* ctx, int nr, int ns, const double* r, const double* s,
                    const double* xrm, const double* xrp, const double* xsm, const double* xsp,
                    const double* yrm, const double* yrp, const double* ysm, const double* ysp,
                    double* x, double* y) {
  SEM_ARG(ctx && r && s && xrm && xrp && xsm && xsp && yrm && yrp && ysm && ysp && x && y,
          "NULL argument");
  SEM_ARG(nr >= 2 && ns >= 2, "sem_gordon_hall: need nr, ns >= 2");
  const long n = (long)nr * ns;
  hipLaunchKernelGGL(k_gordon_hall, dim3(nblk(n)), dim3(256), 0, ctx->stream, nr, ns, r, s, xrm,
                     xrp, xsm, xsp, x);
  hipLaunchKernelGGL(k_gordon_hall, dim3(nblk(n)), dim3(256), 0, ctx->stream, nr, ns, r, s, yrm,
                     yrp, ysm, ysp, y);
  SEM_HIP(hipGetLastError());
  return SEM_OK;
}

int sem_interp_points(sem_ctx* ctx, int nr, int ns, int Ex, int Ey, const double* u,
                      const double* x, const double* y, const double* rx, const double* ry,
                      const double* sx, const double* sy, int npts, const double* xp,
                      const double* yp, double tol, int maxiter, double* out, int* elem) {
  SEM_ARG(ctx && u && x && y && rx && ry && sx && sy && out && (npts == 0 || (xp && yp)),
          "NULL argument");
  SEM_ARG(nr >= 2 && ns >= 2 && nr <= kMaxP && ns <= kMaxP,
          "sem_interp_points: need 2 <= nr, ns <= %d", kMaxP);
  SEM_ARG(Ex >= 1 && Ey >= 1 && npts >= 0 && tol > 0.0 && maxiter >= 0,
          "sem_interp_points: bad sizes or limits");
  if (npts == 0) return SEM_OK;
  hipStream_t st = ctx->stream;
  // GLL nodes and barycentric weights of both directions, the element boxes
  std::vector<double> host(4 * (size_t)kMaxP, 0.0);
  {
    std::vector<double> w(kMaxP);
    gll(nr, host.data(), w.data());
    gll(ns, host.data() + kMaxP, w.data());
    for (int d = 0; d < 2; ++d) {
      const int n = d ? ns : nr;
      const double* z = host.data() + d * kMaxP;
      double* a = host.data() + (2 + d) * kMaxP;
      for (int i = 0; i < n; ++i) {
        double v = 1.0;
        for (int j = 0; j < n; ++j)
          if (j != i) v *= z[i] - z[j];
        a[i] = 1.0 / v;
      }
    }
  }
  const size_t boxb = sizeof(double) * 4 * (size_t)Ex * Ey;
  SEM_TRY(ensure_ws(ctx, sizeof(double) * host.size() + boxb));
  double* tab = static_cast<double*>(ctx->ws);
  double* box = tab + host.size();
  SEM_HIP(hipMemcpyAsync(tab, host.data(), sizeof(double) * host.size(), hipMemcpyHostToDevice, st));
  const int L = nr * Ex;
  hipLaunchKernelGGL(k_elem_box, dim3(nblk((long)Ex * Ey)), dim3(256), 0, st, x, y, nr, ns, Ex,
                     Ey, L, box);
  InterpArgs g{};
  g.u = u;
  g.x = x;
  g.y = y;
  g.rx = rx;
  g.ry = ry;
  g.sx = sx;
  g.sy = sy;
  g.zr = tab;
  g.zs = tab + kMaxP;
  g.ar = tab + 2 * kMaxP;
  g.as = tab + 3 * kMaxP;
  g.nr = nr;
  g.ns = ns;
  g.Ex = Ex;
  g.Ey = Ey;
  g.L = L;
  g.xp = xp;
  g.yp = yp;
  g.npts = npts;
  g.tol = tol;
  g.maxiter = maxiter;
  g.box = box;
  g.out = out;
  g.elem = elem;
  hipLaunchKernelGGL(k_interp_points, dim3(nblk(npts, 64)), dim3(64), 0, st, g);
  SEM_HIP(hipGetLastError());
  // the host table and the boxes live in the shared workspace until the kernels ran
  SEM_HIP(hipStreamSynchronize(st));
  return SEM_OK;
}

int sem_mesh_interp(sem_ctx* ctx, const sem_mesh* m, const double* u, int npts, const double* xp,
                    const double* yp, double tol, int maxiter, double* out, int* elem) {
  SEM_ARG(ctx && m, "NULL argument");
  return sem_interp_points(ctx, m->nr, m->ns, m->Ex, m->Ey, u, m->f[SEM_X], m->f[SEM_Y],
                           m->f[SEM_RX], m->f[SEM_RY], m->f[SEM_SX], m->f[SEM_SY], npts, xp, yp,
                           tol, maxiter, out, elem);
}

}  // extern "C"

// Element operators for any nr x ns blocks (the reference's Mesh(nr, ns, ...)
// takes any order in each direction, mesh.jl:66-133), where the fused strip
// kernels need nr == ns <= 16: one workgroup per element with the block, the
// fluxes and Dr, Ds in LDS, then QQ' and the mask through the sorted map. And
// PCG with such an operator (pcg.jl:16-60; host-driven, device vectors).
#include <algorithm>
#include <cmath>

#include "sem_device.h"

namespace sem {
namespace {

inline int nblk(long n, int bs = 256) { return (int)((n + bs - 1) / bs); }

constexpr int kGenThreads = 256;

// H = nu [Dr'(G11 ur + G12 us) + (G12 ur + G22 us) Ds] + k B u per element
// (laplace, lapl.jl:70-81; hlmz.jl:12-19): ur = Dr u, us = u Ds'.
__global__ void __launch_bounds__(kGenThreads)
k_ax_generic(const double* __restrict__ u, double* __restrict__ y,
             const double* __restrict__ G11, const double* __restrict__ G12,
             const double* __restrict__ G22, const double* __restrict__ B,
             const double* __restrict__ Dr, const double* __restrict__ Ds, double nu,
             const double* __restrict__ nu_f, double kc, const double* __restrict__ k_f,
             int has_mass, int nr, int ns, int Ex, int L) {
  extern __shared__ double sh[];
  const int ne = nr * ns;
  double* su = sh;               // u block (a + nr b)
  double* swr = su + ne;         // wr
  double* sws = swr + ne;        // ws
  double* sdr = sws + ne;        // Dr (nr x nr, column-major)
  double* sds = sdr + nr * nr;   // Ds (ns x ns)
  const int e = blockIdx.x;
  const int ex = e % Ex, ey = e / Ex;
  const long base = (long)ex * nr + (long)L * ((long)ey * ns);
  for (int q = threadIdx.x; q < nr * nr; q += kGenThreads) sdr[q] = Dr[q];
  for (int q = threadIdx.x; q < ns * ns; q += kGenThreads) sds[q] = Ds[q];
  for (int q = threadIdx.x; q < ne; q += kGenThreads) {
    const int a = q % nr, b = q / nr;
    su[q] = u[base + a + (long)L * b];
  }
  __syncthreads();
  for (int q = threadIdx.x; q < ne; q += kGenThreads) {
    const int a = q % nr, b = q / nr;
    double ur = 0.0, us = 0.0;
    for (int k = 0; k < nr; ++k) ur += sdr[a + nr * k] * su[k + nr * b];
    for (int k = 0; k < ns; ++k) us += su[a + nr * k] * sds[b + ns * k];
    const long t = base + a + (long)L * b;
    swr[q] = G11[t] * ur + G12[t] * us;
    sws[q] = G12[t] * ur + G22[t] * us;
  }
  __syncthreads();
  for (int q = threadIdx.x; q < ne; q += kGenThreads) {
    const int a = q % nr, b = q / nr;
    double p1 = 0.0, p2 = 0.0;
    for (int k = 0; k < nr; ++k) p1 += sdr[k + nr * a] * swr[k + nr * b];
    for (int k = 0; k < ns; ++k) p2 += sws[a + nr * k] * sds[k + ns * b];
    const long t = base + a + (long)L * b;
    double h = (nu_f ? nu_f[t] : nu) * (p1 + p2);
    if (has_mass) h = h + (k_f ? k_f[t] : kc) * (B[t] * su[q]);
    y[t] = h;
  }
}

__global__ void k_copy_scale(const double* __restrict__ a, const double* __restrict__ d,
                             double* __restrict__ z, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) z[t] = d ? d[t] * a[t] : a[t];
}

// x += alpha p, r -= alpha w
__global__ void k_gen_xr(double* __restrict__ x, double* __restrict__ r,
                         const double* __restrict__ p, const double* __restrict__ w, double alpha,
                         long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) {
    x[t] = x[t] + alpha * p[t];
    r[t] = r[t] - alpha * w[t];
  }
}

}  // namespace

int launch_apply_generic(hipStream_t st, const sem_mesh* m, const double* u, double* y,
                         const double* G11, const double* G12, const double* G22,
                         const double* B, double nu, const double* nu_f, double k,
                         const double* k_f, int has_mass, int assemble, int bits) {
  const int nr = m->nr, ns = m->ns;
  const size_t lds = sizeof(double) * (3 * (size_t)nr * ns + (size_t)nr * nr + (size_t)ns * ns);
  if (lds > 64 * 1024) {
    set_error("element operators support nr * ns blocks up to 64 KiB of LDS (got %d x %d)", nr,
              ns);
    return SEM_ERR_UNSUPPORTED;
  }
  hipLaunchKernelGGL(k_ax_generic, dim3(m->Ex * m->Ey), dim3(kGenThreads), lds, st, u, y, G11,
                     G12, G22, B, m->Dr, m->Ds, nu, nu_f, k, k_f, has_mass, nr, ns, m->Ex, m->L);
  SEM_HIP(hipGetLastError());
  if (assemble) SEM_TRY(launch_gsmap_masked(st, m->map, m->layout(), bits, y, y));
  return SEM_OK;
}

// pcg.jl:16-60 with a generic operator: x0 = 0, stop when norm(r, Inf) <= tol,
// inner products weighted by the mesh's mult, z = d .* r (or r); the scalars
// are read on the host every iteration.
int pcg_generic(sem_ctx* ctx, const sem_operator* op, const double* d, const double* b, double* x,
                double tol, int maxiter, sem_pcg_info* info) {
  SEM_ARG(op->precision == SEM_F64, "pcg: meshes with nr != ns or nr > %d need SEM_F64", kMaxP);
  const sem_mesh* m = op->mesh;
  const long n = (long)m->nl;
  hipStream_t st = ctx->stream;
  const size_t vb = ((sizeof(double) * n + 255) / 256) * 256;
  constexpr int kParts = 256;
  SEM_TRY(ensure_ws(ctx, 4 * vb + sizeof(double) * (kParts + 1)));
  SEM_TRY(ensure_pinned(ctx, 128));
  char* w0 = static_cast<char*>(ctx->ws);
  double* r = reinterpret_cast<double*>(w0);
  double* p = reinterpret_cast<double*>(w0 + vb);
  double* w = reinterpret_cast<double*>(w0 + 2 * vb);
  double* z = reinterpret_cast<double*>(w0 + 3 * vb);
  double* part = reinterpret_cast<double*>(w0 + 4 * vb);
  double* hv = reinterpret_cast<double*>(static_cast<char*>(ctx->pinned) + 64);
  const double* mult = m->f[SEM_MULT];
  auto fetch = [&](double* out) -> int {
    SEM_HIP(hipMemcpyAsync(hv, part + kParts, sizeof(double), hipMemcpyDeviceToHost, st));
    SEM_HIP(hipStreamSynchronize(st));
    *out = *hv;
    return SEM_OK;
  };
  auto dot = [&](const double* a, const double* c, double* out) -> int {
    SEM_TRY(launch_dot(st, n, a, c, mult, part, kParts, part + kParts));
    return fetch(out);
  };
  auto norm = [&](const double* a, double* out) -> int {
    SEM_TRY(launch_norm_inf(st, n, a, part, kParts, part + kParts));
    return fetch(out);
  };
  SEM_HIP(hipEventRecord(ctx->ev0, st));
  SEM_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, st));
  SEM_HIP(hipMemcpyAsync(r, b, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
  double res = 0.0, t = 0.0, t_prev = 1.0;
  SEM_TRY(norm(r, &res));
  if (tol < 0.0) tol = -tol * res;   // relative stop test
  int k = 0;
  bool conv = true;
  while (res > tol) {
    if (k == maxiter) {
      conv = false;
      break;
    }
    ++k;
    hipLaunchKernelGGL(k_copy_scale, dim3(nblk(n)), dim3(256), 0, st, r, d, z, n);
    SEM_TRY(dot(r, z, &t));
    if (k == 1) {
      SEM_HIP(hipMemcpyAsync(p, z, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    } else {
      SEM_TRY(launch_axpby(st, n, 1.0, z, t / t_prev, p, p));
    }
    SEM_TRY(sem_operator_apply(ctx, op, p, w));
    double pw = 0.0;
    SEM_TRY(dot(p, w, &pw));
    hipLaunchKernelGGL(k_gen_xr, dim3(nblk(n)), dim3(256), 0, st, x, r, p, w, t / pw, n);
    SEM_HIP(hipGetLastError());
    SEM_TRY(norm(r, &res));
    t_prev = t;
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
    set_error("pcg: residual is not finite after %d iterations", k);
    return SEM_ERR_NONFINITE;
  }
  return (conv && res <= tol) ? SEM_OK : SEM_WARN_MAXITER;
}

}  // namespace sem

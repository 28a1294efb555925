// Fused Jacobi-PCG for the SEM operator, restating pcg.jl:16-60 on the GPU.
//
// Per iteration k (all scalars stay on the device):
//   apply kernel (k_ax MODE 2): checks the stop test on the partials of the
//       previous update (norm(r, Inf) <= tol, or k-1 == maxiter), forms
//       t = <r, h>, beta = t / t_prev, u = h + beta u (h = d .* r or r),
//       w = M .* QQ'(A u) except at workgroup seams, and the partials of
//       <u, w>_mult, which equal sum(u .* H) over the unassembled element
//       outputs H because u is continuous and masked;
//   update kernel: finishes the seam copies of w on the fly (same order as
//       the seam kernel of the plain apply), alpha = t / <u, w>,
//       x += alpha u, r -= alpha w, and the partials of <r, h> and
//       norm(r, Inf) for the next iteration.
// Two launches per iteration.
// Partials are summed in a fixed order by every block that needs them, so
// every block sees bit-identical scalars and the solve is deterministic.
// The host launches iterations in growing batches and reads the done flag
// between batches; iterations launched after convergence exit at once.
#include "sem_device.h"

namespace sem {
namespace {

template <typename T>
__global__ void __launch_bounds__(256)
k_pcg_init(const T* __restrict__ b, T* __restrict__ x, T* __restrict__ r,
           const T* __restrict__ d, int P, int L, int NC, int per_x, int per_y, long n,
           double* __restrict__ part, int n3) {
  __shared__ double red[256];
  double rho = 0.0, mx = 0.0;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long)gridDim.x * blockDim.x) {
    const int i = (int)(t % L);
    const int j = (int)(t / L);
    const T bv = b[t];
    x[t] = T(0);
    r[t] = bv;
    const T z = d ? d[t] * bv : bv;
    rho += (double)bv * (double)z * (mult1(i, P, L, per_x) * mult1(j, P, NC, per_y));
    mx = nanmax(mx, fabs((double)bv));
  }
  rho = block_sum(rho, red, 256);
  mx = block_max(mx, red, 256);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = rho;
    part[n3 + blockIdx.x] = mx;
  }
}

template <typename T>
__global__ void __launch_bounds__(256)
k_pcg_update(T* __restrict__ x, T* __restrict__ r, const T* __restrict__ p,
             const T* __restrict__ w, const T* __restrict__ d, const double* __restrict__ Q,
             int nQ, PcgState* __restrict__ st, int k, double* __restrict__ part, int n3,
             SeamGeom sg, long n) {
  const int P = sg.P, L = sg.L, NC = sg.NC, per_x = sg.per_x, per_y = sg.per_y;
  __shared__ double red[256];
  if (block_done(st, red)) return;
  const double pw = reduce_sum_array(Q, nQ, red, 256);
  const T alpha = T(st->t[k & 1] / pw);
  double rho = 0.0, mx = 0.0;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long)gridDim.x * blockDim.x) {
    const int i = (int)(t % L);
    const int j = (int)(t / L);
    x[t] = x[t] + alpha * p[t];
    const T rv = r[t] - alpha * seam_final(w, i, j, sg);
    r[t] = rv;
    const T z = d ? d[t] * rv : rv;
    rho += (double)rv * (double)z * (mult1(i, P, L, per_x) * mult1(j, P, NC, per_y));
    mx = nanmax(mx, fabs((double)rv));
  }
  rho = block_sum(rho, red, 256);
  mx = block_max(mx, red, 256);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = rho;
    part[n3 + blockIdx.x] = mx;
  }
}

// Element diagonal of nu A + k B (nu, k scalars or fields):
// nu [sum_i Dr_ia^2 G11_ib + 2 Dr_aa Ds_bb G12_ab + sum_j Ds_jb^2 G22_aj] + k B_ab
__global__ void k_jacobi_local(const double* __restrict__ G11, const double* __restrict__ G12,
                               const double* __restrict__ G22, const double* __restrict__ B,
                               const double* __restrict__ Dr, const double* __restrict__ Ds,
                               double nu, const double* __restrict__ nu_f, double kc,
                               const double* __restrict__ k_f, int has_mass, int P, int L, long n,
                               double* __restrict__ dl) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = (int)(t % L);
  const int j = (int)(t / L);
  const int a = i % P, i0 = i - a;
  const int b = j % P, j0 = j - b;
  double t1 = 0, t3 = 0;
  for (int q = 0; q < P; ++q) {
    const double dr = Dr[q + P * a];
    t1 += dr * dr * G11[(i0 + q) + (long)L * j];
    const double ds = Ds[q + P * b];
    t3 += ds * ds * G22[i + (long)L * (j0 + q)];
  }
  const double t2 = 2.0 * Dr[a + P * a] * Ds[b + P * b] * G12[t];
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

template <typename T>
int pcg_run_t(sem_ctx* ctx, const sem_operator* op, const T* d, const T* b, T* x, double tol,
              int maxiter, sem_pcg_info* info) {
  const sem_mesh* m = op->mesh;
  const Layout g = m->layout();
  const long n = (long)m->nl;
  const AxPlan pl = make_plan(g, true);
  const int n3 = (int)std::min<long>(512, std::max<long>(1, (n + 2047) / 2048));
  const int nQ = pl.nk1;
  SeamGeom sg{g.P, g.L, g.NC, pl.CX * g.P, pl.nstrips, pl.S * g.P, pl.nsegs, g.per_x, g.per_y,
              op->desc.bc_bits};
  // workspace: r, p, w | part[2][2*n3] | Q[nQ] | state
  const size_t vb = ((sizeof(T) * n + 255) / 256) * 256;
  const size_t pb = sizeof(double) * 4 * n3;
  const size_t qb = ((sizeof(double) * nQ + 255) / 256) * 256;
  SEM_TRY(ensure_ws(ctx, 3 * vb + pb + qb + 256));
  char* w0 = static_cast<char*>(ctx->ws);
  T* r = reinterpret_cast<T*>(w0);
  T* p = reinterpret_cast<T*>(w0 + vb);
  T* w = reinterpret_cast<T*>(w0 + 2 * vb);
  double* part = reinterpret_cast<double*>(w0 + 3 * vb);
  double* Q = reinterpret_cast<double*>(w0 + 3 * vb + pb);
  PcgState* state = reinterpret_cast<PcgState*>(w0 + 3 * vb + pb + qb);
  SEM_TRY(ensure_pinned(ctx, sizeof(PcgState)));
  hipStream_t st = ctx->stream;

  ApplyArgs aa{};
  aa.u = nullptr;
  aa.y = w;
  aa.G11 = op->G11;
  aa.G12 = op->G12;
  aa.G22 = op->G22;
  aa.B = op->B;
  aa.Dr = sizeof(T) == 8 ? (const void*)m->Dr : (const void*)m->Dr32;
  aa.Ds = sizeof(T) == 8 ? (const void*)m->Ds : (const void*)m->Ds32;
  aa.nu_f = op->nu_f;
  aa.k_f = op->k_f;
  aa.nu = op->desc.nu;
  aa.k = op->desc.k;
  aa.has_mass = op->has_mass ? 1 : 0;
  aa.bc = op->desc.bc_bits;
  aa.assemble = 1;
  aa.mfma = op->precision == SEM_F16_MFMA     ? 1
            : op->precision == SEM_BF16_MFMA  ? 2
            : op->precision == SEM_F16X3_MFMA ? 3
            : op->precision == SEM_BF16X3_MFMA ? 4
                                               : 0;

  SEM_HIP(hipEventRecord(ctx->ev0, st));
  SEM_HIP(hipMemsetAsync(state, 0, sizeof(PcgState), st));
  hipLaunchKernelGGL((k_pcg_init<T>), dim3(n3), dim3(256), 0, st, b, x, r, d, g.P, g.L, g.NC,
                     g.per_x, g.per_y, n, part, n3);
  SEM_HIP(hipGetLastError());
  PcgState* hs = static_cast<PcgState*>(ctx->pinned);
  int k = 1, batch = 4;
  const long kmax = (long)maxiter + 1;
  for (;;) {
    for (int q = 0; q < batch && k <= kmax; ++q, ++k) {
      PcgLaunch pc{};
      pc.r = r;
      pc.d = d;
      pc.p = p;
      pc.dev.part_prev = part + 2 * n3 * ((k - 1) & 1);
      pc.dev.n3 = n3;
      pc.dev.Q = Q;
      pc.dev.state = state;
      pc.dev.tol = tol;
      pc.dev.maxiter = maxiter;
      pc.dev.k = k;
      SEM_TRY(launch_apply_pcg<T>(st, g, aa, pc));
      hipLaunchKernelGGL((k_pcg_update<T>), dim3(n3), dim3(256), 0, st, x, r, p, w, d, Q, nQ,
                         state, k, part + 2 * n3 * (k & 1), n3, sg, n);
    }
    SEM_HIP(hipGetLastError());
    SEM_HIP(hipMemcpyAsync(hs, state, sizeof(PcgState), hipMemcpyDeviceToHost, st));
    SEM_HIP(hipStreamSynchronize(st));
    if (hs->done) break;
    if (k > kmax) {
      set_error("pcg: iteration limit passed without the stop test firing");
      return SEM_ERR_BREAKDOWN;
    }
    batch = std::min(batch * 2, 64);
  }
  SEM_HIP(hipEventRecord(ctx->ev1, st));
  SEM_HIP(hipEventSynchronize(ctx->ev1));
  float ms = 0.f;
  SEM_HIP(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
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
                     m->L, (long)m->nl, dl);
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
  if (op->precision == SEM_F64)
    return pcg_run_t<double>(ctx, op, static_cast<const double*>(d),
                             static_cast<const double*>(b), static_cast<double*>(x), tol, maxiter,
                             info);
  return pcg_run_t<float>(ctx, op, static_cast<const float*>(d), static_cast<const float*>(b),
                          static_cast<float*>(x), tol, maxiter, info);
}

}  // namespace sem

// Mixed-precision solve: FP64 defect correction around a low-precision PCG.
//
// Outer (FP64): r = b - A x with the fused FP64 operator (true residual),
// stop when norm(r, Inf) <= tol. Inner: the normalised defect r / norm(r)
// is solved with the low-precision operator's fused PCG to relative
// tolerance inner_tol, and x += norm(r) * e. The inner solves start from
// zero, as pcg! does (pcg.jl:71).
#include "sem_device.h"

namespace sem {
namespace {

__global__ void k_resid(const double* __restrict__ b, double* __restrict__ r, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) r[t] = b[t] - r[t];
}

__global__ void k_scale_to_f32(const double* __restrict__ r, float* __restrict__ r32, double s,
                               long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) r32[t] = (float)(r[t] * s);
}

__global__ void k_axpy_f32(double* __restrict__ x, const float* __restrict__ e, double s,
                           long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) x[t] = x[t] + s * (double)e[t];
}

inline int nblk(long n, int bs = 256) { return (int)((n + bs - 1) / bs); }

}  // namespace
}  // namespace sem

using namespace sem;

extern "C" int sem_pcg_mixed(sem_ctx* ctx, const sem_operator* op64, const sem_operator* op_lo,
                             const double* d64, const double* b, double* x, double tol,
                             double inner_tol, int max_outer, int max_inner,
                             sem_pcg_info* info) {
  SEM_ARG(ctx && op64 && op_lo && b && x, "NULL argument");
  SEM_ARG(op64->precision == SEM_F64, "sem_pcg_mixed: op64 must be SEM_F64");
  SEM_ARG(op_lo->precision != SEM_F64, "sem_pcg_mixed: op_lo must be a low-precision operator");
  SEM_ARG(op64->mesh == op_lo->mesh, "sem_pcg_mixed: operators on different meshes");
  SEM_ARG(op64->desc.assemble && op_lo->desc.assemble, "sem_pcg_mixed: operators must assemble");
  SEM_ARG(max_outer >= 1 && max_inner >= 1 && inner_tol > 0 && tol >= 0,
          "sem_pcg_mixed: bad limits");
  const sem_mesh* m = op64->mesh;
  const long n = (long)m->nl;
  hipStream_t st = ctx->stream;
  // buffers owned by this call (the fused PCG uses ctx->ws)
  double* r64 = nullptr;
  float *r32 = nullptr, *e32 = nullptr, *d32 = nullptr;
  double* res_d = nullptr;
  double* part = nullptr;
  auto cleanup = [&]() {
    hipStreamSynchronize(st);
    hipFree(r64);
    hipFree(r32);
    hipFree(e32);
    hipFree(d32);
    hipFree(part);
  };
  if (hipMalloc(&r64, sizeof(double) * n) != hipSuccess ||
      hipMalloc(&r32, sizeof(float) * n) != hipSuccess ||
      hipMalloc(&e32, sizeof(float) * n) != hipSuccess ||
      hipMalloc(&part, sizeof(double) * 260) != hipSuccess ||
      (d64 && hipMalloc(&d32, sizeof(float) * n) != hipSuccess)) {
    set_error("sem_pcg_mixed: hipMalloc failed");
    cleanup();
    return SEM_ERR_HIP;
  }
  res_d = part + 256;
  if (d64) {
    int s = launch_convert_f64_f32(st, n, d64, d32);
    if (s != SEM_OK) {
      cleanup();
      return s;
    }
  }
  SEM_HIP(hipMemsetAsync(x, 0, sizeof(double) * n, st));
  hipEvent_t e0, e1;
  SEM_HIP(hipEventCreate(&e0));
  SEM_HIP(hipEventCreate(&e1));
  SEM_HIP(hipEventRecord(e0, st));
  int outer = 0, inner_total = 0, status = SEM_WARN_MAXITER;
  double res = 0.0;
  double* hres = nullptr;
  if (ensure_pinned(ctx, sizeof(PcgState)) != SEM_OK) {
    cleanup();
    return SEM_ERR_HIP;
  }
  hres = static_cast<double*>(ctx->pinned);
  for (;;) {
    // true residual in FP64
    int s = sem_operator_apply(ctx, op64, x, r64);
    if (s != SEM_OK) {
      cleanup();
      return s;
    }
    hipLaunchKernelGGL(k_resid, dim3(nblk(n)), dim3(256), 0, st, b, r64, n);
    s = launch_norm_inf(st, n, r64, part, 256, res_d);
    if (s != SEM_OK) {
      cleanup();
      return s;
    }
    SEM_HIP(hipMemcpyAsync(hres, res_d, sizeof(double), hipMemcpyDeviceToHost, st));
    SEM_HIP(hipStreamSynchronize(st));
    res = *hres;
    if (!(res > tol)) {
      status = (res <= tol) ? SEM_OK : SEM_ERR_NONFINITE;
      break;
    }
    if (outer == max_outer) break;
    ++outer;
    hipLaunchKernelGGL(k_scale_to_f32, dim3(nblk(n)), dim3(256), 0, st, r64, r32, 1.0 / res, n);
    sem_pcg_info ii{};
    s = pcg_run(ctx, op_lo, d64 ? SEM_PRECOND_DIAG : SEM_PRECOND_IDENTITY, d32, r32, e32,
                inner_tol, max_inner, &ii);
    if (s < 0) {
      cleanup();
      return s;
    }
    inner_total += ii.iters;
    hipLaunchKernelGGL(k_axpy_f32, dim3(nblk(n)), dim3(256), 0, st, x, e32, res, n);
  }
  SEM_HIP(hipEventRecord(e1, st));
  SEM_HIP(hipEventSynchronize(e1));
  float ms = 0;
  hipEventElapsedTime(&ms, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (info) {
    info->iters = inner_total;
    info->inner_total = inner_total;
    info->outer = outer;
    info->res = res;
    info->converged = status == SEM_OK;
    info->elapsed_ms = ms;
  }
  cleanup();
  if (status == SEM_ERR_NONFINITE) set_error("sem_pcg_mixed: residual is not finite");
  return status;
}

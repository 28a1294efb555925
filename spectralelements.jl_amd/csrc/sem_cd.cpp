// Native time loop of the convection-diffusion equation: step!(cdn)
// (convectionDiffusion.jl:148-155) repeated on the host without returning to
// the caller between steps, so the only host wait per step is the PCG stop
// test. Each step:
//   updateHist!(fld) (mesh.jl:199-215): uh rotates, uh[0] <- u; the cached
//       advect(uh_i) rotate with it;
//   updateHist!(tstep) (time.jl:101-112): t shifts, t[0] = t[1] + dt, and
//       bdfExtK(t, k) (time.jl:31-53) gives a, b;
//   makeRHS! (convectionDiffusion.jl:93-110) with f = 0 and ub = 0:
//       rhs = QQ' mask(sum_i -b[1+i] B uh_i + sum_i a_i exH_i),
//       exH_i = -advect(uh_i), computed once per history (sem_combine);
//   solve! (convectionDiffusion.jl:112-122): pcg!(u, rhs, opLHS; opM, mult)
//       with opLHS = mask(QQ'(hlmz(., nu, b[0]))) and the assembled Jacobi
//       diagonal of that operator (one operator per b[0], cached), stop test
//       norm(r, Inf) <= tol * norm(rhs, Inf) formed on the device;
//   optionally J += dt/2 sum(B .* u .* u) accumulated on the device.
// Every launch is the one the step-by-step host code issues, in the same
// order, so both give identical bits.
#include <algorithm>
#include <new>
#include <vector>

#include "sem_common.h"

struct sem_cd {
  const sem_mesh* mV = nullptr;
  const sem_mesh* mD = nullptr;
  int bc_bits = 0;
  int k = 2;
  double nu = 0.0;
  const double* nu_f = nullptr;
  const double* vx = nullptr;
  const double* vy = nullptr;
  double dt = 0.0;
  double tol = 1e-12;
  int maxiter = 0;
  struct Op {
    double b0;
    sem_operator* op;
    double* dinv;
  };
  std::vector<Op> ops;
  double* rhs = nullptr;
};

using namespace sem;

namespace {

int cd_op(sem_ctx* ctx, sem_cd* cd, double b0, sem_cd::Op** out) {
  for (auto& o : cd->ops)
    if (o.b0 == b0) {
      *out = &o;
      return SEM_OK;
    }
  sem_op_desc d{};
  d.nu = cd->nu;
  d.nu_f = cd->nu_f;
  d.k = b0;
  d.bc_bits = cd->bc_bits;
  d.assemble = 1;
  sem_cd::Op o{b0, nullptr, nullptr};
  SEM_TRY(sem_operator_create(ctx, cd->mV, &d, SEM_F64, &o.op));
  if (hipMalloc(&o.dinv, sizeof(double) * cd->mV->nl) != hipSuccess) {
    sem_operator_destroy(o.op);
    set_error("sem_cd: hipMalloc of the Jacobi diagonal failed");
    return SEM_ERR_HIP;
  }
  const int s = sem_operator_jacobi(ctx, o.op, o.dinv);
  if (s != SEM_OK) {
    hipFree(o.dinv);
    sem_operator_destroy(o.op);
    return s;
  }
  cd->ops.push_back(o);
  *out = &cd->ops.back();
  return SEM_OK;
}

}  // namespace

extern "C" {

int sem_cd_create(sem_ctx* ctx, const sem_mesh* mV, const sem_mesh* mD, int bc_bits, int k,
                  double nu, const double* nu_f, const double* vx, const double* vy, double dt,
                  double tol, int maxiter, sem_cd** out) {
  SEM_ARG(ctx && mV && mD && vx && vy && out, "sem_cd_create: NULL argument");
  SEM_ARG(k >= 1 && k <= 4, "sem_cd_create: order k in [1, 4] (got %d)", k);
  SEM_ARG(dt > 0.0 && tol >= 0.0 && maxiter >= 0, "sem_cd_create: bad dt, tol or maxiter");
  sem_cd* cd = new (std::nothrow) sem_cd();
  SEM_ARG(cd, "out of host memory");
  cd->mV = mV;
  cd->mD = mD;
  cd->bc_bits = bc_bits;
  cd->k = k;
  cd->nu = nu;
  cd->nu_f = nu_f;
  cd->vx = vx;
  cd->vy = vy;
  cd->dt = dt;
  cd->tol = tol;
  cd->maxiter = maxiter;
  if (hipMalloc(&cd->rhs, sizeof(double) * mV->nl) != hipSuccess) {
    delete cd;
    set_error("sem_cd_create: hipMalloc of rhs failed");
    return SEM_ERR_HIP;
  }
  *out = cd;
  return SEM_OK;
}

int sem_cd_destroy(sem_cd* cd) {
  if (!cd) return SEM_OK;
  hipDeviceSynchronize();
  for (auto& o : cd->ops) {
    hipFree(o.dinv);
    sem_operator_destroy(o.op);
  }
  hipFree(cd->rhs);
  delete cd;
  return SEM_OK;
}

int sem_cd_steps(sem_ctx* ctx, sem_cd* cd, int nsteps, double* u, double** uh, double** adv,
                 int* adv_valid, double* t, double* J_dev, double* a_out, double* b_out,
                 int64_t* iters_total) {
  SEM_ARG(ctx && cd && u && uh && adv && adv_valid && t, "sem_cd_steps: NULL argument");
  SEM_ARG(nsteps >= 0, "sem_cd_steps: nsteps < 0");
  const int k = cd->k;
  const sem_mesh* m = cd->mV;
  const int64_t n = m->nl;
  hipStream_t st = ctx->stream;
  std::vector<double> a(k), b(k + 1);
  const double* xs[8];
  double coef[8];
  int w[8];
  int64_t iters = 0;
  int status = SEM_OK;
  for (int s = 0; s < nsteps; ++s) {
    // updateHist!(fld): uh[0] <- u, the others shift; cached advects shift
    double* last = uh[k - 1];
    double* alast = adv[k - 1];
    for (int i = k - 1; i > 0; --i) {
      uh[i] = uh[i - 1];
      adv[i] = adv[i - 1];
      adv_valid[i] = adv_valid[i - 1];
    }
    SEM_HIP(hipMemcpyAsync(last, u, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    uh[0] = last;
    adv[0] = alast;
    adv_valid[0] = 0;
    // updateHist!(tstep)
    for (int i = k; i > 0; --i) t[i] = t[i - 1];
    t[0] = t[1] + cd->dt;
    SEM_TRY(sem_bdfextk(k + 1, t, k, a.data(), b.data()));
    if (a_out)
      for (int i = 0; i < k; ++i) a_out[(int64_t)s * k + i] = a[i];
    if (b_out)
      for (int i = 0; i <= k; ++i) b_out[(int64_t)s * (k + 1) + i] = b[i];
    // makeRHS!
    int nt = 0;
    for (int i = 0; i < k; ++i) {
      if (!adv_valid[i] && a[i] != 0.0) {
        SEM_TRY(sem_advect(ctx, m, cd->mD, uh[i], cd->vx, cd->vy, adv[i]));
        adv_valid[i] = 1;
      }
      if (b[1 + i] != 0.0) {
        coef[nt] = -b[1 + i];
        xs[nt] = uh[i];
        w[nt++] = 1;
      }
      if (a[i] != 0.0) {
        coef[nt] = -a[i];
        xs[nt] = adv[i];
        w[nt++] = 0;
      }
    }
    SEM_TRY(sem_combine(ctx, m, nt, coef, xs, w, cd->bc_bits, 1, cd->rhs));
    // solve!
    sem_cd::Op* op = nullptr;
    SEM_TRY(cd_op(ctx, cd, b[0], &op));
    sem_pcg_info info{};
    const int maxit = cd->maxiter > 0 ? cd->maxiter : (int)std::min<int64_t>(n, 2147483647);
    const int ps = sem_pcg(ctx, op->op, SEM_PRECOND_DIAG, op->dinv, cd->rhs, u, -cd->tol, maxit,
                           &info);
    if (ps < 0) return ps;
    if (ps == SEM_WARN_MAXITER) status = SEM_WARN_MAXITER;
    iters += info.iters;
    if (J_dev) {
      SEM_TRY(ensure_ws(ctx, sizeof(double) * 256));
      SEM_TRY(launch_dot_acc(st, n, u, u, m->f[SEM_B], 0.5 * cd->dt,
                             static_cast<double*>(ctx->ws), 256, J_dev));
    }
  }
  if (iters_total) *iters_total = iters;
  return status;
}

}  // extern "C"

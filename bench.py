"""Benchmark of the SEM operator-apply-and-solve path on one MI355X.

Headline workload: 256x256 quadrilateral elements, N=7 (8x8 GLL nodes per
element), deformed mesh x = r + 0.1 sin(pi r) sin(pi s), y = s + 0.1 sin(pi r)
sin(pi s), Helmholtz (-Lap + 1) u with homogeneous Dirichlet data, FP64. A
step is one operator apply y = M .* QQ'(A u + B u) over the whole mesh
(4,194,304 local / 3,214,849 global nodes). ``value`` is global DOF/s
(unique nodes x applies / s). The same line carries the roofline of the
apply, the FP64 Jacobi-PCG time to 1e-10 relative residual, the mixed
precision solve, and the CPU baseline (the oracle's reference-faithful apply
with the dense QQ' factors of mesh.jl:81-82, timed on a bounded sample).

usage: python bench.py [--gpus 1] [--steps 1000] [--warmup 20] [--no-pcg]
                       [--no-mixed] [--no-adjoint] [--no-cd] [--cd-steps 2000]
                       [--no-cpu] [--rotate 2]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "spectralelements.jl_amd"))
sys.path.insert(0, ROOT)

METRIC = "A·u+QQ^T GDOF/s & PCG time-to-1e-10 residual, FP64 vs mixed, 1 MI355X"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def global_normal_field(se, m, seed):
    """f ~ N(0,1) i.i.d. per global node, scattered to every local copy."""
    return se.workloads.global_normal_field(m, seed)


def bench_apply(se, args):
    import torch
    n, E = 8, 256
    sets = []
    for r in range(args.rotate):
        m = se.Mesh(n, n, E, E, (False, False), se.deform_sinsin(0.1))
        M = se.generateMask("DDDD", m)
        op = se.SEMOperator(1.0, 1.0, m, M)
        u = global_normal_field(se, m, 2 + r)
        y = se.empty_field(*m.shape)
        sets.append((m, op, u, y))
    m = sets[0][0]
    st = torch.cuda.current_stream()
    for i in range(args.warmup):
        _, op, u, y = sets[i % len(sets)]
        op(u, out=y)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(st)
    for i in range(args.steps):
        _, op, u, y = sets[i % len(sets)]
        op(u, out=y)
    e1.record(st)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    # same loop on one buffer set (working set 201 MB, can sit in the 256 MB MALL)
    _, op, u, y = sets[0]
    e0.record(st)
    for i in range(args.steps):
        op(u, out=y)
    e1.record(st)
    torch.cuda.synchronize()
    ms_one = e0.elapsed_time(e1)
    return m, ms / args.steps, ms_one / args.steps, sets


def bench_pcg(se, m, sets):
    import torch
    _, op, _, _ = sets[0]
    f = global_normal_field(se, m, 2)
    M = op.M
    b = se.gatherScatter(se.mask(se.mass(f, m), M), m)
    tol = 1e-10 * float(b.abs().max())
    d = op.jacobi()
    info = se.PCGInfo()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x = se.pcg(b, op, opM=se.Diagonal(d), mult=m.mult, tol=tol, info=info)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    r = b - op(x)
    true_res = float(r.abs().max()) / float(b.abs().max())
    out = {"iters": info.iters, "time_s": wall, "device_ms": info.elapsed_ms,
           "ms_per_iter": info.elapsed_ms / max(info.iters, 1),
           "rel_res_recursive": info.res / float(b.abs().max()),
           "rel_res_true": true_res, "converged": info.converged,
           "precond": "jacobi", "tol_rel": 1e-10}
    out["mg"] = bench_pcg_mg(se, m, op, b, tol, x, wall)
    return out


def bench_pcg_mg(se, m, op, b, tol, x_jacobi, t_jacobi, reps=3):
    """The same solve with the multigrid V-cycle as opM: hierarchy set-up
    (meshes, operators, diagonals, eigenvalue estimates) timed separately
    from the solve; the solve is the median of reps runs."""
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    P = se.MG(op)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    walls = []
    for _ in range(reps):
        info = se.PCGInfo()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x = se.pcg(b, op, opM=P, mult=m.mult, tol=tol, info=info)
        torch.cuda.synchronize()
        walls.append(time.perf_counter() - t0)
    wall = sorted(walls)[len(walls) // 2]
    bmax = float(b.abs().max())
    # the V-cycle in FP32 (the PCG in FP64, flexible beta)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    P32 = se.MG(op, precision="f32")
    torch.cuda.synchronize()
    t_setup32 = time.perf_counter() - t0
    walls32 = []
    for _ in range(reps):
        info32 = se.PCGInfo()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        x32 = se.pcg(b, op, opM=P32, mult=m.mult, tol=tol, info=info32)
        torch.cuda.synchronize()
        walls32.append(time.perf_counter() - t0)
    wall32 = sorted(walls32)[len(walls32) // 2]
    fp32_cycle = {"iters": info32.iters, "setup_s": t_setup32, "solve_s": wall32,
                  "time_to_solution_s": t_setup32 + wall32,
                  "speedup_vs_fp64_cycle_solve": wall / wall32,
                  "rel_res_true": float((b - op(x32)).abs().max()) / bmax,
                  "converged": info32.converged,
                  "rel_diff_vs_fp64_cycle": float((x32 - x).abs().max()) / float(x.abs().max())}
    del P32, x32
    return {"levels": [list(lv[:3]) for lv in P.levels()], "iters": info.iters,
            "setup_s": t_setup, "solve_s": wall, "device_ms": info.elapsed_ms,
            "ms_per_iter": info.elapsed_ms / max(info.iters, 1),
            "time_to_solution_s": t_setup + wall,
            "speedup_vs_jacobi_solve": t_jacobi / wall,
            "speedup_vs_jacobi_incl_setup": t_jacobi / (t_setup + wall),
            "rel_res_true": float((b - op(x)).abs().max()) / bmax,
            "converged": info.converged,
            "rel_diff_vs_jacobi": float((x - x_jacobi).abs().max()) / float(x_jacobi.abs().max()),
            "smoother": "Chebyshev-Jacobi degree 3, range 20, coarsest degree 8",
            "fp32_cycle": fp32_cycle}


# (contraction variant, inner_tol): every variant runs at the configuration's
# inner tolerance 1e-5 and at the residual-reduction factor per FP64
# replacement that was fastest for it on this workload
MIXED_VARIANTS = (("f32", 1e-2), ("f16", 1e-2), ("bf16", 1e-2), ("f16x3", 1e-2), ("bf16x3", 1e-2))


def bench_mixed(se):
    """128x128 elements, N=15 (16x16 GLL blocks), deformed, Helmholtz: FP64
    Jacobi-PCG to 1e-10 relative residual against the mixed-precision solver
    (FP32 inner PCG with the FP32 or FP16/BF16-input MFMA operator, FP64
    residual replacement) for every contraction variant; each reports the
    true FP64 residual it reached and its error against the FP64 solve."""
    import torch
    m = se.Mesh(16, 16, 128, 128, (False, False), se.deform_sinsin(0.1))
    M = se.generateMask("DDDD", m)
    op64 = se.SEMOperator(1.0, 1.0, m, M)
    f = global_normal_field(se, m, 3)
    b = se.gatherScatter(se.mask(se.mass(f, m), M), m)
    bmax = float(b.abs().max())
    tol = 1e-10 * bmax
    d = op64.jacobi()
    out = {"workload": "Helmholtz (nu=1, k=1), 128x128 elements, N=15, deformed, "
                       "f~N(0,1) per global node seed 3, Jacobi, tol 1e-10 relative"}
    info = se.PCGInfo()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    x64 = se.pcg(b, op64, opM=se.Diagonal(d), mult=m.mult, tol=tol, info=info)
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t0
    # reference solution: FP64 multigrid-preconditioned PCG to 1e-13 relative
    # residual (the FP64 GPU solvers match the oracle in the parity tests)
    xref = se.pcg(b, op64, opM=se.MG(op64), mult=m.mult, tol=1e-13 * bmax)
    xrmax = float(xref.abs().max())
    out["fp64"] = {"iters": info.iters, "time_s": t64,
                   "us_per_iter": info.elapsed_ms / max(info.iters, 1) * 1e3,
                   "rel_res_true": float((b - op64(x64)).abs().max()) / bmax,
                   "rel_err_vs_ref": float((x64 - xref).abs().max()) / xrmax}
    # FP64 with the multigrid V-cycle as opM, same tolerance (setup timed apart)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    mg = se.MG(op64)
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    info = se.PCGInfo()
    t0 = time.perf_counter()
    xmg = se.pcg(b, op64, opM=mg, mult=m.mult, tol=tol, info=info)
    torch.cuda.synchronize()
    t_mg = time.perf_counter() - t0
    out["fp64_mg"] = {"iters": info.iters, "setup_s": t_setup, "time_s": t_mg,
                      "time_to_solution_s": t_setup + t_mg, "speedup_vs_fp64": t64 / t_mg,
                      "rel_res_true": float((b - op64(xmg)).abs().max()) / bmax,
                      "rel_err_vs_ref": float((xmg - xref).abs().max()) / xrmax}
    del mg, xmg
    # mixed: the same PCG in FP64 with the V-cycle in FP32
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    mg = se.MG(op64, precision="f32")
    torch.cuda.synchronize()
    t_setup32 = time.perf_counter() - t0
    info = se.PCGInfo()
    t0 = time.perf_counter()
    xmg = se.pcg(b, op64, opM=mg, mult=m.mult, tol=tol, info=info)
    torch.cuda.synchronize()
    t_mg32 = time.perf_counter() - t0
    out["mg_fp32_cycle"] = {"iters": info.iters, "setup_s": t_setup32, "time_s": t_mg32,
                            "time_to_solution_s": t_setup32 + t_mg32,
                            "speedup_vs_fp64": t64 / t_mg32, "speedup_vs_fp64_mg": t_mg / t_mg32,
                            "rel_res_true": float((b - op64(xmg)).abs().max()) / bmax,
                            "rel_err_vs_ref": float((xmg - xref).abs().max()) / xrmax}
    del mg, xmg
    for prec, delta in MIXED_VARIANTS:
        op_lo = se.SEMOperator(1.0, 1.0, m, M, precision=prec)
        for key, dl in (("", delta), ("_inner_tol_1e-5", 1e-5)):
            info = se.PCGInfo()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            x = se.pcg_mixed(b, op64, op_lo, d=d, tol=tol, inner_tol=dl, max_outer=200,
                             info=info)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            out[prec + key] = {
                "inner_tol": dl, "converged": info.converged, "replacements": info.outer,
                "inner_iters": info.iters, "time_s": dt,
                "us_per_inner_iter": info.elapsed_ms / max(info.iters, 1) * 1e3,
                "speedup_vs_fp64": t64 / dt if info.converged else None,
                "rel_res_true": float((b - op64(x)).abs().max()) / bmax,
                "rel_err_vs_fp64": float((x - x64).abs().max()) / float(x64.abs().max()),
                "rel_err_vs_ref": float((x - xref).abs().max()) / xrmax}
        del op_lo
    return out


def calibrate_copy(nbytes=1 << 30):
    """A plain device copy of known size (nbytes read + nbytes written), run
    under the same rocprofv3 counters as the apply to scale FETCH_SIZE /
    WRITE_SIZE to bytes."""
    import torch
    n = nbytes // 8
    x = torch.randn(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    log(f"copy calibration: 5 copies of {nbytes} bytes")


def measured_traffic():
    """HBM bytes per apply from the committed rocprofv3 counter summary
    (FETCH_SIZE + WRITE_SIZE of the apply kernels, scaled by the copy
    calibration), or None when no summary is present."""
    path = os.path.join(ROOT, "profiles", "apply_256x256_N7_hbm_traffic.md")
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_apply:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None


def bench_adjoint(se, n_dirs=20):
    """Steady variable-coefficient diffusion -div(nu grad u) = 1, 128x128
    elements, N=7, FP64: nu = 1 + 0.5 tanh(16 Fourier modes, seed 41), u* the
    solution for the seed-42 coefficient, J = 1/2 sum_local B (u - u*)^2.
    Times the gradient (forward PCG + adjoint PCG + dJ/dnu kernel) and checks
    it against central differences (h = 1e-5) on n_dirs continuous random
    directions (seeds 43..)."""
    import torch
    from semhip import workloads as w
    m = se.Mesh(8, 8, 128, 128, (False, False), se.deform_sinsin(0.1))
    M = se.generateMask("DDDD", m)
    nu = w.coefficient_field(m, 41)
    f = torch.ones_like(nu)
    prob = se.VarDiffusion(m, M, f)
    info = se.PCGInfo()
    solver = se.diffusion_solver(tol=1e-13, info=info)
    ustar = se.linsolve(w.coefficient_field(m, 42), prob, solver).detach()

    def J_of(v):
        u = se.linsolve(v, prob, solver)
        e = u - ustar
        return 0.5 * torch.sum(m.B * e * e)

    v = nu.clone().requires_grad_(True)
    J_of(v).backward()                       # warm-up (operator builds, caches)
    torch.cuda.synchronize()
    v = nu.clone().requires_grad_(True)
    t0 = time.perf_counter()
    J = J_of(v)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    it_fwd = info.iters
    J.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    it_adj = info.iters
    g = v.grad
    worst = 0.0
    with torch.no_grad():
        for s in range(n_dirs):
            d = w.continuous(w.global_normal_field(m, 43 + s), m)
            h = 1e-5
            fd = (float(J_of(nu + h * d)) - float(J_of(nu - h * d))) / (2 * h)
            ad = float(torch.sum(g * d))
            worst = max(worst, abs(fd - ad) / abs(ad))
    return {"workload": "-div(nu grad u)=1, 128x128 elements, N=7, deformed, nu=1+0.5tanh(16 "
                        "Fourier modes, seed 41), u* from seed 42, J=1/2 sum B (u-u*)^2",
            "J": float(J.detach()), "forward_s": t1 - t0, "adjoint_s": t2 - t1,
            "gradient_s": t2 - t0, "iters_forward": it_fwd, "iters_adjoint": it_adj,
            "tol_rel": 1e-13, "fd_dirs": n_dirs, "fd_h": 1e-5,
            "fd_max_rel_err": worst, "fd_target": 1e-6}


def bench_cd(se, nsteps=2000, fd_dirs=1):
    """cd2d: 64x64 elements, N=9 (dealiased p=15), v = (-y, x), nu = 1e-3,
    Gaussian blob sigma 0.1 at (0.5, 0), BDF2/EXT2 (BDF1/EXT1 start), dt =
    1e-3, nsteps steps by count, FP64. Times the forward run and the adjoint
    gradient dJ/du0 of J = sum_n dt/2 sum B (u^n)^2 with sqrt(N) recompute
    checkpointing, and checks it against central differences."""
    import torch
    from semhip import workloads as w
    mV = se.Mesh(10, 10, 64, 64, (False, False), se.deform_sinsin(0.1))
    mD = se.Mesh(15, 15, 64, 64, (False, False), se.deform_sinsin(0.1))
    fld = se.Field("DDDD", mV, 2)
    cdn = se.ConvectionDiffusion("cd2d", fld, -mV.y, mV.x, se.TimeStepper(0.0, nsteps * 1e-3, 1e-3, 2),
                                 mD, nu=1e-3, tol=1e-13)
    u0 = w.gaussian_blob(mV)
    se.cd_objective(cdn, u0, 3)              # warm-up
    torch.cuda.synchronize()
    cdn.iters_total = 0
    t0 = time.perf_counter()
    J = se.cd_objective(cdn, u0, nsteps)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    it_fwd = cdn.iters_total
    t0 = time.perf_counter()
    r = se.cd_gradient(cdn, u0, nsteps)
    torch.cuda.synchronize()
    t_adj = time.perf_counter() - t0
    worst = 0.0
    for s in range(fd_dirs):
        d = w.global_normal_field(mV, 5 + s)
        h = 1e-3
        fd = (se.cd_objective(cdn, u0 + h * d, nsteps) - se.cd_objective(cdn, u0 - h * d, nsteps)) / (2 * h)
        ad = float(torch.sum(r.grad * d))
        worst = max(worst, abs(fd - ad) / abs(ad))
    return {"workload": "cd2d 64x64 elements, N=9, p_d=15, v=(-y,x), nu=1e-3, dt=1e-3, "
                        f"{nsteps} steps, blob sigma 0.1",
            "J": J, "forward_s": t_fwd, "ms_per_step": t_fwd / nsteps * 1e3,
            "pcg_iters_per_step": it_fwd / nsteps, "gradient_s": t_adj,
            "gradient_over_forward": t_adj / t_fwd, "snapshots": r.snapshots,
            "states_held": r.peak_states, "steps_forward_total": r.steps_forward,
            "adjoint_pcg_iters": r.iters_adj, "tol_rel": 1e-13, "fd_dirs": fd_dirs,
            "fd_max_rel_err": worst}


def cpu_baseline(budget_s=15.0):
    """The oracle's reference-faithful apply (ABu reshape-GEMMs and the dense
    QQtx * u * QQty' of gatherScatter.jl:13) at the headline size."""
    import numpy as np
    from oracle import sem_oracle as o
    mo = o.Mesh(8, 8, 256, 256, (False, False), o.deform_sinsin(0.1), dense_gs=True)
    Mo = o.generateMask("DDDD", mo)
    u = np.random.default_rng(2).standard_normal(mo.x.shape)
    o.opLHS(u, 1.0, 1.0, mo, Mo)
    reps, t0 = 0, time.perf_counter()
    while reps < 3 or (time.perf_counter() - t0 < budget_s and reps < 200):
        o.opLHS(u, 1.0, 1.0, mo, Mo)
        reps += 1
    dt = (time.perf_counter() - t0) / reps
    try:
        import threadpoolctl
        th = max(i.get("num_threads", 1) for i in threadpoolctl.threadpool_info())
    except Exception:
        th = os.cpu_count()
    # Jacobi-PCG on the CPU: the 4x4-element N=8 Poisson solve to 1e-10
    # relative, whole; the 256x256 N=7 Helmholtz solve timed over 3
    # iterations and scaled to the GPU solve's iteration count
    p4 = o.Mesh(9, 9, 4, 4, (False, False), o.fixU, dense_gs=True)
    M4 = o.generateMask("DDDD", p4)
    f4 = 2 * np.pi ** 2 * np.sin(np.pi * p4.x) * np.sin(np.pi * p4.y)
    b4 = p4.gs(M4 * o.mass(f4, p4))
    info4 = o.PCGInfo()
    t0 = time.perf_counter()
    o.pcg(b4, lambda v: o.opLHS(v, 1.0, 0.0, p4, M4), mult=p4.mult,
          tol=1e-10 * np.abs(b4).max(), info=info4)
    t_small = time.perf_counter() - t0
    b = mo.gs(Mo * o.mass(np.random.default_rng(3).standard_normal(mo.x.shape), mo))
    dinv = o.jacobi_diag(mo, 1.0, 1.0, Mo)
    t0 = time.perf_counter()
    o.pcg(b, lambda v: o.opLHS(v, 1.0, 1.0, mo, Mo), opM=lambda r: dinv * r, mult=mo.mult,
          tol=0.0, maxiter=3)
    t_it = (time.perf_counter() - t0) / 3
    model = ""
    try:
        with open("/proc/cpuinfo") as fh:
            model = next((ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name")),
                         "")
    except OSError:
        pass
    return {"value": mo.nglobal / dt / 1e9, "unit": "GDOF/s", "cores": th, "kind": "port",
            "cpu_model": model,
            "sample": f"{reps} applies of the reference-faithful NumPy/OpenBLAS apply "
                      f"(dense QQ' GEMMs) at 256x256 N=7, {dt * 1e3:.1f} ms each",
            "pcg_poisson_4x4_N8_s": t_small, "pcg_poisson_4x4_N8_iters": info4.iters,
            "pcg_helmholtz_256x256_N7_s_per_iter": t_it}


def probe(se, kind):
    """Short fixed workloads for counter collection (rocprofv3 --pmc):
    'mfma': 11 assembled applies of each 16x16-block operator form (FP64,
    FP32, FP16, BF16, FP16 x3, BF16 x3) on 128x128 elements, N=15;
    'pcg': 100 Jacobi-PCG iterations at 256x256, N=7 (FP64) and at 128x128,
    N=15 (FP32); 'cd': 60 cd2d steps (64x64, N=9) in the library's loop;
    'mg': the multigrid hierarchy set up and the 256x256, N=7 Helmholtz solve
    to 1e-10 relative, twice."""
    import warnings
    import torch
    from semhip import workloads as W
    warnings.simplefilter("ignore")
    if kind == "mfma":
        m = se.Mesh(16, 16, 128, 128, (False, False), se.deform_sinsin(0.1))
        M = se.generateMask("DDDD", m)
        u = W.global_normal_field(m, 2)
        for prec in ["f64", "f32", "f16", "bf16", "f16x3", "bf16x3"]:
            op = se.SEMOperator(1.0, 1.0, m, M, precision=prec)
            uu = u if prec == "f64" else u.float()
            y = op(uu)
            for _ in range(10):
                op(uu, out=y)
            torch.cuda.synchronize()
            log(f"probe mfma {prec} done")
    elif kind == "pcg":
        m = se.Mesh(8, 8, 256, 256, (False, False), se.deform_sinsin(0.1))
        M = se.generateMask("DDDD", m)
        op = se.SEMOperator(1.0, 1.0, m, M)
        b = W.helmholtz_rhs(se, m, M, seed=3)
        se.pcg(b, op, opM=se.Diagonal(op.jacobi()), mult=m.mult, tol=0.0, maxiter=100)
        torch.cuda.synchronize()
        log("probe pcg fp64 done")
        m2 = se.Mesh(16, 16, 128, 128, (False, False), se.deform_sinsin(0.1))
        M2 = se.generateMask("DDDD", m2)
        op32 = se.SEMOperator(1.0, 1.0, m2, M2, precision="f32")
        op64 = se.SEMOperator(1.0, 1.0, m2, M2)
        b2 = W.helmholtz_rhs(se, m2, M2, seed=3).float()
        se.pcg(b2, op32, opM=se.Diagonal(op64.jacobi().float()), mult=m2.mult, tol=0.0,
               maxiter=100)
        torch.cuda.synchronize()
        log("probe pcg fp32 done")
    elif kind == "cd":
        mV = se.Mesh(10, 10, 64, 64, (False, False), se.deform_sinsin(0.1))
        mD = se.Mesh(15, 15, 64, 64, (False, False), se.deform_sinsin(0.1))
        fld = se.Field("DDDD", mV, 2)
        cdn = se.ConvectionDiffusion("cd2d", fld, -mV.y, mV.x,
                                     se.TimeStepper(0.0, 60e-3, 1e-3, 2), mD, nu=1e-3, tol=1e-13)
        se.cd_objective(cdn, W.gaussian_blob(mV), 60)
        torch.cuda.synchronize()
        log("probe cd done")
    elif kind == "mg":
        m = se.Mesh(8, 8, 256, 256, (False, False), se.deform_sinsin(0.1))
        M = se.generateMask("DDDD", m)
        op = se.SEMOperator(1.0, 1.0, m, M)
        b = W.helmholtz_rhs(se, m, M, seed=3)
        tol = 1e-10 * float(b.abs().max())
        for rep in range(2):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            P = se.MG(op)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            info = se.PCGInfo()
            se.pcg(b, op, opM=P, mult=m.mult, tol=tol, info=info)
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            log(f"probe mg {rep}: setup {1e3 * (t1 - t0):.1f} ms, solve {1e3 * (t2 - t1):.1f} ms, "
                f"{info.iters} iterations")
            del P


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--rotate", type=int, default=2)
    ap.add_argument("--no-pcg", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-cd", action="store_true")
    ap.add_argument("--cd-steps", type=int, default=2000)
    ap.add_argument("--probe", choices=["mfma", "pcg", "cd", "mg"],
                    help="run only a short counter-collection workload and exit")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain 1 GiB device copy (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only; multi-GPU runs are not supported (the "
              "operator would shard in y-strips of elements with one halo exchange per "
              "apply and one 3-scalar allreduce per PCG iteration)", file=sys.stderr)
        sys.exit(2)
    import torch
    import semhip as se
    torch.cuda.set_device(0)
    if args.probe:
        probe(se, args.probe)
        return
    log("apply")
    m, ms, ms_one, sets = bench_apply(se, args)
    nl, ng = m.nlocal, m.nglobal
    alg_bytes = 48 * nl
    achieved = alg_bytes / (ms * 1e-3) / 1e9
    res = {
        "metric": METRIC, "value": ng / (ms * 1e-3) / 1e9, "unit": "GDOF/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: u ~ N(0,1) i.i.d. per global node (seeds 2..), scattered to local copies",
        "config": {"workload": "FP64 Helmholtz apply y = M.*QQ'(A u + B u), 256x256 elements, "
                               "N=7, deformed sin-sin mesh, Dirichlet on all sides",
                   "elements": [256, 256], "N": 7, "n_local": nl, "n_global": ng,
                   "buffer_sets_rotated": args.rotate},
        "gdofs_local": nl / (ms * 1e-3) / 1e9,
        "ms_per_step_single_set": ms_one,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(),
                     "traffic_source": "profiles/apply_256x256_N7_hbm_traffic.md (rocprofv3 "
                                       "--pmc FETCH_SIZE, WRITE_SIZE; copy-calibrated)",
                     "kernel": "fused apply (k_ax strip kernel + k_seams), 48 B per local node"},
    }
    if args.calibrate_copy:
        calibrate_copy()
    if not args.no_pcg:
        log("pcg")
        res["pcg_fp64"] = bench_pcg(se, m, sets)
        it = res["pcg_fp64"]
        res["pcg_fp64"]["achieved_GBps_104B"] = 104 * nl / (it["ms_per_iter"] * 1e-3) / 1e9
        # bytes the two launches move per node: apply 88 B, update 32 B
        res["pcg_fp64"]["moved_GBps_120B"] = 120 * nl / (it["ms_per_iter"] * 1e-3) / 1e9
    del sets
    torch.cuda.empty_cache()
    if not args.no_mixed:
        log("mixed")
        res["mixed"] = bench_mixed(se)
    if not args.no_adjoint:
        log("adjoint (steady variable-coefficient)")
        res["adjoint"] = bench_adjoint(se)
        torch.cuda.empty_cache()
    if not args.no_cd:
        log("cd2d forward + checkpointed adjoint")
        res["cd2d"] = bench_cd(se, args.cd_steps)
        torch.cuda.empty_cache()
    if not args.no_cpu:
        log("cpu baseline")
        res["cpu_baseline"] = cpu_baseline()
        cb = res["cpu_baseline"]
        if "pcg_fp64" in res:
            cb["pcg_helmholtz_256x256_N7_extrapolated_s"] = (
                cb["pcg_helmholtz_256x256_N7_s_per_iter"] * res["pcg_fp64"]["iters"])
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()

"""The allocation-free PCG entry point (sem_pcg_noalloc) called through the
C ABI against sem_pcg on the same inputs: identical iterates and iteration
counts (same kernels and reduction order over a fixed scratch array instead
of the context workspace), on two calls in a row, the second on a smaller
mesh; an oversized mesh is rejected before any launch."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _log(msg):
    print(f"[case] {msg}", file=sys.stderr, flush=True)


@pytest.fixture(scope="module")
def se():
    import semhip
    return semhip


def _solve(se, entry, op, d, b, tol, dtype):
    import torch
    from semhip import _lib as L
    x = se.empty_field(*op.msh.shape, dtype=dtype)
    x.fill_(7.0)                                   # the solver ignores the incoming x
    pi = L.sem_pcg_info()
    st = L.call(entry, se.ctx(), op.h, L.SEM_PRECOND_DIAG, C.c_void_p(d.data_ptr()),
                C.c_void_p(b.data_ptr()), C.c_void_p(x.data_ptr()), float(tol), 100000,
                C.byref(pi))
    torch.cuda.synchronize()
    return st, x, pi


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_noalloc_matches_pcg_two_calls(se, prec):
    import torch
    from semhip import workloads as W
    dtype = torch.float64 if prec == "f64" else torch.float32
    for (n, E) in ((8, 96), (6, 40)):                     # second call: a smaller input
        _log(f"noalloc {prec} n={n} E={E}x{E}")
        m = se.Mesh(n, n, E, E, (False, False), se.deform_sinsin(0.1))
        M = se.generateMask("DDDD", m)
        op = se.SEMOperator(1.0, 1.0, m, M, precision=prec)
        b = W.helmholtz_rhs(se, m, M, seed=E).to(dtype)
        d = se.SEMOperator(1.0, 1.0, m, M).jacobi().to(dtype)
        tol = (1e-10 if prec == "f64" else 1e-5) * float(b.abs().max())
        st0, x0, i0 = _solve(se, "sem_pcg", op, d, b, tol, dtype)
        st1, x1, i1 = _solve(se, "sem_pcg_noalloc", op, d, b, tol, dtype)
        assert st0 == st1 == 0
        assert i0.converged == i1.converged == 1
        assert i0.iters == i1.iters and i0.res == i1.res
        assert torch.equal(x0, x1)


def test_noalloc_rejects_oversized_mesh(se):
    from semhip import _lib as L
    from semhip import workloads as W
    _log("noalloc oversized")
    m = se.Mesh(8, 8, 256, 256, (False, False), se.deform_sinsin(0.1))
    M = se.generateMask("DDDD", m)
    op = se.SEMOperator(1.0, 1.0, m, M)
    b = W.helmholtz_rhs(se, m, M, seed=2)
    d = op.jacobi()
    with pytest.raises(L.SemError, match="UNSUPPORTED"):
        _solve(se, "sem_pcg_noalloc", op, d, b, 1e-10, b.dtype)

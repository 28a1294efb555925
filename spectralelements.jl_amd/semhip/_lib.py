"""ctypes binding of libsemhip.so (include/sem_abi.h).

The library is required: importing the operators without it raises, there is
no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import threading
import warnings

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libsemhip.so")

SEM_OK, SEM_WARN_MAXITER, SEM_WARN_STALLED = 0, 1, 2
SEM_ERR = {-1: "SEM_ERR_ARG", -2: "SEM_ERR_HIP", -3: "SEM_ERR_UNSUPPORTED",
           -4: "SEM_ERR_NONFINITE", -5: "SEM_ERR_BREAKDOWN"}
SEM_F64, SEM_F32, SEM_F16_MFMA, SEM_BF16_MFMA = 0, 1, 2, 3
SEM_DEFORM_NONE, SEM_DEFORM_SINSIN, SEM_DEFORM_AFFINE01 = 0, 1, 2
SEM_PRECOND_IDENTITY, SEM_PRECOND_DIAG, SEM_PRECOND_DIAG_LOCAL = 0, 1, 2
SEM_PW_MUL, SEM_PW_DIV, SEM_PW_SCALAR_DIV, SEM_PW_DIV_SCALAR, SEM_PW_SCALE = 0, 1, 2, 3, 4
FIELD_IDS = {n: i for i, n in enumerate(
    ["x", "y", "Jac", "Jaci", "rx", "ry", "sx", "sy", "B", "Bi", "G11", "G12", "G22", "mult"])}


class SemError(RuntimeError):
    """A libsemhip call returned an error status."""

    def __init__(self, status: int, msg: str):
        super().__init__(f"{SEM_ERR.get(status, status)}: {msg}")
        self.status = status


class sem_op_desc(C.Structure):
    _fields_ = [("nu", C.c_double), ("k", C.c_double),
                ("nu_f", C.c_void_p), ("k_f", C.c_void_p),
                ("G11", C.c_void_p), ("G12", C.c_void_p), ("G22", C.c_void_p), ("B", C.c_void_p),
                ("bc_bits", C.c_int), ("assemble", C.c_int)]


class sem_mg_params(C.Structure):
    _fields_ = [("degree", C.c_int), ("ratio", C.c_double), ("degree_coarse", C.c_int),
                ("npow", C.c_int), ("safety", C.c_double), ("max_levels", C.c_int),
                ("precision", C.c_int)]


class sem_pcg_info(C.Structure):
    _fields_ = [("iters", C.c_int), ("converged", C.c_int), ("res", C.c_double),
                ("elapsed_ms", C.c_double), ("outer", C.c_int), ("inner_total", C.c_int)]


_vp, _i, _d, _i64 = C.c_void_p, C.c_int, C.c_double, C.c_int64
_pvp = C.POINTER(C.c_void_p)
_dp = C.POINTER(C.c_double)

# name -> argtypes (all return int unless noted)
SIGNATURES = {
    "sem_abi_version": [],
    "sem_ctx_create": [_i, _vp, _pvp],
    "sem_ctx_destroy": [_vp],
    "sem_ctx_set_stream": [_vp, _vp],
    "sem_ctx_synchronize": [_vp],
    "sem_gausslobatto": [_i, _dp, _dp],
    "sem_derivmat": [_i, _dp, _dp],
    "sem_interpmat": [_i, _dp, _i, _dp, _dp],
    "sem_semmesh": [_i, _i, _dp, _dp],
    "sem_mesh_create": [_vp, _i, _i, _i, _i, _i, _i, _i, _d, _pvp],
    "sem_mesh_set_coords": [_vp, _vp, _vp, _vp],
    "sem_mesh_destroy": [_vp],
    "sem_mesh_sizes": [_vp, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i), C.POINTER(_i),
                       C.POINTER(_i64), C.POINTER(_i64)],
    "sem_mesh_field": [_vp, _i, _pvp],
    "sem_mesh_dmat": [_vp, _pvp, _pvp],
    "sem_mesh_gll": [_vp, _pvp, _pvp, _pvp, _pvp],
    "sem_mesh_l2g": [_vp, _pvp],
    "sem_mask_bits": [_vp, C.c_char_p, C.POINTER(_i)],
    "sem_mask_fill": [_vp, _vp, _i, _vp],
    "sem_gsmap_create": [_vp, _i64, _vp, _i64, _pvp],
    "sem_gsmap_destroy": [_vp],
    "sem_gsmap_apply": [_vp, _vp, _vp, _vp],
    "sem_gather_scatter": [_vp, _vp, _vp, _vp],
    "sem_gather_scatter_f32": [_vp, _vp, _vp, _vp],
    "sem_mask": [_vp, _i64, _vp, _vp, _vp],
    "sem_mask_bc": [_vp, _vp, _i, _vp, _vp],
    "sem_mass": [_vp, _vp, _vp, _vp],
    "sem_abu": [_vp, _i, _i, _vp, _vp, _i, _i, _vp, _i, _i, _i, _vp],
    "sem_lapl": [_vp, _vp, _vp, _vp],
    "sem_hlmz": [_vp, _vp, _vp, _d, _vp, _d, _vp, _vp],
    "sem_operator_create": [_vp, _vp, C.POINTER(sem_op_desc), _i, _pvp],
    "sem_operator_destroy": [_vp],
    "sem_operator_precision": [_vp, C.POINTER(_i)],
    "sem_operator_apply": [_vp, _vp, _vp, _vp],
    "sem_operator_jacobi": [_vp, _vp, _vp],
    "sem_pcg": [_vp, _vp, _i, _vp, _vp, _vp, _d, _i, C.POINTER(sem_pcg_info)],
    "sem_pcg_noalloc": [_vp, _vp, _i, _vp, _vp, _vp, _d, _i, C.POINTER(sem_pcg_info)],
    "sem_pcg_mixed": [_vp, _vp, _vp, _vp, _vp, _vp, _d, _d, _i, _i, C.POINTER(sem_pcg_info)],
    "sem_mg_default_params": [C.POINTER(sem_mg_params)],
    "sem_mg_create": [_vp, _vp, C.POINTER(sem_mg_params), _pvp],
    "sem_mg_destroy": [_vp],
    "sem_mg_num_levels": [_vp, C.POINTER(_i)],
    "sem_mg_level_info": [_vp, _i, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i), _dp],
    "sem_mg_set_lmax": [_vp, _i, _d],
    "sem_mg_vcycle": [_vp, _vp, _vp, _vp],
    "sem_pcg_mg": [_vp, _vp, _vp, _vp, _vp, _d, _i, C.POINTER(sem_pcg_info)],
    "sem_gordon_hall": [_vp, _i, _i] + [_vp] * 12,
    "sem_interp_points": [_vp, _i, _i, _i, _i] + [_vp] * 7 + [_i, _vp, _vp, _d, _i, _vp, _vp],
    "sem_mesh_interp": [_vp, _vp, _vp, _i, _vp, _vp, _d, _i, _vp, _vp],
    "sem_wdot": [_vp, _vp, _vp, _vp, _dp],
    "sem_norm_inf": [_vp, _i64, _vp, _dp],
    "sem_dot": [_vp, _i64, _vp, _vp, _vp, _dp],
    "sem_bdfextk": [_i, _vp, _i, _vp, _vp],
    "sem_cd_create": [_vp, _vp, _vp, _i, _i, _d, _vp, _vp, _vp, _d, _d, _i, _pvp],
    "sem_cd_destroy": [_vp],
    "sem_cd_steps": [_vp, _vp, _i, _vp, _pvp, _pvp, C.POINTER(C.c_int), _vp, _vp, _vp, _vp,
                     C.POINTER(C.c_int64)],
    "sem_dot_acc": [_vp, _i64, _vp, _vp, _vp, C.c_double, _vp],
    "sem_combine": [_vp, _vp, _i, _dp, _pvp, C.POINTER(C.c_int), _i, _i, _vp],
    "sem_vardiff_factors": [_vp, _vp, _vp, _vp, _vp, _vp],
    "sem_gprod": [_vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "sem_axpby": [_vp, _i64, _d, _vp, _d, _vp, _vp],
    "sem_pointwise": [_vp, _i, _i64, _d, _vp, _vp, _vp],
    "sem_grad": [_vp, _vp, _vp, _vp, _vp],
    "sem_grad_t": [_vp, _vp, _vp, _vp, _vp],
    "sem_adjoint_dnu": [_vp, _vp, _vp, _vp, _vp],
    "sem_advect": [_vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "sem_advect_t": [_vp, _vp, _vp, _vp, _vp, _vp, _vp],
}

_lib = None
_lock = threading.Lock()


def load(path: str = None):
    """Loads libsemhip.so (the in-tree build, or the file named by the
    environment variable SEMHIP_LIB) and declares every entry point; raises
    if it is missing."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    path = path or os.environ.get("SEMHIP_LIB") or LIB_PATH
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise RuntimeError(
                f"libsemhip.so not found at {path}; build it with "
                "`python spectralelements.jl_amd/build_lib.py` (or __graft_entry__.build())")
        lib = C.CDLL(path)
        for name, args in SIGNATURES.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = C.c_int
        lib.sem_last_error.argtypes = []
        lib.sem_last_error.restype = C.c_char_p
        _lib = lib
        return lib


def check(status: int, what: str = "") -> int:
    if status < 0:
        msg = load().sem_last_error().decode(errors="replace")
        raise SemError(status, f"{what}: {msg}" if what else msg)
    return status


def call(name: str, *args) -> int:
    return check(getattr(load(), name)(*args), name)


class StallWarning(RuntimeWarning):
    """The mixed-precision solver stopped making progress before tol."""


def stall_warning(res: float):
    warnings.warn(f"pcg_mixed stalled: res:{res} (the low-precision operator is too "
                  "inaccurate for this problem's conditioning)", StallWarning, stacklevel=3)


def maxiter_warning(res: float):
    # pcg.jl:39 prints "warning: res:" and returns the current iterate
    warnings.warn(f"warning: res:{res}", RuntimeWarning, stacklevel=3)

"""ctypes binding of libgsg.so (the C ABI declared in include/gsg.h).

Loading fails loudly when the library is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIBPATH = Path(os.environ.get("GSG_LIBRARY", Path(__file__).resolve().parent / "libgsg.so"))

GSG_OK = 0
GSG_EINVAL = -1
GSG_ENOMEM = -2
GSG_EDEVICE = -3
GSG_ESTATE = -4
GSG_EFULL = -5
GSG_ENONFINITE = -6

SCHEME = {"sparse": 0, "full": 1}
KIND = {"deriv": 0, "grad_sum": 1, "laplacian": 2, "wave": 3, "transport": 4, "csr": 5}
PREC = {"fp64": 0, "fp32": 1}
MAX_D = 8

i32, i64, dbl = C.c_int32, C.c_int64, C.c_double
dp = C.POINTER(C.c_double)
ip32 = C.POINTER(C.c_int32)
ip64 = C.POINTER(C.c_int64)
vp = C.c_void_p


class OpDesc(C.Structure):
    _fields_ = [("D", i32), ("k", i32), ("n", i32), ("scheme", i32), ("kind", i32), ("dir", i32),
                ("precision", i32), ("reserved", i32), ("coef", dbl * MAX_D)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", i64), ("precision", i32), ("steps_per_tick", i32), ("max_resident", i32),
                ("reserved", i32)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", i64), ("admitted", i64), ("completed", i64), ("queued", i64), ("resident", i64),
                ("ticks", i64), ("dof_updates", dbl), ("busy_seconds", dbl), ("resident_bytes", i64)]


class VlasovDesc(C.Structure):
    _fields_ = [("Dx", i32), ("k", i32), ("n", i32), ("scheme", i32), ("precision", i32), ("reserved", i32)]


class KernelStat(C.Structure):
    _fields_ = [("D", i32), ("k", i32), ("precision", i32), ("sys", i32), ("variant", i32), ("reserved", i32),
                ("launches", i64), ("total_ms", dbl),
                ("algo_bytes", dbl), ("dofs", i64)]


class GsgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libgsg error {code}: {msg}")
        self.code = code


# (name, restype, argtypes) of every exported entry point of include/gsg.h
SIGNATURES = [
    ("gsg_last_error", C.c_char_p, []),
    ("gsg_version", C.c_char_p, []),
    ("gsg_device_count", C.c_int, []),
    ("gsg_get_size", i64, [i32, i32, i32, i32]),
    ("gsg_v2dref", C.c_int, [i32, i32, i32, i32, ip32, ip32, ip32, i64]),
    ("gsg_dg_basis", C.c_int, [i32, dp]),
    ("gsg_dlf_1d", C.c_int, [i32, i32, dp, i64]),
    ("gsg_d_matrix_csc", C.c_int, [i32, i32, i32, i32, i32, ip64, ip64, dp, ip64]),
    ("gsg_transform_1d", C.c_int, [i32, i32, i32, i32, dp, i64]),
    ("gsg_points_1d", C.c_int, [i32, i32, dp, i64]),
    ("gsg_tune_set", C.c_int, [i32, i32, i32, i32, i32]),
    ("gsg_tune_get", C.c_int, [i32, i32, i32, i32]),
    ("gsg_tune_lanes", C.c_int, [i32]),
    ("gsg_op_create", C.c_int, [C.POINTER(OpDesc), C.POINTER(vp)]),
    ("gsg_csc_create", C.c_int, [i64, i64, ip64, ip64, dp, i64, i32, C.POINTER(vp)]),
    ("gsg_transform_create", C.c_int, [i32, i32, i32, i32, i32, i32, C.POINTER(vp)]),
    ("gsg_transform_csc", C.c_int, [i32, i32, i32, i32, i32, i32, ip64, ip64, dp, ip64]),
    ("gsg_vlasov_create", C.c_int, [C.POINTER(VlasovDesc), dp, i64, C.POINTER(vp)]),
    ("gsg_op_len", i64, [vp]),
    ("gsg_op_apply", C.c_int, [vp, dp, dp, i64]),
    ("gsg_op_apply_device", C.c_int, [vp, vp, vp, i64, vp]),
    ("gsg_op_rk4", C.c_int, [vp, dp, i64, dbl, dbl, i64, i64, dp]),
    ("gsg_op_rk4_device", C.c_int, [vp, vp, i64, dbl, i64, vp]),
    ("gsg_op_rk4_device_noalloc", C.c_int, [vp, vp, i64, dbl, i64, vp]),
    ("gsg_op_destroy", C.c_int, [vp]),
    ("gsg_batch_create", C.c_int, [i32, C.POINTER(vp)]),
    ("gsg_batch_add", C.c_int, [vp, C.POINTER(OpDesc), dp, i64, dbl, i64, C.POINTER(i32)]),
    ("gsg_batch_run", C.c_int, [vp]),
    ("gsg_batch_step", C.c_int, [vp, i64]),
    ("gsg_batch_get", C.c_int, [vp, i32, dp, i64]),
    ("gsg_batch_dofs", i64, [vp]),
    ("gsg_batch_profile", C.c_int, [vp, i32]),
    ("gsg_batch_schedule", C.c_int, [vp, i32]),
    ("gsg_tune_schedule", C.c_int, [i32]),
    ("gsg_batch_kernel_stats", C.c_int, [vp, C.c_void_p, i32, C.POINTER(i32)]),
    ("gsg_batch_destroy", C.c_int, [vp]),
    ("gsg_server_start", C.c_int, [C.POINTER(ServerConfig), C.POINTER(vp)]),
    ("gsg_server_submit", C.c_int, [vp, C.POINTER(OpDesc), dp, i64, dbl, i64, ip64]),
    ("gsg_server_poll", C.c_int, [vp, i64, ip32]),
    ("gsg_server_wait", C.c_int, [vp, i64, dp, i64, ip64]),
    ("gsg_server_stats_get", C.c_int, [vp, C.POINTER(ServerStats)]),
    ("gsg_server_stop", C.c_int, [vp]),
    ("gsg_proj_create", C.c_int, [i32, i32, i32, i32, i32, C.POINTER(vp)]),
    ("gsg_proj_len", i64, [vp]),
    ("gsg_proj_cells", i64, [vp]),
    ("gsg_proj_points_per_cell", i64, [vp]),
    ("gsg_proj_nodes", C.c_int, [vp, i64, i64, vp, i64, vp]),
    ("gsg_proj_nodes_1d", C.c_int, [vp, i64, i64, vp, i64, vp]),
    ("gsg_proj_contract", C.c_int, [vp, i64, i64, vp, i64, vp, i64, vp]),
    ("gsg_proj_tensor", C.c_int, [vp, C.POINTER(vp), vp, i64, vp]),
    ("gsg_proj_destroy", C.c_int, [vp]),
]

_lib = None


def lib():
    """The loaded libgsg (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIBPATH.exists():
            raise ImportError(f"{LIBPATH} is missing: build it with __graft_entry__.build() "
                              "(hipcc --offload-arch=gfx950); there is no CPU fallback")
        L = C.CDLL(str(LIBPATH))
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    """Raise GsgError for a negative status code."""
    if rc < 0:
        raise GsgError(rc, lib().gsg_last_error().decode(errors="replace"))
    return rc


def dptr(a: np.ndarray):
    return a.ctypes.data_as(dp)


def f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


def make_desc(D, k, n, kind, scheme="sparse", precision="fp64", dir=0, coef=None) -> OpDesc:
    if scheme not in SCHEME:
        raise ValueError(f"scheme must be 'sparse' or 'full', got {scheme!r}")
    if precision not in PREC:
        raise ValueError(f"precision must be 'fp64' or 'fp32', got {precision!r}")
    d = OpDesc()
    d.D, d.k, d.n = int(D), int(k), int(n)
    d.scheme = SCHEME[scheme]
    d.kind = KIND[kind]
    d.dir = int(dir)
    d.precision = PREC[precision]
    if coef is not None:
        c = list(coef)
        if len(c) != D:
            raise ValueError(f"need {D} coefficients, got {len(c)}")
        for i, v in enumerate(c):
            d.coef[i] = float(v)
    return d

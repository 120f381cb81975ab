"""ctypes wrapper of the FP64 CPU oracle (oracle/gsg_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.

Each function names the reference function it restates (file:line in
GalerkinSparseGrids.jl).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_LIBPATH = _HERE / "libgsg_oracle.so"
_lib = None

i64 = C.c_int64
dp = C.POINTER(C.c_double)
ip32 = C.POINTER(C.c_int32)
ip64 = C.POINTER(C.c_int64)


def build(force: bool = False) -> Path:
    """Compile the oracle with gcc (oracle/Makefile)."""
    if force or not _LIBPATH.exists() or _LIBPATH.stat().st_mtime < (_HERE / "gsg_oracle.c").stat().st_mtime:
        subprocess.run(["make", "-s", "-B" if force else "-s", "libgsg_oracle.so"], cwd=_HERE, check=True)
    return _LIBPATH


def lib():
    global _lib
    if _lib is None:
        if not _LIBPATH.exists():
            build()
        L = C.CDLL(str(_LIBPATH))
        L.gsgo_dg_basis.argtypes = [C.c_int, dp]
        L.gsgo_legendre.argtypes = [C.c_int, dp]
        L.gsgo_hier2pos.argtypes = [C.c_int, C.c_int, dp]
        L.gsgo_pos_dlf.argtypes = [C.c_int, C.c_int, dp]
        L.gsgo_hier_dlf.argtypes = [C.c_int, C.c_int, dp]
        L.gsgo_get_size.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
        L.gsgo_get_size.restype = i64
        L.gsgo_v2d.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, ip32, ip32, ip32]
        L.gsgo_d2v_index.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, ip32, ip32, ip32, i64, ip64]
        L.gsgo_eval_v.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, dp, dp, i64]
        L.gsgo_eval_basis.argtypes = [C.c_int, C.c_int, C.c_int, dp, dp, i64]
        L.gsgo_d_matrix.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.gsgo_d_matrix.restype = C.c_void_p
        L.gsgo_laplacian.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
        L.gsgo_laplacian.restype = C.c_void_p
        L.gsgo_wave_rhs.argtypes = [C.c_void_p]
        L.gsgo_wave_rhs.restype = C.c_void_p
        L.gsgo_csc_free.argtypes = [C.c_void_p]
        L.gsgo_csc_nnz.argtypes = [C.c_void_p]
        L.gsgo_csc_nnz.restype = i64
        L.gsgo_csc_rows.argtypes = [C.c_void_p]
        L.gsgo_csc_rows.restype = i64
        L.gsgo_csc_export.argtypes = [C.c_void_p, ip64, ip64, dp]
        L.gsgo_spmv.argtypes = [C.c_void_p, dp, dp]
        L.gsgo_sys_wave.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
        L.gsgo_sys_wave.restype = C.c_void_p
        L.gsgo_sys_transport.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, dp]
        L.gsgo_sys_transport.restype = C.c_void_p
        L.gsgo_sys_len.argtypes = [C.c_void_p]
        L.gsgo_sys_len.restype = i64
        L.gsgo_sys_rhs.argtypes = [C.c_void_p, dp, dp]
        L.gsgo_sys_rk4.argtypes = [C.c_void_p, dp, C.c_double, i64, C.c_int]
        L.gsgo_sys_free.argtypes = [C.c_void_p]
        L.gsgo_proj1d_npts.argtypes = [C.c_int, C.c_int, C.c_int]
        L.gsgo_proj1d_npts.restype = i64
        L.gsgo_proj1d_nodes.argtypes = [C.c_int, C.c_int, C.c_int, dp]
        L.gsgo_proj1d.argtypes = [C.c_int, C.c_int, C.c_int, dp, dp]
        L.gsgo_tensor_construct.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.POINTER(dp), dp]
        L.gsgo_energy.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, dp]
        L.gsgo_energy.restype = C.c_double
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(dp)


def _full(scheme):
    if scheme not in ("sparse", "full"):
        raise ValueError(f"unknown scheme {scheme!r}")
    return 1 if scheme == "full" else 0


# -- 1-D basis ---------------------------------------------------------------
def legendre(K=10):
    """dg_basis.jl:88-94 legendre(K): (K+1) x 2(K+1) coefficient rows."""
    out = np.zeros((K + 1, 2 * (K + 1)))
    lib().gsgo_legendre(K, _p(out))
    return out


def dg_basis(k):
    """dg_basis.jl:185-191 dg_basis(k): k x 2k rows (x^i part, then sgn(x)x^i part)."""
    out = np.zeros((k, 2 * k))
    if lib().gsgo_dg_basis(k, _p(out)):
        raise ValueError(k)
    return out


def eval_v(k, level, cell, mode, x):
    """dg_methods.jl:27-36 v(k, level, cell, mode, x)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    lib().gsgo_eval_v(k, level, cell, mode, _p(x), _p(y), x.size)
    return y


def eval_basis(level, cell, mode, x):
    """1d_dg_functions.jl:91-93 basis(level, cell, mode, x)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    lib().gsgo_eval_basis(level, cell, mode, _p(x), _p(y), x.size)
    return y


def hier2pos(k, n):
    """1d_dg_functions.jl:242-263 hier2pos(k, n) as a dense matrix."""
    N = k << n
    Q = np.zeros((N, N))
    lib().gsgo_hier2pos(k, n, _p(Q))
    return Q


def pos_dlf(k, n):
    """1d_derivative.jl:108-111 periodic_pos_DLF_matrix(k, n) (already transposed)."""
    N = k << n
    A = np.zeros((N, N))
    lib().gsgo_pos_dlf(k, n, _p(A))
    return A


def hier_dlf(k, n):
    """1d_derivative.jl:113-117 periodic_hier_DLF_matrix(k, n) = Q'(AQ)."""
    N = k << n
    Dm = np.zeros((N, N))
    if lib().gsgo_hier_dlf(k, n, _p(Dm)):
        raise ValueError((k, n))
    return Dm


# -- layout ------------------------------------------------------------------
def get_size(D, k, n, scheme="sparse"):
    """dg_vmethods.jl:35-45 get_size."""
    return int(lib().gsgo_get_size(D, k, n, _full(scheme)))


def v2dref(D, k, n, scheme="sparse"):
    """dg_vmethods.jl:123-142 V2Dref: (levels, cells, modes), each N x D, 1-based."""
    N = get_size(D, k, n, scheme)
    lev = np.zeros((N, D), np.int32)
    cel = np.zeros((N, D), np.int32)
    mod = np.zeros((N, D), np.int32)
    lib().gsgo_v2d(D, k, n, _full(scheme), lev.ctypes.data_as(ip32), cel.ctypes.data_as(ip32),
                   mod.ctypes.data_as(ip32))
    return lev, cel, mod


def d2v_index(D, k, n, lev, cel, mod, scheme="sparse"):
    """dg_vmethods.jl:102-121 D2Vref lookup (0-based result, -1 if cut)."""
    lev = np.ascontiguousarray(lev, np.int32)
    cel = np.ascontiguousarray(cel, np.int32)
    mod = np.ascontiguousarray(mod, np.int32)
    out = np.zeros(lev.shape[0], np.int64)
    lib().gsgo_d2v_index(D, k, n, _full(scheme), lev.ctypes.data_as(ip32), cel.ctypes.data_as(ip32),
                         mod.ctypes.data_as(ip32), lev.shape[0], out.ctypes.data_as(ip64))
    return out


# -- operators -----------------------------------------------------------------
def _csc_to_scipy(h):
    import scipy.sparse as sp
    L = lib()
    m = L.gsgo_csc_rows(h)
    nnz = L.gsgo_csc_nnz(h)
    colptr = np.zeros(m + 1, np.int64)
    rowval = np.zeros(max(nnz, 1), np.int64)
    nzval = np.zeros(max(nnz, 1))
    L.gsgo_csc_export(h, colptr.ctypes.data_as(ip64), rowval.ctypes.data_as(ip64), _p(nzval))
    L.gsgo_csc_free(h)
    return sp.csc_matrix((nzval[:nnz], rowval[:nnz], colptr), shape=(m, m))


def d_matrix(D, d, k, n, scheme="sparse"):
    """multidim_derivative.jl:61-65 D_matrix(D, d, k, n; scheme), d 1-based -> scipy CSC."""
    h = lib().gsgo_d_matrix(D, d, k, n, _full(scheme))
    if not h:
        raise ValueError((D, d, k, n))
    return _csc_to_scipy(h)


def grad_matrix(D, k, n, scheme="sparse"):
    """multidim_derivative.jl:67-69."""
    return [d_matrix(D, d, k, n, scheme) for d in range(1, D + 1)]


def laplacian_matrix(D, k, n, scheme="sparse"):
    """multidim_derivative.jl:71-79."""
    h = lib().gsgo_laplacian(D, k, n, _full(scheme))
    return _csc_to_scipy(h)


class System:
    """Assembled CSC evolution system: wave (pdes.jl:22-70) or transport
    (-sum_d a_d D_d, the pattern of pdes.jl:179,186,191)."""

    def __init__(self, kind, D, k, n, scheme="sparse", a=None):
        L = lib()
        self.kind, self.D, self.k, self.n, self.scheme = kind, D, k, n, scheme
        if kind == "wave":
            self.h = L.gsgo_sys_wave(D, k, n, _full(scheme))
        elif kind == "transport":
            av = np.ascontiguousarray(a, np.float64)
            if av.size != D:
                raise ValueError("advection vector length must equal D")
            self.h = L.gsgo_sys_transport(D, k, n, _full(scheme), _p(av))
        else:
            raise ValueError(kind)
        if not self.h:
            raise ValueError((kind, D, k, n))
        self.len = int(L.gsgo_sys_len(self.h))

    def rhs(self, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty_like(x)
        lib().gsgo_sys_rhs(self.h, _p(x), _p(y))
        return y

    def rk4(self, y0, dt, nsteps, nthreads=1):
        y = np.array(y0, dtype=np.float64, copy=True)
        lib().gsgo_sys_rk4(self.h, _p(y), float(dt), int(nsteps), int(nthreads))
        return y

    def close(self):
        if self.h:
            lib().gsgo_sys_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# -- initial conditions ----------------------------------------------------------
def proj1d(k, n, f, panels=4):
    """vcoeffs_DG(1, k, n, f) (dg_vmethods.jl:149-179) by composite Gauss-Legendre on
    each smooth piece; f is a vectorised callable on [0,1]."""
    L = lib()
    npts = L.gsgo_proj1d_npts(k, n, panels)
    x = np.zeros(npts)
    L.gsgo_proj1d_nodes(k, n, panels, _p(x))
    fv = np.ascontiguousarray(f(x), np.float64)
    out = np.zeros(k << n)
    L.gsgo_proj1d(k, n, panels, _p(fv), _p(out))
    return out


def tensor_construct(D, k, n, coeffs_1d, scheme="sparse"):
    """tensor_construct.jl:57-63 (vector form)."""
    arrs = [np.ascontiguousarray(c, np.float64) for c in coeffs_1d]
    ptrs = (dp * D)(*[_p(a) for a in arrs])
    out = np.zeros(get_size(D, k, n, scheme))
    lib().gsgo_tensor_construct(D, k, n, _full(scheme), ptrs, _p(out))
    return out


def cos_coeffs(k, n, m, phase=0.0, A=1.0, scheme="sparse"):
    """examples/traveling_wave.jl:18-48: A cos(2 pi m.x + phase) as a sum of separable
    sine/cosine products (even numbers of sines, sign (-1)^(#sines/2))."""
    D = len(m)
    wn = 2 * np.pi * np.asarray(m, dtype=np.float64)
    sines = [proj1d(k, n, (lambda x, i=i: np.sin(wn[i] * x + (phase if i == 0 else 0.0)))) for i in range(D)]
    coss = [proj1d(k, n, (lambda x, i=i: np.cos(wn[i] * x + (phase if i == 0 else 0.0)))) for i in range(D)]
    ans = np.zeros(get_size(D, k, n, scheme))
    for code in range(1 << D):
        sc = [(code >> i) & 1 for i in range(D)]
        ns = sum(sc)
        if ns % 2 == 1:
            continue
        sign = 1.0 if ns % 4 == 0 else -1.0
        arr = [sines[i] if sc[i] else coss[i] for i in range(D)]
        ans += sign * tensor_construct(D, k, n, arr, scheme)
    return A * ans


def energy(D, k, n, y, scheme="sparse"):
    """pdes.jl:258-273 energy_func for one state [u; v]."""
    y = np.ascontiguousarray(y, np.float64)
    return float(lib().gsgo_energy(D, k, n, _full(scheme), _p(y)))

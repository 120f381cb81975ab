"""Device operators replacing the SparseMatrixCSC objects of the reference.

D_matrix / grad_matrix / laplacian_matrix (multidim_derivative.jl:61-79) and the
wave RHS of wave_data (pdes.jl:22-49) return `Operator` objects here.  `op @ x`
(or `op * x`) is the reference's `A*x`: numpy vectors go through host copies
(gsg_op_apply), float64 torch tensors on the GPU stay on the device
(gsg_op_apply_device on torch's current stream).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L


def _is_device_tensor(x):
    return hasattr(x, "data_ptr") and getattr(x, "is_cuda", False)


class Operator:
    """A linear operator on vhier vectors, resident on the GPU."""

    def __init__(self, handle, kind, D=None, k=None, n=None, scheme="sparse", precision="fp64", coef=None, dir=None):
        self._h = handle
        self.kind, self.D, self.k, self.n = kind, D, k, n
        self.scheme, self.precision = scheme, precision
        self.coef, self.dir = coef, dir
        self.len = int(L.check(L.lib().gsg_op_len(handle)))

    @property
    def shape(self):
        return (self.len, self.len)

    def apply(self, x):
        if self._h is None:
            raise ValueError("operator has been closed")
        if _is_device_tensor(x):
            import torch
            if x.dtype != torch.float64 or not x.is_contiguous() or x.numel() != self.len:
                raise ValueError("device input must be a contiguous float64 tensor of the operator's length")
            y = torch.empty_like(x)
            stream = torch.cuda.current_stream(x.device).cuda_stream
            L.check(L.lib().gsg_op_apply_device(self._h, C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()),
                                                self.len, C.c_void_p(stream)))
            return y
        xv = L.f64(x)
        if xv.shape != (self.len,):
            raise ValueError(f"vector of length {self.len} expected, got shape {xv.shape}")
        y = np.empty_like(xv)
        L.check(L.lib().gsg_op_apply(self._h, L.dptr(xv), L.dptr(y), self.len))
        return y

    __matmul__ = apply
    __mul__ = apply

    def rk4(self, y0, dt, nsteps, save_every=0):
        """Fixed-step classical RK4 of y' = op y; returns the final state, and the
        snapshots every `save_every` steps when save_every > 0."""
        if _is_device_tensor(y0):
            import torch
            if y0.dtype != torch.float64 or not y0.is_contiguous() or y0.numel() != self.len:
                raise ValueError("device state must be a contiguous float64 tensor of the operator's length")
            stream = torch.cuda.current_stream(y0.device).cuda_stream
            L.check(L.lib().gsg_op_rk4_device(self._h, C.c_void_p(y0.data_ptr()), self.len, float(dt), int(nsteps),
                                              C.c_void_p(stream)))
            return y0
        y = np.array(L.f64(y0), copy=True)
        if y.shape != (self.len,):
            raise ValueError(f"state of length {self.len} expected")
        if save_every > 0:
            nsnap = int(nsteps) // int(save_every)
            snaps = np.zeros((max(nsnap, 1), self.len))
            L.check(L.lib().gsg_op_rk4(self._h, L.dptr(y), self.len, 0.0, float(dt), int(nsteps), int(save_every),
                                       L.dptr(snaps)))
            return y, snaps[:nsnap]
        L.check(L.lib().gsg_op_rk4(self._h, L.dptr(y), self.len, 0.0, float(dt), int(nsteps), 0, None))
        return y

    def close(self):
        if self._h is not None:
            L.lib().gsg_op_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __repr__(self):
        return f"Operator({self.kind}, D={self.D}, k={self.k}, n={self.n}, {self.scheme}, {self.precision}, len={self.len})"


def _create(desc, **meta):
    h = L.vp()
    L.check(L.lib().gsg_op_create(C.byref(desc), C.byref(h)))
    return Operator(h, **meta)


def D_matrix(D, d, k, n, scheme="sparse", precision="fp64"):
    """multidim_derivative.jl:61-65 D_matrix(D, d, k, n; scheme); d is 1-based."""
    if not 1 <= d <= D:
        raise ValueError("d must be in 1..D")
    desc = L.make_desc(D, k, n, "deriv", scheme, precision, dir=d)
    return _create(desc, kind="deriv", D=D, k=k, n=n, scheme=scheme, precision=precision, dir=d)


def grad_matrix(D, k, n, scheme="sparse", precision="fp64"):
    """multidim_derivative.jl:67-69."""
    return [D_matrix(D, d, k, n, scheme, precision) for d in range(1, D + 1)]


def laplacian_matrix(D, k, n, scheme="sparse", precision="fp64"):
    """multidim_derivative.jl:71-79 (applied as sum_d D_d (D_d x), never assembled)."""
    desc = L.make_desc(D, k, n, "laplacian", scheme, precision)
    return _create(desc, kind="laplacian", D=D, k=k, n=n, scheme=scheme, precision=precision)


def grad_sum_matrix(D, k, n, coef, scheme="sparse", precision="fp64"):
    """sum_d coef[d] D_d (grad_matrix contracted with a vector)."""
    desc = L.make_desc(D, k, n, "grad_sum", scheme, precision, coef=coef)
    return _create(desc, kind="grad_sum", D=D, k=k, n=n, scheme=scheme, precision=precision, coef=list(coef))


def wave_matrix(D, k, n, scheme="sparse", precision="fp64"):
    """The RHS [[0, I], [L, 0]] of wave_data (pdes.jl:22-49), length 2N."""
    desc = L.make_desc(D, k, n, "wave", scheme, precision)
    return _create(desc, kind="wave", D=D, k=k, n=n, scheme=scheme, precision=precision)


def transport_matrix(D, k, n, a, scheme="sparse", precision="fp64"):
    """-sum_d a_d D_d: the transport RHS u_t = -a . grad u."""
    desc = L.make_desc(D, k, n, "transport", scheme, precision, coef=a)
    return _create(desc, kind="transport", D=D, k=k, n=n, scheme=scheme, precision=precision, coef=list(a))


def csc_matrix(A, precision="fp64"):
    """Import a square sparse matrix (scipy, or a (m, n, colptr, rowval, nzval)
    1-based triplet as vlasov_evolve writes it, pdes.jl:149-155)."""
    if isinstance(A, tuple):
        m, n, colptr, rowval, nzval = A
        colptr = np.ascontiguousarray(colptr, np.int64)
        rowval = np.ascontiguousarray(rowval, np.int64)
        nzval = L.f64(nzval)
    else:
        import scipy.sparse as sp
        A = sp.csc_matrix(A)
        A.sort_indices()
        m, n = A.shape
        colptr = np.ascontiguousarray(A.indptr, np.int64) + 1
        rowval = np.ascontiguousarray(A.indices, np.int64) + 1
        nzval = L.f64(A.data)
    h = L.vp()
    L.check(L.lib().gsg_csc_create(int(m), int(n), colptr.ctypes.data_as(L.ip64), rowval.ctypes.data_as(L.ip64),
                                   L.dptr(nzval), int(nzval.size), L.PREC[precision], C.byref(h)))
    return Operator(h, kind="csr", precision=precision)


# -- host-side tables (no GPU needed) ---------------------------------------------
def dg_basis(k):
    """dg_basis.jl:185-191: k x 2k coefficient rows."""
    out = np.zeros((k, 2 * k))
    L.check(L.lib().gsg_dg_basis(k, L.dptr(out)))
    return out


def periodic_DLF_matrix(k, n):
    """1d_derivative.jl:136-148 periodic_DLF_matrix(k, n; basis="hier"), dense."""
    N1 = k << n
    out = np.zeros((N1, N1))
    L.check(L.lib().gsg_dlf_1d(k, n, L.dptr(out), N1 * N1))
    return out


def D_matrix_csc(D, d, k, n, scheme="sparse"):
    """D_matrix assembled as a scipy CSC matrix (the reference's return type)."""
    import scipy.sparse as sp
    nnz = np.zeros(1, np.int64)
    lib = L.lib()
    L.check(lib.gsg_d_matrix_csc(D, d, k, n, L.SCHEME[scheme], None, None, None, nnz.ctypes.data_as(L.ip64)))
    N = int(L.check(lib.gsg_get_size(D, k, n, L.SCHEME[scheme])))
    colptr = np.zeros(N + 1, np.int64)
    rowval = np.zeros(max(int(nnz[0]), 1), np.int64)
    nzval = np.zeros(max(int(nnz[0]), 1))
    L.check(lib.gsg_d_matrix_csc(D, d, k, n, L.SCHEME[scheme], colptr.ctypes.data_as(L.ip64),
                                 rowval.ctypes.data_as(L.ip64), L.dptr(nzval), nnz.ctypes.data_as(L.ip64)))
    m = int(nnz[0])
    return sp.csc_matrix((nzval[:m], rowval[:m] - 1, colptr - 1), shape=(N, N))

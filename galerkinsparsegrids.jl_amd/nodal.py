"""Nodal / point bases, the sparse-grid basis transforms and the Vlasov solver.

Mirrors 1d_nodal_basis.jl (transform_1D, the collocation points), multidim_nodal_basis.jl
(transform, make_modal2point_matrices, make_point2modal_matrices),
basic_function_exact_coeffs.jl (get_xi_point, get_r2_modal, get_r2_point) and
pdes.jl:131-227 (vlasov_evolve).

The reference assembles each D-dimensional transform as a restricted Kronecker product
(a SparseMatrixCSC).  Here a transform is a device `Operator` applied matrix-free: the
1-D transform along direction 1, then 2, ..., D.  The 1-D transforms are triangular in
level (modal -> nodal and nodal -> modal only take inputs of equal or finer level,
nodal -> points and points -> nodal only equal or coarser), so every intermediate index
of the sweep stays on the sparse grid and the sweep equals the restricted product.

The Vlasov right-hand side runs on the device as one operator (gsg_vlasov_create):
the 2 Dx directional derivatives in one launch, the transforms, the pointwise products
and the transforms back.  vlasov_evolve integrates it with fixed-step RK4 where the
reference calls ODE.jl's adaptive ode45/ode78.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib as L
from .ics import get_xi_modal, tensor_construct
from .operators import Operator

BASIS = {"modal": 0, "nodal": 1, "points": 2}
NODAL_K = (2, 3, 5)


def _check_k(k):
    if k not in NODAL_K:
        raise ValueError("the nodal basis exists for k in {2, 3, 5} (1d_nodal_basis.jl:47)")


def transform_1D(k, n, frm, to):
    """1d_nodal_basis.jl:383-414 transform_1D(k, n, from, to) for the modal, nodal and
    points bases, as a scipy CSC matrix (entries below 1e-12 dropped, as the reference's
    thresholds do)."""
    import scipy.sparse as sp
    _check_k(k)
    if (frm, to) in (("modal", "points"), ("points", "modal")):
        mid = [transform_1D(k, n, frm, "nodal"), transform_1D(k, n, "nodal", to)]
        P = (mid[1] @ mid[0]).tocsc()
        P.data[np.abs(P.data) < 1e-12] = 0.0
        P.eliminate_zeros()
        return P
    if frm not in BASIS or to not in BASIS:
        raise ValueError(f"unknown basis in {frm!r} -> {to!r}")
    N1 = k << n
    out = np.zeros((N1, N1))
    L.check(L.lib().gsg_transform_1d(k, n, BASIS[frm], BASIS[to], L.dptr(out), N1 * N1))
    return sp.csc_matrix(out)


def eval_points_1D(k, n):
    """x coordinates of the 1-D collocation points in point-index order
    (eval_points_1D, 1d_nodal_basis.jl:180-211): per cell its k - 1 nodes, then the
    cell's extra point (x = 1 at level 0, the cell centre above)."""
    _check_k(k)
    x = np.zeros(k << n)
    L.check(L.lib().gsg_points_1d(k, n, L.dptr(x), k << n))
    return x


class Composed:
    """Two device operators applied in turn: `second @ (first @ x)` (the reference's
    modal <-> points transforms, multidim_nodal_basis.jl:116-119)."""

    def __init__(self, first, second):
        self.first, self.second = first, second
        self.len = first.len

    @property
    def shape(self):
        return (self.len, self.len)

    def apply(self, x):
        return self.second.apply(self.first.apply(x))

    __matmul__ = apply
    __mul__ = apply

    def close(self):
        self.first.close()
        self.second.close()


def transform(D, k, n, frm, to, scheme="sparse"):
    """multidim_nodal_basis.jl:111-122 transform(D, k, n, from, to; scheme) as a device
    operator (FP64)."""
    _check_k(k)
    if sorted([frm, to]) == ["modal", "points"]:
        return Composed(transform(D, k, n, frm, "nodal", scheme), transform(D, k, n, "nodal", to, scheme))
    if frm not in BASIS or to not in BASIS:
        raise ValueError(f"unknown basis in {frm!r} -> {to!r}")
    h = L.vp()
    L.check(L.lib().gsg_transform_create(D, k, n, L.SCHEME[scheme], BASIS[frm], BASIS[to], C.byref(h)))
    return Operator(h, kind="transform", D=D, k=k, n=n, scheme=scheme)


def transform_csc(D, k, n, frm, to, scheme="sparse"):
    """transform(D, k, n, from, to; scheme) assembled as a scipy CSC matrix: the matrix
    the device operator of `transform` applies (the reference's restricted Kronecker
    product, multidim_nodal_basis.jl:83-122, without its 1e-12 drop of small products).
    Assembled on the host (gsg_transform_csc); modal <-> points is the product of the two
    factors, as the reference's is."""
    import scipy.sparse as sp
    _check_k(k)
    if sorted([frm, to]) == ["modal", "points"]:
        return (transform_csc(D, k, n, "nodal", to, scheme) @ transform_csc(D, k, n, frm, "nodal", scheme)).tocsc()
    if frm not in BASIS or to not in BASIS:
        raise ValueError(f"unknown basis in {frm!r} -> {to!r}")
    lib = L.lib()
    nnz = np.zeros(1, np.int64)
    args = (D, k, n, L.SCHEME[scheme], BASIS[frm], BASIS[to])
    L.check(lib.gsg_transform_csc(*args, None, None, None, nnz.ctypes.data_as(L.ip64)))
    N = int(L.check(lib.gsg_get_size(D, k, n, L.SCHEME[scheme])))
    m = int(nnz[0])
    colptr = np.zeros(N + 1, np.int64)
    rowval = np.zeros(max(m, 1), np.int64)
    nzval = np.zeros(max(m, 1))
    L.check(lib.gsg_transform_csc(*args, colptr.ctypes.data_as(L.ip64), rowval.ctypes.data_as(L.ip64), L.dptr(nzval),
                                  nnz.ctypes.data_as(L.ip64)))
    return sp.csc_matrix((nzval[:m], rowval[:m] - 1, colptr - 1), shape=(N, N))


def _csc_entries(name, A):
    """The (m, n, colptr, rowval, nzval) triplet of pdes.jl:149-155, 1-based like the
    SparseMatrixCSC fields the reference writes."""
    A = A.tocsc()
    A.sort_indices()
    return {f"{name}.m": np.int64(A.shape[0]), f"{name}.n": np.int64(A.shape[1]),
            f"{name}.colptr": A.indptr.astype(np.int64) + 1, f"{name}.rowval": A.indices.astype(np.int64) + 1,
            f"{name}.nzval": np.asarray(A.data, np.float64)}


def save_vlasov_output(path, D, k, n, times, sols, scheme="sparse"):
    """What vlasov_evolve writes to vlasov.h5 (pdes.jl:143-163, 217-223), as one .npz
    file (numpy's zip of .npy arrays; HDF5 is not available here) under the same names:
    dimensions, order, levels; m2n, n2p, p2n, n2m and Ds[1] .. Ds[2D] as
    <name>.m / .n / .colptr / .rowval / .nzval (1-based); times; f_modal.000001 ...
    (one per output time, 1-based like the reference's)."""
    from .operators import D_matrix_csc
    out = {"dimensions": np.int64(D), "order": np.int64(k), "levels": np.int64(n)}
    for name, (a, b) in (("m2n", ("modal", "nodal")), ("n2p", ("nodal", "points")), ("p2n", ("points", "nodal")),
                         ("n2m", ("nodal", "modal"))):
        out.update(_csc_entries(name, transform_csc(2 * D, k, n, a, b, scheme)))
    for d in range(1, 2 * D + 1):
        out.update(_csc_entries(f"Ds[{d}]", D_matrix_csc(2 * D, d, k, n, scheme)))
    out["times"] = np.asarray(times, np.float64)
    for i, sol in enumerate(sols, start=1):
        out[f"f_modal.{i:06d}"] = np.asarray(sol, np.float64)
    np.savez(path, **out)


def load_vlasov_output(path):
    """Read a save_vlasov_output file (no pickles): a dict of its arrays, with the
    matrices also as scipy CSC under their names ('m2n', 'Ds[1]', ...)."""
    import scipy.sparse as sp
    with np.load(path, allow_pickle=False) as z:
        out = {key: z[key] for key in z.files}
    for key in [k_ for k_ in out if k_.endswith(".colptr")]:
        name = key[:-len(".colptr")]
        m, nn = int(out[f"{name}.m"]), int(out[f"{name}.n"])
        out[name] = sp.csc_matrix((out[f"{name}.nzval"], out[f"{name}.rowval"] - 1, out[f"{name}.colptr"] - 1),
                                  shape=(m, nn))
    return out


def make_modal2point_matrices(D, k, n, scheme="sparse"):
    """multidim_nodal_basis.jl:124-129: (m2n, n2p)."""
    return transform(D, k, n, "modal", "nodal", scheme), transform(D, k, n, "nodal", "points", scheme)


def make_point2modal_matrices(D, k, n, scheme="sparse"):
    """multidim_nodal_basis.jl:131-136: (p2n, n2m)."""
    return transform(D, k, n, "points", "nodal", scheme), transform(D, k, n, "nodal", "modal", scheme)


def get_xi_point(D, i, k, n, m2n, n2p):
    """basic_function_exact_coeffs.jl:38-42: n2p * (m2n * get_xi_modal(D, i, k, n))."""
    return n2p @ (m2n @ get_xi_modal(D, i, k, n))


def get_xi2_modal(D, i, k, n):
    """basic_function_exact_coeffs.jl:20-29: coefficients of (2 x_i - 1)^2."""
    N1 = k << n
    one = np.zeros(N1)
    one[0] = 1.0
    c2 = np.zeros(N1)
    if N1 > 2:  # the reference sets vhier index 3 whatever k is
        c2[2] = 2.0 / (3.0 * math.sqrt(5.0))
    c2 += one / 3.0
    return tensor_construct(D, k, n, [c2 if d == i - 1 else one for d in range(D)])


def get_r2_modal(D, k, n):
    """basic_function_exact_coeffs.jl:32-35."""
    return sum(get_xi2_modal(D, i, k, n) for i in range(1, D + 1))


def get_r2_point(D, k, n, m2n, n2p):
    """basic_function_exact_coeffs.jl:46-53 (negative values clipped to 0)."""
    v = n2p @ (m2n @ get_r2_modal(D, k, n))
    v[v < 0] = 0.0
    return v


def vlasov_operator(D, k, n, F_point=None, scheme="sparse", precision="fp64"):
    """The right-hand side of vlasov_evolve's steprule (pdes.jl:174-192) on the 2 D
    dimensional phase space as a device operator: `op @ f_modal` is steprule(t, f_modal).
    F_point: D point-basis vectors of length get_size(2 D, k, n), or None for free
    streaming.  precision 'fp32' applies the derivative in float (transforms, products
    and the RK state stay FP64)."""
    _check_k(k)
    d = L.VlasovDesc()
    d.Dx, d.k, d.n = int(D), int(k), int(n)
    d.scheme = L.SCHEME[scheme]
    d.precision = L.PREC[precision]
    h = L.vp()
    if F_point is None or len(F_point) == 0:
        L.check(L.lib().gsg_vlasov_create(C.byref(d), None, 0, C.byref(h)))
    else:
        if len(F_point) != D:
            raise ValueError(f"F_point needs {D} vectors")
        F = np.ascontiguousarray(np.concatenate([L.f64(f) for f in F_point]))
        L.check(L.lib().gsg_vlasov_create(C.byref(d), L.dptr(F), F.size, C.byref(h)))
    return Operator(h, kind="vlasov", D=2 * D, k=k, n=n, scheme=scheme, precision=precision)


def vlasov_evolve(D, k, n, m2n, n2p, p2n, n2m, f0_modal, F_point, time0, time1, nout=2, order="rk4",
                  scheme="sparse", dt=None, precision="fp64", output=None):
    """pdes.jl:131-227 vlasov_evolve(D, k, n, m2n, n2p, p2n, n2m, f0_modal, F_point,
    time0, time1, nout; order, scheme).  Returns (times, [f_modal at each time]) at
    nout evenly spaced times, like the ODE.jl solution the reference writes out.

    The transforms are built inside the Vlasov operator; m2n ... n2m may be None or the
    operators of make_modal2point_matrices / make_point2modal_matrices (checked for
    matching sizes only).  Fixed-step RK4 with a step of at most dt (default: the
    explicit limit 0.1 * 2^-n); the reference's order "45"/"78" raise ValueError.
    output: a path to write the reference's vlasov.h5 contents to (operators, times and
    the f_modal snapshots, save_vlasov_output), or None to write nothing."""
    if order != "rk4":
        raise ValueError(f"order {order!r}: this build integrates with fixed-step RK4 (order='rk4')")
    if nout < 2:
        raise ValueError("nout must be >= 2")
    op = vlasov_operator(D, k, n, F_point, scheme, precision)
    try:
        for M in (m2n, n2p, p2n, n2m):
            if M is not None and M.len != op.len:
                raise ValueError("transform operators do not match get_size(2 D, k, n)")
        f = np.array(L.f64(f0_modal), copy=True)
        if f.size != op.len:
            raise ValueError("f0_modal does not match get_size(2 D, k, n)")
        span = (time1 - time0) / (nout - 1)
        h = 0.1 * 2.0 ** -n if dt is None else float(dt)
        steps = max(1, int(math.ceil(abs(span) / h - 1e-9)))
        times = [time0 + i * span for i in range(nout)]
        if span == 0:
            sols = [f.copy() for _ in range(nout)]
        else:
            _, snaps = op.rk4(f, span / steps, steps * (nout - 1), save_every=steps)
            sols = [L.f64(f0_modal).copy()] + [s.copy() for s in snaps]
        if output is not None:
            save_vlasov_output(output, D, k, n, times, sols, scheme)
        return times, sols
    finally:
        op.close()


__all__ = ["transform_1D", "eval_points_1D", "transform", "make_modal2point_matrices", "make_point2modal_matrices",
           "get_xi_point", "get_xi2_modal", "get_r2_modal", "get_r2_point", "vlasov_operator", "vlasov_evolve",
           "Composed", "transform_csc", "save_vlasov_output", "load_vlasov_output"]

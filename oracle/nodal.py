"""FP64 numpy restatement of the reference's nodal / point bases, the D-dimensional basis
transforms and the Vlasov right-hand side.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.

Follows 1d_nodal_basis.jl (nodal functions, collocation points, 1-D transforms),
multidim_nodal_basis.jl (the restricted Kronecker product `transform`, with its
`relevant_cell` overlap filter and atol drop) and pdes.jl:131-227 (vlasov_evolve's
steprule).  The layout, hier2pos, the position basis and the derivative matrices come
from the C oracle (oracle/gsg_oracle.c via oracle/oracle.py).

Pins (tests/test_oracle_nodal.py): the reference's own inverse-pair tests
(test/transformations.jl:16-77), point values of exactly represented polynomials
(get_xi_point, basic_function_exact_coeffs.jl:36-42), the interpolation property of the
nodal basis, the tensor-product structure of the D-dimensional transforms, and Vlasov
right-hand sides that are exact tensor products (f depending on x only or on v only).
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp

from . import oracle as O

NODAL_K = (2, 3, 5)


def threshold(M, atol):
    """additional_tools.jl:23-42: entries with |x| < atol set to zero."""
    M = np.array(M, dtype=np.float64, copy=True)
    M[np.abs(M) < atol] = 0.0
    return M


# -- nodal functions (1d_nodal_basis.jl:46-115) -------------------------------------
def lag_nodal(k, mode, x):
    """1d_nodal_basis.jl:46-68: Lagrange polynomials on the level-0 nodes; 0 outside [0,1]."""
    if x < 0 or x > 1:
        return 0.0
    if k == 2:
        return x if mode == 1 else 1 - x
    if k == 3:
        return (2 * (x - 0.5) * (x - 1), -4 * x * (x - 1), 2 * x * (x - 0.5))[mode - 1]
    if k == 5:
        a, b = 32.0 / 3.0, -128.0 / 3.0
        return (a * (x - 0.25) * (x - 0.5) * (x - 0.75) * (x - 1),
                b * x * (x - 0.5) * (x - 0.75) * (x - 1),
                64 * x * (x - 0.25) * (x - 0.75) * (x - 1),
                b * x * (x - 0.25) * (x - 0.5) * (x - 1),
                a * x * (x - 0.25) * (x - 0.5) * (x - 0.75))[mode - 1]
    raise ValueError("k must be 2, 3 or 5")


def h_nodal(k, mode, x):
    """1d_nodal_basis.jl:71-104: the level >= 1 functions on their reference cell [-1, 1]."""
    if k == 2:
        if mode == 1:
            return max(0.0, 1 - abs(x))
        if abs(x) > 1 or x == 0:
            return 0.0
        return x + 1 if x < 0 else 0.0
    if k == 3:
        if mode == 1:
            return max(0.0, -4 * x * (x + 1))
        if mode == 2:
            return max(0.0, -4 * x * (x - 1))
        if abs(x) > 1 or x == 0:
            return 0.0
        return 2 * (x + 1) * (x + 0.5) if x < 0 else 0.0
    if k == 5:
        b = -128.0 / 3.0
        if mode == 1:
            return 0.0 if (x < -1 or x > 0) else b * x * (x + 1) * (x + 0.5) * (x + 0.25)
        if mode == 2:
            return 0.0 if (x < -1 or x > 0) else b * x * (x + 1) * (x + 0.5) * (x + 0.75)
        if mode == 3:
            return 0.0 if (x > 1 or x < 0) else b * x * (x - 1) * (x - 0.5) * (x - 0.75)
        if mode == 4:
            return 0.0 if (x > 1 or x < 0) else b * x * (x - 1) * (x - 0.5) * (x - 0.25)
        if abs(x) > 1 or x == 0:
            return 0.0
        return 32.0 / 3.0 * (x + 1) * (x + 0.75) * (x + 0.5) * (x + 0.25) if x < 0 else 0.0
    raise ValueError("k must be 2, 3 or 5")


def v_nodal(k, level, cell, mode, x):
    """1d_nodal_basis.jl:107-113 (level 0-based, cell and mode 1-based)."""
    if level == 0:
        return lag_nodal(k, mode, x)
    return h_nodal(k, mode, (1 << level) * x - (2 * cell - 1))


def _funcs(k, n):
    """(level, cell, mode) of the 1-D nodal / modal functions in vhier order."""
    return [(l, c, m) for l in range(n + 1) for c in range(1, (1 << max(0, l - 1)) + 1) for m in range(1, k + 1)]


# -- collocation points (1d_nodal_basis.jl:180-211) -----------------------------------
def points(k, n):
    """The collocation points in point-index order as (x, side): side +1 evaluates just
    right of x (x + 1e-16), -1 just left (x - 1e-16).  Per cell: the k-1 nodes, then the
    cell's extra point (x = 1 at level 0, the cell centre above)."""
    out = []
    for l in range(n + 1):
        if l == 0:
            nodes = [j / (k - 1) for j in range(k - 1)]
        else:
            nodes = [(2 * j + 1) / (2 * (k - 1)) for j in range(k - 1)]
        nc = 1 << max(0, l - 1)
        for c in range(1, nc + 1):
            for node in nodes:
                out.append(((c - 1 + node) / nc, 1))
            out.append((1.0, -1) if l == 0 else ((c - 0.5) / nc, -1))
    return out


def points_x(k, n):
    return np.array([x for x, _ in points(k, n)])


def nodal2points_1d(k, n, atol=1e-15):
    """1d_nodal_basis.jl:215-235: P[i, j] = nodal function j at point i."""
    pts = points(k, n)
    fs = _funcs(k, n)
    P = np.zeros((len(pts), len(fs)))
    for j, (l, c, m) in enumerate(fs):
        for i, (x, side) in enumerate(pts):
            P[i, j] = v_nodal(k, l, c, m, x + 1e-16 if side > 0 else x - 1e-16)
    return threshold(P, atol)


def points2nodal_1d(k, n, atol=1e-15):
    """1d_nodal_basis.jl:240-243."""
    return threshold(np.linalg.inv(nodal2points_1d(k, n)), atol)


_GL = np.polynomial.legendre.leggauss(12)


def nodal2pos_1d(k, n, atol=1e-15):
    """1d_nodal_basis.jl:290-374: M[i, j] = integral of nodal function j times position
    basis function i (level-n Legendre cell i // k, mode i % k), over each half of the
    position cell (the integrand is a polynomial there: 12-point Gauss-Legendre is exact)."""
    fs = _funcs(k, n)
    N = k << n
    M = np.zeros((N, len(fs)))
    h = 1.0 / (1 << n)
    for j, (l, c, m) in enumerate(fs):
        nc = 1 << max(0, l - 1)
        lo, hi = (c - 1) / nc, c / nc
        for pc in range(1, (1 << n) + 1):
            a, b = (pc - 1) * h, pc * h
            if a > hi or b < lo:
                continue
            for s0, s1 in ((a, a + h / 2), (a + h / 2, b)):
                xs = s0 + (s1 - s0) * (_GL[0] + 1) / 2
                ws = _GL[1] * (s1 - s0) / 2
                fv = np.array([v_nodal(k, l, c, m, x) for x in xs])
                for pm in range(1, k + 1):
                    M[(pc - 1) * k + pm - 1, j] += np.sum(ws * fv * O.eval_basis(n, pc, pm, xs))
    M[np.abs(M) <= np.finfo(float).eps * atol] = 0.0
    return threshold(M, atol)


def pos2nodal_1d(k, n, atol=1e-12):
    """1d_nodal_basis.jl:376-379."""
    return threshold(np.linalg.inv(nodal2pos_1d(k, n, atol=atol)), atol)


def transform_1d(k, n, frm, to, atol=1e-12):
    """1d_nodal_basis.jl:383-414 transform_1D(k, n, from, to) (dense)."""
    if k not in NODAL_K:
        raise ValueError("the nodal basis exists for k in {2, 3, 5}")
    Q = O.hier2pos(k, n)
    key = (frm, to)
    if key == ("nodal", "pos"):
        return nodal2pos_1d(k, n, atol)
    if key == ("pos", "nodal"):
        return pos2nodal_1d(k, n, atol)
    if key == ("nodal", "points"):
        return nodal2points_1d(k, n, atol)
    if key == ("points", "nodal"):
        return points2nodal_1d(k, n, atol)
    if key == ("nodal", "modal"):
        return threshold(Q.T @ nodal2pos_1d(k, n), atol)
    if key == ("modal", "nodal"):
        return threshold(pos2nodal_1d(k, n) @ Q, atol)
    if key == ("modal", "pos"):
        return Q
    if key == ("pos", "modal"):
        return Q.T
    if key == ("points", "pos"):
        return threshold(nodal2pos_1d(k, n) @ points2nodal_1d(k, n), atol)
    if key == ("pos", "points"):
        return threshold(nodal2points_1d(k, n) @ pos2nodal_1d(k, n), atol)
    if key == ("modal", "points"):
        return threshold(transform_1d(k, n, "nodal", "points", atol) @ transform_1d(k, n, "modal", "nodal", atol), atol)
    if key == ("points", "modal"):
        return threshold(transform_1d(k, n, "nodal", "modal", atol) @ transform_1d(k, n, "points", "nodal", atol), atol)
    raise ValueError(f"no transform {frm} -> {to}")


# -- D-dimensional transforms (multidim_nodal_basis.jl:19-141) -------------------------
def index_1d(k, lev, cel, mod):
    """get_index_1D (additional_tools.jl:14-20) from 1-based (level, cell, mode): the
    0-based position in the 1-D vhier vector."""
    L = lev - 1
    base = np.where(L == 0, 0, 1 << np.maximum(L - 1, 0))
    return (base + cel - 1) * k + mod - 1


def _support(lev, cel):
    L = lev - 1
    nc = (1 << np.maximum(L - 1, 0)).astype(np.float64)
    return (cel - 1) / nc, cel / nc


def transform(D, k, n, mat_1d, scheme="sparse", atol=1e-12):
    """multidim_nodal_basis.jl:83-109: T[i, j] = prod_d mat_1d[i_d, j_d] over grid
    indices i, j whose cells overlap in every direction (relevant_cell, :19-36), entries
    below atol dropped.  Built column block by column block; scipy CSC."""
    lev, cel, mod = O.v2dref(D, k, n, scheme)
    N = lev.shape[0]
    idx = np.stack([index_1d(k, lev[:, d], cel[:, d], mod[:, d]) for d in range(D)], 1)
    lo, hi = _support(lev, cel)
    rows, cols, vals = [], [], []
    B = max(1, 4_000_000 // N)
    for j0 in range(0, N, B):
        j1 = min(N, j0 + B)
        V = np.ones((N, j1 - j0))
        for d in range(D):
            V *= mat_1d[idx[:, d][:, None], idx[j0:j1, d][None, :]]
            li, hi_i = lo[:, d][:, None], hi[:, d][:, None]
            lj, hj = lo[j0:j1, d][None, :], hi[j0:j1, d][None, :]
            rel = ((li <= lj) & (hi_i >= hj)) | ((li >= lj) & (hi_i <= hj))
            V *= rel
        r, c = np.nonzero(np.abs(V) >= atol)
        rows.append(r)
        cols.append(c + j0)
        vals.append(V[r, c])
    return sp.csc_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(N, N))


def transform_dd(D, k, n, frm, to, scheme="sparse", atol=1e-12):
    """multidim_nodal_basis.jl:111-122 transform(D, k, n, from, to)."""
    if sorted([frm, to]) == ["modal", "points"]:
        T1 = transform_dd(D, k, n, frm, "nodal", scheme, atol)
        T2 = transform_dd(D, k, n, "nodal", to, scheme, atol)
        P = (T2 @ T1).tocsc()
        P.data[np.abs(P.data) < atol] = 0.0
        P.eliminate_zeros()
        return P
    return transform(D, k, n, transform_1d(k, n, frm, to, atol), scheme, atol)


def make_modal2point_matrices(D, k, n, scheme="sparse"):
    """multidim_nodal_basis.jl:124-129."""
    return transform_dd(D, k, n, "modal", "nodal", scheme), transform_dd(D, k, n, "nodal", "points", scheme)


def make_point2modal_matrices(D, k, n, scheme="sparse"):
    """multidim_nodal_basis.jl:131-136."""
    return transform_dd(D, k, n, "points", "nodal", scheme), transform_dd(D, k, n, "nodal", "modal", scheme)


def xi_modal_1d(k, n):
    """get_xi_modal(1, 1, k, n) (basic_function_exact_coeffs.jl:11-17): 1/sqrt(3) on the
    second level-0 mode."""
    v = np.zeros(k << n)
    v[1] = 1.0 / np.sqrt(3.0)
    return v


def one_modal_1d(k, n):
    v = np.zeros(k << n)
    v[0] = 1.0
    return v


class Vlasov:
    """pdes.jl:131-227 vlasov_evolve's steprule on the 2 Dx-dimensional phase space
    (x_1..x_Dx, v_1..v_Dx): f' = n2m p2n (-sum_d v_d .* n2p m2n D_d f
    + sum_d F_d .* n2p m2n D_{Dx+d} f), the operators assembled as the reference does."""

    def __init__(self, Dx, k, n, F_point=None, scheme="sparse"):
        D = 2 * Dx
        self.Dx, self.k, self.n, self.scheme = Dx, k, n, scheme
        self.Ds = O.grad_matrix(D, k, n, scheme)
        self.m2n, self.n2p = make_modal2point_matrices(D, k, n, scheme)
        self.p2n, self.n2m = make_point2modal_matrices(D, k, n, scheme)
        one, xi = one_modal_1d(k, n), xi_modal_1d(k, n)
        self.v_modal = [O.tensor_construct(D, k, n, [xi if j - Dx == i else one for j in range(D)], scheme)
                        for i in range(Dx)]
        self.v_point = [self.n2p @ (self.m2n @ v) for v in self.v_modal]
        self.F_point = None if F_point is None else [np.asarray(F, np.float64) for F in F_point]
        self.N = self.Ds[0].shape[0]

    def rhs(self, f):
        Dx = self.Dx
        dx = [self.n2p @ (self.m2n @ (self.Ds[d] @ f)) for d in range(Dx)]
        c = -sum(self.v_point[d] * dx[d] for d in range(Dx))
        if self.F_point is not None:
            dp = [self.n2p @ (self.m2n @ (self.Ds[Dx + d] @ f)) for d in range(Dx)]
            c = c + sum(self.F_point[d] * dp[d] for d in range(Dx))
        return self.n2m @ (self.p2n @ c)

    def rk4(self, f0, dt, nsteps):
        """Fixed-step classical RK4 of f' = rhs(f)."""
        y = np.array(f0, np.float64, copy=True)
        for _ in range(int(nsteps)):
            k1 = self.rhs(y)
            k2 = self.rhs(y + dt / 2 * k1)
            k3 = self.rhs(y + dt / 2 * k2)
            k4 = self.rhs(y + dt * k3)
            y = y + dt / 6 * (k1 + 2 * k2 + 2 * k3 + k4)
        return y

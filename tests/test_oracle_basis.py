"""Pins of the oracle's 1-D layer against mathematics and the reference's tests.

Each check is chosen so that a plausible mistake (dropped term, wrong sign, wrong
index, transposed operand) fails it; none retypes the oracle's formula.
"""
import numpy as np
import pytest
from numpy.polynomial import legendre as npl

GL_X, GL_W = npl.leggauss(40)


def _pair_eval(row, k, x):
    """sum_i x^i (a_i + sgn(x) b_i) for a coefficient row [a_0..a_{k-1}, b_0..b_{k-1}]."""
    return np.polyval(row[:k][::-1], x) + np.sign(x) * np.polyval(row[k:][::-1], x)


def test_legendre_matches_closed_form(O):
    """dg_basis.jl:88-94 legendre(K): rows are the orthonormal Legendre polynomials
    sqrt((2m+1)/2) P_m (textbook), and carry no sgn(x) part."""
    Lc = O.legendre(10)
    for m in range(7):
        ref = npl.leg2poly([0] * m + [1]) * np.sqrt((2 * m + 1) / 2)
        np.testing.assert_allclose(Lc[m, :m + 1], ref, atol=1e-13, rtol=1e-12)
        assert np.all(Lc[m, m + 1:11] == 0.0)
        assert np.all(Lc[m, 11:] == 0.0)


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
def test_multiwavelets_orthonormal_with_vanishing_moments(O, k):
    """dg_basis(k) (dg_basis.jl:185-191): k orthonormal functions on [-1,1] orthogonal
    to all polynomials of degree < k (Alpert multiwavelets)."""
    B = O.dg_basis(k)
    G = np.zeros((k, k))
    M = np.zeros((k, k))
    for a, b in [(-1.0, 0.0), (0.0, 1.0)]:
        x = 0.5 * (a + b) + 0.5 * (b - a) * GL_X
        w = 0.5 * (b - a) * GL_W
        V = np.array([_pair_eval(B[i], k, x) for i in range(k)])
        G += (V * w) @ V.T
        M += (V * w) @ np.array([x ** p for p in range(k)]).T
    # the reference's classical Gram-Schmidt loses ~1e-12 at k = 5
    assert np.max(np.abs(G - np.eye(k))) < 1e-11
    assert np.max(np.abs(M)) < 1e-12


def test_haar_closed_form(O):
    """k = 1: the single multiwavelet is sgn(x)/sqrt(2)."""
    np.testing.assert_allclose(O.dg_basis(1), [[0.0, 1 / np.sqrt(2)]], atol=1e-16)


def test_v_level0_and_support(O):
    """dg_methods.jl:27-36: level 0 is sqrt(2) Phat_{m-1}(2x-1); level >= 1 functions
    vanish outside their cell."""
    x = np.linspace(0.01, 0.99, 57)
    for m in range(1, 4):
        ref = np.sqrt(2) * npl.legval(2 * x - 1, [0] * (m - 1) + [1]) * np.sqrt((2 * m - 1) / 2)
        np.testing.assert_allclose(O.eval_v(3, 0, 1, m, x), ref, atol=1e-13)
    vv = O.eval_v(3, 3, 2, 2, x)  # level 3, cell 2: support [1/4, 1/2]
    assert np.all(vv[(x < 0.25) | (x > 0.5)] == 0.0)
    assert np.any(vv[(x > 0.25) & (x < 0.5)] != 0.0)


@pytest.mark.parametrize("k", [2, 3, 5])
def test_hier2pos_orthogonal(O, k):
    """test/transformations.jl:45-53: norm(pos2m*m2pos - I) < 1e-15 * 10^k at n = 5."""
    Q = O.hier2pos(k, 5)
    I = np.eye(Q.shape[0])
    assert np.linalg.norm(Q.T @ Q - I) < 1e-15 * 10 ** k
    assert np.linalg.norm(Q @ Q.T - I) < 1e-15 * 10 ** k


def test_hier2pos_independent_quadrature(O):
    """hier2pos entries equal <basis_pos, v_hier> computed here from numpy Legendre
    polynomials and 40-point Gauss rules."""
    k, n = 3, 3
    Q = O.hier2pos(k, n)
    nc = 1 << n
    for j in [0, 2, 4, 7, 12, k * 6 + 1]:
        c = j // k
        mode = j % k + 1
        level = 0 if c == 0 else int(c).bit_length()
        cell = 1 if c == 0 else c - (1 << (level - 1)) + 1
        for f in range(nc):
            a, b = f / nc, (f + 1) / nc
            x = 0.5 * (a + b) + 0.5 * (b - a) * GL_X
            w = 0.5 * (b - a) * GL_W
            hv = O.eval_v(k, level, cell, mode, x)
            for p in range(1, k + 1):
                y = nc * x - f
                pos = np.sqrt(2) * npl.legval(2 * y - 1, [0] * (p - 1) + [1]) * np.sqrt((2 * p - 1) / 2) * 2 ** (n / 2)
                val = np.dot(w, pos * hv)
                if abs(val) <= 1e-12:
                    val = 0.0
                assert abs(Q[f * k + p - 1, j] - val) < 1e-13


def _exact_pos_dlf(k, n):
    """Central-flux weak derivative in the level-n Legendre basis, from closed forms:
    interior -int phi_i' phi_j and the averaged traces at both cell ends, periodic.
    Returned as the reference returns it (periodic_pos_DLF_matrix = (-Dint + LF)')."""
    nc = 1 << n
    N = k * nc
    P = lambda m, t: npl.legval(t, [0] * (m - 1) + [1]) * np.sqrt((2 * m - 1) / 2)
    scale = np.sqrt(2) * 2 ** (n / 2)
    right = np.array([scale * P(m, 1.0) for m in range(1, k + 1)])   # value at the right end
    left = np.array([scale * P(m, -1.0) for m in range(1, k + 1)])   # value at the left end
    # int_cell phi_i phi_j' dx = 2^(n+1) int_{-1}^{1} Phat_i Phat_j'
    Dint = np.zeros((k, k))
    for i in range(1, k + 1):
        for j in range(1, k + 1):
            dP = npl.legder([0] * (j - 1) + [1]) * np.sqrt((2 * j - 1) / 2)
            Dint[i - 1, j - 1] = 2 ** (n + 1) * np.dot(GL_W, P(i, GL_X) * npl.legval(GL_X, dP))
    A = np.zeros((N, N))  # A[(c1,m1),(c2,m2)] before the transpose
    for c2 in range(nc):
        for c1 in range(nc):
            # average of phi_{c1} at the left and right ends of cell c2
            LF1 = 0.5 * ((right if c1 == (c2 - 1) % nc else 0) + (left if c1 == c2 else 0))
            LF2 = 0.5 * ((right if c1 == c2 else 0) + (left if c1 == (c2 + 1) % nc else 0))
            LF1 = np.broadcast_to(LF1, (k,))
            LF2 = np.broadcast_to(LF2, (k,))
            blk = np.outer(LF2, right) - np.outer(LF1, left)
            if c1 == c2:
                blk = blk - Dint
            A[c1 * k:(c1 + 1) * k, c2 * k:(c2 + 1) * k] += blk
    return A.T


@pytest.mark.parametrize("k,n", [(1, 3), (2, 3), (3, 4), (4, 2)])
def test_pos_dlf_matches_closed_form(O, k, n):
    """periodic_pos_DLF_matrix (1d_derivative.jl:108-111) equals the closed-form
    central-flux operator up to the reference's +-5e-16 trace offsets."""
    A = O.pos_dlf(k, n)
    E = _exact_pos_dlf(k, n)
    assert np.max(np.abs(A - E)) < 1e-11 * np.max(np.abs(E))


@pytest.mark.parametrize("k,n", [(1, 4), (2, 5), (3, 5), (4, 4)])
def test_hier_dlf_skew_and_constant(O, k, n):
    """Central flux + periodic + orthonormal basis: D = -D'; D * constant = 0."""
    Dm = O.hier_dlf(k, n)
    s = np.max(np.abs(Dm))
    assert np.max(np.abs(Dm + Dm.T)) < 1e-12 * s
    assert np.max(np.abs(Dm[:, 0])) < 1e-12 * s  # column 0 = the constant function


@pytest.mark.parametrize("l", [2, 3, 4, 5])
def test_derivative_accuracy_bound(O, l):
    """test/differentiation.jl:13-25: d/dx cos(2 pi x), k = 3, squared L2 error of the
    reconstructed derivative < 2^-(k+l-2).  Computed by Parseval: the error splits into
    |D c - c'|^2 (inside the space) plus |f'|^2 - |c'|^2 (projection residual)."""
    k = 3
    Dm = O.hier_dlf(k, l)
    c = O.proj1d(k, l, lambda x: np.cos(2 * np.pi * x))
    cp = O.proj1d(k, l, lambda x: -2 * np.pi * np.sin(2 * np.pi * x))
    err2 = np.sum((Dm @ c - cp) ** 2) + (2 * np.pi ** 2 - np.sum(cp ** 2))
    assert err2 < 1.0 / (1 << (k + l - 2))
    # and the method converges: finer grids are more accurate
    if l < 5:
        Dm2 = O.hier_dlf(k, l + 1)
        c2 = O.proj1d(k, l + 1, lambda x: np.cos(2 * np.pi * x))
        cp2 = O.proj1d(k, l + 1, lambda x: -2 * np.pi * np.sin(2 * np.pi * x))
        assert np.sum((Dm2 @ c2 - cp2) ** 2) < np.sum((Dm @ c - cp) ** 2)


def _coupling_mask(k, n):
    """Independent statement of the support rule: hierarchical functions v_i, v_j
    couple iff a jump point of one lies in the closed support of the other (jump points:
    level 0 -> {0}, level >= 1 -> cell edges and midpoint, periodic)."""
    def info(i):
        c = i // k
        if c == 0:
            return 0, 0.0, 1.0, [0.0]
        l = int(c).bit_length()
        cell = c - (1 << (l - 1))
        h = 1.0 / (1 << (l - 1))
        a, b = cell * h, (cell + 1) * h
        return l, a, b, [a % 1.0, (a + b) / 2, b % 1.0]

    N = k << n
    inf = [info(i) for i in range(N)]

    def touches(p, a, b):
        return any(a - 1e-12 <= q <= b + 1e-12 for q in (p, p + 1.0, p - 1.0))

    M = np.zeros((N, N), bool)
    for i in range(N):
        _, ai, bi, ji = inf[i]
        for j in range(N):
            _, aj, bj, jj = inf[j]
            M[i, j] = any(touches(p, ai, bi) for p in jj) or any(touches(p, aj, bj) for p in ji)
    return M


@pytest.mark.parametrize("k,n", [(2, 4), (3, 4), (1, 5)])
def test_coupling_rule(O, k, n):
    """Entries outside the jump-point support rule are cancellation noise (the kernel
    skips them); the rule is symmetric."""
    Dm = O.hier_dlf(k, n)
    M = _coupling_mask(k, n)
    assert np.array_equal(M, M.T)
    assert np.max(np.abs(Dm[~M]), initial=0.0) < 5e-14 * np.max(np.abs(Dm))


def test_golden_fixtures_reproduce(O):
    """The committed oracle fixtures are reproduced bit for bit."""
    from pathlib import Path
    z = np.load(Path(__file__).parent / "golden" / "oracle_golden.npz")
    for k in range(1, 5):
        np.testing.assert_array_equal(O.dg_basis(k), z[f"dg_basis_k{k}"])
    np.testing.assert_array_equal(O.hier_dlf(2, 3), z["hier_dlf_k2_n3"])
    np.testing.assert_array_equal(O.hier_dlf(3, 2), z["hier_dlf_k3_n2"])
    np.testing.assert_array_equal(O.pos_dlf(2, 2), z["pos_dlf_k2_n2"])
    np.testing.assert_array_equal(O.hier2pos(2, 2), z["hier2pos_k2_n2"])

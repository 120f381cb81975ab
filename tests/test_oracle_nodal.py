"""The nodal-basis oracle (oracle/nodal.py) pinned, and the library's host-side 1-D
transform tables checked against it (no GPU needed).

Pins: the reference's own inverse-pair tests (test/transformations.jl:16-77); point
values of polynomials the DG space holds exactly (the identity behind get_xi_point,
basic_function_exact_coeffs.jl:36-42); the tensor-product structure of the
D-dimensional restricted products; and Vlasov right-hand sides that are exact tensor
products (f depending on x only, or on v only, with a force linear in x).
"""
import numpy as np
import pytest

from conftest import rel

N5 = 5


@pytest.fixture(scope="module")
def ON():
    from oracle import nodal
    return nodal


@pytest.mark.parametrize("k", [2, 3, 5])
def test_1d_inverse_pairs_reference_bounds(ON, k):
    """test/transformations.jl:16-53: |A B - I| < 1e-15 10^k for the inverse pairs."""
    I = np.eye(k << N5)
    for a, b in [("points", "nodal"), ("points", "modal"), ("nodal", "modal")]:
        T1 = ON.transform_1d(k, N5, a, b)
        T2 = ON.transform_1d(k, N5, b, a)
        assert np.linalg.norm(T1 @ T2 - I) < 1e-15 * 10 ** k
        assert np.linalg.norm(T2 @ T1 - I) < 1e-15 * 10 ** k
    Q = ON.transform_1d(k, N5, "modal", "pos")
    assert np.linalg.norm(Q.T @ Q - I) < 1e-15 * 10 ** k


def test_2d_inverse_pairs_reference_bounds(ON):
    """test/transformations.jl:56-77 (D = 2, k = 3, n = 5): < 1e-10."""
    D, k, n = 2, 3, 5
    for a, b in [("points", "nodal"), ("points", "modal"), ("nodal", "modal")]:
        T1 = ON.transform_dd(D, k, n, a, b)
        T2 = ON.transform_dd(D, k, n, b, a)
        I = np.eye(T1.shape[0])
        assert np.linalg.norm((T1 @ T2).toarray() - I) < 1e-10
        assert np.linalg.norm((T2 @ T1).toarray() - I) < 1e-10


def _polys(k):
    out = [lambda x: np.ones_like(x), lambda x: 2 * x - 1]
    if k >= 3:
        out.append(lambda x: (2 * x - 1) ** 2 - 0.3 * x)
    if k >= 5:
        out.append(lambda x: x ** 4 - 2 * x ** 3 + 0.5)
    return out


@pytest.mark.parametrize("k", [2, 3, 5])
@pytest.mark.parametrize("n", [0, 2, 4])
def test_1d_modal_to_points_evaluates_polynomials(ON, O, k, n):
    """m2p applied to the modal coefficients of a polynomial of degree < k gives its
    values at the collocation points (which pins the point order and both transforms),
    and p2m takes the values back."""
    x = ON.points_x(k, n)
    m2p = ON.transform_1d(k, n, "modal", "points")
    p2m = ON.transform_1d(k, n, "points", "modal")
    for f in _polys(k):
        c = O.proj1d(k, n, f)
        np.testing.assert_allclose(m2p @ c, f(x), rtol=0, atol=1e-11)
        np.testing.assert_allclose(p2m @ f(x), c, rtol=0, atol=1e-11)


@pytest.mark.parametrize("k", [2, 3, 5])
def test_xi_point_is_2x_minus_1(ON, k):
    """get_xi_point(1, 1, k, n) = n2p m2n get_xi_modal = 2 x - 1 at the points."""
    n = 4
    v = ON.transform_1d(k, n, "nodal", "points") @ (ON.transform_1d(k, n, "modal", "nodal") @ ON.xi_modal_1d(k, n))
    np.testing.assert_allclose(v, 2 * ON.points_x(k, n) - 1, rtol=0, atol=1e-12)


def test_nodal_basis_interpolates_at_level0_nodes(ON):
    """The level-0 nodal functions are the Lagrange polynomials of the level-0 nodes:
    1 at their own node, 0 at the others."""
    for k in (2, 3, 5):
        nodes = np.linspace(0, 1, k)
        M = np.array([[ON.lag_nodal(k, m, x) for x in nodes] for m in range(1, k + 1)])
        assert np.allclose(np.sort(M, axis=None)[-k:], 1.0) and np.isclose(np.abs(M).sum(), k)


# The collocation points of eval_points_1D (1d_nodal_basis.jl:180-211), written out by
# hand from the reference's description (1d_nodal_basis.jl:28-38) as (x, side): +1 is
# evaluated just right of x, -1 just left.  Per cell: its k - 1 nodes in increasing x,
# then the extra point (x = 1 at level 0, the cell centre above, both from the left).
_POINTS_BY_HAND = {
    (2, 2): [(0, 1), (1, -1), (.5, 1), (.5, -1), (.25, 1), (.25, -1), (.75, 1), (.75, -1)],
    (3, 2): [(0, 1), (.5, 1), (1, -1), (.25, 1), (.75, 1), (.5, -1),
             (.125, 1), (.375, 1), (.25, -1), (.625, 1), (.875, 1), (.75, -1)],
    (5, 1): [(0, 1), (.25, 1), (.5, 1), (.75, 1), (1, -1), (.125, 1), (.375, 1), (.625, 1), (.875, 1), (.5, -1)],
}


@pytest.mark.parametrize("k,n", sorted(_POINTS_BY_HAND))
def test_collocation_points_by_hand(ON, k, n):
    """The oracle's point list (the point-basis wire format of F_point) equals the list
    written out by hand."""
    assert [(float(x), int(sd)) for x, sd in ON.points(k, n)] == [(float(x), sd) for x, sd in _POINTS_BY_HAND[(k, n)]]


@pytest.mark.parametrize("k,n", [(2, 4), (3, 4), (5, 3)])
def test_collocation_points_cover_the_fine_grid(ON, k, n):
    """The right-evaluated points of levels 0..n are the (k - 1) 2^n multiples of
    1 / ((k - 1) 2^n) in [0, 1), each once; the left-evaluated ones are x = 1 and the
    cell centres of every level >= 1 (the jump points of the level's last mode)."""
    pts = ON.points(k, n)
    right = sorted(x for x, sd in pts if sd > 0)
    m = (k - 1) << n
    np.testing.assert_array_equal(right, np.arange(m) / m)
    left = sorted(x for x, sd in pts if sd < 0)
    centres = [1.0] + [(2 * c + 1) / (2 << l) for l in range(n) for c in range(1 << l)]
    np.testing.assert_array_equal(left, sorted(centres))


@pytest.mark.parametrize("k,n", [(2, 3), (3, 3), (5, 2)])
def test_nodal_basis_hierarchical_interpolation(ON, k, n):
    """1d_nodal_basis.jl:69-113 with the points of :180-211.  A nodal function of level
    l >= 1 (cell c, mode m < k) is 1 at the m-th node of its own cell and 0 at the
    cell's other nodes; the last mode is 1 just left of the cell centre and 0 at the
    nodes.  Every function vanishes at the other cells of its level and at every point
    of a coarser level, so nodal2points is block lower triangular in level.  Within a
    cell the block (points x modes) is the identity, except that for k = 2 the single
    node is the centre, where the hat (mode 1) is also 1.  The level-0 block is the
    identity for k = 3, 5 and the swap for k = 2, whose mode 1 is x and mode 2 is 1 - x
    (1d_nodal_basis.jl:51-52) while the points are 0 and then 1."""
    P = ON.nodal2points_1d(k, n)
    blocks = [(0, 1)] + [(l, c) for l in range(1, n + 1) for c in range(1, (1 << (l - 1)) + 1)]
    for b, (l, c) in enumerate(blocks):
        E = np.eye(k)
        if k == 2:
            E = np.array([[0.0, 1.0], [1.0, 0.0]]) if l == 0 else np.array([[1.0, 0.0], [1.0, 1.0]])
        first = next(i for i, (ll, _) in enumerate(blocks) if ll == l)
        last = max(i for i, (ll, _) in enumerate(blocks) if ll == l)
        others = np.r_[0:b * k, (b + 1) * k:(last + 1) * k]
        for m in range(k):
            col = P[:, b * k + m]
            # points are evaluated at x -+ 1e-16, which leaves residues up to 1.8e-15
            # (above the reference's 1e-15 drop) where a function vanishes exactly
            np.testing.assert_allclose(col[b * k:(b + 1) * k], E[:, m], rtol=0, atol=1e-14)
            assert np.max(np.abs(col[others]), initial=0.0) < 1e-14, (l, c, m)
        assert first <= b <= last


@pytest.mark.parametrize("D,k,n,scheme", [(2, 3, 3, "sparse"), (3, 2, 3, "sparse"), (2, 5, 2, "full")])
def test_dd_transform_of_tensor_polynomial(ON, O, D, k, n, scheme):
    """The D-dimensional m2p of a product of polynomials of degree < k is the product of
    their values at each grid point's per-direction collocation points (pins the
    restricted product's indexing and its overlap filter)."""
    lev, cel, mod = O.v2dref(D, k, n, scheme)
    idx = np.stack([ON.index_1d(k, lev[:, d], cel[:, d], mod[:, d]) for d in range(D)], 1)
    x1 = ON.points_x(k, n)
    fs = _polys(k)
    pick = [fs[(d + 1) % len(fs)] for d in range(D)]
    c = O.tensor_construct(D, k, n, [O.proj1d(k, n, f) for f in pick], scheme)
    want = np.prod([pick[d](x1[idx[:, d]]) for d in range(D)], axis=0)
    m2n = ON.transform_dd(D, k, n, "modal", "nodal", scheme)
    n2p = ON.transform_dd(D, k, n, "nodal", "points", scheme)
    np.testing.assert_allclose(n2p @ (m2n @ c), want, rtol=0, atol=1e-10)
    p2n = ON.transform_dd(D, k, n, "points", "nodal", scheme)
    n2m = ON.transform_dd(D, k, n, "nodal", "modal", scheme)
    np.testing.assert_allclose(n2m @ (p2n @ want), c, rtol=0, atol=1e-10)


def test_vlasov_rhs_exact_tensor_cases(ON, O):
    """f = g(x): f' = -(2v - 1) dg/dx exactly, i.e. -tensor(D g, xi); with f = h(v)
    and the force F = 2x - 1: f' = F dh/dv = tensor(xi, D h).  D is the 1-D DG operator."""
    k, n = 3, 3
    g = O.proj1d(k, n, lambda x: np.sin(2 * np.pi * x) + 0.3 * np.cos(4 * np.pi * x))
    h = O.proj1d(k, n, lambda v: np.exp(-8 * (v - 0.5) ** 2))
    one, xi = ON.one_modal_1d(k, n), ON.xi_modal_1d(k, n)
    Dm = O.hier_dlf(k, n)
    V = ON.Vlasov(1, k, n)
    r = V.rhs(O.tensor_construct(2, k, n, [g, one]))
    assert rel(r, -O.tensor_construct(2, k, n, [Dm @ g, xi])) < 1e-11
    Fp = V.n2p @ (V.m2n @ O.tensor_construct(2, k, n, [xi, one]))
    VF = ON.Vlasov(1, k, n, F_point=[Fp])
    r = VF.rhs(O.tensor_construct(2, k, n, [one, h]))
    assert rel(r, O.tensor_construct(2, k, n, [xi, Dm @ h])) < 1e-11
    # free streaming of f = h(v) is stationary
    assert np.linalg.norm(V.rhs(O.tensor_construct(2, k, n, [one, h]))) < 1e-12


def test_oracle_planted_mistake_is_caught(ON, O):
    """A dropped minus sign in the Vlasov combination fails the exact-tensor pin."""
    k, n = 2, 3
    g = O.proj1d(k, n, lambda x: np.sin(2 * np.pi * x))
    one, xi = ON.one_modal_1d(k, n), ON.xi_modal_1d(k, n)
    V = ON.Vlasov(1, k, n)
    f = O.tensor_construct(2, k, n, [g, one])
    dx = V.n2p @ (V.m2n @ (V.Ds[0] @ f))
    wrong = V.n2m @ (V.p2n @ (V.v_point[0] * dx))
    want = -O.tensor_construct(2, k, n, [O.hier_dlf(k, n) @ g, xi])
    assert rel(wrong, want) > 1.0
    assert rel(V.rhs(f), want) < 1e-11


# -- the library's host tables (C ABI, no GPU compute) ----------------------------------
BAS = {"modal": 0, "nodal": 1, "points": 2}


@pytest.mark.parametrize("k", [2, 3, 5])
@pytest.mark.parametrize("n", [0, 1, 3, 5])
def test_library_transform_1d_matches_oracle(g, ON, k, n):
    for a, b in [("modal", "nodal"), ("nodal", "modal"), ("nodal", "points"), ("points", "nodal")]:
        A = g.transform_1D(k, n, a, b).toarray()
        R = ON.transform_1d(k, n, a, b)
        assert np.array_equal(A != 0, R != 0), (a, b)
        np.testing.assert_allclose(A, R, rtol=0, atol=1e-12 * max(1.0, np.abs(R).max()))
    A = g.transform_1D(k, n, "modal", "points").toarray()
    np.testing.assert_allclose(A, ON.transform_1d(k, n, "modal", "points"), rtol=0, atol=1e-11)
    np.testing.assert_array_equal(g.eval_points_1D(k, n), ON.points_x(k, n))


def test_library_transform_argument_errors(g):
    from galerkinsparsegrids_jl_amd import _lib as L
    out = np.zeros(64)
    lib = L.lib()
    assert lib.gsg_transform_1d(4, 2, 0, 1, L.dptr(out), 16 * 16) < 0  # k = 4 has no nodal basis
    assert lib.gsg_transform_1d(2, 2, 0, 2, L.dptr(out), 64) < 0  # modal -> points is composed
    assert lib.gsg_transform_1d(2, 2, 0, 1, L.dptr(out), 63) < 0  # wrong length
    assert lib.gsg_points_1d(3, 13, L.dptr(out), 3 << 13) < 0
    with pytest.raises(ValueError):
        g.transform_1D(4, 2, "modal", "nodal")


def test_library_vlasov_and_transform_argument_errors(g):
    """Invalid descriptors are rejected before any device work (no GPU needed)."""
    import ctypes as C
    from galerkinsparsegrids_jl_amd import _lib as L
    lib = L.lib()
    h = L.vp()
    assert lib.gsg_transform_create(2, 4, 3, 0, 0, 1, C.byref(h)) < 0  # k = 4: no nodal basis
    assert lib.gsg_transform_create(2, 3, 3, 0, 0, 2, C.byref(h)) < 0  # modal -> points is composed
    assert lib.gsg_transform_create(9, 3, 3, 0, 0, 1, C.byref(h)) < 0  # D > 8
    assert lib.gsg_transform_create(2, 3, 13, 0, 0, 1, C.byref(h)) < 0  # n > 12
    d = L.VlasovDesc()
    d.Dx, d.k, d.n, d.scheme, d.precision = 5, 2, 3, 0, 0  # 2 Dx = 10 > 8
    assert lib.gsg_vlasov_create(C.byref(d), None, 0, C.byref(h)) < 0
    d.Dx, d.k = 2, 4  # k = 4
    assert lib.gsg_vlasov_create(C.byref(d), None, 0, C.byref(h)) < 0
    d.k, d.precision = 2, 7  # unknown precision
    assert lib.gsg_vlasov_create(C.byref(d), None, 0, C.byref(h)) < 0
    assert b"precision" in lib.gsg_last_error()
    with pytest.raises(ValueError):
        g.vlasov_operator(2, 4, 3)
    with pytest.raises(ValueError):
        g.transform(2, 4, 3, "modal", "nodal")


def test_library_projection_argument_errors(g):
    """gsg_proj_create validates before touching the device (no GPU needed)."""
    import ctypes as C
    from galerkinsparsegrids_jl_amd import _lib as L
    lib = L.lib()
    h = L.vp()
    assert lib.gsg_proj_create(0, 2, 3, 0, 4, C.byref(h)) < 0    # D = 0
    assert lib.gsg_proj_create(2, 11, 3, 0, 4, C.byref(h)) < 0   # k > 10
    assert lib.gsg_proj_create(2, 2, 3, 0, 0, C.byref(h)) < 0    # q = 0
    assert lib.gsg_proj_create(2, 2, 3, 5, 4, C.byref(h)) < 0    # unknown scheme
    assert lib.gsg_proj_create(8, 2, 3, 0, 8, C.byref(h)) < 0    # 16^8 points per cell
    assert b"lower q" in lib.gsg_last_error()
    assert lib.gsg_proj_len(None) < 0


@pytest.mark.parametrize("D,k,n,scheme", [(2, 3, 4, "sparse"), (3, 2, 3, "sparse"), (2, 2, 3, "full"),
                                          (3, 5, 2, "sparse"), (1, 2, 0, "sparse"), (1, 5, 4, "sparse")])
def test_library_transform_csc_matches_oracle(g, ON, D, k, n, scheme):
    """gsg_transform_csc (the matrix the device transform applies) against the oracle's
    restricted Kronecker product (multidim_nodal_basis.jl:83-122).  The oracle drops
    products below 1e-12; the export keeps them, so its pattern contains the oracle's."""
    for a, b in [("modal", "nodal"), ("nodal", "points"), ("points", "nodal"), ("nodal", "modal"),
                 ("modal", "points"), ("points", "modal")]:
        A = g.transform_csc(D, k, n, a, b, scheme)
        R = ON.transform_dd(D, k, n, a, b, scheme)
        assert A.shape == R.shape
        scale = max(1.0, abs(R).max())
        assert abs(A - R).max() < 1e-12 * scale, (a, b)
        Rp = R.copy()
        Rp.data[:] = 1.0
        Ap = A.copy()
        Ap.data[:] = 1.0
        assert (Rp - Rp.multiply(Ap)).nnz == 0, (a, b)  # every oracle entry is in the export


def test_vlasov_output_file_round_trip(g, tmp_path):
    """save_vlasov_output writes the names of vlasov.h5 (pdes.jl:145-163, 217-223): the
    dimensions, the 1-based CSC triplets of m2n, n2p, p2n, n2m and Ds[1..2D], times and
    f_modal.%06d; load_vlasov_output reads them back without pickles."""
    D, k, n = 1, 2, 3
    N = g.get_size(2 * D, k, n)
    rng = np.random.default_rng(0)
    sols = [rng.standard_normal(N) for _ in range(3)]
    path = tmp_path / "vlasov.npz"
    g.save_vlasov_output(path, D, k, n, [0.0, 0.5, 1.0], sols)
    out = g.load_vlasov_output(path)
    assert int(out["dimensions"]) == D and int(out["order"]) == k and int(out["levels"]) == n
    for name in ("m2n", "n2p", "p2n", "n2m", "Ds[1]", "Ds[2]"):
        for field in ("m", "n", "colptr", "rowval", "nzval"):
            assert f"{name}.{field}" in out
        assert out[f"{name}.colptr"][0] == 1 and out[f"{name}.rowval"].min() >= 1
    assert abs(out["m2n"] - g.transform_csc(2 * D, k, n, "modal", "nodal")).max() == 0.0
    assert abs(out["Ds[2]"] - g.D_matrix_csc(2 * D, 2, k, n)).max() == 0.0
    np.testing.assert_array_equal(out["times"], [0.0, 0.5, 1.0])
    for i, s in enumerate(sols, start=1):
        np.testing.assert_array_equal(out[f"f_modal.{i:06d}"], s)


def test_library_transform_csc_argument_errors(g):
    import ctypes as C
    from galerkinsparsegrids_jl_amd import _lib as L
    lib = L.lib()
    nnz = np.zeros(1, np.int64)
    pn = nnz.ctypes.data_as(L.ip64)
    assert lib.gsg_transform_csc(2, 4, 3, 0, 0, 1, None, None, None, pn) < 0   # k = 4: no nodal basis
    assert lib.gsg_transform_csc(2, 3, 3, 0, 0, 2, None, None, None, pn) < 0   # modal -> points is composed
    assert lib.gsg_transform_csc(9, 3, 3, 0, 0, 1, None, None, None, pn) < 0   # D > 8
    assert lib.gsg_transform_csc(2, 3, 3, 0, 0, 1, None, None, None, None) < 0  # no nnz
    assert lib.gsg_transform_csc(2, 3, 3, 0, 0, 1, None, None, None, pn) == 0
    out = np.zeros(int(nnz[0]))
    bad = nnz.copy()
    bad[0] -= 1
    assert lib.gsg_transform_csc(2, 3, 3, 0, 0, 1, None, None, L.dptr(out), bad.ctypes.data_as(L.ip64)) < 0
    assert b"nnz" in lib.gsg_last_error()
    with pytest.raises(ValueError):
        g.transform_csc(2, 4, 3, "modal", "nodal")

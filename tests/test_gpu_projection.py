"""Projection of initial data on the device (vcoeffs_DG, dg_vmethods.jl:149-179) against
the oracle.

Pins: separable f, whose coefficients are the tensor product of 1-D projections
(tensor_construct.jl:57-63 over the oracle's composite Gauss-Legendre proj1d); the
non-separable cos(2 pi m.x + phase), which examples/traveling_wave.jl:18-48 expands
exactly into separable products (the oracle's cos_coeffs); polynomials, which the
Gauss-Legendre rule integrates exactly (the constant is e_1, 2 x_i - 1 is
get_xi_modal, basic_function_exact_coeffs.jl:5-17).  Tolerance: 1e-12 relative L2
(1e-11 for the expanded cosine, whose 1-D pieces carry their own quadrature error),
with q = 12 nodes per half cell where f oscillates up to cos(4 pi x): the rule's own
error is then ~1e-19 (with q = 9 it is ~4e-13 on the level-0 halves).
At the headline grid (D = 4, k = 4, n = 9) the separable check is size-independent.
"""
import math

import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu


def _proj(g, gpu, D, k, n, f, scheme="sparse", q=None, **kw):
    y = g.vcoeffs_DG_device(D, k, n, f, scheme=scheme, q=q, device=gpu, **kw)
    import torch
    torch.cuda.synchronize()
    return y.cpu().numpy()


SEP_CASES = [(1, 4, 6, "sparse"), (2, 3, 4, "sparse"), (3, 2, 5, "sparse"), (2, 5, 3, "full"), (4, 3, 4, "sparse"),
             (5, 2, 3, "sparse"), (3, 1, 4, "sparse"), (2, 3, 0, "sparse"), (3, 2, 2, "full"), (2, 10, 2, "sparse")]


@pytest.mark.parametrize("D,k,n,scheme", SEP_CASES)
def test_separable_matches_tensor_of_1d_projections(g, gpu, O, D, k, n, scheme):
    import torch
    print(f"projection separable D={D} k={k} n={n} {scheme}", flush=True)
    fs_np = [lambda x: np.sin(2 * np.pi * x) + 0.5, lambda x: np.cos(4 * np.pi * x), lambda x: np.exp(-4 * (x - 0.3) ** 2)]
    fs_t = [lambda x: torch.sin(2 * math.pi * x) + 0.5, lambda x: torch.cos(4 * math.pi * x),
            lambda x: torch.exp(-4 * (x - 0.3) ** 2)]
    pick = [(2 * d) % 3 for d in range(D)]  # cos(4 pi x) in direction 3 only: its coefficients start at level 2

    def f(X):
        out = fs_t[pick[0]](X[0])
        for d in range(1, D):
            out = out * fs_t[pick[d]](X[d])
        return out
    want = O.tensor_construct(D, k, n, [O.proj1d(k, n, fs_np[pick[d]]) for d in range(D)], scheme)
    got = _proj(g, gpu, D, k, n, f, scheme, q=12)  # rule error ~1e-19 for these frequencies
    assert got.shape == want.shape and np.linalg.norm(want) > 1e-2
    # k >= 8: the multiwavelets evaluated from their monomial coefficients (degree k - 1)
    # lose ~1e-12 absolute; the 1-D host rule at q = 16 and q = 20 differs by 4e-12 at k = 10
    assert rel(got, want) < (1e-11 if k >= 8 else 1e-12)


@pytest.mark.parametrize("m,k,n", [([1, 2], 3, 4), ([2, 1, 1], 2, 4), ([1, 1, 2, 1], 4, 3)])
def test_nonseparable_cosine_matches_expansion(g, gpu, O, m, k, n):
    import torch
    D = len(m)
    print(f"projection cosine m={m} k={k} n={n}", flush=True)
    want = O.cos_coeffs(k, n, m, phase=0.3)

    def f(X):
        arg = sum(2 * math.pi * m[d] * X[d] for d in range(D))
        return torch.cos(arg + 0.3)
    assert rel(_proj(g, gpu, D, k, n, f, q=12), want) < 1e-11


def test_polynomials_are_exact(g, gpu, O):
    import torch
    from oracle import nodal as ON
    D, k, n = 3, 3, 4
    one = np.zeros(g.get_size(D, k, n))
    one[0] = 1.0
    got = _proj(g, gpu, D, k, n, lambda X: torch.ones_like(X[0]), q=2)
    assert np.abs(got - one).max() < 1e-14
    o1, xi = ON.one_modal_1d(k, n), ON.xi_modal_1d(k, n)
    for i in range(D):
        want = O.tensor_construct(D, k, n, [xi if d == i else o1 for d in range(D)])
        assert rel(_proj(g, gpu, D, k, n, lambda X, i=i: 2 * X[i] - 1, q=2), want) < 1e-13
    # x_1^2 x_3 (degree 2 per direction, k = 3): the 1-D projections are exact
    want = O.tensor_construct(D, k, n, [O.proj1d(k, n, lambda x: x * x), o1, O.proj1d(k, n, lambda x: x)])
    assert rel(_proj(g, gpu, D, k, n, lambda X: X[0] ** 2 * X[2], q=3), want) < 1e-13


def test_chunking_is_bit_exact(g, gpu):
    """Any split of the cells into calls gives the same bits: cells are independent."""
    import torch
    D, k, n = 3, 2, 4
    f = lambda X: torch.exp(torch.sin(2 * math.pi * X[0]) * X[1]) + X[2] ** 3
    a = _proj(g, gpu, D, k, n, f)
    P3 = (2 * (k + 4)) ** 3
    for mp in (1, 7 * P3, 64 * P3):
        np.testing.assert_array_equal(_proj(g, gpu, D, k, n, f, max_points=mp), a)


def test_matches_host_rule(g, gpu):
    """The device rule is the host vcoeffs_DG's rule (same nodes and weights)."""
    import torch
    D, k, n = 2, 4, 4
    fh = lambda X: np.exp(np.sin(2 * np.pi * X[0]) * np.cos(2 * np.pi * X[1]))
    ft = lambda X: torch.exp(torch.sin(2 * math.pi * X[0]) * torch.cos(2 * math.pi * X[1]))
    assert rel(_proj(g, gpu, D, k, n, ft, q=6), g.vcoeffs_DG(D, k, n, fh, q=6)) < 1e-13
    assert rel(_proj(g, gpu, D, k, n, ft, scheme="full", q=6), g.vcoeffs_DG(D, k, n, fh, scheme="full", q=6)) < 1e-13
    y = g.vcoeffs_DG(D, k, n, ft, q=6, device=gpu)  # the device= form of vcoeffs_DG
    assert torch.is_tensor(y) and y.device.type == "cuda"


def test_headline_grid_separable(g, gpu, O):
    """D = 4, k = 4, n = 9 (6,815,744 DoFs, 4.3e9 nodes at q = 10)."""
    import torch
    D, k, n = 4, 4, 9
    print("projection headline grid", flush=True)
    f = lambda X: torch.sin(2 * math.pi * X[0]) * torch.cos(2 * math.pi * X[1]) * (1 + X[2]) * torch.exp(-X[3])
    got = _proj(g, gpu, D, k, n, f, q=10)
    want = O.tensor_construct(D, k, n, [O.proj1d(k, n, lambda x: np.sin(2 * np.pi * x)),
                                        O.proj1d(k, n, lambda x: np.cos(2 * np.pi * x)),
                                        O.proj1d(k, n, lambda x: 1 + x), O.proj1d(k, n, lambda x: np.exp(-x))])
    assert got.size == 6815744
    # q = 10: the default q = k + 4 = 8 leaves 8e-12 of truncation on the level-1
    # coefficients of cos(2 pi x) (the host rule alike); at q = 10 the 1-D rule's error is
    # at rounding (3e-15)
    err = rel(got, want)
    print(f"headline grid: rel-L2 {err:.2e}, max abs {np.abs(got - want).max():.2e}", flush=True)
    assert err < 1e-12


def test_bad_f_is_rejected(g, gpu):
    import torch
    with pytest.raises(ValueError):
        g.vcoeffs_DG_device(2, 2, 3, lambda X: X[0][:5], device=gpu)
    with pytest.raises(ValueError):
        g.vcoeffs_DG_device(2, 2, 3, lambda X: np.zeros(X.shape[1]), device=gpu)


def test_projection_abi_argument_errors(g, gpu):
    """gsg_proj_nodes / gsg_proj_contract reject ranges and lengths that do not match the
    handle before launching anything."""
    import ctypes as C
    import torch
    from galerkinsparsegrids_jl_amd import _lib as L
    lib = L.lib()
    h = L.vp()
    assert lib.gsg_proj_create(2, 2, 3, 0, 2, C.byref(h)) == 0
    try:
        N, ncells, ppc = lib.gsg_proj_len(h), lib.gsg_proj_cells(h), lib.gsg_proj_points_per_cell(h)
        assert (N, ppc) == (g.get_size(2, 2, 3), 16) and N == ncells * 4
        X = torch.empty((2, ncells * ppc), dtype=torch.float64, device=gpu)
        fv = torch.zeros(ncells * ppc, dtype=torch.float64, device=gpu)
        out = torch.empty(N, dtype=torch.float64, device=gpu)
        px, pf, po = C.c_void_p(X.data_ptr()), C.c_void_p(fv.data_ptr()), C.c_void_p(out.data_ptr())
        assert lib.gsg_proj_nodes(h, 0, ncells + 1, px, X.shape[1], None) < 0       # past the last cell
        assert lib.gsg_proj_nodes(h, -1, 1, px, X.shape[1], None) < 0               # negative start
        assert lib.gsg_proj_nodes(h, 0, ncells, px, X.shape[1] - 1, None) < 0       # ld too small
        assert lib.gsg_proj_contract(h, 0, 2, pf, 2 * ppc - 1, po, N, None) < 0     # fvals_len
        assert lib.gsg_proj_contract(h, 0, 2, pf, 2 * ppc, po, N - 1, None) < 0     # out_len
        assert lib.gsg_proj_contract(h, ncells - 1, 2, pf, 2 * ppc, po, N, None) < 0
        assert b"cell range" in lib.gsg_last_error()
        # a valid call on zero data gives zeros
        out.fill_(1.0)
        assert lib.gsg_proj_contract(h, 0, ncells, pf, ncells * ppc, po, N, None) == 0
        torch.cuda.synchronize()
        assert float(out.abs().max()) == 0.0
    finally:
        lib.gsg_proj_destroy(h)


TENSOR_CASES = [(1, 3, 5, "sparse"), (2, 3, 4, "sparse"), (3, 2, 5, "full"), (4, 4, 6, "sparse"), (5, 2, 3, "sparse"),
                (2, 1, 0, "sparse"), (8, 1, 2, "sparse"), (4, 4, 9, "sparse")]


@pytest.mark.parametrize("D,k,n,scheme", TENSOR_CASES)
def test_tensor_construct_device_is_bit_exact(g, gpu, O, D, k, n, scheme):
    """tensor_construct on the device equals the oracle's bit for bit (same products in
    the same order); (4, 4, 9) is the headline grid."""
    print(f"tensor_construct D={D} k={k} n={n} {scheme}", flush=True)
    rng = np.random.default_rng(D * 100 + k * 10 + n)
    vecs = [rng.standard_normal(k << n) for _ in range(D)]
    got = g.tensor_construct(D, k, n, vecs, scheme, device=gpu).cpu().numpy()
    np.testing.assert_array_equal(got, O.tensor_construct(D, k, n, vecs, scheme))


def test_cos_coeffs_on_device(g, gpu, O):
    """examples/traveling_wave.jl:18-48 with the products on the device: equal to the host
    form bit for bit, and to the oracle's expansion within 1e-13."""
    for m, k, n in [([1, 2], 3, 4), ([2, -1, 1], 2, 5), ([1, 1, 2, 1], 4, 4)]:
        dev = g.cos_coeffs(k, n, m, phase=0.3, device=gpu).cpu().numpy()
        np.testing.assert_array_equal(dev, g.cos_coeffs(k, n, m, phase=0.3))
        assert rel(dev, O.cos_coeffs(k, n, m, phase=0.3)) < 1e-13
    s = g.sin_coeffs(3, 4, [1, 2], device=gpu).cpu().numpy()
    assert rel(s, O.cos_coeffs(3, 4, [1, 2], phase=-np.pi / 2)) < 1e-13


def test_tensor_construct_device_rejects_bad_vectors(g, gpu):
    with pytest.raises(ValueError):
        g.tensor_construct(2, 3, 4, [np.zeros(48), np.zeros(47)], device=gpu)
    with pytest.raises(ValueError):
        g.tensor_construct(3, 3, 4, [np.zeros(48)] * 2, device=gpu)


def test_broadcast_nodes_give_the_same_bits(g, gpu):
    """broadcast=True hands f per-direction node factors; elementwise f gives the same
    values at the same nodes, so the coefficients are bit-identical to the default form."""
    import torch

    def f(X):
        return torch.cos(2 * math.pi * (X[0] + 2 * X[1] - X[2]) + 0.3) * torch.exp(-(X[0] - 0.5) ** 2) + X[1] * X[2]
    for D3, k, n, scheme, mp in [(3, 3, 4, "sparse", 1 << 24), (3, 2, 3, "full", 7 * 12 ** 3), (3, 4, 5, "sparse", 1)]:
        a = g.vcoeffs_DG_device(D3, k, n, f, scheme, device=gpu, max_points=mp)
        b = g.vcoeffs_DG_device(D3, k, n, f, scheme, device=gpu, max_points=mp, broadcast=True)
        torch.cuda.synchronize()
        assert torch.equal(a, b), (D3, k, n, scheme)
    # a constant result broadcasts too
    c = g.vcoeffs_DG_device(2, 3, 3, lambda X: torch.tensor(1.0, device=gpu), device=gpu, broadcast=True)
    assert abs(float(c[0]) - 1.0) < 1e-14 and float(c[1:].abs().max()) < 1e-14
    with pytest.raises(ValueError):
        g.vcoeffs_DG_device(2, 3, 3, lambda X: torch.zeros(5, device=gpu), device=gpu, broadcast=True)

"""Pins of the oracle's multi-D layer: layout, assembly, Laplacian, RK4, energy."""
import itertools

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import rel


def _levels(D, n, scheme):
    for rev in itertools.product(range(1, n + 2), repeat=D):
        lev = rev[::-1]  # dimension 1 fastest
        if scheme == "sparse" and sum(lev) > n + D:
            continue
        yield lev


def _v2d_python(D, k, n, scheme):
    """Direct statement of dg_vmethods.jl:123-142 (CartesianIndices order)."""
    out = []
    for lev in _levels(D, n, scheme):
        cells = [1 << max(0, l - 2) for l in lev]
        for crev in itertools.product(*[range(1, c + 1) for c in cells[::-1]]):
            for mrev in itertools.product(*[range(1, k + 1)] * D):
                out.append((lev, crev[::-1], mrev[::-1]))
    return out


@pytest.mark.parametrize("D,k,n,scheme", [(2, 3, 5, "sparse"), (3, 2, 3, "sparse"), (2, 2, 2, "full"),
                                          (1, 3, 4, "sparse"), (4, 1, 3, "sparse")])
def test_layout_matches_cartesian_order(O, D, k, n, scheme):
    ref = _v2d_python(D, k, n, scheme)
    lev, cel, mod = O.v2dref(D, k, n, scheme)
    assert len(ref) == O.get_size(D, k, n, scheme)
    got = [(tuple(a), tuple(b), tuple(c)) for a, b, c in zip(lev.tolist(), cel.tolist(), mod.tolist())]
    assert got == ref
    idx = O.d2v_index(D, k, n, lev, cel, mod, scheme)
    np.testing.assert_array_equal(idx, np.arange(len(ref)))


def test_get_size_known_values(O):
    """Sizes quoted for the benchmark configurations."""
    assert O.get_size(2, 3, 5) == 1008
    assert O.get_size(3, 3, 7) == 45792
    assert O.get_size(4, 2, 8) == 167936
    assert O.get_size(2, 3, 5, "full") == (3 * 32) ** 2


def _bruteforce_d(O, D, d, k, n, scheme):
    """I x ... x D_1D (at d) x ... x I on the full tensor grid, restricted to the
    sparse index set: A[r, c] = D1[i_d(r), i_d(c)] prod_{e != d} delta(i_e(r), i_e(c))."""
    D1 = O.hier_dlf(k, n)
    lev, cel, mod = O.v2dref(D, k, n, scheme)
    l0 = lev - 1
    base = np.where(l0 == 0, 0, 1 << np.maximum(l0 - 1, 0))
    i1 = k * (base + cel - 1) + mod - 1  # N x D 1-D indices
    N = i1.shape[0]
    rows, cols, vals = [], [], []
    key_other = [tuple(np.delete(i1[r], d - 1)) for r in range(N)]
    groups = {}
    for r in range(N):
        groups.setdefault(key_other[r], []).append(r)
    for members in groups.values():
        for r in members:
            for c in members:
                v = D1[i1[r, d - 1], i1[c, d - 1]]
                if v != 0.0:
                    rows.append(r)
                    cols.append(c)
                    vals.append(v)
    return sp.csc_matrix((vals, (rows, cols)), shape=(N, N))


@pytest.mark.parametrize("D,k,n,scheme", [(2, 2, 3, "sparse"), (3, 2, 2, "sparse"), (2, 3, 2, "full"),
                                          (3, 1, 3, "sparse")])
def test_d_matrix_equals_restricted_kronecker(O, D, k, n, scheme):
    """multidim_derivative.jl:15-58 assembly == brute-force restricted Kronecker
    product, value for value."""
    for d in range(1, D + 1):
        A = O.d_matrix(D, d, k, n, scheme)
        B = _bruteforce_d(O, D, d, k, n, scheme)
        assert (A != B).nnz == 0


def test_d_matrix_commutes_with_axis_permutation(O):
    """Relabelling axes 1 <-> 2 maps D_1 to D_2."""
    D, k, n = 2, 2, 3
    lev, cel, mod = O.v2dref(D, k, n)
    perm = O.d2v_index(D, k, n, lev[:, ::-1].copy(), cel[:, ::-1].copy(), mod[:, ::-1].copy())
    P = sp.csr_matrix((np.ones(len(perm)), (np.arange(len(perm)), perm)))
    D1 = O.d_matrix(D, 1, k, n)
    D2 = O.d_matrix(D, 2, k, n)
    assert abs(P @ D1 @ P.T - D2).max() == 0.0


def test_laplacian_is_sum_of_squares(O):
    D, k, n = 2, 2, 4
    Lp = O.laplacian_matrix(D, k, n)
    ref = sum(A @ A for A in O.grad_matrix(D, k, n))
    assert abs(Lp - ref).max() <= 1e-12 * abs(ref).max()


def test_wave_and_transport_rhs(O):
    D, k, n = 2, 2, 3
    N = O.get_size(D, k, n)
    rng = np.random.default_rng(1)
    y = rng.standard_normal(2 * N)
    w = O.System("wave", D, k, n).rhs(y)
    Lp = O.laplacian_matrix(D, k, n)
    np.testing.assert_array_equal(w[:N], y[N:])
    assert rel(w[N:], Lp @ y[:N]) < 1e-14
    a = [0.7, -0.3]
    u = rng.standard_normal(N)
    t = O.System("transport", D, k, n, a=a).rhs(u)
    Ds = O.grad_matrix(D, k, n)
    assert rel(t, -(a[0] * (Ds[0] @ u) + a[1] * (Ds[1] @ u))) < 1e-14


def test_rk4_is_the_classical_polynomial(O):
    """Fixed-step RK4 = (I + hA + (hA)^2/2 + (hA)^3/6 + (hA)^4/24)^steps, and it
    approaches exp(TA)."""
    from scipy.linalg import expm
    D, k, n, a = 1, 2, 3, [1.0]
    s = O.System("transport", D, k, n, a=a)
    N = s.len
    A = np.column_stack([s.rhs(e) for e in np.eye(N)])
    rng = np.random.default_rng(2)
    u0 = rng.standard_normal(N)
    h, steps = 1e-3, 25
    hA = h * A
    R = np.eye(N) + hA + hA @ hA / 2 + hA @ hA @ hA / 6 + hA @ hA @ hA @ hA / 24
    ref = np.linalg.matrix_power(R, steps) @ u0
    got = s.rk4(u0, h, steps)
    assert rel(got, ref) < 1e-13
    assert rel(got, expm(steps * h * A) @ u0) < 1e-6


def test_wave_energy_conservation_2d(O):
    """test/solvers.jl:54-77 shape (D=2, k=3, n=5, sparse, sin*sin, v0 = 0, T = 1):
    energy decreases monotonically by less than 1e-8 and sqrt(E) = sqrt(2) pi +- 1e-4.
    The reference integrates with ode78/ode45; fixed-step RK4 meets the same bounds
    with dt = 2^-5 / 20."""
    D, k, n = 2, 3, 5
    s1 = O.proj1d(k, n, lambda x: np.sin(2 * np.pi * x))
    u0 = O.tensor_construct(D, k, n, [s1, s1])
    y = np.concatenate([u0, np.zeros_like(u0)])
    sysw = O.System("wave", D, k, n)
    dt = 2.0 ** -n / 20
    E = [O.energy(D, k, n, y)]
    for _ in range(4):
        y = sysw.rk4(y, dt, 160)
        E.append(O.energy(D, k, n, y))
    assert all(E[i] > E[i + 1] for i in range(len(E) - 1))
    assert E[0] - E[-1] < 1e-8
    for e in E:
        assert abs(np.sqrt(e) - np.sqrt(2) * np.pi) < 1e-4


def test_transport_exact_solution(O):
    """u_t + a . grad u = 0 on the periodic box: after time T the solution is
    u0(x - aT); the discrete solution reaches it to discretisation accuracy, and the
    error falls at least 4x per level."""
    D, k = 2, 3
    a = [1.0, 0.5]
    T = 0.25
    f = lambda x: np.cos(2 * np.pi * x)
    g = lambda x: np.sin(2 * np.pi * x)
    errs = []
    for n in (4, 5):
        steps = 25 << n
        u0 = O.tensor_construct(D, k, n, [O.proj1d(k, n, f), O.proj1d(k, n, g)])
        s = O.System("transport", D, k, n, a=a)
        uT = s.rk4(u0, T / steps, steps)
        ex = O.tensor_construct(D, k, n, [O.proj1d(k, n, lambda x: f(x - a[0] * T)),
                                         O.proj1d(k, n, lambda x: g(x - a[1] * T))])
        errs.append(rel(uT, ex))
    assert errs[0] < 5e-3
    assert errs[1] < errs[0] / 4


def test_golden_multid(O):
    from pathlib import Path
    z = np.load(Path(__file__).parent / "golden" / "oracle_golden.npz")
    for d in (1, 2):
        A = O.d_matrix(2, d, 2, 3)
        np.testing.assert_array_equal(A.indptr, z[f"d_matrix_D2_k2_n3_d{d}_colptr"])
        np.testing.assert_array_equal(A.indices, z[f"d_matrix_D2_k2_n3_d{d}_rowval"])
        np.testing.assert_array_equal(A.data, z[f"d_matrix_D2_k2_n3_d{d}_nzval"])
    sysw = O.System("wave", 2, 3, 5)
    np.testing.assert_array_equal(sysw.rk4(z["wave_D2_k3_n5_y0"], 0.1 * 2.0 ** -5, 20), z["wave_D2_k3_n5_y20"])
    syst = O.System("transport", 3, 2, 4, a=[1.0, 0.5, 0.25])
    np.testing.assert_array_equal(syst.rhs(z["transport_D3_k2_n4_u0"]), z["transport_D3_k2_n4_rhs"])

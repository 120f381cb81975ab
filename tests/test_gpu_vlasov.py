"""Device basis transforms and the Vlasov right-hand side / RK4 against the oracle
(oracle/nodal.py: the reference's restricted Kronecker products and steprule).

Tolerances: FP64 paths within 1e-12 relative L2 of the oracle; the FP32-derivative
Vlasov operator within 1e-6.  At the full phase-space size of the Vlasov workload
(2 Dx = 4, k = 2, n = 8) the checks are size-independent: inverse pairs, exact tensor
right-hand sides and eager-vs-captured equality.
"""
import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ON():
    from oracle import nodal
    return nodal


TRANSFORM_CASES = [(1, 2, 4, "sparse"), (2, 3, 5, "sparse"), (3, 5, 3, "sparse"), (4, 2, 4, "sparse"),
                   (4, 3, 3, "sparse"), (2, 2, 3, "full"), (5, 2, 2, "sparse"), (7, 2, 1, "sparse")]


@pytest.mark.parametrize("D,k,n,scheme", TRANSFORM_CASES)
def test_transforms_match_restricted_kronecker(g, gpu, ON, D, k, n, scheme):
    print(f"transform D={D} k={k} n={n} {scheme}", flush=True)
    rng = np.random.default_rng(D * 100 + k * 10 + n)
    for a, b in [("modal", "nodal"), ("nodal", "points"), ("points", "nodal"), ("nodal", "modal"),
                 ("modal", "points"), ("points", "modal")]:
        T = ON.transform_dd(D, k, n, a, b, scheme)
        op = g.transform(D, k, n, a, b, scheme)
        x = rng.standard_normal(T.shape[0])
        assert op.len == T.shape[0]
        assert rel(op @ x, T @ x) < 1e-12, (a, b)
        op.close()


def test_transform_device_tensors(g, gpu, ON):
    import torch
    D, k, n = 3, 3, 3
    op = g.transform(D, k, n, "modal", "nodal")
    x = np.random.default_rng(5).standard_normal(op.len)
    y = op @ torch.from_numpy(x).to(gpu)
    torch.cuda.synchronize()
    assert rel(y.cpu().numpy(), ON.transform_dd(D, k, n, "modal", "nodal") @ x) < 1e-12


def test_transform_inverse_pairs_at_workload_size(g, gpu):
    """2 Dx = 4, k = 2, n = 8 (N = 167,936): p2n n2p = I and n2m m2n = I on the grid."""
    D, k, n = 4, 2, 8
    m2n, n2p = g.make_modal2point_matrices(D, k, n)
    p2n, n2m = g.make_point2modal_matrices(D, k, n)
    x = np.random.default_rng(1).standard_normal(m2n.len)
    assert m2n.len == 167936
    assert rel(p2n @ (n2p @ x), x) < 1e-10
    assert rel(n2m @ (m2n @ x), x) < 1e-10
    assert rel(m2n @ (n2m @ x), x) < 1e-10
    for op in (m2n, n2p, p2n, n2m):
        op.close()


@pytest.mark.parametrize("D,k,n", [(5, 3, 3), (4, 5, 3), (7, 2, 4)])
def test_transform_inverse_pairs_large_cells(g, gpu, D, k, n):
    """Large cells, at sizes where the oracle's dense products are too slow:
    p2n n2p = I and n2m m2n = I on the grid. Cells of more than 128 values
    (k^D = 243, 625) are never split over a workgroup, so their lanes take lists longer
    than a round (4 entries at k > 2) in several rounds. 128-value cells (D = 7, k = 2)
    split their long lists over 2 lanes a DoF and read the others from the fixed-width
    table."""
    print(f"transform inverse pairs D={D} k={k} n={n}", flush=True)
    m2n, n2p = g.make_modal2point_matrices(D, k, n)
    p2n, n2m = g.make_point2modal_matrices(D, k, n)
    x = np.random.default_rng(D * 10 + k).standard_normal(m2n.len)
    assert rel(p2n @ (n2p @ x), x) < 1e-10
    assert rel(n2m @ (m2n @ x), x) < 1e-10
    assert rel(m2n @ (n2m @ x), x) < 1e-10
    for op in (m2n, n2p, p2n, n2m):
        op.close()


def test_xi_point_on_device(g, gpu, ON):
    """get_xi_point(D, i, k, n) = 2 x_i - 1 at each grid point's collocation point."""
    from oracle import oracle as O
    D, k, n = 3, 5, 3
    m2n, n2p = g.make_modal2point_matrices(D, k, n)
    lev, cel, mod = O.v2dref(D, k, n)
    x1 = ON.points_x(k, n)
    for i in range(1, D + 1):
        xi = x1[ON.index_1d(k, lev[:, i - 1], cel[:, i - 1], mod[:, i - 1])]
        np.testing.assert_allclose(g.get_xi_point(D, i, k, n, m2n, n2p), 2 * xi - 1, rtol=0, atol=1e-11)


# the last n2p sweep combines Dx vectors without a force and 2 Dx with one: 1, 2, 3, 4, 6 and 8
VLASOV_CASES = [(1, 3, 4, False), (1, 5, 3, True), (2, 2, 3, False), (2, 2, 3, True), (2, 3, 2, True),
                (3, 2, 2, False), (3, 2, 2, True), (4, 2, 1, False), (4, 2, 1, True)]


def _force(ON, V, k, n, Dx):
    from oracle import oracle as O
    one, xi = ON.one_modal_1d(k, n), ON.xi_modal_1d(k, n)
    Fs = []
    for d in range(Dx):
        c = O.tensor_construct(2 * Dx, k, n, [xi if j == d else one for j in range(2 * Dx)])
        Fs.append(-(d + 1.5) * (V.n2p @ (V.m2n @ c)))  # a force linear in x_d
    return Fs


@pytest.mark.parametrize("Dx,k,n,force", VLASOV_CASES)
def test_vlasov_rhs_matches_oracle(g, gpu, ON, Dx, k, n, force):
    print(f"vlasov rhs Dx={Dx} k={k} n={n} force={force}", flush=True)
    V = ON.Vlasov(Dx, k, n)
    F = _force(ON, V, k, n, Dx) if force else None
    if force:
        V = ON.Vlasov(Dx, k, n, F_point=F)
    f = np.random.default_rng(Dx * 10 + k).standard_normal(V.N)
    want = V.rhs(f)
    op = g.vlasov_operator(Dx, k, n, F)
    assert rel(op @ f, want) < 1e-12
    op.close()
    op = g.vlasov_operator(Dx, k, n, F, precision="fp32")
    assert rel(op @ f, want) < 1e-6
    op.close()


@pytest.mark.parametrize("Dx,k,n,force", [(1, 3, 4, True), (2, 2, 3, True)])
def test_vlasov_rk4_matches_oracle(g, gpu, ON, Dx, k, n, force):
    from oracle import oracle as O
    print(f"vlasov rk4 Dx={Dx} k={k} n={n}", flush=True)
    V = ON.Vlasov(Dx, k, n)
    F = _force(ON, V, k, n, Dx) if force else None
    V = ON.Vlasov(Dx, k, n, F_point=F)
    D = 2 * Dx
    f0 = O.tensor_construct(D, k, n, [O.proj1d(k, n, (lambda x, d=d: np.exp(-10 * (x - 0.4 - 0.05 * d) ** 2)))
                                      for d in range(D)])
    dt, steps = 0.05 * 2.0 ** -n, 20
    want = V.rk4(f0, dt, steps)
    op = g.vlasov_operator(Dx, k, n, F)
    assert rel(op.rk4(f0, dt, steps), want) < 1e-12
    op.close()
    op = g.vlasov_operator(Dx, k, n, F, precision="fp32")
    assert rel(op.rk4(f0, dt, steps), want) < 1e-6
    op.close()


def test_vlasov_exact_tensor_rhs_at_workload_size(g, gpu, ON):
    """Free streaming at 2 Dx = 4, k = 2, n = 8: f = g(x_1) gives f' = -tensor(D g, 1, xi, 1)."""
    from oracle import oracle as O
    k, n = 2, 8
    gx = O.proj1d(k, n, lambda x: np.sin(2 * np.pi * x) + 0.2 * np.cos(6 * np.pi * x))
    one, xi = ON.one_modal_1d(k, n), ON.xi_modal_1d(k, n)
    f = O.tensor_construct(4, k, n, [gx, one, one, one])
    want = -O.tensor_construct(4, k, n, [O.hier_dlf(k, n) @ gx, one, xi, one])
    op = g.vlasov_operator(2, k, n)
    assert rel(op @ f, want) < 1e-10
    op.close()


def test_vlasov_captured_steps_equal_eager(g, gpu):
    """24 steps in one call (replayed captured steps) equal 24 single-step calls."""
    Dx, k, n = 2, 2, 5
    op = g.vlasov_operator(Dx, k, n, precision="fp32")
    f0 = np.random.default_rng(3).standard_normal(op.len) * 1e-2
    dt = 0.05 * 2.0 ** -n
    a = op.rk4(f0, dt, 24)
    b = f0.copy()
    for _ in range(24):
        b = op.rk4(b, dt, 1)
    np.testing.assert_array_equal(a, b)
    c = op.rk4(f0, dt, 24)  # replay of the same captured steps
    np.testing.assert_array_equal(a, c)
    d = op.rk4(f0, dt / 2, 48)  # a new dt re-captures
    assert rel(d, a) < 1e-3
    op.close()


def test_vlasov_evolve_api(g, gpu, ON):
    from oracle import oracle as O
    Dx, k, n = 1, 3, 4
    m2n, n2p = g.make_modal2point_matrices(2 * Dx, k, n)
    p2n, n2m = g.make_point2modal_matrices(2 * Dx, k, n)
    f0 = O.tensor_construct(2, k, n, [O.proj1d(k, n, lambda x: np.exp(-20 * (x - 0.5) ** 2))] * 2)
    Fp = [-(g.get_xi_point(2, 1, k, n, m2n, n2p))]
    times, sols = g.vlasov_evolve(Dx, k, n, m2n, n2p, p2n, n2m, f0, Fp, 0.0, 0.05, 3, dt=0.01)
    assert len(times) == 3 and len(sols) == 3 and np.allclose(times, [0, 0.025, 0.05])
    V = ON.Vlasov(Dx, k, n, F_point=Fp)
    assert rel(sols[1], V.rk4(f0, 0.025 / 3, 3)) < 1e-12
    assert rel(sols[2], V.rk4(f0, 0.025 / 3, 6)) < 1e-12
    with pytest.raises(ValueError):
        g.vlasov_evolve(Dx, k, n, m2n, n2p, p2n, n2m, f0, Fp, 0.0, 0.05, 3, order="45")


def test_exported_transforms_are_the_device_operators(g, gpu):
    """transform_csc is the matrix the device transform applies (FP64, 1e-12)."""
    D, k, n = 3, 3, 3
    x = np.random.default_rng(11).standard_normal(g.get_size(D, k, n))
    for a, b in [("modal", "nodal"), ("nodal", "points"), ("points", "nodal"), ("nodal", "modal")]:
        op = g.transform(D, k, n, a, b)
        assert rel(g.transform_csc(D, k, n, a, b) @ x, op @ x) < 1e-12, (a, b)
        op.close()


def test_vlasov_evolve_writes_output(g, gpu, tmp_path):
    from oracle import oracle as O
    Dx, k, n = 1, 2, 4
    f0 = O.tensor_construct(2, k, n, [O.proj1d(k, n, lambda x: np.exp(-20 * (x - 0.5) ** 2))] * 2)
    path = tmp_path / "vlasov.npz"
    times, sols = g.vlasov_evolve(Dx, k, n, None, None, None, None, f0, [], 0.0, 0.02, 3, dt=0.005, output=path)
    out = g.load_vlasov_output(path)
    np.testing.assert_array_equal(out["times"], times)
    for i, s in enumerate(sols, start=1):
        np.testing.assert_array_equal(out[f"f_modal.{i:06d}"], s)
    # the written operators reproduce the device right-hand side (free streaming)
    N = f0.size
    v = out["n2p"] @ (out["m2n"] @ g.get_xi_modal(2, 2, k, n))
    rhs = out["n2m"] @ (out["p2n"] @ (-(v * (out["n2p"] @ (out["m2n"] @ (out["Ds[1]"] @ f0))))))
    op = g.vlasov_operator(Dx, k, n)
    assert op.len == N and rel(op @ f0, rhs) < 1e-12
    op.close()

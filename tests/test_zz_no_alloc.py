"""gsg_op_rk4_device_noalloc (RK4 with its state in a fixed device array, no
allocation) against gsg_op_rk4_device: bit for bit, on two calls in a row with the
second on a smaller grid (the array keeps the first call's contents), in both
precisions, through the C ABI; a state larger than the array is refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _noalloc(L, op, y, dt, steps, stream):
    return L.lib().gsg_op_rk4_device_noalloc(op._h, C.c_void_p(y.data_ptr()), op.len, float(dt), int(steps),
                                            C.c_void_p(stream))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_noalloc_equals_rk4_device(g, gpu, precision):
    import torch
    from galerkinsparsegrids_jl_amd import _lib as L
    print(f"[gpu] rk4 without allocation, {precision}", flush=True)
    rng = np.random.default_rng(17)
    cases = [("transport", 4, 4, 6, 30), ("wave", 3, 3, 4, 25), ("transport", 2, 2, 4, 40)]  # each smaller
    for kind, D, k, n, steps in cases:
        a = list(rng.uniform(-1, 1, D))
        op = g.wave_matrix(D, k, n, precision=precision) if kind == "wave" else \
            g.transport_matrix(D, k, n, a, precision=precision)
        y0 = rng.standard_normal(op.len)
        dt = 0.05 * 2.0 ** -n / max(1.0, sum(abs(v) for v in a))
        ya = torch.tensor(y0, device=gpu)
        op.rk4(ya, dt, steps)  # gsg_op_rk4_device
        yb = torch.tensor(y0, device=gpu)
        stream = torch.cuda.current_stream(gpu).cuda_stream
        L.check(_noalloc(L, op, yb, dt, steps, stream))
        np.testing.assert_array_equal(yb.cpu().numpy(), ya.cpu().numpy())
        op.close()


def test_noalloc_rejects_large_state(g, gpu):
    import torch
    from galerkinsparsegrids_jl_amd import _lib as L
    op = g.transport_matrix(4, 4, 9, [1.0, 0.5, -0.5, 0.25], precision="fp32")  # 6.8 M DoFs, 191 MB of state
    y = torch.zeros(op.len, dtype=torch.float64, device=gpu)
    rc = _noalloc(L, op, y, 1e-4, 1, torch.cuda.current_stream(gpu).cuda_stream)
    assert rc == L.GSG_EINVAL
    assert "no-allocation scratch" in L.lib().gsg_last_error().decode()
    op.close()

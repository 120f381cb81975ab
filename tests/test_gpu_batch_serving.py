"""Batched solves and the continuous-batching server on the GPU: every result equals
the same solve run alone (bit for bit: the kernels' arithmetic does not depend on the
other solves in the launch), and the FP64 oracle."""
import time

import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu

SOLVES = [
    # kind, D, k, n, nsteps
    ("transport", 2, 3, 5, 30),
    ("transport", 3, 2, 6, 17),
    ("wave", 2, 2, 4, 25),
    ("transport", 4, 2, 5, 30),
    ("transport", 2, 4, 6, 9),
    ("wave", 3, 3, 3, 30),
    ("transport", 3, 4, 4, 30),
]


def _mk(g, kind, D, k, n, seed):
    rng = np.random.default_rng(seed)
    N = g.get_size(D, k, n)
    a = list(rng.uniform(-1, 1, D))
    y0 = rng.standard_normal(2 * N if kind == "wave" else N)
    dt = 0.05 * 2.0 ** -n / max(1.0, sum(abs(v) for v in a))
    return a, y0, dt


def _alone(g, kind, D, k, n, a, y0, dt, steps, precision):
    op = g.wave_matrix(D, k, n, precision=precision) if kind == "wave" else g.transport_matrix(D, k, n, a, precision=precision)
    y = op.rk4(y0, dt, steps)
    op.close()
    return y


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_batch_equals_alone(g, O, gpu, precision):
    print(f"[gpu] batch of {len(SOLVES)} uneven solves {precision}", flush=True)
    b = g.Batch(precision)
    cases = []
    for s, (kind, D, k, n, steps) in enumerate(SOLVES):
        a, y0, dt = _mk(g, kind, D, k, n, s)
        slot = b.add(kind, D, k, n, y0, dt, steps, a=a if kind == "transport" else None)
        cases.append((slot, kind, D, k, n, a, y0, dt, steps))
    b.run()
    for slot, kind, D, k, n, a, y0, dt, steps in cases:
        yb = b.get(slot)
        ya = _alone(g, kind, D, k, n, a, y0, dt, steps, precision)
        np.testing.assert_array_equal(yb, ya)
        if precision == "fp64":
            sysO = O.System("wave", D, k, n) if kind == "wave" else O.System("transport", D, k, n, a=a)
            assert rel(yb, sysO.rk4(y0, dt, steps)) < 1e-12
    b.close()


def test_batch_step_partial(g, gpu):
    b = g.Batch("fp64")
    a, y0, dt = _mk(g, "transport", 2, 3, 4, 11)
    s0 = b.add("transport", 2, 3, 4, y0, dt, 10, a=a)
    b.step(4)
    y4 = b.get(s0)
    b.step(100)
    y10 = b.get(s0)
    np.testing.assert_array_equal(y4, _alone(g, "transport", 2, 3, 4, a, y0, dt, 4, "fp64"))
    np.testing.assert_array_equal(y10, _alone(g, "transport", 2, 3, 4, a, y0, dt, 10, "fp64"))


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_server_result_alone_equals_in_mix(g, gpu, precision):
    """Requests arriving while others run are admitted and batched; each result is
    identical to the same solve run alone."""
    print(f"[gpu] server mix {precision}", flush=True)
    reqs = []
    with g.Server(precision=precision, steps_per_tick=7) as srv:
        for s, (kind, D, k, n, steps) in enumerate(SOLVES * 2):
            a, y0, dt = _mk(g, kind, D, k, n, 100 + s)
            rid = srv.submit(kind, D, k, n, y0, dt, steps, a=a if kind == "transport" else None)
            reqs.append((rid, kind, D, k, n, a, y0, dt, steps))
            time.sleep(0.002 * (s % 3))
        for rid, kind, D, k, n, a, y0, dt, steps in reqs:
            y, lat = srv.wait(rid)
            assert lat > 0
            np.testing.assert_array_equal(y, _alone(g, kind, D, k, n, a, y0, dt, steps, precision))
        st = srv.stats()
        assert st["completed"] == len(reqs) and st["queued"] == 0 and st["resident"] == 0
        assert st["dof_updates"] > 0


def _until(srv, rid, status, timeout=60.0):
    t0 = time.time()
    while srv.poll(rid) != status:
        assert time.time() - t0 < timeout, f"request {rid} never became {status}"
        time.sleep(0.002)


def _resident_bytes(g, kind, D, k, n, copies=1):
    """The server's charge (resident_bytes) with `copies` long solves of one grid resident:
    the states plus the grid's operator tables, counted once."""
    with g.Server(precision="fp64", steps_per_tick=1) as srv:
        ids = []
        for c in range(copies):
            a, y0, dt = _mk(g, kind, D, k, n, 300 + c)
            ids.append(srv.submit(kind, D, k, n, y0, dt, 100000, a=a))
        for rid in ids:
            _until(srv, rid, "running")
        return srv.stats()["resident_bytes"]


def test_fp64_segmented_kernel_batch(g, O, gpu):
    """An FP64 k = 2 class of at least 64K cells runs the segmented kernel (several cells a
    wavefront); a single solve of the grid (10.5K cells) runs the plan kernel.  Eight solves
    of D = 4, k = 2, n = 8 in one batch: each equals its solve run alone to 1e-13, the
    batch's class is the tiled variant, and one solve matches the oracle's RK4."""
    print("[gpu] fp64 segmented kernel batch", flush=True)
    D, k, n, steps = 4, 2, 8, 6
    N = g.get_size(D, k, n)
    b = g.Batch("fp64")
    inputs = []
    for s in range(8):
        a, y0, dt = _mk(g, "transport", D, k, n, 400 + s)
        b.add("transport", D, k, n, y0, dt, steps, a=a)
        inputs.append((a, y0, dt))
    b.profile(True)
    b.run()
    stats = [st for st in b.kernel_stats() if st["D"] == D and st["k"] == k]
    assert stats and all(st["variant"] == 2 for st in stats), stats  # GSG_KERNEL_TILE
    for slot, (a, y0, dt) in enumerate(inputs):
        alone = _alone(g, "transport", D, k, n, a, y0, dt, steps, "fp64")
        assert rel(b.get(slot), alone) < 1e-13, slot
    a, y0, dt = inputs[0]
    ref = O.System("transport", D, k, n, a=list(a)).rk4(y0, dt, steps, nthreads=8)
    assert rel(b.get(0), ref) < 1e-12
    assert N * 8 // 16 >= 65536  # the class holds enough cells for the segmented kernel
    b.close()


def test_server_charges_grid_tables_once(g, gpu):
    """The HBM budget counts a grid's operator tables with its first admitted solve:
    two resident solves of one grid cost one state more than one solve."""
    print("[gpu] server table charge", flush=True)
    D, k, n = 3, 3, 5
    N = g.get_size(D, k, n)
    one = _resident_bytes(g, "transport", D, k, n, 1)
    two = _resident_bytes(g, "transport", D, k, n, 2)
    state = two - one
    assert state == 4 * N * 8  # FP64 transport: y, acc and two stage inputs
    assert one > state  # + the tables


def test_server_small_request_passes_blocked_one(g, gpu):
    """Admission is first fit: a request that cannot fit beside the resident ones does
    not hold back a later one that does (for a bounded number of passes)."""
    print("[gpu] server first fit", flush=True)
    big = ("transport", 4, 3, 5)
    budget = _resident_bytes(g, *big)  # the big solve fits alone, exactly
    with g.Server(precision="fp64", hbm_budget_bytes=budget, steps_per_tick=4) as srv:
        a, y0, dt = _mk(g, "transport", 2, 3, 5, 31)
        long_id = srv.submit("transport", 2, 3, 5, y0, dt, 100000, a=a)
        _until(srv, long_id, "running")
        ab, yb, dtb = _mk(g, *big, 32)
        big_id = srv.submit(*big, yb, dtb, 5, a=ab)
        a3, y3, dt3 = _mk(g, "transport", 2, 2, 4, 33)
        small_id = srv.submit("transport", 2, 2, 4, y3, dt3, 8, a=a3)
        y, _ = srv.wait(small_id)
        np.testing.assert_array_equal(y, _alone(g, "transport", 2, 2, 4, a3, y3, dt3, 8, "fp64"))
        assert srv.poll(big_id) == "queued"  # still waiting for the long solve's memory


def test_server_request_blocked_by_earlier_tables_fails(g, gpu):
    """With nothing resident, a request that does not fit beside the operator tables of
    grids served before (they stay resident and stay charged) fails with GSG_EFULL
    instead of waiting for memory that never frees."""
    from galerkinsparsegrids_jl_amd import _lib
    print("[gpu] server: request blocked by earlier tables", flush=True)
    A = ("transport", 2, 3, 5)
    B = ("transport", 3, 4, 5)
    one_a = _resident_bytes(g, *A)
    tables_a = one_a - 4 * g.get_size(*A[1:]) * 8
    one_b = _resident_bytes(g, *B)
    assert tables_a > 8 and one_b > one_a
    with g.Server(precision="fp64", hbm_budget_bytes=one_b + tables_a - 8, steps_per_tick=4) as srv:
        a, y0, dt = _mk(g, *A, 51)
        srv.wait(srv.submit(*A, y0, dt, 4, a=a))  # A's tables stay charged
        ab, yb, dtb = _mk(g, *B, 52)
        rid = srv.submit(*B, yb, dtb, 4, a=ab)  # fits the budget alone, not beside A's tables
        with pytest.raises(_lib.GsgError) as e:
            srv.wait(rid)
        assert e.value.code == _lib.GSG_EFULL


def test_server_admission_by_footprint(g, gpu):
    """A small HBM budget admits one request at a time; all still complete in order,
    and a request larger than the budget is refused at submission."""
    from galerkinsparsegrids_jl_amd import _lib
    D, k, n = 3, 3, 5
    one = _resident_bytes(g, "transport", D, k, n)
    budget = one + one // 4  # fits one FP64 transport solve and the grid's tables, not two
    with g.Server(precision="fp64", hbm_budget_bytes=budget, steps_per_tick=4) as srv:
        ids = []
        for s in range(3):
            a, y0, dt = _mk(g, "transport", D, k, n, 200 + s)
            ids.append((srv.submit("transport", D, k, n, y0, dt, 12, a=a), a, y0, dt))
        for rid, a, y0, dt in ids:
            y, _ = srv.wait(rid)
            np.testing.assert_array_equal(y, _alone(g, "transport", D, k, n, a, y0, dt, 12, "fp64"))
        big = np.zeros(g.get_size(3, 3, 8))
        with pytest.raises(_lib.GsgError) as e:
            srv.submit("transport", 3, 3, 8, big, 1e-3, 5, a=[1, 1, 1])
        assert e.value.code == _lib.GSG_EFULL


def test_server_nonfinite_state_fails_request(g, gpu):
    from galerkinsparsegrids_jl_amd import _lib
    D, k, n = 2, 2, 3
    with g.Server(precision="fp64") as srv:
        y0 = np.ones(g.get_size(D, k, n))
        y0[3] = np.nan
        rid = srv.submit("transport", D, k, n, y0, 1e-3, 3, a=[1.0, 1.0])
        a, y1, dt = _mk(g, "transport", D, k, n, 9)
        ok = srv.submit("transport", D, k, n, y1, dt, 3, a=a)
        with pytest.raises(_lib.GsgError) as e:
            srv.wait(rid)
        assert e.value.code == _lib.GSG_ENONFINITE
        y, _ = srv.wait(ok)
        np.testing.assert_array_equal(y, _alone(g, "transport", D, k, n, a, y1, dt, 3, "fp64"))


@pytest.mark.parametrize("with_wave", [False, True])
def test_fused_schedule_equals_streams(g, gpu, with_wave):
    """One launch per stage for every class (the fused schedule) gives the states of the
    per-class launches on four streams bit for bit, over eager and captured steps and a
    switch of schedule in the middle; classes without a fused body (a plan-kernel class,
    D = 5) launch beside the fused one."""
    print(f"[gpu] fused schedule, wave={with_wave}", flush=True)
    shapes = [("transport", 4, 4, 4), ("transport", 4, 3, 4), ("transport", 3, 3, 5), ("transport", 2, 2, 4),
              ("transport", 5, 2, 3), ("transport", 3, 2, 5), ("transport", 2, 4, 5)]
    if with_wave:
        shapes += [("wave", 2, 3, 4), ("wave", 4, 3, 3), ("wave", 3, 4, 3)]
    rng = np.random.default_rng(77)
    solves = []
    for kind, D, k, n in shapes:
        a, y0, dt = _mk(g, kind, D, k, n, int(rng.integers(1000)))
        solves.append((kind, D, k, n, a, y0, dt))
    out = {}
    for order in (("streams", "streams"), ("fused", "fused"), ("streams", "fused")):
        b = g.Batch("fp32")
        for kind, D, k, n, a, y0, dt in solves:
            b.add(kind, D, k, n, y0, dt, 1000, a=a if kind == "transport" else None)
        b.schedule(order[0])
        b.step(3)    # eager
        b.step(17)   # captured steps
        b.schedule(order[1])
        b.step(9)
        out[order] = [b.get(i) for i in range(len(solves))]
        b.close()
    for order in (("fused", "fused"), ("streams", "fused")):
        for i in range(len(solves)):
            np.testing.assert_array_equal(out[order][i], out[("streams", "streams")][i])


def test_tile_alt_equals_tile(g, gpu):
    """The segmented kernel with the other number of entries per load round gives the
    same states bit for bit: a direction's entries accumulate in list order whatever the
    round size."""
    from galerkinsparsegrids_jl_amd import tune
    print("[gpu] tile_alt vs tile", flush=True)
    shapes = [("transport", 4, 3, 5), ("transport", 3, 3, 4), ("transport", 2, 4, 5), ("transport", 4, 2, 5),
              ("wave", 3, 2, 4), ("wave", 2, 3, 4)]
    rng = np.random.default_rng(91)
    solves = [(kind, D, k, n) + _mk(g, kind, D, k, n, int(rng.integers(1000))) for kind, D, k, n in shapes]
    out = {}
    try:
        for var in ("tile", "tile_alt"):
            for _, D, k, n, *_ in solves:
                tune.set_variant(D, k, n, "fp32", var)
            assert all(tune.get_variant(D, k, n, "fp32") == var for _, D, k, n, *_ in solves)
            b = g.Batch("fp32")
            for kind, D, k, n, a, y0, dt in solves:
                b.add(kind, D, k, n, y0, dt, 23, a=a if kind == "transport" else None)
            b.run()
            out[var] = [b.get(i) for i in range(len(solves))]
            b.close()
    finally:
        for _, D, k, n, *_ in solves:
            tune.set_variant(D, k, n, "fp32", "auto")
    for x, y in zip(out["tile"], out["tile_alt"]):
        np.testing.assert_array_equal(x, y)


def test_server_stop_with_blocked_waiters(g, gpu):
    """stop() while other threads wait on a running and a queued request: the waiters
    return 'server stopped' (GSG_ESTATE), nothing is freed under the worker, and stop
    returns once they have left."""
    import threading
    from galerkinsparsegrids_jl_amd import _lib
    print("[gpu] server stop with blocked waiters", flush=True)
    D, k, n = 4, 3, 6
    N = g.get_size(D, k, n)
    srv = g.Server(precision="fp64", steps_per_tick=4, hbm_budget_bytes=N * 8 * 6)  # one resident at a time
    ids = []
    for s in range(2):
        a, y0, dt = _mk(g, "transport", D, k, n, 31 + s)
        ids.append(srv.submit("transport", D, k, n, y0, dt, 200000, a=a))
    codes = []

    def waiter(rid):
        try:
            srv.wait(rid)
            codes.append("done")
        except _lib.GsgError as e:
            codes.append(e.code)

    threads = [threading.Thread(target=waiter, args=(r,)) for r in ids]
    for t in threads:
        t.start()
    deadline = time.time() + 30
    while srv.poll(ids[0]) != "running" and time.time() < deadline:
        time.sleep(0.01)
    assert srv.poll(ids[0]) == "running" and srv.poll(ids[1]) == "queued"
    time.sleep(0.2)
    srv.stop()
    for t in threads:
        t.join(timeout=30)
    assert not any(t.is_alive() for t in threads)
    assert sorted(codes) == [_lib.GSG_ESTATE, _lib.GSG_ESTATE]


_GRAPH_SCRIPT = r"""
import sys, numpy as np
sys.path.insert(0, sys.argv[1])
import __graft_entry__ as ge
g = ge.load_package()
out = sys.argv[2]
rng = np.random.default_rng(5)
b = g.Batch("fp32")
sizes = []
for s, (kind, D, k, n) in enumerate([("transport", 4, 4, 4), ("transport", 4, 3, 4), ("wave", 2, 3, 4),
                                     ("transport", 3, 2, 5), ("transport", 2, 2, 4)]):
    N = g.get_size(D, k, n)
    y0 = rng.standard_normal(2 * N if kind == "wave" else N)
    b.add(kind, D, k, n, y0, 0.01 * 2.0 ** -n, 100, a=list(rng.uniform(-1, 1, D)) if kind == "transport" else None)
    sizes.append(y0.size)
snaps = []
for chunk in (3, 1, 9, 2):  # eager first call, then the 1-step and 8-step graphs
    b.step(chunk)
    snaps.append(np.concatenate([b.get(i) for i in range(len(sizes))]))
np.save(out, np.stack(snaps))
"""


def test_graph_replays_equal_eager_launches(g, gpu, tmp_path):
    """The captured RK4 steps (HIP graphs of 1 and 8 steps, replayed, host syncs and
    copies between the calls) give bit-identical states to eager launches (GSG_GRAPH=0)
    at every call."""
    import os
    import subprocess
    import sys
    from conftest import ROOT
    print("[gpu] graph replay vs eager", flush=True)
    script = tmp_path / "graph_case.py"
    script.write_text(_GRAPH_SCRIPT)
    res = {}
    for mode in ("1", "0"):
        env = dict(os.environ, GSG_GRAPH=mode)
        out = tmp_path / f"states_{mode}.npy"
        r = subprocess.run([sys.executable, str(script), str(ROOT), str(out)], env=env, capture_output=True,
                           text=True, timeout=300)
        assert r.returncode == 0, r.stderr[-2000:]
        res[mode] = np.load(out)
    assert res["1"].shape == res["0"].shape and res["1"].shape[0] == 4
    for i in range(4):
        np.testing.assert_array_equal(res["1"][i], res["0"][i])

"""Autotuning of the batched operator path's kernels and launch schedule.

Per grid (D, k, n, precision) `autotune` chooses the kernel variant: the tile variant
(direct, segmented or LDS fiber kernel, by shape; grids with D <= 6, k <= 5,
k^D <= 256 and D k^(D-1) <= 256), the scalar-plan gather kernel (every grid), and for
the FP32 segmented shapes the segmented kernel with the other number of coupled input
cells moved per load round ('tile_alt': 4 where the rule takes 2, 2 where it takes 4).
Each candidate runs inside real batched RK4 steps (transport systems, the kernel time
measured with HIP events around each launch) and the fastest is kept.

Per device `tune_batch` chooses how a batch's classes are launched: the number of
streams the per-class launches spread over (1, 2, 4 or 8) and the schedule (those
streams, or one fused launch per stage for every class), timed on whole batched RK4
steps of a mixed workload.

The choices are written to ``gsg_tune.txt`` beside libgsg.so: one line
``D k n precision variant`` per grid, ``lanes L`` and ``schedule S``; the library
reads that file at first use, so C, Julia and Python callers all get the tuned choice.
`set_variant` overrides one grid for the current process.  This has no counterpart in
the reference, which has a single SpMV path (SparseArrays' CSC product).
"""
from __future__ import annotations

import itertools
import time
from pathlib import Path

import numpy as np

from ._lib import LIBPATH, PREC, check, lib
from .solvers import Batch

AUTO, PLAN, TILE, TILE_ALT = 0, 1, 2, 3
VARIANT = {"auto": AUTO, "plan": PLAN, "tile": TILE, "tile_alt": TILE_ALT}
VARIANT_NAME = {AUTO: "auto", PLAN: "plan", TILE: "tile", TILE_ALT: "tile_alt"}
SCHEDULE = {"auto": 0, "streams": 1, "fused": 2}
TUNE_FILE = LIBPATH.parent / "gsg_tune.txt"


def alt_eligible(D: int, k: int, precision: str) -> bool:
    """Whether the segmented kernel (and so 'tile_alt') serves this shape."""
    return precision == "fp32" and D <= 4 and k <= 4 and tile_eligible(D, k) and k ** (D - 1) < 64


def tile_eligible(D: int, k: int) -> bool:
    """Whether the LDS fiber kernel can serve grids of this (D, k)."""
    return 1 <= D <= 6 and 1 <= k <= 5 and k ** D <= 256 and D * k ** (D - 1) <= 256


def set_variant(D: int, k: int, n: int, precision: str, variant: str) -> None:
    """Use `variant` ('plan', 'tile', 'tile_alt' or 'auto') for operators and batches
    created later."""
    check(lib().gsg_tune_set(D, k, n, PREC[precision], VARIANT[variant]))


def get_variant(D: int, k: int, n: int, precision: str) -> str:
    """The variant ('plan', 'tile' or 'tile_alt') that runs for this grid and precision."""
    return VARIANT_NAME[check(lib().gsg_tune_get(D, k, n, PREC[precision]))]


def set_lanes(lanes: int) -> None:
    """Streams (1..8, 0: the tuned count) for batches built later."""
    check(lib().gsg_tune_lanes(int(lanes)))


def set_schedule(schedule: str) -> None:
    """Process default schedule ('streams', 'fused', 'auto') for batches built later."""
    check(lib().gsg_tune_schedule(SCHEDULE[schedule]))


def _class_ms(precision, D, k, n, copies, steps, seed):
    """Average time of one stage launch of the (D, k) transport class over `steps`
    batched RK4 steps of `copies` independent solves."""
    from .layout import get_size

    rng = np.random.default_rng(seed)
    N = get_size(D, k, n)
    b = Batch(precision)
    for _ in range(copies):
        a = rng.uniform(-1.0, 1.0, D)
        b.add("transport", D, k, n, rng.standard_normal(N), 1e-3 * 2.0 ** -n, steps + 1, a=a)
    b.step(1)  # first step builds and uploads; untimed
    b.profile(True)
    b.step(steps)
    ms = sum(s["total_ms"] for s in b.kernel_stats()) / sum(s["launches"] for s in b.kernel_stats())
    b.close()
    return ms


def autotune(grids=None, precisions=("fp32", "fp64"), steps=6, min_dofs=3_000_000, path=TUNE_FILE,
             write=True, verbose=False):
    """Time the variants of every tile-eligible grid in `grids` ((D, k, n) tuples): plan,
    tile, and tile_alt where the segmented kernel serves the grid; keep the fastest.  Each
    measurement advances enough copies of a transport solve to cover at least `min_dofs`
    DoFs, so the launch fills the GPU as in a busy batch; tile_alt must also beat tile on
    one copy alone (a grid's entry also serves its single solves).

    Returns a list of dicts (D, k, n, precision, variant and the <variant>_ms timed).  With
    `write`, the table is written to `path` (default: beside libgsg.so), merged with
    the entries already there for other grids.
    """
    from .layout import get_size

    if grids is None:
        grids = [(D, k, n) for D, k in itertools.product((2, 3, 4), (2, 3, 4)) for n in range(4, 9)]
    rows = []
    for (D, k, n), prec in itertools.product(grids, precisions):
        if not tile_eligible(D, k):
            continue
        copies = max(1, -(-int(min_dofs) // get_size(D, k, n)))
        t = {}
        for v in ("plan", "tile") + (("tile_alt",) if alt_eligible(D, k, prec) else ()):
            set_variant(D, k, n, prec, v)
            t[v] = _class_ms(prec, D, k, n, copies, steps, seed=D * 100 + k * 10 + n)
        best = min(t, key=t.get)
        if best == "tile_alt":
            # the table is keyed by grid, and a grid also runs as a single solve: the other
            # round size is kept only if it also wins on one copy of the grid alone
            one = {}
            for v in ("tile", "tile_alt"):
                set_variant(D, k, n, prec, v)
                one[v] = _class_ms(prec, D, k, n, 1, steps, seed=D * 100 + k * 10 + n)
            t.update({f"{v}_single": ms for v, ms in one.items()})
            if one["tile_alt"] >= one["tile"]:
                best = "tile"
        set_variant(D, k, n, prec, best)
        rows.append(dict(D=D, k=k, n=n, precision=prec, variant=best, **{f"{v}_ms": ms for v, ms in t.items()}))
        if verbose:
            print(f"tune D={D} k={k} n={n} {prec}: " + ", ".join(f"{v} {ms:.4f} ms" for v, ms in t.items()) +
                  f" -> {best}", flush=True)
    if write:
        write_table(rows, path)
    return rows


def _batch_ms(solves, steps, warmup=3):
    """ms per batched FP32 RK4 step of `solves` ((D, k, n, a, u0, dt) tuples) under the
    current process settings."""
    b = Batch("fp32")
    for D, k, n, a, u0, dt in solves:
        b.add("transport", D, k, n, u0, dt, 1 << 40, a=a)
    b.step(1)            # builds, uploads, runs eagerly
    b.step(warmup)       # captures the step graphs
    t0 = time.perf_counter()
    b.step(steps)
    ms = 1e3 * (time.perf_counter() - t0) / steps
    b.close()
    return ms


def tune_batch(solves=None, nsolves=256, seed=0, steps=20, path=TUNE_FILE, write=True, verbose=False):
    """Time whole batched FP32 RK4 steps of a mixed workload (default: `nsolves` transport
    solves with (D, k, n) uniform on {2,3,4} x {2,3,4} x {4..8}, random-normal states) for
    1, 2, 4 and 8 streams and for the fused schedule, and keep the fastest lane count and
    schedule.  Returns {"lanes_ms": {...}, "schedule_ms": {...}, "lanes": L, "schedule": S}."""
    from .layout import get_size
    if solves is None:
        rng = np.random.default_rng(seed)
        solves = []
        for _ in range(nsolves):
            D, k, n = int(rng.integers(2, 5)), int(rng.integers(2, 5)), int(rng.integers(4, 9))
            a = rng.uniform(-1, 1, D)
            solves.append((D, k, n, a, rng.standard_normal(get_size(D, k, n)),
                           0.1 * 2.0 ** -n / max(float(np.abs(a).sum()), 1e-3)))
    set_schedule("streams")
    lanes_ms = {}
    for nl in (1, 2, 4, 8):
        set_lanes(nl)
        lanes_ms[nl] = _batch_ms(solves, steps)
        if verbose:
            print(f"tune lanes={nl}: {lanes_ms[nl]:.4f} ms/step", flush=True)
    # more streams only for a clear win (3%): within that margin the difference is noise
    # on one workload and a workload with another class mix can lose 9% (8 streams
    # against 4 on the bench's 256-solve draw)
    best_lanes = min(lanes_ms)
    for nl in sorted(lanes_ms):
        if lanes_ms[nl] < 0.97 * lanes_ms[best_lanes]:
            best_lanes = nl
    set_lanes(best_lanes)
    sched_ms = {"streams": lanes_ms[best_lanes]}
    set_schedule("fused")
    sched_ms["fused"] = _batch_ms(solves, steps)
    best_sched = min(sched_ms, key=sched_ms.get)
    set_schedule(best_sched)
    if verbose:
        print(f"tune schedule: {sched_ms} -> {best_sched}, lanes {best_lanes}", flush=True)
    out = dict(lanes_ms=lanes_ms, schedule_ms=sched_ms, lanes=best_lanes, schedule=best_sched)
    if write:
        write_table([], path, lanes=best_lanes, schedule=best_sched)
    return out


def read_settings(path=TUNE_FILE):
    """The file's "lanes L" and "schedule S" lines as {"lanes": L, "schedule": S}."""
    out = {}
    p = Path(path)
    if p.exists():
        for line in p.read_text().splitlines():
            f = line.split("#", 1)[0].split()
            if len(f) == 2 and f[0] in ("lanes", "schedule"):
                out[f[0]] = int(f[1])
    return out


def read_table(path=TUNE_FILE):
    """Grid entries of a tuning file as {(D, k, n, precision): variant}."""
    out = {}
    p = Path(path)
    if not p.exists():
        return out
    for line in p.read_text().splitlines():
        f = line.split("#", 1)[0].split()
        if len(f) == 5:
            D, k, n, prec, var = map(int, f)
            out[(D, k, n, "fp32" if prec == 1 else "fp64")] = VARIANT_NAME[var]
    return out


def write_table(rows, path=TUNE_FILE, lanes=None, schedule=None):
    """Write (merge) autotune rows, and the lane count and schedule, into the tuning file
    read by libgsg."""
    table = read_table(path)
    settings = read_settings(path)
    if lanes is not None:
        settings["lanes"] = int(lanes)
    if schedule is not None:
        settings["schedule"] = SCHEDULE[schedule] if isinstance(schedule, str) else int(schedule)
    notes = {}
    for r in rows:
        key = (r["D"], r["k"], r["n"], r["precision"])
        table[key] = r["variant"]
        ms = [f"{v.replace('_single', ' alone')} {r[v + '_ms']:.4f} ms"
              for v in ("plan", "tile", "tile_alt", "tile_single", "tile_alt_single") if v + "_ms" in r]
        if ms:
            notes[key] = ", ".join(ms)
    lines = ["# D k n precision(0=fp64,1=fp32) variant(1=plan,2=tile,3=tile_alt); lanes L; "
             "schedule S(1=streams,2=fused)  -- written by tune.autotune / tune.tune_batch",
             f"# {time.strftime('%Y-%m-%d %H:%M:%S')}"]
    for key in ("lanes", "schedule"):
        if key in settings:
            lines.append(f"{key} {settings[key]}")
    for key in sorted(table):
        D, k, n, prec = key
        note = f"  # {notes[key]}" if key in notes else ""
        lines.append(f"{D} {k} {n} {PREC[prec]} {VARIANT[table[key]]}{note}")
    Path(path).write_text("\n".join(lines) + "\n")


__all__ = ["autotune", "tune_batch", "set_variant", "get_variant", "set_lanes", "set_schedule", "tile_eligible",
           "alt_eligible", "read_table", "read_settings", "write_table", "TUNE_FILE"]

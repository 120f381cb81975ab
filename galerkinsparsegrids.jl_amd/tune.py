"""Kernel-variant autotuning for the batched operator path.

Two gfx950 kernels apply the sparse-grid DG operator: the LDS fiber kernel (grids with
D <= 6, k <= 5, k^D <= 256 and D k^(D-1) <= 256) and the scalar-plan gather kernel
(every grid).  Which one is faster depends on (D, k, n) and the precision, so
`autotune` times both inside real batched RK4 steps (transport systems, the kernel time
measured with HIP events around each launch) and records the winner per grid.

The winners are written to ``gsg_tune.txt`` beside libgsg.so, one line
``D k n precision variant`` per grid; the library reads that file at first use, so C,
Julia and Python callers all get the tuned choice.  `set_variant` overrides one grid
for the current process.  This has no counterpart in the reference, which has a single
SpMV path (SparseArrays' CSC product).
"""
from __future__ import annotations

import itertools
import time
from pathlib import Path

import numpy as np

from ._lib import LIBPATH, PREC, check, lib
from .solvers import Batch

AUTO, PLAN, TILE = 0, 1, 2
VARIANT = {"auto": AUTO, "plan": PLAN, "tile": TILE}
VARIANT_NAME = {AUTO: "auto", PLAN: "plan", TILE: "tile"}
TUNE_FILE = LIBPATH.parent / "gsg_tune.txt"


def tile_eligible(D: int, k: int) -> bool:
    """Whether the LDS fiber kernel can serve grids of this (D, k)."""
    return 1 <= D <= 6 and 1 <= k <= 5 and k ** D <= 256 and D * k ** (D - 1) <= 256


def set_variant(D: int, k: int, n: int, precision: str, variant: str) -> None:
    """Use `variant` ('plan', 'tile' or 'auto') for operators and batches created later."""
    check(lib().gsg_tune_set(D, k, n, PREC[precision], VARIANT[variant]))


def get_variant(D: int, k: int, n: int, precision: str) -> str:
    """The variant ('plan' or 'tile') that runs for this grid and precision."""
    return VARIANT_NAME[check(lib().gsg_tune_get(D, k, n, PREC[precision]))]


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
    """Time both kernels for every tile-eligible grid in `grids` ((D, k, n) tuples) and
    keep the faster one.  Each measurement advances enough copies of a transport solve
    to cover at least `min_dofs` DoFs, so the launch fills the GPU as in a busy batch.

    Returns a list of dicts (D, k, n, precision, plan_ms, tile_ms, variant).  With
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
        for v in ("plan", "tile"):
            set_variant(D, k, n, prec, v)
            t[v] = _class_ms(prec, D, k, n, copies, steps, seed=D * 100 + k * 10 + n)
        best = min(t, key=t.get)
        set_variant(D, k, n, prec, best)
        rows.append(dict(D=D, k=k, n=n, precision=prec, plan_ms=t["plan"], tile_ms=t["tile"], variant=best))
        if verbose:
            print(f"tune D={D} k={k} n={n} {prec}: plan {t['plan']:.4f} ms, tile {t['tile']:.4f} ms -> {best}",
                  flush=True)
    if write:
        write_table(rows, path)
    return rows


def read_table(path=TUNE_FILE):
    """Entries of a tuning file as {(D, k, n, precision): variant}."""
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


def write_table(rows, path=TUNE_FILE):
    """Write (merge) autotune rows into the tuning file read by libgsg."""
    table = read_table(path)
    notes = {}
    for r in rows:
        key = (r["D"], r["k"], r["n"], r["precision"])
        table[key] = r["variant"]
        if "plan_ms" in r:
            notes[key] = f"plan {r['plan_ms']:.4f} ms, tile {r['tile_ms']:.4f} ms"
    lines = ["# D k n precision(0=fp64,1=fp32) variant(1=plan,2=tile)  -- written by tune.autotune",
             f"# {time.strftime('%Y-%m-%d %H:%M:%S')}"]
    for key in sorted(table):
        D, k, n, prec = key
        note = f"  # {notes[key]}" if key in notes else ""
        lines.append(f"{D} {k} {n} {PREC[prec]} {VARIANT[table[key]]}{note}")
    Path(path).write_text("\n".join(lines) + "\n")


__all__ = ["autotune", "set_variant", "get_variant", "tile_eligible", "read_table", "write_table", "TUNE_FILE"]

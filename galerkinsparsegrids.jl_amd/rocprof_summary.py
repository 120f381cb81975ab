#!/usr/bin/env python
"""Summaries of rocprofv3 output for the operator kernels.

    python galerkinsparsegrids.jl_amd/rocprof_summary.py copy
        a device-to-device copy of 256 MiB, five times: the calibration run for the
        FETCH_SIZE / WRITE_SIZE counters (profile it with --pmc like the workload)
    python galerkinsparsegrids.jl_amd/rocprof_summary.py vlasov [--steps 50] [fp32] [fp64]
        the Vlasov workload to profile: free streaming on the 2 + 2 dimensional phase
        space, k = 2, n = 8, f0 = Maxwellian in v times (1 + 0.1 cos 2 pi x1), RK4 steps
        on the device
    python galerkinsparsegrids.jl_amd/rocprof_summary.py projection [--n 9]
        the initial-data projection to profile: vcoeffs_DG on the device at D = 4, k = 4
        (default rule q = 8, 16^4 nodes per cell) of a non-separable f evaluated by
        torch, once at n = 5 to warm up, then at n = --n
    python galerkinsparsegrids.jl_amd/rocprof_summary.py stats RUN_kernel_stats.csv
        per-kernel call count and average duration, the RK-stage instantiations of one
        (kernel, D, k, type) folded together (markdown table on stdout)
    python galerkinsparsegrids.jl_amd/rocprof_summary.py traffic --kernel 'k_direct<4,4,float>' \\
            --fetch F.csv --write W.csv --copy-fetch CF.csv --copy-write CW.csv [--json OUT]
        HBM bytes per launch of one kernel: the per-dispatch FETCH_SIZE and WRITE_SIZE
        (KiB) averaged over its dispatches, each scaled by what the counter reads for the
        copy kernel of known size

    python galerkinsparsegrids.jl_amd/rocprof_summary.py pmc RUN_counter_collection.csv \
            --kernel 'k_direct<4,4,double>' [--kernel ...] [--cus 256] [--xcds 8]
        per kernel, the average per dispatch of every counter in the file, and for
        TA_TA_BUSY_sum / TD_TD_BUSY_sum the busy fraction: the sum over the CUs'
        texture units divided by (GRBM_GUI_ACTIVE / XCDs) x CUs; GRBM_GUI_ACTIVE is
        summed over the 8 XCDs (markdown table on stdout)

Kernel names in the CSVs look like `void gsg::(anonymous namespace)::k_direct<4, 4,
float, 2, 1>(gsg::LaunchArgs)`; a kernel is named here as `k_direct<4,4,float>`.
"""
from __future__ import annotations

import argparse
import collections
import csv
import json
import re
import sys

COPY_BYTES = 256 << 20


def short_name(full: str) -> str:
    m = re.search(r"(k_[a-z0-9_]+)<([^>]*)>", full)
    if not m:
        m = re.search(r"(k_[a-z0-9_]+)\(", full)
        return m.group(1) if m else full.split("(")[0].strip()
    args = [a.strip() for a in m.group(2).split(",")]
    keep = 2 if m.group(1) == "k_seg" else 3  # (D, k[, type]); the op / stage arguments fold
    return f"{m.group(1)}<{','.join(args[:keep])}>"


def cmd_copy(_):
    import torch
    x = torch.randn(COPY_BYTES // 4, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(f"copied {COPY_BYTES} bytes 5 times", flush=True)


def cmd_vlasov(a):
    import time
    from pathlib import Path

    import numpy as np
    import torch
    sys.path.insert(0, str(Path(__file__).resolve().parent.parent))
    import __graft_entry__ as ge
    g = ge.load_package()
    torch.cuda.set_device(0)
    k, n = 2, 8
    fx = g.vcoeffs_1d(k, n, lambda x: 1 + 0.1 * np.cos(2 * np.pi * x))
    one = g.vcoeffs_1d(k, n, lambda x: np.ones_like(x))
    mv = g.vcoeffs_1d(k, n, lambda v: np.exp(-0.5 * ((v - 0.5) / 0.1) ** 2) / (0.1 * np.sqrt(2 * np.pi)))
    f0 = g.tensor_construct(4, k, n, [fx, one, mv, mv])
    for prec in a.precision or ["fp32"]:
        op = g.vlasov_operator(2, k, n, None, precision=prec)
        y = torch.tensor(f0, device="cuda:0")
        op.rk4(y, 0.05 * 2.0 ** -n, 2)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        op.rk4(y, 0.05 * 2.0 ** -n, a.steps)
        torch.cuda.synchronize()
        print(f"vlasov {prec}: {1e3 * (time.perf_counter() - t0) / a.steps:.3f} ms per RK4 step", flush=True)
        op.close()


def cmd_projection(a):
    import math
    import time
    from pathlib import Path

    import torch
    sys.path.insert(0, str(Path(__file__).resolve().parent.parent))
    import __graft_entry__ as ge
    g = ge.load_package()
    torch.cuda.set_device(0)

    def f(X):
        return torch.cos(2 * math.pi * (X[0] + 2 * X[1] - X[2] + X[3]) + 0.3) * torch.exp(-(X[0] - 0.5) ** 2 - (X[3] - 0.5) ** 2)
    for n in (5, a.n):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        y = g.vcoeffs_DG_device(4, 4, n, f, device="cuda")
        torch.cuda.synchronize()
        print(f"projection D=4 k=4 n={n}: {y.numel()} coefficients, {y.numel() // 256 * 16 ** 4} nodes, "
              f"{time.perf_counter() - t0:.4f} s", flush=True)


def cmd_stats(a):
    agg = collections.OrderedDict()
    with open(a.csv) as f:
        for r in csv.DictReader(f):
            k = short_name(r["Name"])
            calls = int(r["Calls"])
            tot = float(r["TotalDurationNs"])
            c = agg.setdefault(k, [0, 0.0])
            c[0] += calls
            c[1] += tot
    total = sum(v[1] for v in agg.values())
    print("| kernel | calls | total ms | average us | share |")
    print("|---|---:|---:|---:|---:|")
    for k, (calls, tot) in sorted(agg.items(), key=lambda kv: -kv[1][1]):
        print(f"| `{k}` | {calls} | {tot / 1e6:.3f} | {tot / calls / 1e3:.2f} | {100 * tot / total:.1f}% |")


def _per_dispatch(path, counter, kernel=None):
    vals = []
    with open(path) as f:
        for r in csv.DictReader(f):
            if r["Counter_Name"] != counter:
                continue
            if kernel is not None and short_name(r["Kernel_Name"]) != kernel:
                continue
            vals.append(float(r["Counter_Value"]))
    return vals


def cmd_traffic(a):
    fetch = _per_dispatch(a.fetch, "FETCH_SIZE", a.kernel)
    write = _per_dispatch(a.write, "WRITE_SIZE", a.kernel)
    if not fetch or not write:
        sys.exit(f"no dispatches of {a.kernel} in the counter files")
    # calibration: the largest copy dispatch (the 256 MiB copies)
    cf = max(_per_dispatch(a.copy_fetch, "FETCH_SIZE"))
    cw = max(_per_dispatch(a.copy_write, "WRITE_SIZE"))
    kib = COPY_BYTES / 1024.0
    sf, sw = kib / cf, kib / cw
    fb = sum(fetch) / len(fetch) * 1024.0 * sf
    wb = sum(write) / len(write) * 1024.0 * sw
    out = {"kernel": a.kernel, "dispatches": len(fetch), "fetch_scale": sf, "write_scale": sw,
           "fetch_bytes_per_launch": fb, "write_bytes_per_launch": wb, "bytes_per_launch": fb + wb}
    if a.json:
        extra = json.loads(a.extra) if a.extra else {}
        out.update(extra)
        with open(a.json, "w") as f:
            json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


def cmd_pmc(a):
    per = collections.defaultdict(lambda: collections.defaultdict(list))
    with open(a.csv) as f:
        for r in csv.DictReader(f):
            k = short_name(r["Kernel_Name"])
            if k in a.kernel:
                per[k][r["Counter_Name"]].append(float(r["Counter_Value"]))
    counters = sorted({c for k in per for c in per[k]})
    print("| kernel | dispatches | " + " | ".join(counters) + " |")
    print("|---|---:|" + "---:|" * len(counters))
    for k in a.kernel:
        if k not in per:
            continue
        n = max(len(v) for v in per[k].values())
        avg = {c: sum(v) / len(v) for c, v in per[k].items()}
        print(f"| `{k}` | {n} | " + " | ".join(f"{avg[c]:.4g}" if c in avg else "" for c in counters) + " |")
    gui = "GRBM_GUI_ACTIVE"
    for k in a.kernel:
        if k not in per or gui not in per[k]:
            continue
        avg = {c: sum(v) / len(v) for c, v in per[k].items()}
        for c in ("TA_TA_BUSY_sum", "TD_TD_BUSY_sum"):
            if c in avg:
                cycles = avg[gui] / a.xcds
                print(f"{k}: {c} / (({gui} / {a.xcds}) x {a.cus}) = {avg[c] / (cycles * a.cus):.3f}")


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    sub.add_parser("copy")
    v = sub.add_parser("vlasov")
    v.add_argument("--steps", type=int, default=50)
    v.add_argument("precision", nargs="*", choices=["fp32", "fp64"])
    pj = sub.add_parser("projection")
    pj.add_argument("--n", type=int, default=9)
    s = sub.add_parser("stats")
    s.add_argument("csv")
    t = sub.add_parser("traffic")
    t.add_argument("--kernel", required=True)
    t.add_argument("--fetch", required=True)
    t.add_argument("--write", required=True)
    t.add_argument("--copy-fetch", required=True)
    t.add_argument("--copy-write", required=True)
    t.add_argument("--json")
    t.add_argument("--extra", help="JSON object merged into the output (e.g. the workload)")
    m = sub.add_parser("pmc")
    m.add_argument("csv")
    m.add_argument("--kernel", action="append", required=True)
    m.add_argument("--cus", type=int, default=256)
    m.add_argument("--xcds", type=int, default=8)
    a = ap.parse_args()
    {"copy": cmd_copy, "vlasov": cmd_vlasov, "projection": cmd_projection, "stats": cmd_stats,
     "traffic": cmd_traffic, "pmc": cmd_pmc}[a.cmd](a)


if __name__ == "__main__":
    main()

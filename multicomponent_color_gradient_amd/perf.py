"""Offline MLUPS / roofline reporter over rocprofv3 CSV output.

Reads the files ``rocprofv3 --output-format csv`` writes:
  * ``*_kernel_stats.csv`` (``--kernel-trace --stats``): per-kernel call
    count and durations;
  * ``*_counter_collection.csv``: one row per dispatch and counter, from
    ``--pmc`` of the sized L2 request counters (READ_COUNTERS, WRITE_COUNTERS:
    exact bytes) or of FETCH_SIZE / WRITE_SIZE (KiB, scaled by a probe copy);
  * ``*_kernel_trace.csv`` (``--kernel-trace``): one row per dispatch with its
    grid, which separates the launches of one kernel on different grids.
and turns them into the figures the benchmark reports: bytes per launch
(scaled by a plain copy of known size, see ``copy_probe_scale``), bytes per
lattice update, GB/s and the fraction of the HBM peak.  MLUPS follows
utilities/countmlups.cuh:42-48 (steps * NX*NY*NZ / t / 1e6).

Command line:
  python -m multicomponent_color_gradient_amd.perf stats <kernel_stats.csv>
  python -m multicomponent_color_gradient_amd.perf trace <kernel_trace.csv>
  python -m multicomponent_color_gradient_amd.perf requests <read.csv> <write.csv> --kernel k_fused --nodes N
  python -m multicomponent_color_gradient_amd.perf traffic <fetch.csv> <write.csv> --kernel k_fused --nodes N
"""
from __future__ import annotations

import argparse
import csv
import sys
from collections import defaultdict

HBM_PEAK_GBS = 8000.0
COPY_PROBE_BYTES = 1 << 30
COPY_KERNEL = "__amd_rocclr_copyBuffer"


def mlups(steps: int, cells: int, seconds: float) -> float:
    """Million lattice updates per second (utilities/countmlups.cuh:42-48)."""
    return steps * cells / seconds / 1e6


def roofline(bytes_per_update: float, updates: int, seconds: float, peak_gbs: float = HBM_PEAK_GBS) -> dict:
    """Achieved GB/s of bytes_per_update x updates in seconds, and its fraction of peak_gbs."""
    gbs = bytes_per_update * updates / seconds / 1e9
    return {"achieved": gbs, "peak": peak_gbs, "unit": "GB/s", "frac": gbs / peak_gbs}


def read_kernel_stats(path: str) -> list[dict]:
    """Rows of a kernel_stats.csv: name, calls, total/avg/min/max ns, percentage."""
    out = []
    with open(path, newline="") as fh:
        for r in csv.DictReader(fh):
            out.append({"name": r["Name"], "calls": int(r["Calls"]), "total_ns": float(r["TotalDurationNs"]),
                        "avg_ns": float(r["AverageNs"]), "min_ns": float(r["MinNs"]), "max_ns": float(r["MaxNs"]),
                        "percent": float(r["Percentage"])})
    return out


def trace_stats(path: str) -> list[dict]:
    """Per (kernel, grid size in work-items) from a kernel_trace.csv: calls and
    average / min / max duration in ns, in order of first dispatch."""
    groups: dict = {}
    with open(path, newline="") as fh:
        for r in csv.DictReader(fh):
            grid = int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"])
            d = int(r["End_Timestamp"]) - int(r["Start_Timestamp"])
            key = (r["Kernel_Name"], grid)
            g = groups.setdefault(key, {"name": key[0], "grid": grid, "first": int(r["Dispatch_Id"]), "calls": 0,
                                        "total_ns": 0.0, "min_ns": float("inf"), "max_ns": 0.0})
            g["calls"] += 1
            g["total_ns"] += d
            g["min_ns"] = min(g["min_ns"], d)
            g["max_ns"] = max(g["max_ns"], d)
    out = sorted(groups.values(), key=lambda g: g["first"])
    for g in out:
        g["avg_ns"] = g["total_ns"] / g["calls"]
    return out


def trace_table(rows: list[dict]) -> str:
    lines = ["| kernel | grid (work-items) | calls | avg us | min us | max us |", "|---|---|---|---|---|---|"]
    for r in rows:
        lines.append(f"| `{r['name'][:70]}` | {r['grid']} | {r['calls']} | {r['avg_ns'] / 1e3:.1f} | "
                     f"{r['min_ns'] / 1e3:.1f} | {r['max_ns'] / 1e3:.1f} |")
    return "\n".join(lines)


def read_counters(path: str, grid: int | None = None) -> dict:
    """{kernel name: {dispatch id: {counter: value summed over the rows of that dispatch}}},
    optionally only the dispatches on a grid of ``grid`` work-items (one workload of a
    run that launches the same kernel on several grids)."""
    out: dict = defaultdict(lambda: defaultdict(lambda: defaultdict(float)))
    with open(path, newline="") as fh:
        for r in csv.DictReader(fh):
            if grid is not None and int(r["Grid_Size"]) != grid and not r["Kernel_Name"].startswith(COPY_KERNEL):
                continue
            out[r["Kernel_Name"]][int(r["Dispatch_Id"])][r["Counter_Name"]] += float(r["Counter_Value"])
    return out


def per_dispatch(counters: dict, kernel: str, counter: str) -> list[float]:
    """Values of one counter for every dispatch of the kernels whose name contains ``kernel``, in dispatch order."""
    vals = []
    for name, disp in counters.items():
        if kernel in name:
            vals += [(d, c[counter]) for d, c in disp.items() if counter in c]
    return [v for _, v in sorted(vals)]


def copy_probe_scale(counters: dict, counter: str, nbytes: int = COPY_PROBE_BYTES, min_kib: float = 1e5) -> float:
    """Bytes per counter unit, from the probe's device copies of ``nbytes``.

    The probe (``bench.py --copy-probe-mib``) runs three identical copies; they
    are the first three copy dispatches whose counter exceeds ``min_kib``
    (smaller copies before them belong to the solver's set-up, larger ones
    after them to the adjoint).  A copy reads and writes exactly nbytes, so
    the scale is nbytes / counter."""
    vals = [v for v in per_dispatch(counters, COPY_KERNEL, counter) if v > min_kib][:3]
    if len(vals) != 3 or max(vals) - min(vals) > 1e-3 * max(vals):
        raise ValueError(f"copy probe not found in the {counter} dispatches: {vals}")
    return nbytes / (sum(vals) / 3)


def traffic(fetch: dict, write: dict, kernel: str, nodes: int | None = None) -> dict:
    """Scaled HBM bytes per launch of ``kernel`` from a FETCH_SIZE and a WRITE_SIZE collection."""
    sf = copy_probe_scale(fetch, "FETCH_SIZE")
    sw = copy_probe_scale(write, "WRITE_SIZE")
    f = per_dispatch(fetch, kernel, "FETCH_SIZE")
    w = per_dispatch(write, kernel, "WRITE_SIZE")
    if not f or not w:
        raise ValueError(f"no dispatches of {kernel}")
    rd = sum(f) / len(f) * sf
    wr = sum(w) / len(w) * sw
    out = {"kernel": kernel, "dispatches": len(f), "read_bytes": rd, "write_bytes": wr, "bytes_per_launch": rd + wr,
           "fetch_scale": sf * 1 / 1024, "write_scale": sw * 1 / 1024}
    if nodes:
        out["read_per_node"] = rd / nodes
        out["write_per_node"] = wr / nodes
    return out


# Request-size counters of the L2 -> memory interface (one rocprofv3 --pmc pass
# each for reads and writes, 4 and 2 TCC counters).  On gfx950 the derived
# FETCH_SIZE expects 128-byte reads in TCC_BUBBLE, which stays 0, so it counts
# every 128-byte read as 64 bytes; summing the sized requests needs no scale.
READ_COUNTERS = ("TCC_EA0_RDREQ_32B_sum", "TCC_EA0_RDREQ_64B_sum", "TCC_EA0_RDREQ_128B_sum", "TCC_EA0_RDREQ_sum")
WRITE_COUNTERS = ("TCC_EA0_WRREQ_64B_sum", "TCC_EA0_WRREQ_sum")


def read_bytes(c: dict) -> float:
    """Bytes read by one dispatch from its sized read requests."""
    return 32.0 * c["TCC_EA0_RDREQ_32B_sum"] + 64.0 * c["TCC_EA0_RDREQ_64B_sum"] + 128.0 * c["TCC_EA0_RDREQ_128B_sum"]


def write_bytes(c: dict) -> float:
    """Bytes written by one dispatch: 64-byte requests, the rest 32-byte."""
    n64 = c["TCC_EA0_WRREQ_64B_sum"]
    return 64.0 * n64 + 32.0 * (c["TCC_EA0_WRREQ_sum"] - n64)


def _dispatch_values(counters: dict, kernel: str, fn, need) -> list[float]:
    vals = []
    for name, disp in counters.items():
        if kernel in name:
            vals += [(d, fn(c)) for d, c in disp.items() if all(k in c for k in need)]
    return [v for _, v in sorted(vals)]


def probe_copy_check(rd: dict, wr: dict, nbytes: int = COPY_PROBE_BYTES, tol: float = 0.01) -> dict:
    """The probe's copies of ``nbytes`` (``bench.py --copy-probe-mib``): the
    dispatches of the copy kernel that read within ``tol`` of nbytes.  Each
    must also write within ``tol`` of nbytes; returns their read/write ratios."""
    reads = [b for b in _dispatch_values(rd, COPY_KERNEL, read_bytes, READ_COUNTERS[:3]) if abs(b / nbytes - 1) < tol]
    writes = [b for b in _dispatch_values(wr, COPY_KERNEL, write_bytes, WRITE_COUNTERS) if abs(b / nbytes - 1) < tol]
    if not reads or len(writes) != len(reads):
        raise ValueError(f"probe copies of {nbytes} bytes not found: reads {reads}, writes {writes}")
    return {"copies": len(reads), "read_ratio": sum(reads) / len(reads) / nbytes,
            "write_ratio": sum(writes) / len(writes) / nbytes}


def request_traffic(rd: dict, wr: dict, kernel: str, nodes: int | None = None) -> dict:
    """Bytes per launch of ``kernel`` from the sized-request counters (a read
    and a write collection), with the probe copies as the check of the counting."""
    check = probe_copy_check(rd, wr)
    r = _dispatch_values(rd, kernel, read_bytes, READ_COUNTERS[:3])
    w = _dispatch_values(wr, kernel, write_bytes, WRITE_COUNTERS)
    if not r or not w:
        raise ValueError(f"no dispatches of {kernel}")
    rb, wb = sum(r) / len(r), sum(w) / len(w)
    out = {"kernel": kernel, "dispatches": len(r), "read_bytes": rb, "write_bytes": wb, "bytes_per_launch": rb + wb,
           "probe_copy": check}
    if nodes:
        out["read_per_node"] = rb / nodes
        out["write_per_node"] = wb / nodes
    return out


def stats_table(rows: list[dict]) -> str:
    lines = ["| kernel | calls | avg us | min us | max us | % |", "|---|---|---|---|---|---|"]
    for r in rows:
        lines.append(f"| `{r['name'][:70]}` | {r['calls']} | {r['avg_ns'] / 1e3:.1f} | {r['min_ns'] / 1e3:.1f} | "
                     f"{r['max_ns'] / 1e3:.1f} | {r['percent']:.2f} |")
    return "\n".join(lines)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m multicomponent_color_gradient_amd.perf")
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("stats")
    a.add_argument("csv")
    t = sub.add_parser("trace")
    t.add_argument("csv")
    b = sub.add_parser("traffic")
    b.add_argument("fetch_csv")
    b.add_argument("write_csv")
    b.add_argument("--kernel", required=True)
    b.add_argument("--nodes", type=int, default=None)
    q = sub.add_parser("requests")
    q.add_argument("read_csv")
    q.add_argument("write_csv")
    q.add_argument("--kernel", required=True)
    q.add_argument("--nodes", type=int, default=None)
    q.add_argument("--grid", type=int, default=None, help="only dispatches on a grid of this many work-items")
    args = ap.parse_args(argv)
    if args.cmd == "stats":
        print(stats_table(read_kernel_stats(args.csv)))
    elif args.cmd == "trace":
        print(trace_table(trace_stats(args.csv)))
    elif args.cmd == "requests":
        t = request_traffic(read_counters(args.read_csv, args.grid), read_counters(args.write_csv, args.grid),
                            args.kernel, args.nodes)
        for k, v in t.items():
            print(f"{k}: {v}")
    else:
        t = traffic(read_counters(args.fetch_csv), read_counters(args.write_csv), args.kernel, args.nodes)
        for k, v in t.items():
            print(f"{k}: {v}")
    return 0


if __name__ == "__main__":
    sys.exit(main())

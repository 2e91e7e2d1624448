"""Mean profiles of the round jet under each population storage, and their
differences: FP16 and BF16 offsets against FP32 on the same jet.

``run_profiles(case, storage, steps)`` steps one solver, with the radial
statistics accumulated from ``case.NSTATS_START_STEP`` every
``case.NSTATS_SAMPLE`` steps (kernels.cu:188-233, main.cu:56-62). It returns
the mean fields per (y, r-bin) (post/mean_profile_io.py:182-210), the FP64
totals (red mass, blue mass, momentum) and the device time per step.

``compare(a, b)`` gives the errors of run ``a`` against run ``b``:
  uy          max over sampled bins |<u_y>_a - <u_y>_b| / max |<u_y>_b|
  centreline  the same over the r-bin 0 of every y (the jet axis)
  phi         max over sampled bins |<phi>_a - <phi>_b| (phi is O(1))
  mass        |M_a - M_b| / M_b, M = sum (rho_r + rho_b) in FP64

Command line (one GPU):
  python -m multicomponent_color_gradient_amd.jet_profiles [--NX 256 --NY 512 --NZ 256 --radius 16
      --steps 40000 --stats-start 10000 --stats-every 10 --storages fp32,fp16,bf16 --out DIR]
prints a progress line per chunk of steps on stderr and one JSON line with,
per storage, ms/step, MLUPS and its errors against the first storage named.
``--out`` also saves each run's radial sums as ``<storage>.npz``.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from .case import McgCase, perturbed_jet_case
from .post import averaged_radial_fields
from .solver import Solver


def run_profiles(case: McgCase, storage: str, steps: int, chunk: int = 2000, log=None) -> dict:
    """Jet initialisation, ``steps`` steps; radial sums, mean fields, totals and timing."""
    s = Solver(case, storage=storage)
    try:
        s.init_jet()
        ms_total = 0.0
        done = 0
        t0 = time.perf_counter()
        while done < steps:
            n = min(chunk, steps - done)
            ms_total += s.time_steps(n)
            done += n
            if log is not None:
                print(f"{storage}: step {done}/{steps} ({time.perf_counter() - t0:.1f} s)", file=log, flush=True)
        s.check_health()
        radial = s.radial()
        info = s.stats_info()
        sums = s.global_sums()
    finally:
        s.close()
    return {"storage": storage, "radial": radial, "mean": averaged_radial_fields(radial), "samples": info["samples"],
            "sums": sums, "ms_per_step": ms_total / max(steps, 1)}


def compare(a: dict, b: dict) -> dict:
    """Errors of run ``a`` against run ``b`` (both from ``run_profiles``)."""
    ca, cb = a["radial"]["count"], b["radial"]["count"]
    if not np.array_equal(ca, cb):
        raise ValueError("the runs sampled different bins")
    mask = cb > 0
    if not mask.any():
        raise ValueError("no statistics samples")
    ua, ub = a["mean"]["uy"], b["mean"]["uy"]
    scale = float(np.max(np.abs(ub[mask])))
    axis = mask[:, 0]
    out = {"uy": float(np.max(np.abs(ua[mask] - ub[mask])) / scale),
           "centreline": float(np.max(np.abs(ua[axis, 0] - ub[axis, 0])) / scale) if axis.any() else None}
    if "phi" in a["mean"] and "phi" in b["mean"]:
        out["phi"] = float(np.max(np.abs(a["mean"]["phi"][mask] - b["mean"]["phi"][mask])))
    ma, mb = float(a["sums"][0] + a["sums"][1]), float(b["sums"][0] + b["sums"][1])
    out["mass"] = abs(ma - mb) / mb
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m multicomponent_color_gradient_amd.jet_profiles")
    ap.add_argument("--NX", type=int, default=256)
    ap.add_argument("--NY", type=int, default=512)
    ap.add_argument("--NZ", type=int, default=256)
    ap.add_argument("--radius", type=float, default=16.0)
    ap.add_argument("--steps", type=int, default=40000)
    ap.add_argument("--stats-start", type=int, default=10000)
    ap.add_argument("--stats-every", type=int, default=10)
    ap.add_argument("--storages", default="fp32,fp16,bf16")
    ap.add_argument("--chunk", type=int, default=2000, help="steps between progress lines")
    ap.add_argument("--out", default="", help="directory for <storage>.npz radial sums")
    args = ap.parse_args(argv)
    case = perturbed_jet_case(args.NX, args.NY, args.NZ, jet_radius=args.radius)
    case.NSTATS_START_STEP = args.stats_start
    case.NSTATS_SAMPLE = args.stats_every
    storages = [x for x in args.storages.split(",") if x]
    runs = []
    for st in storages:
        r = run_profiles(case, st, args.steps, args.chunk, log=sys.stderr)
        if args.out:
            os.makedirs(args.out, exist_ok=True)
            np.savez(os.path.join(args.out, f"{st}.npz"), **r["radial"], sums=r["sums"])
        runs.append(r)
    cells = case.NX * case.NY * case.NZ
    out = {"case": f"perturbed jet {case.NX}x{case.NY}x{case.NZ}, R={case.jet_radius:g}, U={case.jet_velocity:g}, "
                   f"tau {case.taur:g}/{case.taub:g}, A={case.A_override:g}, beta={case.beta_recolor:g}",
           "steps": args.steps, "stats_start": args.stats_start, "stats_every": args.stats_every, "runs": {}}
    for r in runs:
        e = {"ms_per_step": round(r["ms_per_step"], 5), "mlups": round(cells / r["ms_per_step"] / 1e3, 2),
             "samples": r["samples"], "mass": float(r["sums"][0] + r["sums"][1])}
        if r is not runs[0]:
            e[f"vs_{runs[0]['storage']}"] = compare(r, runs[0])
        out["runs"][r["storage"]] = e
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

"""The reference's own jet case (constants.cuh: 128x400x128, D = 14, U = 0.05,
Re 5000, We 2000, beta 0.8, sharp inlet, radial statistics every step from
step 6000) under a chosen storage and with optional noise, and its centre-line
summary: the numbers of the reference's resultsJET/profile_summary.txt
(virtual origin and spreading rates, post/plot_profiles.py:276-308).

At tau = 0.50042 the jet's transition is seeded by whatever noise the run
carries, so this module perturbs the red populations f_r,i by the factor
1 + amp u, u uniform in [-1, 1] (numpy default_rng(seed)): at the initial state,
and optionally again every ``every`` steps, which stands in for the continual
rounding noise of 16-bit storage.

Command line (one GPU):
  python -m multicomponent_color_gradient_amd.reference_runs --out DIR [--storage fp32] [--steps 72000]
      [--amp 1e-4] [--seed 1] [--every 0]
writes DIR/JET_VTK/Re5000_We2000/mean_profiles and prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import io as mio
from . import post
from .case import reference_jet
from .solver import Solver


def perturb(s: Solver, amp: float, rng) -> None:
    fr, fb = s.populations()
    fr *= (1.0 + amp * rng.uniform(-1.0, 1.0, fr.shape)).astype(np.float32)
    s.set_populations(fr, fb)


def run(out: str, storage: str = "fp32", steps: int = 72000, amp: float = 0.0, seed: int = 1, every: int = 0,
        log=sys.stderr) -> dict:
    c = reference_jet()
    c.write_vti_output = 0
    case_dir, mean_dir, _ = mio.output_dirs(c, out)
    mio.write_case_info(c, case_dir)
    s = Solver(c, storage=storage)
    s.init_jet()
    rng = np.random.default_rng(seed)
    if amp > 0:
        perturb(s, amp, rng)
    done, t0 = 0, time.perf_counter()
    chunk = every if every > 0 else 6000
    while done < steps:
        n = min(chunk, steps - done)
        s.step(n)
        done += n
        if every > 0 and amp > 0 and done < steps:
            perturb(s, amp, rng)
        print(f"reference case {storage} amp={amp:g} every={every}: step {done}/{steps} "
              f"({time.perf_counter() - t0:.0f} s)", file=log, flush=True)
    s.check_health()
    s.write_radial(mean_dir)
    s.close()
    r = post.centerline_summary(mean_dir)
    return {"storage": storage, "amp": amp, "seed": seed, "every": every, "steps": steps, "B": r["B"], "y0": r["y0"],
            "S": [x["S"] for x in r["slices"]], "r_half_15D": r["slices"][-1]["r_half"],
            "samples": r["radial_sample_count"]}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m multicomponent_color_gradient_amd.reference_runs")
    ap.add_argument("--out", required=True)
    ap.add_argument("--storage", default="fp32")
    ap.add_argument("--steps", type=int, default=72000)
    ap.add_argument("--amp", type=float, default=0.0, help="relative noise on the red populations")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--every", type=int, default=0, help="perturb again every this many steps (0: initial state only)")
    args = ap.parse_args(argv)
    os.makedirs(args.out, exist_ok=True)
    print(json.dumps(run(args.out, args.storage, args.steps, args.amp, args.seed, args.every)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

"""Population census: how many of the stored populations are negative, per
direction, at chosen steps of a run.

The populations f_i of the reference's collision (lbm.cuh:243-244,290-291)
are not bounded below: at tau near 1/2 the factor (1 - omega) of the
non-equilibrium part is close to -1, and a strongly sheared node can store
g_i = f_r,i + f_b,i < 0.  The 16-bit storage keeps the sign of both colours
(the minority, the colour of smaller magnitude, is stored signed); this census
counts how often each case exercises that.

``negative_counts(solver)`` returns, over the interior nodes and per direction
i = 0..26: g_i < 0, f_r,i < 0, f_b,i < 0, and the nodes whose minority colour
is negative.

Command line (one GPU):
  python -m multicomponent_color_gradient_amd.census --case reference|jet256 [--storage fp32]
      [--steps 72000] [--every 12000]
prints one JSON line per sample.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from .case import perturbed_jet_case, reference_jet
from .solver import Solver


def negative_counts(s: Solver) -> dict:
    fr, fb = s.populations()
    inner = (slice(None), slice(1, -1), slice(1, -1), slice(1, -1))
    fr, fb = fr[inner], fb[inner]
    g = fr + fb
    red_minor = np.abs(fr) < np.abs(fb)
    minority = np.where(red_minor, fr, fb)
    per_dir = lambda a: np.count_nonzero(a.reshape(27, -1), axis=1).tolist()  # noqa: E731
    # D3Q27 weights in the direction order of stencil_ct.cuh:9-16: rest, 6 faces, 12 edges, 8 corners
    w = np.array([8 / 27] + [2 / 27] * 6 + [1 / 54] * 12 + [1 / 216] * 8)
    out = {"nodes": int(np.prod(g.shape[1:])), "g_negative": per_dir(g < 0), "fr_negative": per_dir(fr < 0),
           "fb_negative": per_dir(fb < 0), "minority_negative": per_dir(minority < 0),
           "g_min_over_w": float(np.min(g.reshape(27, -1).min(axis=1) / w))}
    for k in ("g_negative", "fr_negative", "fb_negative", "minority_negative"):
        out[k + "_total"] = int(sum(out[k]))
    return out


def make_case(name: str):
    if name == "reference":
        c = reference_jet()
        c.write_vti_output = 0
        return c
    if name == "jet256":
        c = perturbed_jet_case(256, 512, 256, jet_radius=16.0)
        c.NSTATS_START_STEP = 1 << 30
        return c
    raise SystemExit(f"unknown case {name}")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m multicomponent_color_gradient_amd.census")
    ap.add_argument("--case", default="reference", help="reference (128x400x128, Re 5000) or jet256")
    ap.add_argument("--storage", default="fp32")
    ap.add_argument("--steps", type=int, default=72000)
    ap.add_argument("--every", type=int, default=12000)
    args = ap.parse_args(argv)
    c = make_case(args.case)
    s = Solver(c, storage=args.storage)
    s.init_jet()
    done, t0 = 0, time.perf_counter()
    while done < args.steps:
        n = min(args.every, args.steps - done)
        s.step(n)
        done += n
        s.check_health()
        r = negative_counts(s)
        r.update({"case": args.case, "storage": args.storage, "step": done})
        print(json.dumps(r), flush=True)
        print(f"census {args.case} {args.storage}: step {done}/{args.steps} ({time.perf_counter() - t0:.0f} s), "
              f"negative g {r['g_negative_total']} of {27 * r['nodes']}", file=sys.stderr, flush=True)
    s.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())

"""FP64 central differences of the rising bubble's centroid height against the
GPU adjoint, with the branch switches that limit them.  Test infrastructure
(a checker): it runs the FP64 oracle beside the product's GPU adjoint.

Case: the buoyant rising bubble (x/z walls, gravity on red, tau 0.55/1.0,
A = 1e-3, R = 16 at y = 64 on 192x384x192 by default).  Objective: the centroid
height J = sum y rho_b / sum rho_b of the final populations
(multicomponent_color_gradient_amd.rising_bubble).  For each parameter p (A or
gravity) and relative step h, the oracle runs p (1 + h) and p (1 - h) in
lockstep; after every step it compares the branches the two collisions took
(Oracle.branch_masks): nodes that switch between bulk and interface
(lbm.cuh:214-216) and interface nodes whose |F| <= 1e-4 cut differs
(forces.cuh:61-64, collisionOperators.cuh:36-39).  J is only piecewise smooth
across such switches, so their number bounds what a central difference can
check.

Command line (one GPU, the oracle on the host cores):
  python -m oracle.rising_bubble_fd [--steps 400] [--param A|gravity] [--h 1e-3,1e-4]
prints a progress line every 50 steps and one JSON line per (parameter, h).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from multicomponent_color_gradient_amd import rising_bubble as rb  # noqa: E402
from multicomponent_color_gradient_amd.case import rising_bubble_case  # noqa: E402
from oracle.oracle import Oracle  # noqa: E402

PARAMS = {"A": "A_override", "gravity": "gravity"}


def fd64(case, name: str, h_rel: float, steps: int, log=sys.stderr) -> dict:
    """Central difference of J in FP64 along parameter ``name`` with relative
    step h_rel, and the per-step branch switches between the two runs."""
    field = PARAMS[name]
    p0 = getattr(case, field)
    h = h_rel * p0
    runs = []
    for sign in (1.0, -1.0):
        c = case.copy()
        setattr(c, field, p0 + sign * h)
        o = Oracle(c)
        o.init_bubble_tanh()
        runs.append(o)
    iface, cut = [], []
    t0 = time.perf_counter()
    for n in range(steps):
        for o in runs:
            o.step(1)
        mp, mm = runs[0].branch_masks(), runs[1].branch_masks()
        diff = mp ^ mm
        iface.append(int(np.count_nonzero(diff & 1)))
        # the |F| cut compared where both runs took the interface branch
        cut.append(int(np.count_nonzero((diff & 2) & (mp & mm & 1) * 2)))
        if n % 50 == 49:
            print(f"fd64 {name} h={h_rel:g}: step {n + 1}/{steps} ({time.perf_counter() - t0:.0f} s), "
                  f"switches so far: interface {sum(iface)}, |F| cut {sum(cut)}", file=log, flush=True)
    J = [rb.centroid_height(o.populations()[1]) for o in runs]
    for o in runs:
        o.close()
    iface_a, cut_a = np.array(iface), np.array(cut)
    return {"param": name, "h_rel": h_rel, "h": h, "J_plus": J[0], "J_minus": J[1], "fd": (J[0] - J[1]) / (2 * h),
            "switch_interface_total": int(iface_a.sum()), "switch_interface_max_per_step": int(iface_a.max()),
            "steps_with_interface_switch": int(np.count_nonzero(iface_a)),
            "switch_cut_total": int(cut_a.sum()), "switch_cut_max_per_step": int(cut_a.max()),
            "steps_with_cut_switch": int(np.count_nonzero(cut_a))}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m oracle.rising_bubble_fd")
    ap.add_argument("--NX", type=int, default=192)
    ap.add_argument("--NY", type=int, default=384)
    ap.add_argument("--NZ", type=int, default=192)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--every", type=int, default=200, help="checkpoint interval of the GPU adjoint")
    ap.add_argument("--param", default="A,gravity", help="comma-separated: A, gravity")
    ap.add_argument("--h", default="1e-3", help="comma-separated relative steps")
    args = ap.parse_args(argv)
    case = rising_bubble_case(args.NX, args.NY, args.NZ)
    t0 = time.perf_counter()
    g = rb.gradient(case, args.steps, args.every)
    adj = {"A": g["dA"], "gravity": g["dgravity"]}
    print(f"gpu adjoint: J {g['J']:.12g} dJ/dA {g['dA']:.9e} dJ/dgravity {g['dgravity']:.9e} "
          f"({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
    for name in args.param.split(","):
        for h in (float(x) for x in args.h.split(",")):
            r = fd64(case, name, h, args.steps)
            r.update({"case": f"rising bubble {case.NX}x{case.NY}x{case.NZ}", "steps": args.steps, "J_gpu": g["J"],
                      "adjoint": adj[name], "rel_diff": abs(adj[name] - r["fd"]) / abs(r["fd"]) if r["fd"] else None})
            print(json.dumps(r), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

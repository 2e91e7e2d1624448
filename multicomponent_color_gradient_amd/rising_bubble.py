"""Buoyant rising bubble: the gradient of the bubble's centroid height at the
final step with respect to the surface-tension amplitude A and the body force
(gravity), by the GPU adjoint, with central finite differences of the FP32
forward beside it.

Case (``case.rising_bubble_case``): no-slip walls in x and z, periodic y, a
blue tanh bubble of radius R at y0 in red, the body force F_y = -gravity rho_r
on the red (heavy) component, per-node viscosity between taur (red) and taub
(blue), A given directly.

Objective: J = sum_x y rho_b(x) / sum_x rho_b(x) over the interior nodes after
the last step, rho_b = sum_i f_b,i (the populations of the final state).

Command line (one GPU):
  python -m multicomponent_color_gradient_amd.rising_bubble [--steps 8000 --every 200 --fd-rel 1e-3]
prints one JSON line: J, dJ/dA and dJ/dgravity from the adjoint and from FD,
and the device times.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from .case import McgCase, rising_bubble_case
from .solver import Solver


def centroid_height(fb: np.ndarray) -> float:
    """J from the blue populations (27, NZ, NY, NX) of the final state."""
    rb = fb[:, 1:-1, 1:-1, 1:-1].astype(np.float64).sum(0)
    y = np.arange(1, rb.shape[1] + 1, dtype=np.float64).reshape(1, -1, 1)
    return float((rb * y).sum() / rb.sum())


def centroid_seed(fb: np.ndarray) -> np.ndarray:
    """dJ/df_b (27, NZ, NY, NX): (y - J)/sum rho_b at every interior node, every direction."""
    rb = fb[:, 1:-1, 1:-1, 1:-1].astype(np.float64).sum(0)
    M = rb.sum()
    y = np.arange(1, rb.shape[1] + 1, dtype=np.float64).reshape(1, -1, 1)
    J = float((rb * y).sum() / M)
    lam = np.zeros(fb.shape, np.float32)
    lam[:, 1:-1, 1:-1, 1:-1] = ((y - J) / M).astype(np.float32)
    return lam


def forward_J(case: McgCase, steps: int) -> float:
    s = Solver(case)
    s.init_bubble_tanh()
    s.step(steps)
    J = centroid_height(s.populations()[1])
    s.close()
    return J


def gradient(case: McgCase, steps: int, every: int, max_states: int = 0) -> dict:
    """J, dJ/dA, dJ/dbeta, dJ/dgravity by the adjoint, and the device times.
    max_states > 0 caps the stored states (binomial recomputation schedule)."""
    s = Solver(case)
    s.init_bubble_tanh()
    s.adjoint_set_checkpoint_limit(max_states)
    t0 = time.perf_counter()
    s.adjoint_record(steps, every)
    record_s = time.perf_counter() - t0
    fr, fb = s.populations()
    J = centroid_height(fb)
    lam_b = centroid_seed(fb)
    del fr
    s.adjoint_backward(np.zeros_like(lam_b), lam_b)
    g = s.adjoint_gradients()
    out = {"J": J, "dA": g["dA"], "dbeta": g["dbeta"], "dgravity": g["dgravity"], "record_s": record_s,
           "backward_ms": s.adjoint_last_ms(), "recomputed_steps": s.adjoint_recomputed_steps()}
    s.close()
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m multicomponent_color_gradient_amd.rising_bubble")
    ap.add_argument("--NX", type=int, default=192)
    ap.add_argument("--NY", type=int, default=384)
    ap.add_argument("--NZ", type=int, default=192)
    ap.add_argument("--steps", type=int, default=8000)
    ap.add_argument("--every", type=int, default=200, help="checkpoint interval K")
    ap.add_argument("--fd-rel", type=float, default=1e-3, help="relative FD step (0: no FD)")
    ap.add_argument("--max-states", type=int, default=0,
                    help="cap on the stored states (0: one per segment); the sweep recomputes the rest")
    args = ap.parse_args(argv)
    case = rising_bubble_case(args.NX, args.NY, args.NZ)
    g = gradient(case, args.steps, args.every, args.max_states)
    out = {"case": f"rising bubble {case.NX}x{case.NY}x{case.NZ}, R={case.bubble_radius:g} at y={case.bubble_y0:g}, "
                   f"gravity {case.gravity:g}, tau {case.taur:g}/{case.taub:g}, A={case.A_override:g}",
           "steps": args.steps, "checkpoint_every": args.every, **{k: g[k] for k in ("J", "dA", "dgravity", "dbeta")},
           "max_states": args.max_states, "recomputed_steps": g["recomputed_steps"],
           "record_s": round(g["record_s"], 3), "backward_ms": round(g["backward_ms"], 3)}
    if args.fd_rel > 0:
        for name, key in (("A_override", "dA"), ("gravity", "dgravity")):
            p0 = getattr(case, name)
            h = args.fd_rel * p0
            cp, cm = case.copy(), case.copy()
            setattr(cp, name, p0 + h)
            setattr(cm, name, p0 - h)
            fd = (forward_J(cp, args.steps) - forward_J(cm, args.steps)) / (2 * h)
            out[f"fd_{key}"] = fd
            out[f"rel_err_{key}"] = abs(g[key] - fd) / abs(fd) if fd != 0 else None
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

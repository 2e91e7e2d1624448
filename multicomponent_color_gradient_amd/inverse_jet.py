"""Inverse jet problem: the gradient of a time-averaged centre-line u_y with
respect to the inlet velocity of every jet column, by the GPU adjoint.

The control is a per-column offset du (3 x NZ x NX, component-major u_x, u_y,
u_z) added to the inlet velocity of jet/boundary_conditions.cuh:45-47 (here
with the tanh profile and seeded transverse perturbation of the perturbed-jet
case).  The objective is

    J = mean over the last ``last`` steps of the mean over y in [y0, y1] of
        u_y on the centre line,

the centre line being the 2 x 2 (or 1) nodes nearest to (jet_x0, jet_z0).  It
is a u_y functional (``Solver.adjoint_set_uy_seed``), accumulated by the
forward steps in FP64; the adjoint (``adjoint_record`` / ``adjoint_backward``)
returns dJ/d(du) per column in FP64 (``adjoint_inlet_gradient``).

Command line (one GPU):
  python -m multicomponent_color_gradient_amd.inverse_jet [--steps 2000 --last 500 --y0 150 --y1 200]
prints one JSON line: J, the gradient's norm and its projections on seeded
random directions, and the forward / backward device times.  ``--fd-h h``
adds central differences (J(h d) - J(-h d)) / 2h of the FP32 forward along
each direction d, and their relative differences from the projections.  Over
2000 steps the FP32 forward is not smooth enough for these to check the
adjoint closer than a few per cent (they differ by up to 5% between h = 1e-3
and 1e-2); central differences of the FP64 oracle with h = 1e-4 agree with the
adjoint to 2e-5.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from .case import McgCase, perturbed_jet_case
from .solver import Solver


def inverse_jet_case(NX: int = 128, NY: int = 256, NZ: int = 128, jet_radius: float = 8.0) -> McgCase:
    """The round jet with the tanh inlet and the seeded 1% transverse perturbation."""
    return perturbed_jet_case(NX, NY, NZ, jet_radius=jet_radius)


def centre_columns(case: McgCase) -> list[tuple[int, int]]:
    """(x, z) of the nodes nearest to the jet axis: 2 per direction when the axis lies between nodes."""
    def near(c):
        lo = int(np.floor(c))
        return [lo] if abs(c - lo) < 1e-9 else [lo, lo + 1]
    return [(x, z) for z in near(case.jet_z0) for x in near(case.jet_x0)]


def centreline_weights(case: McgCase, y0: int, y1: int, last: int) -> np.ndarray:
    """w (NZ, NY, NX) with sum_x w u_y summed over ``last`` steps = the objective J."""
    cols = centre_columns(case)
    w = np.zeros((case.NZ, case.NY, case.NX), np.float32)
    v = 1.0 / (len(cols) * (y1 - y0 + 1) * last)
    for x, z in cols:
        w[z, y0:y1 + 1, x] = v
    return w


def jet_mask(case: McgCase) -> np.ndarray:
    """Jet columns (inner (x, z) within jet_radius of the axis), shape (NZ, NX)."""
    z, x = np.meshgrid(np.arange(case.NZ), np.arange(case.NX), indexing="ij")
    inner = (x >= 1) & (x <= case.NX - 2) & (z >= 1) & (z <= case.NZ - 2)
    return inner & ((x - case.jet_x0) ** 2 + (z - case.jet_z0) ** 2 <= case.jet_radius ** 2)


def random_directions(case: McgCase, n: int, seed: int) -> list[np.ndarray]:
    """n seeded standard-normal directions over the jet columns (3, NZ, NX), scaled to max |d| = 1."""
    rng = np.random.default_rng(seed)
    jet = jet_mask(case)
    out = []
    for _ in range(n):
        d = rng.standard_normal((3, case.NZ, case.NX)) * jet
        out.append(d / np.max(np.abs(d)))
    return out


def _solver(case: McgCase, du, w, window) -> Solver:
    s = Solver(case)
    s.init_jet()
    if du is not None:
        s.set_inlet_offset(du)
    s.adjoint_set_uy_seed(w, *window)
    return s


def functional(case: McgCase, steps: int, last: int, y0: int, y1: int, du=None) -> float:
    """J after a plain forward run of ``steps`` steps from the jet initialisation."""
    w = centreline_weights(case, y0, y1, last)
    s = _solver(case, du, w, (steps - last, steps - 1))
    s.step(steps)
    J = s.uy_functional()
    s.close()
    return J


def gradient(case: McgCase, steps: int, last: int, y0: int, y1: int, checkpoint_every: int = 200, du=None,
             max_states: int = 0) -> dict:
    """J and dJ/d(du) per column (3, NZ, NX) by the adjoint, plus dJ/djet_velocity,
    dJ/dperturbation_amp and the device times of the forward and backward sweeps.
    max_states > 0 caps the stored states (binomial recomputation schedule)."""
    w = centreline_weights(case, y0, y1, last)
    s = _solver(case, du, w, (steps - last, steps - 1))
    s.adjoint_set_checkpoint_limit(max_states)
    t0 = time.perf_counter()
    s.adjoint_record(steps, checkpoint_every)
    record_s = time.perf_counter() - t0
    J = s.uy_functional()
    zero = np.zeros((27, case.NZ, case.NY, case.NX), np.float32)
    s.adjoint_backward(zero, zero)
    bwd_ms = s.adjoint_last_ms()
    recomputed = s.adjoint_recomputed_steps()
    (dU, damp), din = s.adjoint_inlet_gradient()
    s.close()
    return {"J": J, "dinlet": din, "dU": dU, "damp": damp, "record_s": record_s, "backward_ms": bwd_ms,
            "recomputed_steps": recomputed}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m multicomponent_color_gradient_amd.inverse_jet")
    ap.add_argument("--NX", type=int, default=128)
    ap.add_argument("--NY", type=int, default=256)
    ap.add_argument("--NZ", type=int, default=128)
    ap.add_argument("--radius", type=float, default=8.0)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--last", type=int, default=500)
    ap.add_argument("--y0", type=int, default=150)
    ap.add_argument("--y1", type=int, default=200)
    ap.add_argument("--every", type=int, default=200, help="checkpoint interval K")
    ap.add_argument("--max-states", type=int, default=0,
                    help="cap on the stored states (0: one per segment); the sweep recomputes the rest")
    ap.add_argument("--directions", type=int, default=8)
    ap.add_argument("--fd-h", type=float, default=0.0,
                    help="central-difference step along each direction (0: no finite differences)")
    ap.add_argument("--seed", type=int, default=5)
    args = ap.parse_args(argv)
    case = inverse_jet_case(args.NX, args.NY, args.NZ, args.radius)
    g = gradient(case, args.steps, args.last, args.y0, args.y1, args.every, max_states=args.max_states)
    dirs = random_directions(case, args.directions, args.seed)
    out = {"case": f"perturbed jet {case.NX}x{case.NY}x{case.NZ}, R={case.jet_radius:g}", "steps": args.steps,
           "window": [args.steps - args.last, args.steps - 1], "y": [args.y0, args.y1],
           "jet_columns": int(jet_mask(case).sum()), "J": g["J"], "grad_norm": float(np.linalg.norm(g["dinlet"])),
           "dJ_dU": g["dU"], "dJ_damp": g["damp"],
           "projections": [float(np.sum(g["dinlet"] * d)) for d in dirs],
           "checkpoint_every": args.every, "max_states": args.max_states, "recomputed_steps": g["recomputed_steps"],
           "record_s": round(g["record_s"], 3), "backward_ms": round(g["backward_ms"], 3)}
    if args.fd_h > 0:
        h = args.fd_h
        fd = []
        for d in dirs:
            jp = functional(case, args.steps, args.last, args.y0, args.y1, du=(h * d).astype(np.float32))
            jm = functional(case, args.steps, args.last, args.y0, args.y1, du=(-h * d).astype(np.float32))
            fd.append((jp - jm) / (2 * h))
        out["fd_h"] = h
        out["fd"] = fd
        out["fd_rel"] = [abs(p - f) / abs(f) if f != 0 else None for p, f in zip(out["projections"], fd)]
        out["fd_max_diff_over_max_fd"] = float(np.max(np.abs(np.array(out["projections"]) - np.array(fd)))
                                               / np.max(np.abs(fd)))
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())

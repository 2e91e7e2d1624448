"""Throughput benchmark of the D3Q27 colour-gradient step on one MI355X.

Workload (default): the round jet on 256x512x256 nodes (ghost layers
included), inlet disk R=16 lu with the tanh-edged top-hat U_in=0.05 (w=2) and
the seeded uniform transverse perturbation of 1% of U_in (SplitMix64, seed
20261203), ambient fluid at rest, tau_r/tau_b = 0.55/0.6 (the collision relaxes
with tau_b, the reference's single rate), A=1e-3, beta=0.7, FP32 populations,
radial statistics of u_y, u_r and phi sampled every 10 steps from the first
step (so their cost is inside the timed window).  A step is one full time step
(moments, colour gradient, collision, perturbation, recolouring, streaming,
inlet, outlet, statistics when due) over the whole grid, state resident in HBM.

MLUPS follows utilities/countmlups.cuh:42-48 (steps * NX*NY*NZ / t, ghost
layers counted); the interior-only figure is reported too.  Defaults: 200
warm-up steps, then 1000 timed steps in 5 windows of 200 (HIP events on the
solver's stream around each); `value` is the rate over all 1000, and the
median window's rate is reported beside it.  Each solver first times the
candidate z segmentations of its step kernel and keeps the fastest (untimed
steps after the warm-up until it has chosen; `config.launch` gives the shape
and how many steps the choice took); `--segments n` fixes the shape instead.  The roofline
uses 432 algorithmic bytes per interior lattice update (27 populations x 2
colours read + written, FP32; 216 B with 16-bit storage) against 8 TB/s;
`hbm_ceiling` gives what plain read, write and copy kernels reach on the same
device, and the step's measured HBM traffic per second against the copy.

The CPU baseline runs the FP64 OpenMP oracle (oracle/, a restatement of the
reference step) on the same grid for a bounded number of steps.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload jet256|jet128|bubble64]
                       [--storage fp32|fp16|bf16] [--kernel fused|twopass]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_LU = {"fp32": 432, "fp16": 216, "bf16": 216}
# the timed steps are split into this many windows; the JSON line also gives
# each window's ms/step and the MLUPS of the median window
WINDOWS = 5


def make_case(name: str):
    import multicomponent_color_gradient_amd as m

    if name == "jet256":
        c = m.perturbed_jet_case(256, 512, 256, jet_radius=16.0)
    elif name == "jet128":
        c = m.perturbed_jet_case(128, 256, 128, jet_radius=8.0)
    elif name == "bubble64":
        c = m.periodic_bubble_case(64, radius=12.0, tau=0.8, A=1e-3, beta=0.7)
        return c, "bubble_tanh"
    else:
        raise SystemExit(f"unknown workload {name}")
    c.NSTATS_START_STEP = 0
    c.NSTATS_SAMPLE = 10
    return c, "jet"


def load_traffic(workload: str, storage: str):
    p = os.path.join(ROOT, "profiles", f"traffic_{workload}_{storage}.json")
    if not os.path.exists(p):
        return None
    with open(p) as fh:
        d = json.load(fh)
    return d.get("bytes_per_launch")


def copy_probe(nbytes: int, reps: int = 3) -> None:
    """Device-to-device hipMemcpy of nbytes, reps times: the runtime's blit copy kernel
    reads and writes exactly nbytes per launch."""
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    if hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) or hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)):
        raise RuntimeError("copy probe: hipMalloc failed")
    try:
        if hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)):
            raise RuntimeError("copy probe: hipMemset failed")
        for _ in range(reps):
            if hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3):  # hipMemcpyDeviceToDevice
                raise RuntimeError("copy probe: hipMemcpy failed")
        if hip.hipDeviceSynchronize():
            raise RuntimeError("copy probe: hipDeviceSynchronize failed")
    finally:
        hip.hipFree(a)
        hip.hipFree(b)


def settle_launch(s, limit: int = 1000) -> int:
    """Untimed steps until the solver's launch-shape timings have chosen: the
    plain steps' (about 2 steps per candidate shape) and, where the case samples
    statistics, the statistics steps' (one statistics step per candidate, then
    12 more); returns the steps taken."""
    sampled = s.case.NSTATS_START_STEP <= s.step_count + limit

    def settled(info):
        return info["tuned"] and (info["stats"]["tuned"] or not sampled)

    n = 0
    while not settled(s.launch_info()) and n < limit:
        s.step(4)
        n += 4
    s.synchronize()
    return n


PROBE_KINDS = {"read": 0, "write": 1, "copy": 2}


def hbm_ceiling(full: bool, nbytes: int = 4 << 30, reps: int = 10) -> dict:
    """The bandwidth plain streams reach on this device (mcg_hbm_probe): read-only,
    write-only and copy over ``nbytes`` buffers, for each access width (4, 8, 16
    bytes per lane), cached or non-temporal stores and workgroups per CU; the
    best of ``reps`` launches per variant.  ``full`` sweeps every variant, else
    only 16-byte accesses with 8 workgroups per CU."""
    import ctypes

    import multicomponent_color_gradient_amd as m

    lib = m.load_library()
    rows = []
    widths = (4, 8, 16) if full else (16,)
    per_cu = (2, 4, 8, 16) if full else (8,)
    for kind, k in PROBE_KINDS.items():
        for vb in widths:
            for nt in ((0, 1) if kind != "read" else (0,)):
                for bpc in per_cu:
                    g = ctypes.c_double()
                    rc = lib.mcg_hbm_probe(k, vb, nt, bpc, ctypes.c_uint64(nbytes), reps, ctypes.byref(g))
                    if rc != 0:
                        raise RuntimeError(f"hbm probe failed: {lib.mcg_last_error().decode()}")
                    rows.append({"kind": kind, "vec_bytes": vb, "nontemporal": nt, "blocks_per_cu": bpc,
                                 "gbs": round(g.value, 1)})
                    print(f"hbm probe: {kind} {vb} B/lane nt={nt} {bpc} wg/CU: {g.value:.1f} GB/s",
                          file=sys.stderr, flush=True)
    best = {kind: max((r for r in rows if r["kind"] == kind), key=lambda r: r["gbs"]) for kind in PROBE_KINDS}
    return {"buffer_bytes": nbytes, "reps": reps, "best_gbs": {k: v["gbs"] for k, v in best.items()},
            "best": best, "rows": rows}


def adjoint_ratio(steps: int, every: int) -> dict:
    """Adjoint/forward step-cost ratio on the 192x384x192 rising bubble (walls in
    x/z, periodic y, R=16 at y=64, gravity 1e-5 on red, tau 0.55/1.0, A=1e-3,
    beta 0.7): device time of the backward sweep (checkpoint recomputation +
    adjoint kernels) for the centroid-height objective over the device time of
    the same number of plain forward steps."""
    import multicomponent_color_gradient_amd as m
    from multicomponent_color_gradient_amd.rising_bubble import centroid_seed

    c = m.rising_bubble_case(192, 384, 192)
    s = m.Solver(c)
    s.init_bubble_tanh()
    s.step(10)
    settle_launch(s)
    fwd_ms = s.time_steps(steps)
    s.adjoint_record(steps, every)
    fb = s.populations()[1]
    lam = centroid_seed(fb)
    del fb
    s.adjoint_backward(lam * 0, lam)
    bwd_ms = s.adjoint_last_ms()
    s.close()
    return {"workload": "rising bubble 192x384x192, walls x/z, R=16 at y=64, gravity 1e-5, tau 0.55/1.0, A=1e-3, "
                        "beta 0.7; centroid-height objective", "steps": steps,
            "checkpoint_every": every, "forward_ms": round(fwd_ms, 3), "backward_ms": round(bwd_ms, 3),
            "ratio": round(bwd_ms / fwd_ms, 3),
            "forward_mlups": round(steps * c.NX * c.NY * c.NZ / (fwd_ms / 1e3) / 1e6, 2)}


def adjoint_jet_ratio(steps: int, every: int) -> dict:
    """The same ratio on a 128x256x128 perturbed jet (R=8, tanh inlet, 1%
    transverse perturbation): inlet and convective outlet differentiated, the
    outlet state carried, a u_y seed on the centre line over every step."""
    import numpy as np

    import multicomponent_color_gradient_amd as m

    c = m.perturbed_jet_case(128, 256, 128, jet_radius=8.0)
    s = m.Solver(c)
    s.init_jet()
    s.step(10)
    settle_launch(s)
    fwd_ms = s.time_steps(steps)
    w = np.zeros(s.shape, np.float32)
    w[int(c.jet_z0), 1:-1, int(c.jet_x0)] = 1.0
    n0 = s.step_count
    s.adjoint_set_uy_seed(w, n0, n0 + steps - 1)
    s.adjoint_record(steps, every)
    lam = np.zeros((27,) + s.shape, np.float32)
    s.adjoint_backward(lam, lam)
    bwd_ms = s.adjoint_last_ms()
    s.close()
    return {"workload": "perturbed jet 128x256x128, R=8, U 0.05, tanh inlet, perturbation 0.01", "steps": steps,
            "checkpoint_every": every, "forward_ms": round(fwd_ms, 3), "backward_ms": round(bwd_ms, 3),
            "ratio": round(bwd_ms / fwd_ms, 3),
            "forward_mlups": round(steps * c.NX * c.NY * c.NZ / (fwd_ms / 1e3) / 1e6, 2)}


def interface_rich(steps: int, warmup: int, kernel: str) -> dict:
    """MLUPS on the interface-rich workload: the 256x512x256 jet grid and
    parameters, initialised with 256 seeded tanh spheres (seed 7, R in [6, 14],
    width 2) so that a large share of the nodes takes the interface branch of
    the collision (the worst ALU mix), FP32 and FP16-offset storage."""
    import numpy as np

    import multicomponent_color_gradient_amd as m

    c = m.perturbed_jet_case(256, 512, 256, jet_radius=16.0)
    c.NSTATS_START_STEP = 0
    c.NSTATS_SAMPLE = 10
    out = {"workload": "jet grid 256x512x256 initialised with 256 tanh spheres (seed 7, R in [6,14], width 2)"}
    cells, interior = c.NX * c.NY * c.NZ, m.interior_nodes(c)
    for storage in ("fp32", "fp16"):
        s = m.Solver(c, storage=storage, kernel=kernel)
        s.init_spheres(256, 7, 6.0, 14.0)
        print(f"bench: interface-rich 256x512x256 storage={storage} warmup={warmup} steps={steps}", file=sys.stderr,
              flush=True)
        s.step(warmup)
        settle_launch(s)
        ms = s.time_steps(steps)
        s.check_health()
        if storage == "fp32":
            # share of the interior nodes on the interface branch (4 rho_r rho_b / rho^2 > 1e-4,
            # lbm.cuh:214-216) in the moments of the last timed step
            f = s.fields()
            rr, rb = f["rhor"][1:-1, 1:-1, 1:-1], f["rhob"][1:-1, 1:-1, 1:-1]
            out["interface_fraction"] = round(float(np.mean(4 * rr * rb / (rr + rb) ** 2 > 1e-4)), 4)
            del f, rr, rb
        s.close()
        avg = ms / steps
        gbs = BYTES_PER_LU[storage] * interior / (avg / 1e3) / 1e9
        out[storage] = {"mlups": round(steps * cells / (ms / 1e3) / 1e6, 2), "ms_per_step": round(avg, 5),
                        "achieved_gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    return out


def cpu_baseline(case, init: str, budget_s: float):
    """FP64 OpenMP oracle on the same grid: 1 warm-up step, then steps until budget_s."""
    from oracle import oracle as orc

    t0 = time.perf_counter()
    o = orc.Oracle(case)
    getattr(o, "init_" + init)()
    o.step(1)
    t_setup = time.perf_counter() - t0
    n = 0
    t1 = time.perf_counter()
    while True:
        o.step(1)
        n += 1
        if time.perf_counter() - t1 >= budget_s or n >= 50:
            break
    dt = time.perf_counter() - t1
    cells = case.NX * case.NY * case.NZ
    o.close()
    return {
        "value": round(n * cells / dt / 1e6, 3),
        "unit": "MLUPS",
        "cores": orc.num_threads(),
        "kind": "port",
        "sample": f"FP64 OpenMP oracle, {case.NX}x{case.NY}x{case.NZ}, 1 warm-up + {n} timed steps "
                  f"({dt:.1f} s, setup {t_setup:.1f} s)",
        "other_grids": cpu_other_grids(),
    }


def cpu_other_grids(budget_s: float = 2.0) -> dict:
    """The oracle's MLUPS on the two smaller configurations: the 64^3 static
    bubble and the 128x256x128 jet (1 warm-up step, then steps for about
    budget_s each)."""
    import multicomponent_color_gradient_amd as m
    from oracle import oracle as orc

    out = {}
    for name, c, init in (("static bubble 64^3", m.periodic_bubble_case(64, radius=12.0, tau=0.8), "bubble_tanh"),
                          ("jet 128x256x128", m.perturbed_jet_case(128, 256, 128, jet_radius=8.0), "jet")):
        o = orc.Oracle(c)
        getattr(o, "init_" + init)()
        o.step(1)
        n, t1 = 0, time.perf_counter()
        while time.perf_counter() - t1 < budget_s and n < 200:
            o.step(1)
            n += 1
        dt = time.perf_counter() - t1
        o.close()
        out[name] = {"mlups": round(n * c.NX * c.NY * c.NZ / dt / 1e6, 3), "steps": n}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--workload", default="jet256")
    ap.add_argument("--storage", default="fp32")
    ap.add_argument("--kernel", default="fused")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint/forward cost measurement")
    ap.add_argument("--no-fp16", action="store_true", help="skip the FP16- and BF16-offset comparison runs")
    ap.add_argument("--no-interface-rich", action="store_true", help="skip the interface-rich (256 spheres) run")
    ap.add_argument("--adjoint-steps", type=int, default=200)
    ap.add_argument("--adjoint-every", type=int, default=200)
    ap.add_argument("--segments", type=int, default=0,
                    help="launch shape of the step kernel (0: the solver times the candidates on its first steps, "
                         "the default; n > 0: about n uniform z segments, fixed; -k: the k-th banded schedule, "
                         "fixed; config.launch.shape of a default run is the value that fixes the shape it chose, "
                         "e.g. to profile it)")
    ap.add_argument("--no-ceiling", action="store_true",
                    help="skip the short read/write/copy bandwidth probe reported beside the step")
    ap.add_argument("--ceiling", action="store_true",
                    help="only measure the HBM bandwidth of plain read, write and copy streams (every probe "
                         "variant) and print it as one JSON line")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="after the timed steps, copy a device buffer of this size 3 times with hipMemcpy "
                         "(a plain copy kernel of known size that calibrates rocprofv3's FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this solver runs on one GPU; a multi-GPU run (y-slab decomposition with halo "
              "exchange over RCCL) is not implemented, so only --gpus 1 is measured", file=sys.stderr)
        sys.exit(2)

    import multicomponent_color_gradient_amd as m

    if args.ceiling:
        print(json.dumps({"hbm_ceiling": hbm_ceiling(full=True)}), flush=True)
        return

    case, init = make_case(args.workload)

    def measure(storage: str) -> dict:
        s = m.Solver(case, storage=storage, kernel=args.kernel)
        s.set_launch_segments(args.segments)
        getattr(s, "init_" + init)()
        print(f"bench: {args.workload} {case.NX}x{case.NY}x{case.NZ} storage={storage} kernel={args.kernel} "
              f"warmup={args.warmup} steps={args.steps}", file=sys.stderr, flush=True)
        s.step(args.warmup)
        # the launch-shape timing finishes before the timed steps, whatever --warmup says
        shape_steps = settle_launch(s) if args.segments == 0 else 0
        s.synchronize()
        # the K timed steps run as WINDOWS consecutive windows, each bracketed by
        # HIP events on the solver's stream; the total is the sum of the windows
        sizes = [args.steps // WINDOWS + (1 if k < args.steps % WINDOWS else 0) for k in range(WINDOWS)]
        t0 = time.perf_counter()
        wms = [s.time_steps(n) for n in sizes if n > 0]
        wall = time.perf_counter() - t0
        ms = sum(wms)
        s.check_health()
        launch = s.launch_info()
        s.close()
        cells = case.NX * case.NY * case.NZ
        interior = m.interior_nodes(case)
        avg_ms = ms / args.steps
        bpl = BYTES_PER_LU[storage]
        achieved = bpl * interior / (avg_ms / 1e3) / 1e9
        per_window = [w / n for w, n in zip(wms, [n for n in sizes if n > 0])]
        return {"mlups": args.steps * cells / (ms / 1e3) / 1e6, "mlups_int": args.steps * interior / (ms / 1e3) / 1e6,
                "avg_ms": avg_ms, "wall": wall, "bpl": bpl, "achieved": achieved,
                "traffic": load_traffic(args.workload, storage),
                "windows": [round(x, 5) for x in per_window], "median_ms": float(sorted(per_window)[len(per_window) // 2]),
                "launch": dict(launch, shape_timing_steps=shape_steps)}

    r = measure(args.storage)
    mlups, mlups_int, avg_ms, wall = r["mlups"], r["mlups_int"], r["avg_ms"], r["wall"]
    bpl, achieved, traffic = r["bpl"], r["achieved"], r["traffic"]
    # the FP16- and BF16-offset comparisons of the metric, on the same workload
    r16 = measure("fp16") if args.storage == "fp32" and not args.no_fp16 else None
    rb16 = measure("bf16") if args.storage == "fp32" and not args.no_fp16 else None
    if args.copy_probe_mib > 0:
        copy_probe(args.copy_probe_mib << 20)

    out = {
        "metric": "MLUPS + % HBM roofline (FP32 vs FP16-offset); adjoint/forward step cost ratio",
        "value": round(mlups, 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(avg_ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if args.storage == "fp32" else f"{args.storage}-storage/f32-math",
        "data": ("synthetic (jet initialisation at rest, tanh inlet with seeded 1% transverse perturbation)"
                 if init == "jet" else "synthetic (static tanh bubble)"),
        "config": {"workload": f"round jet {case.NX}x{case.NY}x{case.NZ}, R={case.jet_radius:g}, U={case.jet_velocity:g}, "
                               f"tanh inlet + 1% perturbation, tau {case.taur:g}/{case.taub:g}, A={case.A_override:g}, "
                               f"beta={case.beta_recolor:g}"
                               if init == "jet" else f"periodic tanh bubble {case.NX}^3",
                   "NX": case.NX, "NY": case.NY, "NZ": case.NZ, "storage": args.storage, "kernel": args.kernel,
                   "stats_every": case.NSTATS_SAMPLE, "mlups_interior": round(mlups_int, 2),
                   "wall_s": round(wall, 4), "windows_ms_per_step": r["windows"],
                   "mlups_median_window": round(case.NX * case.NY * case.NZ / r["median_ms"] / 1e3, 2),
                   "launch": r["launch"]},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "bytes_per_lu": bpl, "kernel_ms": round(avg_ms, 5)},
    }
    for key, rr in (("fp16_offset", r16), ("bf16_offset", rb16)):
        if rr is None:
            continue
        out[key] = {"value": round(rr["mlups"], 2), "unit": "MLUPS", "ms_per_step": round(rr["avg_ms"], 5),
                    "speedup_vs_fp32": round(rr["mlups"] / mlups, 3), "windows_ms_per_step": rr["windows"],
                    "launch": rr["launch"],
                    "mlups_median_window": round(case.NX * case.NY * case.NZ / rr["median_ms"] / 1e3, 2),
                    "roofline": {"achieved": round(rr["achieved"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": round(rr["achieved"] / HBM_PEAK_GBS, 4), "traffic": rr["traffic"],
                                 "bytes_per_lu": rr["bpl"]}}
    if not args.no_interface_rich and args.storage == "fp32" and args.workload == "jet256":
        out["interface_rich"] = interface_rich(args.steps, args.warmup, args.kernel)
    if not args.no_adjoint and args.storage == "fp32":
        out["adjoint"] = adjoint_ratio(args.adjoint_steps, args.adjoint_every)
        out["adjoint_jet"] = adjoint_jet_ratio(args.adjoint_steps, args.adjoint_every)
    if not args.no_ceiling:
        # what plain streams reach on this device, beside what the step moves
        ceil = hbm_ceiling(full=False, reps=5)["best_gbs"]
        moved = traffic / (avg_ms / 1e3) / 1e9 if traffic else None
        out["hbm_ceiling"] = {"unit": "GB/s", "read": ceil["read"], "write": ceil["write"], "copy": ceil["copy"],
                              "step_traffic_gbs": round(moved, 1) if moved else None,
                              "step_traffic_over_copy": round(moved / ceil["copy"], 3) if moved else None,
                              "probe": "16 B per lane, 8 workgroups per CU, 4 GiB buffers, best of 5"}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(case, init, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

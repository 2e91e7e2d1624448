"""The rising-bubble objective (centroid height of the blue component after
the last step) through the GPU adjoint against the FP64 CPU adjoint with the
same seed: dJ/dA, dJ/dbeta, dJ/dgravity and the initial-state gradient.

Case: 24x32x24, walls in x/z, periodic y, R = 5 at y = 10, gravity 1e-4 on
red, tau 0.55/1.0, A = 1e-3, beta 0.7; 40 steps from the tanh initialisation,
checkpoints every 10.  The FP64 adjoint is pinned to finite differences of the
oracle in test_adjoint_rising_cpu.py; here FP32 and FP64 agree to 1e-4 in
dJ/dgravity and dJ/dbeta, and the state gradient after 40 steps to 5e-4 of
its largest entry (measured 2.6e-4).  dJ/dA of this objective is a
sum of node terms of both signs that cancels to -2e-5: the FP32 adjoint
fields carry it to 1.1e-2 (bound 2e-2)."""
import numpy as np
import pytest

import multicomponent_color_gradient_amd as m
from multicomponent_color_gradient_amd.adjoint import cpu_adjoint
from multicomponent_color_gradient_amd.rising_bubble import centroid_height, centroid_seed

pytestmark = pytest.mark.gpu
INNER = (slice(None),) + (slice(1, -1),) * 3


def test_centroid_gradient_matches_fp64_adjoint():
    c = m.rising_bubble_case(24, 32, 24, radius=5.0, y0=10.0, gravity=1e-4)
    steps = 40
    print(f"rising centroid {c.NX}x{c.NY}x{c.NZ} steps={steps}", flush=True)
    s = m.Solver(c)
    s.init_bubble_tanh()
    fr0, fb0 = s.populations()
    fields0 = s.fields()
    s.adjoint_record(steps, 10)
    fr, fb = s.populations()
    lam_b = centroid_seed(fb)
    lam_r = np.zeros_like(lam_b)
    lr, lb, _ = s.adjoint_backward(lam_r, lam_b)
    g = s.adjoint_gradients()
    ref = cpu_adjoint(c, fr0.astype(np.float64), fb0.astype(np.float64), steps, lam_r.astype(np.float64),
                      lam_b.astype(np.float64), fields0=fields0)
    assert abs(centroid_height(ref["fb"]) - centroid_height(fb)) < 1e-6 * centroid_height(fb)  # FP32 vs FP64
    assert abs(g["dgravity"] - ref["dgravity"]) <= 1e-4 * abs(ref["dgravity"]), (g["dgravity"], ref["dgravity"])
    assert abs(g["dbeta"] - ref["dbeta"]) <= 1e-4 * abs(ref["dbeta"]), (g["dbeta"], ref["dbeta"])
    assert abs(g["dA"] - ref["dA"]) <= 2e-2 * abs(ref["dA"]), (g["dA"], ref["dA"])
    scale = np.max(np.abs(ref["lam_fb"][INNER]))
    assert np.max(np.abs(lb[INNER] - ref["lam_fb"][INNER])) <= 5e-4 * scale



def test_centroid_gradient_over_1000_steps():
    """Long horizon: 24x48x24 (R = 5 at y = 12, gravity 1e-4), 1000 steps,
    checkpoints every 100.  The GPU adjoint against the FP64 CPU adjoint (same
    seed): dJ/dA and dJ/dgravity within 1e-4 (measured 1.2e-5, 2e-6).  Against
    central differences of the FP64 oracle: dJ/dA within 1e-4 at h = 1e-5
    relative (measured 5e-8).  J(gravity) is only piecewise smooth here: a
    change of gravity moves the interface, and nodes cross the bulk/interface
    and |F| thresholds (held at the forward state by the adjoint) one step
    earlier or later, so its differences move away from the adjoint as h
    shrinks (1.4e-4 at h = 1e-3, 9e-4 at h = 1e-5); they are printed, not
    asserted."""
    from oracle.oracle import Oracle

    c = m.rising_bubble_case(24, 48, 24, radius=5.0, y0=12.0, gravity=1e-4)
    steps = 1000
    print(f"rising centroid {c.NX}x{c.NY}x{c.NZ} steps={steps}", flush=True)
    s = m.Solver(c)
    s.init_bubble_tanh()
    fr0, fb0 = s.populations()
    fields0 = s.fields()
    s.adjoint_record(steps, 100)
    fb = s.populations()[1]
    lam_b = centroid_seed(fb)
    s.adjoint_backward(np.zeros_like(lam_b), lam_b)
    g = s.adjoint_gradients()
    s.close()
    ref = cpu_adjoint(c, fr0.astype(np.float64), fb0.astype(np.float64), steps, np.zeros(lam_b.shape),
                      lam_b.astype(np.float64), fields0=fields0)
    print(f"dA: GPU {g['dA']:.10e} FP64 adjoint {ref['dA']:.10e}; dgravity: GPU {g['dgravity']:.10e} "
          f"FP64 adjoint {ref['dgravity']:.10e}", flush=True)
    assert abs(g["dA"] - ref["dA"]) <= 1e-4 * abs(ref["dA"]), (g["dA"], ref["dA"])
    assert abs(g["dgravity"] - ref["dgravity"]) <= 1e-4 * abs(ref["dgravity"]), (g["dgravity"], ref["dgravity"])

    def oracle_J(name, value):
        cc = c.copy()
        setattr(cc, name, value)
        o = Oracle(cc)
        o.init_bubble_tanh()
        o.step(steps)
        J = centroid_height(o.populations()[1])
        o.close()
        return J

    def fd(name, rel):
        p0 = getattr(c, name)
        return (oracle_J(name, p0 * (1 + rel)) - oracle_J(name, p0 * (1 - rel))) / (2 * rel * p0)

    fd_a = fd("A_override", 1e-5)
    print(f"dA: FP64 FD (h = 1e-5) {fd_a:.10e}", flush=True)
    assert abs(ref["dA"] - fd_a) <= 1e-4 * abs(fd_a), (ref["dA"], fd_a)
    for rel in (1e-3, 1e-5):
        print(f"dgravity: FP64 FD (h = {rel:g}) {fd('gravity', rel):.10e}", flush=True)


@pytest.mark.skipif(__import__("os").environ.get("MCG_LONG") != "1",
                    reason="96x192x96 over 2000 steps against FP64 differences (minutes of host time): set MCG_LONG=1")
def test_centroid_dA_96_over_2000_steps():
    """96x192x96 (R = 8 at y = 32; gravity 1e-5, tau 0.55/1.0, A = 1e-3 as the
    full-size case), 2000 steps, checkpoints every 200: the GPU adjoint's dJ/dA
    against central differences of the FP64 oracle at h = 1e-5 relative,
    within 1e-4.  (dJ/dgravity differences are not a check here: J(gravity) is
    only piecewise smooth, see test_centroid_gradient_over_1000_steps.)"""
    from oracle.oracle import Oracle

    c = m.rising_bubble_case(96, 192, 96, radius=8.0, y0=32.0)
    steps = 2000
    print(f"rising centroid {c.NX}x{c.NY}x{c.NZ} steps={steps}: adjoint", flush=True)
    s = m.Solver(c)
    s.init_bubble_tanh()
    s.adjoint_record(steps, 200)
    fb = s.populations()[1]
    lam_b = centroid_seed(fb)
    s.adjoint_backward(np.zeros_like(lam_b), lam_b)
    g = s.adjoint_gradients()
    s.close()

    def oracle_J(value):
        cc = c.copy()
        cc.A_override = value
        o = Oracle(cc)
        o.init_bubble_tanh()
        for k in range(steps // 500):
            o.step(500)
            print(f"oracle A={value:.8g}: step {500 * (k + 1)}/{steps}", flush=True)
        J = centroid_height(o.populations()[1])
        o.close()
        return J

    a0 = c.A_override
    fd = (oracle_J(a0 * (1 + 1e-5)) - oracle_J(a0 * (1 - 1e-5))) / (2e-5 * a0)
    print(f"dA: adjoint {g['dA']:.10e} FP64 FD (h = 1e-5) {fd:.10e} rel {abs(g['dA'] - fd) / abs(fd):.2e}", flush=True)
    assert abs(g["dA"] - fd) <= 1e-4 * abs(fd), (g["dA"], fd)

"""The rising-bubble extensions on the GPU (x/z bounce-back walls, body force
on red, per-node viscosity between taur and taub) against the FP64 oracle,
the fused kernel against the two-pass kernels, and 16-bit storage.

Case: a blue tanh bubble (R = 5) in red on 26x34x26, walls in x and z,
periodic y, gravity 1e-4 on red, tau 0.55 (red) / 1.0 (blue), A = 1e-3,
beta = 0.7.  Populations after 10 steps to 1e-5 relative, after 100 steps to
1e-4 (FP32 rounding against FP64)."""
import numpy as np
import pytest

import multicomponent_color_gradient_amd as m
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu
INNER = (slice(None),) + (slice(1, -1),) * 3


def small_case(**kw):
    return m.rising_bubble_case(26, 34, 26, radius=5.0, y0=10.0, gravity=1e-4, **kw)


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("steps,tol", [(10, 1e-5), (100, 1e-4)])
def test_rising_bubble_matches_oracle(steps, tol):
    c = small_case()
    print(f"rising bubble 26x34x26 steps={steps}", flush=True)
    s = m.Solver(c)
    s.init_bubble_tanh()
    s.step(steps)
    o = Oracle(c)
    o.init_bubble_tanh()
    o.step(steps)
    gr, gb = s.populations()
    orr, ob = o.populations()
    assert rel(gr[INNER], orr[INNER]) < tol
    assert rel(gb[INNER], ob[INNER]) < tol
    f, of = s.fields(), o.fields()
    assert np.max(np.abs(f["uy"] - of["uy"])) < 1e-6


def test_rising_bubble_fused_equals_twopass_and_conserves_mass():
    c = small_case()
    out = []
    for kernel in ("fused", "twopass"):
        s = m.Solver(c, kernel=kernel)
        s.init_bubble_tanh()
        s.step(1)
        g0 = s.global_sums()
        s.step(30)
        g1 = s.global_sums()
        assert abs(g1[0] - g0[0]) <= 1e-6 * g0[0] and abs(g1[1] - g0[1]) <= 1e-6 * g0[1]
        out.append(s.populations())
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_rising_bubble_fp16_close_to_fp32():
    c = small_case()
    res = []
    for st in ("fp32", "fp16"):
        s = m.Solver(c, storage=st)
        s.init_bubble_tanh()
        s.step(50)
        res.append(s.fields())
    assert np.max(np.abs(res[1]["uy"] - res[0]["uy"])) < 2e-5
    assert np.max(np.abs(res[1]["rhob"] - res[0]["rhob"])) < 2e-3


@pytest.mark.parametrize("storage", ["fp16", "bf16"])
def test_rising_bubble_16bit_fused_equals_twopass(storage):
    """Walls and the body force (the EXT instance of the fused kernel) with
    16-bit storage: the fused and two-pass kernels store the same stochastically
    rounded words (same random stream per node and step), bit for bit."""
    c = small_case()
    out = []
    for kernel in ("fused", "twopass"):
        s = m.Solver(c, storage=storage, kernel=kernel)
        s.init_bubble_tanh()
        s.step(30)
        out.append(s.populations())
        s.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])

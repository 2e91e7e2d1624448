"""GPU adjoint of rising-bubble steps (x/z walls, body force on red,
per-node viscosity) against the FP64 CPU adjoint, which
test_adjoint_rising_cpu.py pins to central finite differences of the oracle.
Same state and objective as there.  FP32 forward and adjoint, FP64 parameter
sums: relative agreement 1e-4 (dJ/dA, a sum of cancelling node terms, 2e-3).
"""
import numpy as np
import pytest

import multicomponent_color_gradient_amd as m
from multicomponent_color_gradient_amd.adjoint import cpu_adjoint
from test_adjoint_rising_cpu import INNER, STEPS, make_state

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.fixture(scope="module")
def rising():
    case, fr, fb, fields = make_state()
    fr, fb = fr.astype(np.float32).astype(np.float64), fb.astype(np.float32).astype(np.float64)
    rng = np.random.default_rng(1)
    lam = [rng.standard_normal(fr.shape), rng.standard_normal(fb.shape)]
    for a in lam:
        a[:, 0], a[:, -1] = 0, 0
        a[:, :, 0], a[:, :, -1] = 0, 0
        a[:, :, :, 0], a[:, :, :, -1] = 0, 0
    ref = cpu_adjoint(case, fr, fb, STEPS, lam[0], lam[1], fields0=fields)
    return case, fr, fb, lam, ref


@pytest.mark.parametrize("every", [1, 3])
def test_gpu_rising_adjoint_matches_fp64_cpu_adjoint(rising, every):
    case, fr, fb, lam, ref = rising
    print(f"rising adjoint {case.NX}x{case.NY}x{case.NZ} steps={STEPS} every={every}", flush=True)
    s = m.Solver(case)
    s.init_bubble_tanh()
    s.set_populations(fr.astype(np.float32), fb.astype(np.float32))
    s.adjoint_record(STEPS, every)
    gr, gb = s.populations()
    assert rel(gr[INNER], ref["fr"][INNER]) < 1e-5 and rel(gb[INNER], ref["fb"][INNER]) < 1e-5
    lr, lb, (dA, dbeta) = s.adjoint_backward(lam[0], lam[1])
    g = s.adjoint_gradients()
    assert g["dA"] == dA and g["dbeta"] == dbeta
    assert rel(lr[INNER], ref["lam_fr"][INNER]) < 1e-4
    assert rel(lb[INNER], ref["lam_fb"][INNER]) < 1e-4
    assert abs(dbeta - ref["dbeta"]) <= 1e-4 * abs(ref["dbeta"]), (dbeta, ref["dbeta"])
    assert abs(g["dgravity"] - ref["dgravity"]) <= 1e-4 * abs(ref["dgravity"]), (g["dgravity"], ref["dgravity"])
    assert abs(dA - ref["dA"]) <= 2e-3 * abs(ref["dA"]), (dA, ref["dA"])


def test_torch_autograd_rising_matches_the_abi(rising):
    """LBMRisingSteps through torch.autograd: J = <lam, f_N>; gradients equal
    the direct C-ABI adjoint bit for bit (dJ/dgravity included)."""
    import torch

    from multicomponent_color_gradient_amd.adjoint import make_rising_steps_function

    case, fr, fb, lam, _ = rising
    F = make_rising_steps_function()
    fr0 = torch.tensor(fr, dtype=torch.float32, requires_grad=True)
    fb0 = torch.tensor(fb, dtype=torch.float32, requires_grad=True)
    A = torch.tensor(case.A_override, dtype=torch.float64, requires_grad=True)
    beta = torch.tensor(case.beta_recolor, dtype=torch.float64, requires_grad=True)
    g = torch.tensor(case.gravity, dtype=torch.float64, requires_grad=True)
    gr, gb = F.apply(fr0, fb0, A, beta, g, case, STEPS, 2)
    J = (torch.tensor(lam[0], dtype=torch.float32) * gr).sum() + (torch.tensor(lam[1], dtype=torch.float32) * gb).sum()
    J.backward()
    s = m.Solver(case)
    s.init_bubble_tanh()
    s.set_populations(fr.astype(np.float32), fb.astype(np.float32))
    s.adjoint_record(STEPS, 2)
    lr, lb, _ = s.adjoint_backward(lam[0], lam[1])
    d = s.adjoint_gradients()
    assert np.array_equal(fr0.grad.numpy(), lr) and np.array_equal(fb0.grad.numpy(), lb)
    assert float(A.grad) == d["dA"] and float(beta.grad) == d["dbeta"] and float(g.grad) == d["dgravity"]

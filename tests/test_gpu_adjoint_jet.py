"""GPU adjoint of jet steps (C ABI mcg_adjoint_record / mcg_adjoint_backward /
mcg_adjoint_set_uy_seed / mcg_adjoint_inlet_gradient) against the FP64 CPU
jet adjoint, which test_adjoint_jet_cpu.py pins to central finite differences
of the FP64 oracle.  Same state and objective as there: the jet
initialisation with the tanh inlet and the seeded transverse perturbation,
interior populations perturbed at random, J = <lam, f_N> + sum over the
moments passes of steps 2..5 of sum_x w(x) u_y(x).  FP32 forward and adjoint,
FP64 parameter sums: relative agreement 1e-4 with the FP64 adjoint.
"""
import numpy as np
import pytest

import multicomponent_color_gradient_amd as m
from multicomponent_color_gradient_amd.adjoint import cpu_adjoint_jet
from test_adjoint_jet_cpu import INNER, STEPS, UY_STEPS, base_case, initial_state, jet_columns, oracle_J, weights

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.fixture(scope="module")
def jet_setup():
    case = base_case()
    fr, fb, fields = initial_state(case)
    # start from the FP32-rounded state so both adjoints linearise the same forward
    fr, fb = fr.astype(np.float32).astype(np.float64), fb.astype(np.float32).astype(np.float64)
    rng = np.random.default_rng(1)
    lam = [rng.standard_normal(fr.shape) * 0.1, rng.standard_normal(fb.shape) * 0.1]
    for a in lam:
        a[:, 0], a[:, -1] = 0, 0
        a[:, :, 0], a[:, :, -1] = 0, 0
        a[:, :, :, 0], a[:, :, :, -1] = 0, 0
    w = weights(case)
    ref = cpu_adjoint_jet(case, fr, fb, fields, STEPS, lam[0], lam[1], w_uy=w, uy_steps=UY_STEPS)
    return case, fr, fb, lam, w, ref


def run_gpu(case, fr, fb, lam, w, every, uy=UY_STEPS, du=None):
    s = m.Solver(case)
    s.init_jet()
    s.set_populations(fr.astype(np.float32), fb.astype(np.float32))
    if du is not None:
        s.set_inlet_offset(du)
    s.adjoint_set_uy_seed(w, *uy)
    s.adjoint_record(STEPS, every)
    gr, gb = s.populations()
    lr, lb, (dA, dbeta) = s.adjoint_backward(lam[0], lam[1])
    (dU, damp), din = s.adjoint_inlet_gradient()
    return {"fr": gr, "fb": gb, "lam_fr": lr, "lam_fb": lb, "dA": dA, "dbeta": dbeta, "dU": dU, "damp": damp,
            "dinlet": din}


@pytest.mark.parametrize("every", [1, 3, STEPS])
def test_gpu_jet_adjoint_matches_fp64_cpu_adjoint(jet_setup, every):
    case, fr, fb, lam, w, ref = jet_setup
    print(f"jet adjoint {case.NX}x{case.NY}x{case.NZ} steps={STEPS} every={every}", flush=True)
    g = run_gpu(case, fr, fb, lam, w, every)
    assert rel(g["fr"][INNER], ref["fr"][INNER]) < 1e-5
    assert rel(g["fb"][INNER], ref["fb"][INNER]) < 1e-5
    assert rel(g["lam_fr"][INNER], ref["lam_fr"][INNER]) < 1e-4
    assert rel(g["lam_fb"][INNER], ref["lam_fb"][INNER]) < 1e-4
    for k in ("dU", "damp", "dbeta"):
        assert abs(g[k] - ref[k]) <= 1e-4 * abs(ref[k]), (k, g[k], ref[k])
    # dJ/dA sums cancelling node terms: FP32 rounding shows at the 1e-3 level
    assert abs(g["dA"] - ref["dA"]) <= 2e-3 * abs(ref["dA"]), (g["dA"], ref["dA"])
    assert rel(g["dinlet"], ref["dinlet"]) < 1e-4
    assert np.all(g["dinlet"][:, ref["dinlet"][1] == 0] == 0)
    assert np.all(g["lam_fr"][:, 0] == 0) and np.all(g["lam_fb"][:, :, :, -1] == 0)  # ghost slots


def test_gpu_jet_adjoint_without_seed_window(jet_setup):
    """The u_y seed only acts on its steps: a window past the recorded steps
    leaves <lam, f_N> alone, which the CPU adjoint gives with w = None."""
    case, fr, fb, lam, w, _ = jet_setup
    ref = cpu_adjoint_jet(case, fr, fb, initial_state(case)[2], STEPS, lam[0], lam[1])
    g = run_gpu(case, fr, fb, lam, w, 2, uy=(STEPS, STEPS + 4))
    assert rel(g["lam_fr"][INNER], ref["lam_fr"][INNER]) < 1e-4
    assert abs(g["dU"] - ref["dU"]) <= 1e-4 * abs(ref["dU"])


def test_gpu_jet_adjoint_restores_the_forward(jet_setup):
    """After the backward sweep the solver continues from the recorded final
    state: the populations, the inlet/outlet arrays and the outlet state."""
    case, fr, fb, lam, w, _ = jet_setup
    s = m.Solver(case)
    s.init_jet()
    s.set_populations(fr.astype(np.float32), fb.astype(np.float32))
    s.adjoint_record(STEPS, 3)
    f0 = s.fields()
    s.adjoint_backward(lam[0], lam[1])
    f1 = s.fields()
    for k in ("rhob", "uy"):
        assert np.array_equal(f0[k], f1[k])
    s.step(3)
    t = m.Solver(case)
    t.init_jet()
    t.set_populations(fr.astype(np.float32), fb.astype(np.float32))
    t.step(STEPS + 3)
    assert np.array_equal(s.populations()[0], t.populations()[0])
    assert np.array_equal(s.populations()[1], t.populations()[1])
    assert np.array_equal(s.fields()["rhob"], t.fields()["rhob"])


def offset(case, seed=11):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal((3, case.NZ, case.NX)) * 2e-3 * jet_columns(case)).astype(np.float32).astype(np.float64)


def test_gpu_inlet_offset_and_uy_functional_match_the_oracle(jet_setup):
    """Forward with a per-column inlet velocity offset and the u_y functional
    accumulated by the steps, against the FP64 oracle.  Populations to 1e-5;
    the functional is a sum of velocities of both signs, so its FP32 error is
    bounded against its scale sum|w| x (window steps) x U: 1e-5 of that."""
    case, fr, fb, lam, w, _ = jet_setup
    du = offset(case)
    zero = [np.zeros_like(fr), np.zeros_like(fb)]
    j_ref, (gr, gb) = oracle_J(case, fr, fb, zero, w, du=du)
    s = m.Solver(case)
    s.init_jet()
    s.set_populations(fr.astype(np.float32), fb.astype(np.float32))
    s.set_inlet_offset(du)
    s.adjoint_set_uy_seed(w, *UY_STEPS)
    s.step(STEPS)
    pr, pb = s.populations()
    assert rel(pr[INNER], gr[INNER]) < 1e-5 and rel(pb[INNER], gb[INNER]) < 1e-5
    scale = np.sum(np.abs(w)) * (UY_STEPS[1] - UY_STEPS[0] + 1) * case.jet_velocity
    assert abs(s.uy_functional() - j_ref) <= 1e-5 * scale, (s.uy_functional(), j_ref)
    # the two-pass kernel observes the same functional
    t = m.Solver(case, kernel="twopass")
    t.init_jet()
    t.set_populations(fr.astype(np.float32), fb.astype(np.float32))
    t.set_inlet_offset(du)
    t.adjoint_set_uy_seed(w, *UY_STEPS)
    t.step(STEPS)
    assert np.array_equal(t.populations()[0], pr)
    assert t.uy_functional() == s.uy_functional()


def test_gpu_jet_adjoint_with_inlet_offset(jet_setup):
    case, fr, fb, lam, w, _ = jet_setup
    du = offset(case)
    ref = cpu_adjoint_jet(case, fr, fb, initial_state(case)[2], STEPS, lam[0], lam[1], w_uy=w, uy_steps=UY_STEPS,
                          inlet_du=du)
    g = run_gpu(case, fr, fb, lam, w, 4, du=du)
    assert rel(g["lam_fb"][INNER], ref["lam_fb"][INNER]) < 1e-4
    assert rel(g["dinlet"], ref["dinlet"]) < 1e-4
    assert abs(g["dU"] - ref["dU"]) <= 1e-4 * abs(ref["dU"])


def test_torch_autograd_jet_matches_the_abi(jet_setup):
    """LBMJetSteps through torch.autograd: J = <lam, f_N> + 2 J_uy; its
    gradients equal the direct C-ABI adjoint with the u_y seed scaled by 2,
    bit for bit, and J_uy the functional of a plain forward."""
    import torch

    from multicomponent_color_gradient_amd.adjoint import make_jet_steps_function

    case, fr, fb, lam, w, _ = jet_setup
    du = offset(case)
    F = make_jet_steps_function()
    fr0 = torch.tensor(fr, dtype=torch.float32, requires_grad=True)
    fb0 = torch.tensor(fb, dtype=torch.float32, requires_grad=True)
    A = torch.tensor(case.A_override, dtype=torch.float64, requires_grad=True)
    beta = torch.tensor(case.beta_recolor, dtype=torch.float64, requires_grad=True)
    U = torch.tensor(case.jet_velocity, dtype=torch.float64, requires_grad=True)
    amp = torch.tensor(case.perturbation_amp, dtype=torch.float64, requires_grad=True)
    dut = torch.tensor(du, dtype=torch.float64, requires_grad=True)
    gr, gb, juy = F.apply(fr0, fb0, A, beta, U, amp, dut, case, STEPS, 3, torch.tensor(w), UY_STEPS)
    J = (torch.tensor(lam[0], dtype=torch.float32) * gr).sum() + (torch.tensor(lam[1], dtype=torch.float32) * gb).sum()
    (J.double() + 2.0 * juy).backward()
    g = run_gpu(case, fr, fb, lam, 2.0 * w, 3, du=du)
    assert np.array_equal(fr0.grad.numpy(), g["lam_fr"]) and np.array_equal(fb0.grad.numpy(), g["lam_fb"])
    assert float(A.grad) == g["dA"] and float(beta.grad) == g["dbeta"]
    assert float(U.grad) == g["dU"] and float(amp.grad) == g["damp"]
    assert np.array_equal(dut.grad.numpy(), g["dinlet"])
    s = m.Solver(case)
    s.init_jet()
    s.set_populations(fr.astype(np.float32), fb.astype(np.float32))
    s.set_inlet_offset(du)
    s.adjoint_set_uy_seed(w, *UY_STEPS)
    s.step(STEPS)
    assert float(juy.detach()) == s.uy_functional()

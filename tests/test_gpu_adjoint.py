"""GPU adjoint (C ABI mcg_adjoint_record / mcg_adjoint_backward) against the
FP64 CPU adjoint, which test_adjoint_cpu.py pins to central finite
differences of the FP64 oracle.  FP32 forward and adjoint, FP64 parameter
sums: relative agreement 1e-4 with the FP64 adjoint.
"""
import numpy as np
import pytest

import multicomponent_color_gradient_amd as m
from multicomponent_color_gradient_amd.adjoint import bubble_centroid_height, cpu_adjoint, make_steps_function
from test_adjoint_cpu import INNER, make_state

pytestmark = pytest.mark.gpu


def seeded_lambda(shape, seed=1):
    rng = np.random.default_rng(seed)
    lam = [rng.standard_normal(shape), rng.standard_normal(shape)]
    for a in lam:
        a[:, 0], a[:, -1] = 0, 0
        a[:, :, 0], a[:, :, -1] = 0, 0
        a[:, :, :, 0], a[:, :, :, -1] = 0, 0
    return lam


def gpu_solver(case, fr, fb):
    s = m.Solver(case)
    s.init_bubble_tanh()
    s.set_populations(fr.astype(np.float32), fb.astype(np.float32))
    return s


def rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("steps,every", [(4, 2), (10, 3)])
def test_gpu_adjoint_matches_fp64_cpu_adjoint(steps, every):
    case, fr, fb = make_state(N=12)
    # start from the FP32-rounded state so both adjoints linearise the same forward
    fr, fb = fr.astype(np.float32).astype(np.float64), fb.astype(np.float32).astype(np.float64)
    lam = seeded_lambda(fr.shape)
    ref = cpu_adjoint(case, fr, fb, steps, lam[0], lam[1])
    s = gpu_solver(case, fr, fb)
    s.adjoint_record(steps, every)
    gr, gb = s.populations()
    assert rel(gr[INNER], ref["fr"][INNER]) < 1e-5
    lr, lb, (dA, dbeta) = s.adjoint_backward(lam[0], lam[1])
    assert rel(lr[INNER], ref["lam_fr"][INNER]) < 1e-4
    assert rel(lb[INNER], ref["lam_fb"][INNER]) < 1e-4
    assert abs(dA - ref["dA"]) <= 1e-4 * abs(ref["dA"])
    assert abs(dbeta - ref["dbeta"]) <= 1e-4 * abs(ref["dbeta"])
    assert np.all(lr[:, 0] == 0) and np.all(lb[:, :, :, -1] == 0)  # ghost slots


def test_checkpoint_interval_does_not_change_the_adjoint():
    case, fr, fb = make_state(N=12)
    lam = seeded_lambda(fr.shape, 4)
    out = []
    for every in (1, 4, 9):
        s = gpu_solver(case, fr, fb)
        s.adjoint_record(9, every)
        out.append(s.adjoint_backward(lam[0], lam[1]))
    for o in out[1:]:
        assert np.array_equal(o[0], out[0][0]) and np.array_equal(o[1], out[0][1])
        assert abs(o[2][0] - out[0][2][0]) <= 1e-12 * abs(out[0][2][0])


def test_backward_restores_state_and_statistics():
    case, fr, fb = make_state(N=12)
    case.NSTATS_START_STEP = 0
    s = gpu_solver(case, fr, fb)
    s.adjoint_record(6, 2)
    p0 = s.populations()
    r0 = s.radial()
    info0 = s.stats_info()
    f0 = s.fields()
    lam = seeded_lambda(fr.shape)
    s.adjoint_backward(lam[0], lam[1])
    p1 = s.populations()
    assert np.array_equal(p0[0], p1[0]) and np.array_equal(p0[1], p1[1])
    assert s.stats_info() == info0 and np.array_equal(s.radial()["sum_uy"], r0["sum_uy"])
    f1 = s.fields()
    assert np.array_equal(f0["ux"], f1["ux"]) and np.array_equal(f0["rhob"], f1["rhob"])
    # the solver continues as if nothing happened
    s.step(2)
    t = gpu_solver(case, fr, fb)
    t.adjoint_record(6, 2)
    t.step(2)
    assert np.array_equal(s.populations()[0], t.populations()[0])


def test_unsupported_and_misuse_errors():
    jet = m.jet_case(16, 16, 16, jet_radius=3.0)
    s = m.Solver(jet, storage="fp16")
    s.init_jet()
    with pytest.raises(m.McgError):  # 16-bit populations
        s.adjoint_record(2, 1)
    jw = jet.copy()
    jw.wrap_gradient = 1
    s = m.Solver(jw)
    s.init_jet()
    with pytest.raises(m.McgError):  # jet with a wrapped gradient
        s.adjoint_record(2, 1)
    s = m.Solver(jet)
    s.init_bubble()
    with pytest.raises(m.McgError):  # first jet step after a bubble initialisation
        s.adjoint_record(2, 1)
    case, fr, fb = make_state(N=10)
    t = gpu_solver(case, fr, fb)
    lam = seeded_lambda(fr.shape)
    with pytest.raises(m.McgError):  # nothing recorded
        t.adjoint_backward(lam[0], lam[1])
    t.adjoint_record(3, 1)
    t.step(1)
    with pytest.raises(m.McgError):  # advanced after recording
        t.adjoint_backward(lam[0], lam[1])


def test_torch_autograd_centroid_height():
    """Gradients of the blue centroid height after 8 steps through
    torch.autograd equal the direct C-ABI adjoint with the same seed, bit for
    bit; against the FP64 CPU adjoint the state gradient agrees to 1e-4 and
    dJ/dA, dJ/dbeta to 2e-3 (these two are sums of cancelling node terms --
    dJ/dA is 3.7e-6 here -- so FP32 rounding shows at the 1e-3 level)."""
    torch = pytest.importorskip("torch")
    case, fr, fb = make_state(N=12)
    fr32, fb32 = fr.astype(np.float32), fb.astype(np.float32)
    LBMSteps = make_steps_function()
    t_fr = torch.from_numpy(fr32.copy()).requires_grad_(True)
    t_fb = torch.from_numpy(fb32.copy()).requires_grad_(True)
    A = torch.tensor(case.A_override, dtype=torch.float64, requires_grad=True)
    beta = torch.tensor(case.beta_recolor, dtype=torch.float64, requires_grad=True)
    out_r, out_b = LBMSteps.apply(t_fr, t_fb, A, beta, case, 8, 4)
    J = bubble_centroid_height(out_r, out_b)
    J.backward()
    # the seed dJ/df_N of the same objective, then the direct ABI adjoint
    o_b = out_b.detach().clone().requires_grad_(True)
    bubble_centroid_height(out_r.detach(), o_b).backward()
    seed_b = o_b.grad.numpy()
    s = gpu_solver(case, fr32, fb32)
    s.adjoint_record(8, 4)
    lr, lb, (dA, dbeta) = s.adjoint_backward(np.zeros_like(seed_b), seed_b)
    assert np.array_equal(t_fb.grad.numpy(), lb) and np.array_equal(t_fr.grad.numpy(), lr)
    assert abs(A.grad.item() - dA) <= 1e-12 * abs(dA) and abs(beta.grad.item() - dbeta) <= 1e-12 * abs(dbeta)
    ref = cpu_adjoint(case, fr32.astype(np.float64), fb32.astype(np.float64), 8, np.zeros_like(seed_b),
                      seed_b.astype(np.float64))
    assert rel(lb[INNER], ref["lam_fb"][INNER]) < 1e-4
    assert abs(dA - ref["dA"]) <= 2e-3 * abs(ref["dA"])
    assert abs(dbeta - ref["dbeta"]) <= 2e-3 * abs(ref["dbeta"])

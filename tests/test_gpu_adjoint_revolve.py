"""Memory-bounded checkpointing of the adjoint (mcg_adjoint_set_checkpoint_limit):
with fewer stored states than segments the sweep recomputes the missing
segment states on a binomial schedule.  The forward steps are deterministic,
so every gradient must be bitwise that of the unlimited record:
  * rising bubble (walls, body force, per-node viscosity), 24 steps in 12
    segments of 2, limits 6, 3, 2, 1;
  * jet (inlet/outlet arrays, outlet state, u_y seed, inlet gradient), 7 steps
    in 7 segments, limits 3, 1;
  * rising bubble, 25 steps in segments of 4 (the last of one step), limits 3, 1;
and a second sweep over the same record gives the same result.  The number of
recomputed forward steps equals that of the schedule restated below, and the
unlimited record recomputes each step once.
"""
import numpy as np
import pytest

import multicomponent_color_gradient_amd as m
from test_adjoint_jet_cpu import STEPS as JET_STEPS, UY_STEPS, base_case, initial_state, weights
from test_adjoint_rising_cpu import make_state

pytestmark = pytest.mark.gpu


def beta(f, t):
    v = 1.0
    for i in range(1, f + 1):
        v = v * (t + i) / i
    return v


def split(a, b, f):
    L = b - a
    t = 0
    while beta(f, t) < L:
        t += 1
    left = max(1, L - int(min(beta(f - 1, t), L)))
    return a + min(left, L - 1)


def schedule_cost(nsteps, every, limit):
    """Forward steps the sweep recomputes: the schedule of mcg_adjoint_backward
    (first descent stored by the record, binomial splits, linear when no slot
    is free), counted on the host."""
    nseg = -(-nsteps // every)
    nslot = min(nseg, limit) if limit > 0 else nseg
    slot = [-1] * nseg
    if nslot == nseg:
        slot = list(range(nseg))
    else:
        used, a = 0, 0
        slot[0] = used
        used += 1
        while used < nslot and nseg - a > 1:
            a = split(a, nseg, nslot - used)
            slot[a] = used
            used += 1
    free = [v for v in range(nslot) if v not in slot]
    cost = [0]

    def seg_len(k):
        return min(every, nsteps - k * every)

    def rev(sa, sb):
        while sb - sa > 1:
            f = len(free) + sum(1 for j in range(sa + 1, sb) if slot[j] >= 0)
            if f == 0:
                for j in range(sb - 1, sa, -1):
                    cost[0] += (j - sa) * every + seg_len(j)
                break
            mm = split(sa, sb, f)
            if slot[mm] < 0:
                slot[mm] = free.pop()
                cost[0] += (mm - sa) * every
            rev(mm, sb)
            free.append(slot[mm])
            slot[mm] = -1
            sb = mm
        cost[0] += seg_len(sa)

    rev(0, nseg)
    return cost[0]


def test_schedule_model_limits():
    # one state per segment: each step once; a single state: every segment's
    # first state recomputed from step 0 (sum of j*every) plus each segment once
    assert schedule_cost(24, 2, 0) == 24 and schedule_cost(24, 2, 12) == 24
    assert schedule_cost(24, 2, 1) == 24 + sum(2 * j for j in range(1, 12))
    costs = [schedule_cost(24, 2, k) for k in (1, 2, 3, 6, 12)]
    assert costs == sorted(costs, reverse=True)


def sweep(s, lam):
    lr, lb, _ = s.adjoint_backward(lam[0], lam[1])
    g = s.adjoint_gradients()
    return lr, lb, g, s.adjoint_recomputed_steps()


def same(a, b):
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert a[2] == b[2]


def test_rising_bubble_gradients_independent_of_the_checkpoint_limit():
    case, fr, fb, _ = make_state()
    steps, every = 24, 2
    rng = np.random.default_rng(5)
    lam = [rng.standard_normal((27,) + fr.shape[1:]).astype(np.float32) for _ in range(2)]
    res = {}
    for limit in (0, 6, 3, 2, 1):
        print(f"rising bubble {case.NX}x{case.NY}x{case.NZ}: {steps} steps, every {every}, limit {limit}", flush=True)
        s = m.Solver(case)
        s.init_bubble_tanh()
        s.set_populations(fr.astype(np.float32), fb.astype(np.float32))
        s.adjoint_set_checkpoint_limit(limit)
        s.adjoint_record(steps, every)
        res[limit] = sweep(s, lam)
        assert res[limit][3] == schedule_cost(steps, every, limit), (limit, res[limit][3])
        if limit == 2:
            again = sweep(s, lam)
            same(again, res[limit])
        s.close()
    assert res[0][3] == steps
    for limit in (6, 3, 2, 1):
        same(res[limit], res[0])


def test_jet_gradients_independent_of_the_checkpoint_limit():
    case = base_case()
    fr, fb, _ = initial_state(case)
    w = weights(case)
    rng = np.random.default_rng(6)
    lam = [rng.standard_normal(fr.shape).astype(np.float32) * 0.1 for _ in range(2)]
    res = {}
    for limit in (0, 3, 1):
        print(f"jet {case.NX}x{case.NY}x{case.NZ}: {JET_STEPS} steps, every 1, limit {limit}", flush=True)
        s = m.Solver(case)
        s.init_jet()
        s.set_populations(fr.astype(np.float32), fb.astype(np.float32))
        s.adjoint_set_uy_seed(w, *UY_STEPS)
        s.adjoint_set_checkpoint_limit(limit)
        s.adjoint_record(JET_STEPS, 1)
        r = sweep(s, lam)
        (dU, damp), din = s.adjoint_inlet_gradient()
        res[limit] = (r, dU, damp, din)
        assert r[3] == schedule_cost(JET_STEPS, 1, limit), (limit, r[3])
        s.close()
    for limit in (3, 1):
        same(res[limit][0], res[0][0])
        assert res[limit][1:3] == res[0][1:3]
        assert np.array_equal(res[limit][3], res[0][3])


def test_torch_jet_steps_with_a_state_cap():
    """LBMJetSteps with max_states = 2 (trailing argument): the gradients equal
    those of the unlimited record through the C ABI, bit for bit."""
    import torch

    from multicomponent_color_gradient_amd.adjoint import make_jet_steps_function

    case = base_case()
    fr, fb, _ = initial_state(case)
    w = weights(case)
    rng = np.random.default_rng(7)
    lam = [rng.standard_normal(fr.shape).astype(np.float32) * 0.1 for _ in range(2)]
    du = np.zeros((3, case.NZ, case.NX), np.float32)
    print(f"torch jet {case.NX}x{case.NY}x{case.NZ}: {JET_STEPS} steps, every 1, max_states 2", flush=True)
    F = make_jet_steps_function()
    fr0 = torch.tensor(fr, dtype=torch.float32, requires_grad=True)
    fb0 = torch.tensor(fb, dtype=torch.float32, requires_grad=True)
    A = torch.tensor(case.A_override, dtype=torch.float64, requires_grad=True)
    beta = torch.tensor(case.beta_recolor, dtype=torch.float64, requires_grad=True)
    U = torch.tensor(case.jet_velocity, dtype=torch.float64, requires_grad=True)
    amp = torch.tensor(case.perturbation_amp, dtype=torch.float64, requires_grad=True)
    dut = torch.tensor(du, dtype=torch.float64, requires_grad=True)
    gr, gb, juy = F.apply(fr0, fb0, A, beta, U, amp, dut, case, JET_STEPS, 1, torch.tensor(w), UY_STEPS, 2)
    J = (torch.tensor(lam[0]) * gr).sum() + (torch.tensor(lam[1]) * gb).sum()
    (J.double() + juy).backward()
    s = m.Solver(case)
    s.init_jet()
    s.set_populations(fr.astype(np.float32), fb.astype(np.float32))
    s.set_inlet_offset(du)
    s.adjoint_set_uy_seed(w, *UY_STEPS)
    s.adjoint_record(JET_STEPS, 1)
    lr, lb, (dA, dbeta) = s.adjoint_backward(lam[0], lam[1])
    (dU, damp), din = s.adjoint_inlet_gradient()
    s.close()
    assert np.array_equal(fr0.grad.numpy(), lr) and np.array_equal(fb0.grad.numpy(), lb)
    assert float(A.grad) == dA and float(beta.grad) == dbeta
    assert float(U.grad) == dU and float(amp.grad) == damp
    assert np.array_equal(dut.grad.numpy(), din)


def test_checkpoint_limit_misuse():
    case, fr, fb, _ = make_state()
    s = m.Solver(case)
    with pytest.raises(m.McgError):
        s.adjoint_set_checkpoint_limit(-1)
    assert s.adjoint_recomputed_steps() == 0  # no sweep yet
    s.close()


def test_short_last_segment():
    """25 steps with K = 4: seven segments, the last of one step; limits 3 and
    1 give the unlimited record's gradients bit for bit and the schedule's
    recomputation counts."""
    case, fr, fb, _ = make_state()
    steps, every = 25, 4
    rng = np.random.default_rng(8)
    lam = [rng.standard_normal((27,) + fr.shape[1:]).astype(np.float32) for _ in range(2)]
    res = {}
    for limit in (0, 3, 1):
        print(f"rising bubble {case.NX}x{case.NY}x{case.NZ}: {steps} steps, every {every}, limit {limit}", flush=True)
        s = m.Solver(case)
        s.init_bubble_tanh()
        s.set_populations(fr.astype(np.float32), fb.astype(np.float32))
        s.adjoint_set_checkpoint_limit(limit)
        s.adjoint_record(steps, every)
        res[limit] = sweep(s, lam)
        assert res[limit][3] == schedule_cost(steps, every, limit), (limit, res[limit][3])
        s.close()
    for limit in (3, 1):
        same(res[limit], res[0])

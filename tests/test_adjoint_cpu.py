"""FP64 CPU adjoint of the step (C ABI mcg_cpu_adjoint, host code) against the
independent FP64 oracle: its forward must reproduce the oracle's step, and
its gradients must match central finite differences of the oracle's step.

State: a periodic tanh bubble whose populations are perturbed at random, so
that every node has velocity and non-equilibrium stress and the interface is
wide.  Objective J = sum lam_N . f^N over the interior with a fixed random
lam_N; gradients with respect to A, beta and the initial populations.
"""
import numpy as np
import pytest

import multicomponent_color_gradient_amd as m
from multicomponent_color_gradient_amd.adjoint import cpu_adjoint
from oracle.oracle import Oracle

INNER = (slice(None),) + (slice(1, -1),) * 3
STEPS = 4


def make_state(N=10, A=1e-2, beta=0.7, seed=0):
    case = m.periodic_bubble_case(N, radius=2.6, tau=0.8, A=A, beta=beta, width=1.5)
    o = Oracle(case)
    o.init_bubble_tanh()
    fr, fb = o.populations()
    rng = np.random.default_rng(seed)
    fr = fr * (1 + 0.04 * (rng.random(fr.shape) - 0.5))
    fb = fb * (1 + 0.04 * (rng.random(fb.shape) - 0.5))
    return case, fr, fb


def oracle_final(case, fr, fb, steps=STEPS, A=None, beta=None):
    c = case.copy()
    if A is not None:
        c.A_override = A
    if beta is not None:
        c.beta_recolor = beta
    o = Oracle(c)
    o.init_bubble_tanh()
    o.set_populations(fr, fb)
    o.step(steps)
    return o.populations()


def J(lam, f):
    return float(np.sum(lam[0][INNER] * f[0][INNER]) + np.sum(lam[1][INNER] * f[1][INNER]))


@pytest.fixture(scope="module")
def setup():
    case, fr, fb = make_state()
    rng = np.random.default_rng(1)
    lam = (rng.standard_normal(fr.shape), rng.standard_normal(fb.shape))
    for a in lam:
        a[:, 0, :, :] = a[:, -1, :, :] = 0
        a[:, :, 0, :] = a[:, :, -1, :] = 0
        a[:, :, :, 0] = a[:, :, :, -1] = 0
    res = cpu_adjoint(case, fr, fb, STEPS, lam[0], lam[1])
    return case, fr, fb, lam, res


def test_cpu_forward_reproduces_oracle(setup):
    case, fr, fb, lam, res = setup
    gr, gb = oracle_final(case, fr, fb)
    assert np.max(np.abs(res["fr"][INNER] - gr[INNER])) < 1e-13
    assert np.max(np.abs(res["fb"][INNER] - gb[INNER])) < 1e-13


@pytest.mark.parametrize("param", ["A", "beta"])
def test_parameter_gradient_matches_fd(setup, param):
    case, fr, fb, lam, res = setup
    p0 = case.A_override if param == "A" else case.beta_recolor
    h = 1e-5 * p0
    kw_p = {param: p0 + h}
    kw_m = {param: p0 - h}
    fd = (J(lam, oracle_final(case, fr, fb, **kw_p)) - J(lam, oracle_final(case, fr, fb, **kw_m))) / (2 * h)
    adj = res["dA"] if param == "A" else res["dbeta"]
    assert abs(fd) > 1e-3  # the parameter matters for this state
    assert abs(adj - fd) <= 1e-6 * abs(fd), (adj, fd)


@pytest.mark.parametrize("seed", [2, 3])
def test_state_gradient_matches_fd(setup, seed):
    case, fr, fb, lam, res = setup
    rng = np.random.default_rng(seed)
    dr = rng.standard_normal(fr.shape) * fr * 1e-3
    db = rng.standard_normal(fb.shape) * fb * 1e-3
    eps = 1e-4
    jp = J(lam, oracle_final(case, fr + eps * dr, fb + eps * db))
    jm = J(lam, oracle_final(case, fr - eps * dr, fb - eps * db))
    fd = (jp - jm) / (2 * eps)
    adj = float(np.sum(res["lam_fr"][INNER] * dr[INNER]) + np.sum(res["lam_fb"][INNER] * db[INNER]))
    assert abs(adj - fd) <= 1e-6 * abs(fd), (adj, fd)


def test_adjoint_of_zero_steps_is_identity(setup):
    case, fr, fb, lam, _ = setup
    res = cpu_adjoint(case, fr, fb, 0, lam[0], lam[1])
    assert np.array_equal(res["lam_fr"], lam[0]) and np.array_equal(res["lam_fb"], lam[1])
    assert res["dA"] == 0 and res["dbeta"] == 0


def test_unsupported_cases_are_refused():
    case = m.jet_case(12, 12, 12, jet_radius=3.0)
    fr = np.zeros((27, 12, 12, 12))
    with pytest.raises(m.McgError):
        cpu_adjoint(case, fr, fr, 1)

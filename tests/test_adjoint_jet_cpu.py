"""FP64 CPU adjoint of jet steps (inlet and convective outlet differentiated;
C ABI mcg_cpu_adjoint_jet) against the independent FP64 oracle.

State: the jet initialisation with the tanh inlet and the seeded transverse
perturbation, its interior populations perturbed at random so that the whole
box moves.  Objective J = <lam, f_N> + sum over the moments passes of steps
2..5 of sum_x w(x) u_y(x) (w on a centre-line segment, the shape of the
inverse problem's time-averaged centre-line velocity).  Every gradient is
compared with central finite differences of the oracle's step.
"""
import numpy as np
import pytest

import multicomponent_color_gradient_amd as m
from multicomponent_color_gradient_amd.adjoint import cpu_adjoint_jet
from oracle.oracle import Oracle

INNER = (slice(None),) + (slice(1, -1),) * 3
STEPS = 7
UY_STEPS = (2, 5)


def base_case(**kw):
    return m.perturbed_jet_case(14, 16, 14, jet_radius=3.2, perturbation=0.2, seed=99, **kw)


def initial_state(case, seed=0):
    o = Oracle(case)
    o.init_jet()
    fr, fb = o.populations()
    fields = o.fields()
    rng = np.random.default_rng(seed)
    fr = fr * (1 + 0.03 * (rng.random(fr.shape) - 0.5))
    fb = fb * (1 + 0.03 * (rng.random(fb.shape) - 0.5))
    return fr, fb, fields


def weights(case):
    w = np.zeros((case.NZ, case.NY, case.NX))
    xc, zc = int(case.jet_x0), int(case.jet_z0)
    w[zc, 6:12, xc] = 1.0
    w[zc + 1, 6:12, xc] = 0.5
    return w


def oracle_J(case, fr, fb, lam, w, steps=STEPS, du=None, **over):
    c = case.copy()
    for k, v in over.items():
        setattr(c, k, v)
    o = Oracle(c)
    o.init_jet()
    o.set_populations(fr, fb)
    if du is not None:
        o.set_inlet_offset(du)
    J = 0.0
    for k in range(steps):
        o.step(1)
        n = k  # after step k+1 the moment arrays hold the moments pass of step k
        if UY_STEPS[0] <= n <= UY_STEPS[1]:
            J += float(np.sum(w * o.fields()["uy"]))
    gr, gb = o.populations()
    J += float(np.sum(lam[0][INNER] * gr[INNER]) + np.sum(lam[1][INNER] * gb[INNER]))
    return J, (gr, gb)


@pytest.fixture(scope="module")
def setup():
    case = base_case()
    fr, fb, fields = initial_state(case)
    rng = np.random.default_rng(1)
    lam = [rng.standard_normal(fr.shape) * 0.1, rng.standard_normal(fb.shape) * 0.1]
    for a in lam:
        a[:, 0], a[:, -1] = 0, 0
        a[:, :, 0], a[:, :, -1] = 0, 0
        a[:, :, :, 0], a[:, :, :, -1] = 0, 0
    w = weights(case)
    res = cpu_adjoint_jet(case, fr, fb, fields, STEPS, lam[0], lam[1], w_uy=w, uy_steps=UY_STEPS)
    return case, fr, fb, lam, w, res


def test_jet_forward_reproduces_oracle(setup):
    case, fr, fb, lam, w, res = setup
    _, (gr, gb) = oracle_J(case, fr, fb, lam, w)
    assert np.max(np.abs(res["fr"][INNER] - gr[INNER])) < 1e-13
    assert np.max(np.abs(res["fb"][INNER] - gb[INNER])) < 1e-13


@pytest.mark.parametrize("param,key,rel_h", [("A_override", "dA", 1e-5), ("beta_recolor", "dbeta", 1e-5),
                                             ("jet_velocity", "dU", 1e-5), ("perturbation_amp", "damp", 1e-5)])
def test_jet_parameter_gradients_match_fd(setup, param, key, rel_h):
    case, fr, fb, lam, w, res = setup
    p0 = getattr(case, param)
    h = rel_h * p0
    jp, _ = oracle_J(case, fr, fb, lam, w, **{param: p0 + h})
    jm, _ = oracle_J(case, fr, fb, lam, w, **{param: p0 - h})
    fd = (jp - jm) / (2 * h)
    assert abs(fd) > 1e-4
    assert abs(res[key] - fd) <= 2e-6 * abs(fd), (key, res[key], fd)


@pytest.mark.parametrize("seed", [5, 6])
def test_jet_state_gradient_matches_fd(setup, seed):
    case, fr, fb, lam, w, res = setup
    rng = np.random.default_rng(seed)
    dr = rng.standard_normal(fr.shape) * fr * 1e-3
    db = rng.standard_normal(fb.shape) * fb * 1e-3
    eps = 1e-4
    jp, _ = oracle_J(case, fr + eps * dr, fb + eps * db, lam, w)
    jm, _ = oracle_J(case, fr - eps * dr, fb - eps * db, lam, w)
    fd = (jp - jm) / (2 * eps)
    adj = float(np.sum(res["lam_fr"][INNER] * dr[INNER]) + np.sum(res["lam_fb"][INNER] * db[INNER]))
    assert abs(adj - fd) <= 2e-6 * abs(fd), (adj, fd)


def test_inlet_gradient_sums_to_the_velocity_derivative(setup):
    """dJ/dU = sum over jet columns of dJ/de . de/dU; with the perturbation
    switched off de/dU is (0, profile(r), 0), so the per-column gradients must
    sum, weighted by the profile, to dJ/dU."""
    case0, fr, fb, lam, w, _ = setup
    case = case0.copy()
    case.perturbation_amp = 0.0
    res = cpu_adjoint_jet(case, fr, fb, initial_state(case)[2], STEPS, lam[0], lam[1], w_uy=w, uy_steps=UY_STEPS)
    z, x = np.meshgrid(np.arange(case.NZ), np.arange(case.NX), indexing="ij")
    r = np.hypot(x - case.jet_x0, z - case.jet_z0)
    prof = 0.5 * (1 - np.tanh((r - case.jet_radius) / case.interface_width))
    assert abs(float(np.sum(res["dinlet"][1] * prof)) - res["dU"]) <= 1e-10 * abs(res["dU"])
    assert np.all(res["dinlet"][:, r > case.jet_radius] == 0)


def jet_columns(case):
    z, x = np.meshgrid(np.arange(case.NZ), np.arange(case.NX), indexing="ij")
    inner = (x >= 1) & (x <= case.NX - 2) & (z >= 1) & (z <= case.NZ - 2)
    return inner & ((x - case.jet_x0) ** 2 + (z - case.jet_z0) ** 2 <= case.jet_radius ** 2)


def test_inlet_gradient_matches_fd_on_random_directions(setup):
    """dJ/d(du) for the per-column inlet velocity offset (the control of the
    inverse jet problem) against central differences of the FP64 oracle on 8
    seeded random directions over the jet columns, at a nonzero offset."""
    case, fr, fb, lam, w, _ = setup
    jet = jet_columns(case)
    rng = np.random.default_rng(11)
    du0 = rng.standard_normal((3, case.NZ, case.NX)) * 2e-3 * jet
    res = cpu_adjoint_jet(case, fr, fb, initial_state(case)[2], STEPS, lam[0], lam[1], w_uy=w, uy_steps=UY_STEPS,
                          inlet_du=du0)
    assert np.all(res["dinlet"][:, ~jet] == 0)
    for k in range(8):
        d = rng.standard_normal((3, case.NZ, case.NX)) * jet
        h = 1e-4 / np.max(np.abs(d))
        jp, _ = oracle_J(case, fr, fb, lam, w, du=du0 + h * d)
        jm, _ = oracle_J(case, fr, fb, lam, w, du=du0 - h * d)
        fd = (jp - jm) / (2 * h)
        adj = float(np.sum(res["dinlet"] * d))
        assert abs(adj - fd) <= 2e-6 * abs(fd), (k, adj, fd)


def test_uniform_axial_offset_equals_a_faster_jet(setup):
    """Sharp inlet, no perturbation: an offset c on u_y of every jet column is
    the same inlet as jet_velocity U + c (pins the offset in the oracle)."""
    case0, fr, fb, lam, w, _ = setup
    case = case0.copy()
    case.inlet_profile = 0
    case.perturbation_amp = 0.0
    du = np.zeros((3, case.NZ, case.NX))
    du[1][jet_columns(case)] = 0.01
    ja, (ra, ba) = oracle_J(case, fr, fb, lam, w, du=du)
    jb, (rb, bb) = oracle_J(case, fr, fb, lam, w, jet_velocity=case.jet_velocity + 0.01)
    assert np.array_equal(ra, rb) and np.array_equal(ba, bb)
    assert ja == jb

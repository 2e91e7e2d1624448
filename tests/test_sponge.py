"""The sponge layer (utilities/mathUtilities.cuh:48-90): nu(y) = nu (1 + K s^3)
with s = clamp01((y - y_start)/(y_end - y_start)), y_end = NY - 2,
y_start = y_end - sponge_cells + 1, and omega(y) = 1/(1/2 + nu(y)/cs^2) in
the collision (lbm.cuh:211) and the outlet (jet/boundary_conditions.cuh:211).

Pin (oracle, CPU): a periodic box holding a uniform state whose only
non-equilibrium moment is Pi_xz = q/9 (f_i = w_i (1 + q c_x c_z), a pure
second-order Hermite component, so the regularisation keeps it whole).  One
step relaxes it to (1 - omega(y)) of itself at each plane and streams it; the
directions with c_x c_z != 0 carry 2/3 of the w (c_x c_z)^2 weight with c_y = 0
and 1/6 with each c_y = +-1, so after the step

    Pi_xz(y) = (q/9) [2/3 (1 - omega(y)) + 1/6 (1 - omega(y-1)) + 1/6 (1 - omega(y+1))]

exactly.  omega takes three closed-form values that a wrong layer position,
exponent or gain would miss: 1/taub below and at y_start, 1/(1/2 + (1 +
K/8)(taub - 1/2)) at the layer's midpoint and 1/(1/2 + (1 + K)(taub - 1/2))
at y_end.  With A = 0 each colour's mass is conserved exactly whatever omega(y).

GPU: the same pin through the C ABI, and a jet with K > 0 (sponge in front of
the convective outlet) against the FP64 oracle.
"""
import numpy as np
import pytest

import multicomponent_color_gradient_amd as m
from conftest import INNER
from oracle import oracle as orc
from oracle.oracle import Oracle

TAU = 0.7
K = 3.0
CELLS = 9  # y_start = y_end - 8: the midpoint s = 1/2 is a node


def sponge_box(NY=24):
    c = m.periodic_bubble_case(8, radius=2.0, tau=TAU, A=0.0)
    c.NY = NY
    c.sigma = 0.0
    c.sponge_K = K
    c.sponge_cells = CELLS
    return c


def omega_profile(c):
    """omega(y) for y = 0..NY-1 from the three closed-form values and 1/taub
    below the layer; the layer's other nodes interpolate s^3 on a uniform grid."""
    yend = c.NY - 2
    ystart = yend - CELLS + 1
    nu = (TAU - 0.5) / 3.0
    om = np.full(c.NY, 1.0 / TAU)
    for y in range(ystart, c.NY):
        s = min(1.0, (y - ystart) / (yend - ystart))
        om[y] = 1.0 / (0.5 + 3.0 * nu * (1.0 + K * s ** 3))
    mid = ystart + (CELLS - 1) // 2
    assert om[ystart] == 1.0 / TAU
    assert om[mid] == pytest.approx(1.0 / (0.5 + (1.0 + K / 8.0) * (TAU - 0.5)), rel=1e-15)
    assert om[yend] == pytest.approx(1.0 / (0.5 + (1.0 + K) * (TAU - 0.5)), rel=1e-15)
    return om


def shear_state(c, q):
    L = orc.lattice()
    w, cx, cz = L["w"], L["cx"], L["cz"]
    shape = (27, c.NZ, c.NY, c.NX)
    fr = np.broadcast_to((w * (1.0 + q * cx * cz))[:, None, None, None], shape).copy()
    return fr, np.zeros(shape)


def pi_xz_after_one_step(fr):
    L = orc.lattice()
    return np.tensordot(L["cx"] * L["cz"], fr, axes=(0, 0))  # (NZ, NY, NX)


def expected_pi_xz(c, q):
    om = omega_profile(c)
    per = c.NY - 2
    ys = np.arange(1, c.NY - 1)
    up = 1 + (ys - 2) % per  # y - 1 wrapped over the interior planes
    dn = 1 + ys % per        # y + 1 wrapped
    return (q / 9.0) * (2.0 / 3.0 * (1 - om[ys]) + 1.0 / 6.0 * (1 - om[up]) + 1.0 / 6.0 * (1 - om[dn]))


def test_sponge_profile_in_the_oracle_step():
    c = sponge_box()
    q = 1e-2
    o = Oracle(c)
    o.init_bubble_tanh()
    o.set_populations(*shear_state(c, q))
    o.step(1)
    fr, fb = o.populations()
    pxz = pi_xz_after_one_step(fr)[1:-1, 1:-1, 1:-1]
    exp = expected_pi_xz(c, q)
    assert np.max(np.abs(pxz - exp[None, :, None])) <= 1e-16
    assert np.max(np.abs(fb[(slice(None),) + INNER])) == 0.0
    # the midpoint plane's relaxation differs from the s^2 or s^4 profiles by far more than rounding
    om = omega_profile(c)
    assert abs(om[c.NY - 2 - 4] - 1.0 / (0.5 + (1 + K / 4) * (TAU - 0.5))) > 1e-2


def test_sponge_keeps_colour_masses_in_a_periodic_bubble():
    c = m.periodic_bubble_case(18, radius=4.5, tau=0.8, A=0.0)
    c.sigma = 0.0
    c.sponge_K, c.sponge_cells = 5.0, 8
    o = Oracle(c)
    o.init_bubble_tanh()
    o.step(1)
    s0 = o.global_sums()
    o.step(30)
    s1 = o.global_sums()
    assert abs(s1[0] - s0[0]) <= 1e-12 * s0[0] and abs(s1[1] - s0[1]) <= 1e-12 * s0[1]


@pytest.mark.gpu
def test_gpu_sponge_profile(mcg):
    c = sponge_box()
    q = 1e-2
    s = mcg.Solver(c)
    s.init_bubble_tanh()
    fr, fb = shear_state(c, q)
    s.set_populations(fr.astype(np.float32), fb.astype(np.float32))
    s.step(1)
    gr, gb = s.populations()
    pxz = pi_xz_after_one_step(gr.astype(np.float64))[1:-1, 1:-1, 1:-1]
    exp = expected_pi_xz(c, q)
    # FP32 populations of size ~w: the sum over 12 directions rounds to a few 1e-9
    assert np.max(np.abs(pxz - exp[None, :, None])) <= 2e-8
    s.close()


@pytest.mark.gpu
@pytest.mark.parametrize("kernel", ["fused", "twopass"])
def test_gpu_jet_with_sponge_matches_oracle(mcg, kernel):
    """Reference-semantics jet (sharp inlet, reference gradient) with a sponge
    of 10 planes in front of the outlet, K = 20, against the FP64 oracle."""
    c = mcg.jet_case(24, 28, 24, jet_radius=4.0, jet_velocity=0.05, Re=300.0)
    c.sponge_K, c.sponge_cells = 20.0, 10
    s = mcg.Solver(c, kernel=kernel)
    o = Oracle(c)
    s.init_jet()
    o.init_jet()
    for steps, scale in ((10, 1.0), (90, 10.0)):
        s.step(steps)
        o.step(steps)
        g, r = s.fields(), o.fields()
        for k in ("rhor", "rhob", "ux", "uy", "uz"):
            tol = (5e-6 if k.startswith("rho") else 1e-6) * scale
            e = float(np.max(np.abs(g[k].astype(np.float64) - r[k])))
            assert e <= tol, f"{kernel} {k} after {steps}: {e:.3e}"
    # the sponge is active: without it u_y differs by 2.5e-4 (rho_r 5.9e-4),
    # far beyond the bounds above
    c0 = mcg.jet_case(24, 28, 24, jet_radius=4.0, jet_velocity=0.05, Re=300.0)
    o0 = Oracle(c0)
    o0.init_jet()
    o0.step(100)
    assert float(np.max(np.abs(o0.fields()["uy"] - r["uy"]))) > 1e-4
    s.close()

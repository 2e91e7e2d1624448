"""Pins of the oracle's x/z walls (half-way bounce-back) and body force
(F_y = -gravity rho_r on red), the two extensions of the rising-bubble case.

  * body force: the post-collision momentum is j + F at every node, so in a
    periodic box the total y momentum falls by gravity x (red mass) per step,
    exactly (FP64 rounding);
  * walls: bounce-back conserves the mass of each colour exactly;
  * walls + body force, one phase: the steady flow in the square duct matches
    the series solution of the Poisson problem -nu lap u = g (walls half-way
    between the ghost and the first fluid node) to 1% (measured 0.46%);
  * no body force: both extensions off reproduce the periodic step bit for bit.
"""
import numpy as np

import multicomponent_color_gradient_amd as m
from oracle.oracle import Oracle, lattice


def momentum_y(o):
    fr, fb = o.populations()
    cy = np.array(lattice()["cy"], float)
    inner = (slice(None),) + (slice(1, -1),) * 3
    return float(np.tensordot(cy, (fr + fb)[inner], axes=(0, 0)).sum())


def test_body_force_adds_exactly_F_per_step():
    c = m.periodic_bubble_case(12, radius=3.0, tau=0.8)
    c.gravity = 1e-3
    o = Oracle(c)
    o.init_bubble_tanh()
    o.step(1)
    mass_r = float(o.global_sums()[0])
    p0 = momentum_y(o)
    o.step(5)
    p1 = momentum_y(o)
    assert abs((p1 - p0) - (-5 * c.gravity * mass_r)) <= 1e-12 * mass_r


def test_walls_conserve_each_colour():
    c = m.rising_bubble_case(12, 14, 12, radius=3.0, y0=6.0, gravity=0.0)
    o = Oracle(c)
    o.init_bubble_tanh()
    o.step(1)
    g0 = o.global_sums()
    o.step(20)
    g1 = o.global_sums()
    assert abs(g1[0] - g0[0]) <= 1e-12 * g0[0] and abs(g1[1] - g0[1]) <= 1e-12 * g0[1]


def duct_series(x, z, a, b, G, mu, terms=60):
    """Velocity of fully developed flow in the duct |x| < a, |z| < b under the force density G."""
    u = np.zeros(np.broadcast(x, z).shape)
    for k in range(terms):
        n = 2 * k + 1
        u += ((-1) ** k / n ** 3) * (1 - np.cosh(n * np.pi * z / (2 * a)) / np.cosh(n * np.pi * b / (2 * a))) * \
            np.cos(n * np.pi * x / (2 * a))
    return 16 * a * a * G / (mu * np.pi ** 3) * u


def test_square_duct_poiseuille():
    N = 12
    c = m.rising_bubble_case(N, 4, N, radius=0.0, y0=1.0, gravity=1e-5, taur=1.0, taub=1.0)
    c.viscosity_mode = 0
    o = Oracle(c)
    o.init_jet()  # all red at rest (the jet initialisation with y periodic has no inlet)
    o.step(3000)
    uy = o.fields()["uy"][1:-1, 1, 1:-1]  # (z, x) over the fluid nodes
    a = (N - 2) / 2.0  # walls half-way: the duct spans x in [0.5, N - 1.5]
    xs = np.arange(1, N - 1) - (N - 1) / 2.0
    Z, X = np.meshgrid(xs, xs, indexing="ij")
    nu = (1.0 - 0.5) / 3.0
    ref = -duct_series(X, Z, a, a, c.gravity, nu)
    assert np.max(np.abs(uy - ref)) <= 0.01 * np.max(np.abs(ref)), (uy.min(), ref.min())


def test_extensions_off_are_bit_identical():
    c = m.periodic_bubble_case(10, radius=3.0, tau=0.8)
    o1 = Oracle(c)
    o1.init_bubble_tanh()
    o1.step(4)
    c2 = c.copy()
    c2.gravity = 0.0
    c2.wall_xz = 0
    o2 = Oracle(c2)
    o2.init_bubble_tanh()
    o2.step(4)
    assert np.array_equal(o1.populations()[0], o2.populations()[0])

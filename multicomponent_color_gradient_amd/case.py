"""Case descriptions: the run-time form of the reference's constants.cuh.

The reference fixes its case at compile time (constants.cuh:8-76).  Here a
case is an ``McgCase`` struct; the helpers below fill it with the reference's
derivations:

  nu    = 2 R U / Re                       (constants.cuh:60)
  taur  = taub = 1/2 + nu / cs^2           (constants.cuh:62-64)
  sigma = rhob0 U^2 (2 R) / We             (constants.cuh:73)
  A     = sigma / (4 cs^4 taub)            (forces.cuh:87-90,102-103)
"""
from __future__ import annotations

from ._abi import (MCG_INLET_SHARP, MCG_INLET_TANH, MCG_VISC_INTERFACE, MCG_VISC_SINGLE, MCG_Y_JET, MCG_Y_PERIODIC,
                   McgCase)

CS2 = 1.0 / 3.0


def reference_jet() -> McgCase:
    """The reference's own case (constants.cuh:8-76): 128x400x128 jet, R=7, U=0.05, Re 5000, We 2000."""
    c = McgCase()
    c.NX, c.NY, c.NZ = 128, 400, 128
    c.NSTEP, c.NOUTPUT, c.NSTATS_START_STEP, c.NSTATS_SAMPLE = 200000, 2000, 6000, 1
    c.write_vti_output = 1
    c.sponge_cells = 34
    c.y_mode = MCG_Y_JET
    c.wrap_gradient = 0
    c.viscosity_mode = MCG_VISC_SINGLE
    c.inlet_profile = MCG_INLET_SHARP
    c.jet_radius, c.jet_velocity = 7.0, 0.05
    c.jet_x0 = (c.NX - 1) / 2.0
    c.jet_z0 = (c.NZ - 1) / 2.0
    c.bubble_radius = 10.0
    c.bubble_x0 = (c.NX - 1) / 2.0
    c.bubble_y0 = (c.NY - 1) / 2.0
    c.bubble_z0 = (c.NZ - 1) / 2.0
    c.rhor0 = c.rhob0 = 1.0
    c.Re, c.We = 5000.0, 2000.0
    nu = (2.0 * c.jet_radius * c.jet_velocity) / c.Re
    c.taur = c.taub = 0.5 + nu / CS2
    c.sigma = (c.rhob0 * c.jet_velocity * c.jet_velocity * 2.0 * c.jet_radius) / c.We
    c.beta_recolor = 0.80
    c.A_override = 0.0
    c.sponge_K = 0.0
    c.interface_width = 2.0
    c.perturbation_amp = 0.0
    return c


def jet_case(NX: int, NY: int, NZ: int, jet_radius: float, jet_velocity: float = 0.05, Re: float = 5000.0,
             We: float = 2000.0, beta: float = 0.80, **overrides) -> McgCase:
    """A reference-semantics jet on another grid: same derivations as constants.cuh:59-76."""
    c = reference_jet()
    c.NX, c.NY, c.NZ = NX, NY, NZ
    c.jet_radius, c.jet_velocity = float(jet_radius), float(jet_velocity)
    c.jet_x0 = (NX - 1) / 2.0
    c.jet_z0 = (NZ - 1) / 2.0
    c.bubble_x0, c.bubble_y0, c.bubble_z0 = (NX - 1) / 2.0, (NY - 1) / 2.0, (NZ - 1) / 2.0
    c.Re, c.We = float(Re), float(We)
    nu = (2.0 * c.jet_radius * c.jet_velocity) / c.Re
    c.taur = c.taub = 0.5 + nu / CS2
    c.sigma = (c.rhob0 * c.jet_velocity * c.jet_velocity * 2.0 * c.jet_radius) / c.We
    c.beta_recolor = float(beta)
    for k, v in overrides.items():
        setattr(c, k, v)
    return c


PERTURBATION_SEED = 20261203


def perturbed_jet_case(NX: int, NY: int, NZ: int, jet_radius: float, jet_velocity: float = 0.05,
                       perturbation: float = 0.01, seed: int = PERTURBATION_SEED, width: float = 2.0,
                       taur: float = 0.55, taub: float = 0.6, A: float = 1e-3, beta: float = 0.7,
                       **overrides) -> McgCase:
    """Round jet with the tanh-edged top-hat inlet u_y = U (1 - tanh((r - R)/width))/2 and a
    seeded uniform transverse perturbation of amplitude ``perturbation`` x U (MCG_INLET_TANH, an
    extension beyond the reference's sharp inlet).  The collision relaxes with taub everywhere
    (the reference's single rate); taur is kept for the interface-viscosity option."""
    c = jet_case(NX, NY, NZ, jet_radius, jet_velocity)
    c.inlet_profile = MCG_INLET_TANH
    c.interface_width = float(width)
    c.perturbation_amp = float(perturbation)
    c.perturbation_seed = int(seed)
    c.taur, c.taub = float(taur), float(taub)
    c.A_override = float(A)
    c.beta_recolor = float(beta)
    for k, v in overrides.items():
        setattr(c, k, v)
    return c


def periodic_bubble_case(N: int, radius: float, tau: float = 0.8, A: float = 1e-3, beta: float = 0.7,
                         width: float = 2.0, **overrides) -> McgCase:
    """Static-bubble (Laplace) case: N^3 nodes including ghost layers, periodic in
    all directions with a periodic colour gradient, tanh interface, direct A."""
    c = reference_jet()
    c.NX = c.NY = c.NZ = N
    c.y_mode = MCG_Y_PERIODIC
    c.wrap_gradient = 1
    c.taur = c.taub = float(tau)
    c.A_override = float(A)
    c.beta_recolor = float(beta)
    c.interface_width = float(width)
    c.bubble_radius = float(radius)
    c.bubble_x0 = c.bubble_y0 = c.bubble_z0 = (N - 1) / 2.0
    c.jet_x0 = c.jet_z0 = (N - 1) / 2.0
    c.NSTATS_START_STEP = 1 << 30
    for k, v in overrides.items():
        setattr(c, k, v)
    return c


def rising_bubble_case(NX: int = 192, NY: int = 384, NZ: int = 192, radius: float = 16.0, y0: float = 64.0,
                       gravity: float = 1e-5, taur: float = 0.55, taub: float = 1.0, A: float = 1e-3,
                       beta: float = 0.7, width: float = 2.0, **overrides) -> McgCase:
    """Buoyant rising bubble (an extension): no-slip walls in x and z, periodic y,
    a blue tanh bubble of radius R centred at ((NX-1)/2, y0, (NZ-1)/2) in red, the
    body force F_y = -gravity rho_r on the red (heavy) component, per-node
    viscosity between taur (red) and taub (blue), A given directly.  Start it
    with init_bubble_tanh."""
    c = reference_jet()
    c.NX, c.NY, c.NZ = NX, NY, NZ
    c.y_mode = MCG_Y_PERIODIC
    c.wall_xz = 1
    c.wrap_gradient = 1
    c.viscosity_mode = MCG_VISC_INTERFACE
    c.taur, c.taub = float(taur), float(taub)
    c.A_override = float(A)
    c.beta_recolor = float(beta)
    c.gravity = float(gravity)
    c.interface_width = float(width)
    c.bubble_radius = float(radius)
    c.bubble_x0, c.bubble_y0, c.bubble_z0 = (NX - 1) / 2.0, float(y0), (NZ - 1) / 2.0
    c.jet_x0, c.jet_z0 = (NX - 1) / 2.0, (NZ - 1) / 2.0
    c.NSTATS_START_STEP = 1 << 30
    for k, v in overrides.items():
        setattr(c, k, v)
    return c


def derived(c: McgCase) -> dict:
    """Derived constants of a case (constants.cuh:59-76, forces.cuh:87-105)."""
    nu = (c.taub - 0.5) * CS2
    cs4 = CS2 * CS2
    if c.A_override > 0:
        A = c.A_override
    elif c.viscosity_mode == MCG_VISC_INTERFACE:
        A = None  # per node
    else:
        A = c.sigma / (4.0 * cs4 * c.taub)
    return {"nu": nu, "omega": 1.0 / c.taub, "sigma_over_4cs4": c.sigma / (4.0 * cs4), "A": A,
            "NR_BINS": c.NX // 2, "Ncells": c.NX * c.NY * c.NZ}


def interior_nodes(c: McgCase) -> int:
    return (c.NX - 2) * (c.NY - 2) * (c.NZ - 2)

"""The interface-rich initialisation (a field of seeded tanh spheres).

CPU: the sphere list of the C ABI (mcg_sphere_list) against a restatement with
the SplitMix64 pinned in test_oracle_step.py.  GPU: the initial fields against
a numpy evaluation of the field formula from that list, and ten steps from
that state against the FP64 oracle.
"""
import numpy as np
import pytest

import multicomponent_color_gradient_amd as m
from multicomponent_color_gradient_amd.solver import sphere_list
from test_oracle_step import MASK64, splitmix64


def restated_list(seed, n, NX, NY, NZ, rmin, rmax):
    base = splitmix64(seed & 0xFFFFFFFF)
    out = []
    for k in range(n):
        key = splitmix64(base ^ k)
        u = [(splitmix64((key + j) & MASK64) >> 40) / 16777216.0 for j in range(1, 5)]
        out.append([1 + u[0] * (NX - 3), 1 + u[1] * (NY - 3), 1 + u[2] * (NZ - 3), rmin + u[3] * (rmax - rmin)])
    return np.array(out)


def test_sphere_list_matches_the_restatement():
    got = sphere_list(7, 256, 256, 512, 256, 6.0, 14.0)
    assert np.array_equal(got, restated_list(7, 256, 256, 512, 256, 6.0, 14.0))
    assert got[:, 0].min() >= 1 and got[:, 0].max() <= 254 and got[:, 3].min() >= 6 and got[:, 3].max() <= 14


def test_sphere_list_rejects_bad_arguments():
    with pytest.raises(m.McgError):
        sphere_list(7, 0, 32, 32, 32, 6.0, 14.0)
    with pytest.raises(m.McgError):
        sphere_list(7, 4, 32, 32, 32, 6.0, 5.0)


def field_fraction(case, sph):
    z, y, x = np.meshgrid(np.arange(case.NZ), np.arange(case.NY), np.arange(case.NX), indexing="ij")
    frac = np.zeros(x.shape)
    for cx, cy, cz, R in sph:
        r = np.sqrt((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2)
        frac = np.maximum(frac, 0.5 * (1 + np.tanh((R - r) / case.interface_width)))
    return frac


@pytest.mark.gpu
def test_gpu_spheres_init_and_steps_match_the_oracle():
    from oracle.oracle import Oracle

    case = m.perturbed_jet_case(34, 48, 34, jet_radius=4.0)
    n, seed, rmin, rmax = 12, 7, 3.0, 6.0
    print(f"spheres {case.NX}x{case.NY}x{case.NZ} n={n}", flush=True)
    s = m.Solver(case)
    s.init_spheres(n, seed, rmin, rmax)
    frac = field_fraction(case, sphere_list(seed, n, case.NX, case.NY, case.NZ, rmin, rmax))
    # interior: the populations at rest (w_i rho); ghosts: the ghost densities
    # (the exported moments of the interior are zero before the first step)
    fr, fb = s.populations()
    f = s.fields()
    inner3 = (slice(1, -1),) * 3
    rr, rb = fr.astype(np.float64).sum(0)[inner3], fb.astype(np.float64).sum(0)[inner3]
    assert np.max(np.abs(rb - frac[inner3] * case.rhob0)) < 2e-6
    assert np.max(np.abs(rr - (1 - frac[inner3]) * case.rhor0)) < 2e-6
    ghost = np.ones(frac.shape, bool)
    ghost[inner3] = False
    assert np.max(np.abs(f["rhob"][ghost] - frac[ghost] * case.rhob0)) < 2e-6
    interface = 4 * rr * rb / (rr + rb) ** 2 > 1e-4
    assert 0.1 < interface.mean() < 0.9
    # ten steps from this state, GPU against the FP64 oracle
    o = Oracle(case)
    o.init_jet()
    o.set_populations(fr.astype(np.float64), fb.astype(np.float64))
    o.set_densities(f["rhor"].astype(np.float64), f["rhob"].astype(np.float64))
    s.step(10)
    o.step(10)
    gr, gb = s.populations()
    orr, ob = o.populations()
    inner = (slice(None),) + (slice(1, -1),) * 3
    for g, r in ((gr, orr), (gb, ob)):
        assert np.max(np.abs(g[inner] - r[inner])) / np.max(np.abs(r[inner])) < 1e-5

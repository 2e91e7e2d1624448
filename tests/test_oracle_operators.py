"""Pins of the oracle's single-node operators against closed forms.

Every check below is a property the mathematics fixes (lattice isotropy,
Hermite orthogonality, the recursive third-order regularisation, the
moments of the perturbation and recolouring operators), chosen so that a
dropped term, a wrong sign or index, or a wrong coefficient fails.
"""
import itertools

import numpy as np
import pytest

from oracle import oracle as orc

CS2 = 1.0 / 3.0


@pytest.fixture(scope="module")
def L():
    return orc.lattice()


def cvec(L):
    return np.stack([L["cx"], L["cy"], L["cz"]], axis=1).astype(np.float64)


def test_velocity_set(L):
    c = cvec(L)
    assert len({tuple(v) for v in c}) == 27
    assert tuple(c[0]) == (0, 0, 0)
    # opposite directions adjacent: (2k-1, 2k) (stencil_ct.cuh:9-16)
    for k in range(1, 14):
        assert np.array_equal(c[2 * k], -c[2 * k - 1])
    s = (c ** 2).sum(1)
    assert np.all(s[1:7] == 1) and np.all(s[7:19] == 2) and np.all(s[19:] == 3)
    assert sorted(np.where(c[:, 1] == 1)[0].tolist()) == [3, 7, 11, 14, 17, 19, 21, 24, 25]
    assert sorted(np.where(c[:, 1] == -1)[0].tolist()) == [4, 8, 12, 13, 18, 20, 22, 23, 26]


def test_weights_isotropy(L):
    c, w = cvec(L), L["w"]
    assert abs(w.sum() - 1.0) < 1e-15
    assert np.allclose(w @ c, 0, atol=1e-16)
    assert np.allclose(np.einsum("i,ia,ib->ab", w, c, c), CS2 * np.eye(3), atol=1e-15)
    assert np.allclose(np.einsum("i,ia,ib,ic->abc", w, c, c, c), 0, atol=1e-16)
    d = np.eye(3)
    iso4 = CS2 ** 2 * (np.einsum("ab,cd->abcd", d, d) + np.einsum("ac,bd->abcd", d, d) + np.einsum("ad,bc->abcd", d, d))
    assert np.allclose(np.einsum("i,ia,ib,ic,id->abcd", w, c, c, c, c), iso4, atol=1e-15)


def test_perturbation_coefficients(L):
    c, B = cvec(L), L["B"]
    assert abs(B[0] + 10.0 / 27.0) < 1e-16
    assert np.allclose(B[1:], L["w"][1:])
    assert abs(B.sum() - 1.0 / 3.0) < 1e-15
    assert np.allclose(np.einsum("i,ia,ib->ab", B, c, c), CS2 * np.eye(3), atol=1e-15)


def test_inverse_norms(L):
    c = cvec(L)
    n = np.sqrt((c ** 2).sum(1))
    assert L["invc"][0] == 0
    assert np.allclose(L["invc"][1:], 1 / n[1:], rtol=1e-15)


def hermite2(c):
    return {"xx": c[:, 0] ** 2 - CS2, "xy": c[:, 0] * c[:, 1], "yy": c[:, 1] ** 2 - CS2, "yz": c[:, 1] * c[:, 2],
            "zz": c[:, 2] ** 2 - CS2, "xz": c[:, 0] * c[:, 2]}


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_feq_moments(L, seed):
    rng = np.random.default_rng(seed)
    c = cvec(L)
    rho = 0.7 + rng.random()
    u = 0.08 * rng.standard_normal(3)
    f = np.array([orc.feq(i, rho, *u) for i in range(27)])
    assert abs(f.sum() - rho) < 1e-14
    assert np.allclose(f @ c, rho * u, atol=1e-14)
    # second moment: rho (u u + cs^2 I) -- the Hermite expansion is exact to this order
    assert np.allclose(np.einsum("i,ia,ib->ab", f, c, c), rho * (np.outer(u, u) + CS2 * np.eye(3)), atol=1e-14)
    # third moment of D3Q27 at third order: rho (cs^2 (u d + perms) + u u u) for mixed indices
    m3 = np.einsum("i,ia,ib,ic->abc", f, c, c, c)
    for a, b, cc in itertools.product(range(3), repeat=3):
        if len({a, b, cc}) == 1:
            continue  # c^3 = c aliases the xxx moment on D3Q27
        exp = rho * (CS2 * (u[a] * (b == cc) + u[b] * (a == cc) + u[cc] * (a == b)) + u[a] * u[b] * u[cc])
        assert abs(m3[a, b, cc] - exp) < 1e-14, (a, b, cc)
    # at rest the equilibrium is w_i rho
    assert np.allclose([orc.feq(i, rho, 0, 0, 0) for i in range(27)], L["w"] * rho, rtol=1e-15)


def test_feq_axis_flow_closed_form():
    # u = (0, U, 0): f_3 = w1 (1 + 3U + 3U^2) (third-order term vanishes along an axis)
    U = 0.05
    assert abs(orc.feq(3, 1.0, 0.0, U, 0.0) - (2 / 27) * (1 + 3 * U + 3 * U * U)) < 1e-16
    assert abs(orc.feq(4, 1.0, 0.0, U, 0.0) - (2 / 27) * (1 - 3 * U + 3 * U * U)) < 1e-16


@pytest.mark.parametrize("seed", [3, 4, 5])
def test_fneqr_recursive_regularisation(L, seed):
    """sum_i fneqr_i H_i^(n): 0 for n=0,1; Pi^neq for n=2; the recursive
    a3_{abg} = u_a Pi_bg + u_b Pi_ag + u_g Pi_ab (Malaspinas 2015) for n=3."""
    rng = np.random.default_rng(seed)
    c = cvec(L)
    P = 1e-3 * rng.standard_normal(6)  # xx, xy, yy, yz, zz, xz
    u = 0.05 * rng.standard_normal(3)
    f = np.array([orc.fneqr(i, P, *u) for i in range(27)])
    assert abs(f.sum()) < 1e-17
    assert np.allclose(f @ c, 0, atol=1e-17)
    H = hermite2(c)
    names = ["xx", "xy", "yy", "yz", "zz", "xz"]
    for k, n in enumerate(names):
        assert abs(f @ H[n] - P[k]) < 1e-16, n
    Pm = np.array([[P[0], P[1], P[5]], [P[1], P[2], P[3]], [P[5], P[3], P[4]]])
    for a, b, g in [(0, 0, 1), (0, 0, 2), (0, 1, 1), (0, 2, 2), (1, 1, 2), (1, 2, 2), (0, 1, 2)]:
        h3 = c[:, a] * c[:, b] * c[:, g] - CS2 * (c[:, a] * (b == g) + c[:, b] * (a == g) + c[:, g] * (a == b))
        a3 = u[a] * Pm[b, g] + u[b] * Pm[a, g] + u[g] * Pm[a, b]
        assert abs(f @ h3 - a3) < 1e-17, (a, b, g)


@pytest.mark.parametrize("seed", [6, 7])
def test_omega2_moments(L, seed):
    """Perturbation operator: mass and momentum neutral; second moment
    (A/2)|F| (cs^4 (I + 2 n n) - cs^2 I)."""
    rng = np.random.default_rng(seed)
    c, w, B = cvec(L), L["w"], L["B"]
    F = 0.05 * rng.standard_normal(3)
    aF = float(np.linalg.norm(F))
    A = 1e-3
    om = np.array([orc.omega2(i, F, aF, A) for i in range(27)])
    assert abs(om.sum()) < 1e-18
    assert np.allclose(om @ c, 0, atol=1e-18)
    n = F / aF
    exp = 0.5 * A * aF * (CS2 ** 2 * (np.eye(3) + 2 * np.outer(n, n)) - CS2 * np.eye(3))
    assert np.allclose(np.einsum("i,ia,ib->ab", om, c, c), exp, atol=1e-18)
    assert abs(om[0] - (-(A / 2) * aF * B[0])) < 1e-20
    # |F| <= 1e-4: cos^2 rule switched off
    Fs = F / aF * 0.9e-4
    small = np.array([orc.omega2(i, Fs, 0.9e-4, A) for i in range(27)])
    assert np.allclose(small, -(A / 2) * 0.9e-4 * B, rtol=1e-14)


@pytest.mark.parametrize("seed", [8, 9])
def test_recolor_properties(L, seed):
    """Latva-Kokko recolouring: zero sum, antisymmetric in opposite pairs,
    colour flux beta rr rb/rho lambda n with lambda = sum_i w_i c_ix^2/|c_i|."""
    rng = np.random.default_rng(seed)
    c, w = cvec(L), L["w"]
    F = 0.1 * rng.standard_normal(3)
    aF = float(np.linalg.norm(F))
    rr, rb, beta = 0.3 + rng.random(), 0.2 + rng.random(), 0.7
    d = np.array([orc.recolor(i, rr, rb, F, aF, beta) for i in range(27)])
    assert d[0] == 0
    assert abs(d.sum()) < 1e-17
    for k in range(1, 14):
        assert abs(d[2 * k] + d[2 * k - 1]) < 1e-17
    lam = 6 * (2 / 27) * (2 / 6) + 12 * (1 / 54) * (8 / 12) / np.sqrt(2) + 8 * (1 / 216) / np.sqrt(3)
    lam_direct = float(np.sum(w * c[:, 0] ** 2 * L["invc"]))
    assert abs(lam - lam_direct) < 1e-15
    rho = rr + rb
    assert np.allclose(d @ c, beta * rr * rb / rho * lam * F / aF, atol=1e-16)
    # pure phase or weak gradient: no recolouring
    assert all(orc.recolor(i, 1.0, 0.0, F, aF, beta) == 0 for i in range(27))
    assert all(orc.recolor(i, rr, rb, F * 1e-3 / aF, 0.9e-4, beta) == 0 for i in range(27))

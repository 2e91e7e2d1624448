"""16-bit population storage (FP16 / BF16 offsets, FP32 arithmetic).

Encoding per (node, direction): t/w_i - 1 with t = f_r + f_b, and the
minority colour m/w_i (the colour of smaller magnitude, signed) whose mantissa
LSB is the colour flag.  Round-trip bounds (round to nearest; a = t/w_i - 1 is
stored with p = 10 (FP16) or 7 (BF16) mantissa bits, the minority with p - 1):
  FP16: |d t| <= 2^-11 |a| w_i,  |d m| <= 2^-10 |m|  (+ 2^-23 w_i below the normal range)
  BF16: |d t| <= 2^-8 |a| w_i,   |d m| <= 2^-7 |m|
"""
import numpy as np
import pytest

from conftest import INNER

pytestmark = pytest.mark.gpu

BOUNDS = {"fp16": (2.0 ** -11, 2.0 ** -10, 2.0 ** -23), "bf16": (2.0 ** -8, 2.0 ** -7, 0.0)}


def weights():
    from oracle.oracle import lattice

    return lattice()["w"][:, None, None, None]


@pytest.mark.parametrize("storage", ["fp16", "bf16"])
def test_round_trip_bounds_and_pure_phases(mcg, storage):
    """x slabs of the interior: 1-5 pure red, 6-8 pure blue, 9-10 pure red and
    11-12 pure blue with negative populations, 13-14 a negative red minority,
    15-16 a negative blue minority next to a positive majority, the rest mixed."""
    case = mcg.jet_case(24, 16, 18, jet_radius=3.0)
    s = mcg.Solver(case, storage=storage)
    s.init_jet()
    rng = np.random.default_rng(3)
    w = weights()
    shape = s.shape
    pops = (27,) + shape
    frac = rng.random(pops)
    frac[..., :6] = 0.0  # pure red
    frac[..., 6:9] = 1.0  # pure blue
    t = w * (1 + 0.2 * (rng.random(pops) - 0.5))
    fr = t * (1 - frac)
    fb = t * frac
    neg = -w * (0.01 + 0.2 * rng.random(pops))  # negative populations
    fr[..., 9:11], fb[..., 9:11] = neg[..., 9:11], 0.0
    fr[..., 11:13], fb[..., 11:13] = 0.0, neg[..., 11:13]
    fr[..., 13:15] = 0.3 * neg[..., 13:15]  # small negative red next to positive blue
    fb[..., 15:17] = 0.3 * neg[..., 15:17]  # small negative blue next to positive red
    fr, fb = fr.astype(np.float32), fb.astype(np.float32)
    s.set_populations(fr, fb)
    gr, gb = s.populations()
    sl = (slice(None),) + INNER
    bt, bm, floor = BOUNDS[storage]
    tt = (fr + fb).astype(np.float64)
    ww = w * np.ones(shape)
    a = tt / ww - 1.0
    # the FP32 sum fr + fb and the decode's (a + 1) w add a few 1e-7 of |t|
    dt = np.abs((gr + gb).astype(np.float64) - tt)[sl]
    assert np.all(dt <= (bt * np.abs(a) * ww + 3e-7 * np.abs(tt) + floor * ww)[sl])
    red_minor = np.abs(fr) < np.abs(fb)
    mm = np.where(red_minor, fr, fb).astype(np.float64)
    gm = np.where(red_minor, gr, gb).astype(np.float64)
    dm = np.abs(gm - mm)[sl]
    assert np.all(dm <= (bm * np.abs(mm) + floor * ww)[sl] + 1e-12)
    # signs survive, including negative minorities and negative pure phases
    nz = (mm != 0)[sl]
    assert np.all(np.sign(gm[sl][nz]) == np.sign(mm[sl][nz]))
    assert np.any(mm[sl] < 0) and np.any((tt[sl] < 0) & ~nz)
    # pure phases stay exactly pure, whatever the sign of the majority
    inner_x = (slice(None), slice(1, -1), slice(1, -1))
    for x0, x1, zero in ((1, 6, gb), (6, 9, gr), (9, 11, gb), (11, 13, gr)):
        assert np.all(zero[inner_x + (slice(x0, x1),)] == 0), (x0, x1)
    assert np.all(gr[inner_x + (slice(9, 11),)] < 0) and np.all(gb[inner_x + (slice(11, 13),)] < 0)


@pytest.mark.parametrize("storage,tol_u", [("fp16", 2e-4), ("bf16", 2e-3)])
def test_jet_16bit_close_to_fp64_oracle(mcg, storage, tol_u):
    """Velocity error after 100 jet steps, against the FP64 oracle: the FP32 run
    is within 1e-5; 16-bit storage adds its rounding (|d u| <= tol_u)."""
    from oracle.oracle import Oracle

    case = mcg.jet_case(30, 40, 30, jet_radius=5.0, Re=500.0)
    s = mcg.Solver(case, storage=storage)
    s.init_jet()
    s.step(100)
    o = Oracle(case)
    o.init_jet()
    o.step(100)
    g, r = s.fields(), o.fields()
    for k in ("ux", "uy", "uz"):
        e = float(np.max(np.abs(g[k][INNER] - r[k][INNER])))
        assert e <= tol_u, f"{storage} {k}: {e:.3e}"
    e = float(np.max(np.abs(g["rhob"][INNER] - r["rhob"][INNER])))
    assert e <= 20 * tol_u


@pytest.mark.parametrize("storage", ["fp16", "bf16"])
def test_16bit_fused_equals_twopass(mcg, storage):
    case = mcg.jet_case(34, 30, 30, jet_radius=5.0, Re=500.0)
    a = mcg.Solver(case, storage=storage, kernel="fused")
    b = mcg.Solver(case, storage=storage, kernel="twopass")
    a.init_jet()
    b.init_jet()
    a.step(20)
    b.step(20)
    pa, pb = a.populations(), b.populations()
    assert np.array_equal(pa[0], pb[0]) and np.array_equal(pa[1], pb[1])


def mean_profile_errors(mcg, storage, steps=120):
    """Radial mean u_y profile <u_y>(y, r) = sum_uy / count of a perturbed jet
    against the FP64 oracle: max |d <u_y>| / max |<u_y>| over the sampled bins,
    and the relative error of the total mass (rho_r + rho_b)."""
    from oracle.oracle import Oracle

    case = mcg.perturbed_jet_case(34, 48, 34, jet_radius=6.0)
    case.NSTATS_START_STEP = 20
    case.NSTATS_SAMPLE = 2
    s = mcg.Solver(case, storage=storage)
    s.init_jet()
    s.step(steps)
    o = Oracle(case)
    o.init_jet()
    o.step(steps)
    g, r = s.radial(), o.radial()
    assert np.array_equal(g["count"], r["count"])
    mask = r["count"] > 0
    mg = g["sum_uy"][mask] / g["count"][mask]
    mr = r["sum_uy"][mask] / r["count"][mask]
    prof = float(np.max(np.abs(mg - mr)) / np.max(np.abs(mr)))
    gs, os_ = s.global_sums(), o.global_sums()
    mass = float(abs((gs[0] + gs[1]) - (os_[0] + os_[1])) / (os_[0] + os_[1]))
    return prof, mass


@pytest.mark.parametrize("storage,bound", [("fp32", 1e-4), ("fp16", 1e-3), ("bf16", 2e-3)])
def test_jet_mean_profile_error(mcg, storage, bound):
    """Mean u_y profiles of the jet: FP16-offset storage within 1e-3 of the FP64
    oracle (BF16 within 2e-3, FP32 within 1e-4; measured with the stochastic
    rounding of the stores: 5e-6, 5e-5, 6e-4); total mass within 1e-5."""
    prof, mass = mean_profile_errors(mcg, storage)
    assert prof <= bound, f"{storage}: mean-profile error {prof:.3e}"
    assert mass <= 1e-5, f"{storage}: mass error {mass:.3e}"


@pytest.mark.parametrize("storage", ["fp16", "bf16"])
def test_stochastic_rounding_is_unbiased(mcg, storage):
    """One step of a uniform red flow in a periodic box: every node collides
    the same populations to the same x_i and stores x_i on the 16-bit grid,
    rounded up or down with its own random bits.  Per direction the stored
    values are the two grid neighbours of x_i (FP64 oracle), and their
    average over 32^3 nodes equals x_i within 5 standard errors of the
    binomial draw, sigma = g sqrt(p (1 - p) / n) (grid step g, fraction p of
    the way up).  Round to nearest would sit g min(p, 1 - p) away, which the
    test tells apart for most directions."""
    from oracle.oracle import Oracle, feq

    N = 36
    case = mcg.periodic_bubble_case(N, radius=8.0, tau=0.8)
    s = mcg.Solver(case, storage=storage)
    s.init_bubble_tanh()
    u = (0.021, 0.047, -0.033)
    fr = np.empty((27,) + s.shape, np.float32)
    for i in range(27):
        fr[i] = feq(i, 1.0, *u)
    s.set_populations(fr, np.zeros_like(fr))
    r0, b0 = s.populations()
    o = Oracle(case)
    o.init_bubble_tanh()
    o.set_populations(r0, b0)
    o.step(1)
    s.step(1)
    xr, xb = o.populations()
    gr, gb = s.populations()
    core = (slice(None),) + (slice(2, -2),) * 3
    t = (gr + gb)[core].reshape(27, -1).astype(np.float64)
    x = (xr + xb)[core].reshape(27, -1)
    n = t.shape[1]
    assert np.all(gb[core] == 0)
    assert np.ptp(x, axis=1).max() <= 1e-15
    zs, powered = [], 0
    for i in range(27):
        xi = float(x[i, 0])
        vals = np.unique(t[i])
        assert len(vals) <= 2, (i, vals)
        lo, hi = float(vals[0]), float(vals[-1])
        # the GPU collides in FP32: its x_i is within 2.1e-7 relative of the
        # FP64 value (the oracle's FP32 mode on this state); allow 5e-7
        tol = 5e-7 * abs(xi)
        assert lo - tol <= xi <= hi + tol, (i, xi, vals)
        if len(vals) == 1:
            assert abs(lo - xi) <= tol, (i, xi, lo)
            continue
        g = hi - lo
        p = min(max((xi - lo) / g, 0.0), 1.0)
        sigma = g * np.sqrt(max(p * (1 - p), 1.0 / n) / n)
        z = abs(t[i].mean() - xi) / (sigma + tol)
        zs.append(z)
        assert z <= 5.0, f"{storage} direction {i}: mean off by {z:.1f} sigma (p={p:.3f})"
        if g * min(p, 1 - p) > 5 * (sigma + tol):
            powered += 1
    print(f"{storage}: max |mean - x| = {max(zs):.2f} sigma over {len(zs)} directions; "
          f"round to nearest would be >5 sigma off in {powered}", flush=True)
    assert powered >= 10


@pytest.mark.parametrize("storage", ["fp16", "bf16"])
def test_16bit_checkpoint_restart_is_bitwise(mcg, storage):
    """The stochastic rounding draws its bits from (node, step): a restart
    from a checkpoint (which holds the step count) continues the same stream,
    so the restarted run equals the uninterrupted one bit for bit."""
    case = mcg.perturbed_jet_case(34, 30, 30, jet_radius=5.0)
    s = mcg.Solver(case, storage=storage)
    s.init_jet()
    s.step(15)
    buf = s.checkpoint()
    s.step(5)
    a = s.populations()
    s.restore(buf)
    s.step(5)
    b = s.populations()
    s.close()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    t = mcg.Solver(case, storage=storage)  # a fresh solver restored from the same bytes
    t.restore(buf)
    t.step(5)
    c = t.populations()
    t.close()
    assert np.array_equal(a[0], c[0]) and np.array_equal(a[1], c[1])


def test_census_counts_negative_populations(mcg):
    """census.negative_counts against numpy's own count on populations set with
    a known pattern of negatives (FP32 storage keeps every value)."""
    from multicomponent_color_gradient_amd.census import negative_counts

    case = mcg.jet_case(20, 16, 18, jet_radius=3.0)
    s = mcg.Solver(case)
    s.init_jet()
    w = weights()
    shape = (27,) + s.shape
    rng = np.random.default_rng(8)
    fr = (w * (0.5 + rng.random(shape))).astype(np.float32)
    fb = (w * rng.random(shape)).astype(np.float32)
    neg_r = rng.random(shape) < 0.01
    neg_b = rng.random(shape) < 0.02
    fr[neg_r] *= -1.0
    fb[neg_b] *= -3.0
    s.set_populations(fr, fb)
    c = negative_counts(s)
    s.close()
    sl = (slice(None),) + INNER
    r, b = fr[sl], fb[sl]
    minority = np.where(np.abs(r) < np.abs(b), r, b)
    assert c["nodes"] == int(np.prod(r.shape[1:]))
    assert c["fr_negative"] == np.count_nonzero((r < 0).reshape(27, -1), axis=1).tolist()
    assert c["fb_negative"] == np.count_nonzero((b < 0).reshape(27, -1), axis=1).tolist()
    assert c["g_negative"] == np.count_nonzero(((r + b) < 0).reshape(27, -1), axis=1).tolist()
    assert c["minority_negative_total"] == int(np.count_nonzero(minority < 0)) > 0

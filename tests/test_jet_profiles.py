"""Host logic of the storage comparison (jet_profiles.compare) on hand-built
radial sums: no GPU."""
import numpy as np
import pytest

from multicomponent_color_gradient_amd.jet_profiles import compare
from multicomponent_color_gradient_amd.post import averaged_radial_fields


def run(sum_uy, count, sum_phi, mass):
    z = np.zeros_like(sum_uy)
    radial = {"sum_uy": sum_uy, "sum_uy2": z, "sum_ur": z, "sum_ur2": z, "sum_uruy": z,
              "sum_phi": sum_phi, "sum_phi2": z, "count": count}
    return {"radial": radial, "mean": averaged_radial_fields(radial), "sums": np.array([mass, 0.0, 0.0, 0.0, 0.0])}


def test_compare_hand_case():
    # 3 y rows x 2 r-bins, 4 samples per bin except an empty one; means of b:
    # uy [[0.05, 0.02], [0.04, -], [0.03, 0.01]] -> scale 0.05
    count = np.array([[4, 4], [4, 0], [4, 4]], np.uint64)
    ub = np.array([[0.20, 0.08], [0.16, 0.0], [0.12, 0.04]])
    ua = ub.copy()
    ua[1, 0] += 4 * 1e-4     # axis bin of row 1: mean differs by 1e-4
    ua[2, 1] -= 4 * 2.5e-5   # off-axis bin: 2.5e-5
    ua[1, 1] = 123.0         # empty bin: ignored
    pb = np.array([[4.0, -4.0], [4.0, 0.0], [-4.0, -4.0]])
    pa = pb.copy()
    pa[0, 1] += 4 * 3e-3
    a = run(ua, count, pa, 1000.5)
    b = run(ub, count, pb, 1000.0)
    e = compare(a, b)
    assert e["uy"] == pytest.approx(1e-4 / 0.05, rel=1e-9)
    assert e["centreline"] == pytest.approx(1e-4 / 0.05, rel=1e-9)
    assert e["phi"] == pytest.approx(3e-3, rel=1e-9)
    assert e["mass"] == pytest.approx(0.5 / 1000.0, rel=1e-12)
    # identical runs compare to zero
    assert compare(b, b) == {"uy": 0.0, "centreline": 0.0, "phi": 0.0, "mass": 0.0}


def test_compare_off_axis_difference_leaves_the_centreline():
    count = np.ones((2, 3), np.uint64)
    ub = np.array([[0.05, 0.03, 0.01], [0.04, 0.02, 0.0]])
    ua = ub.copy()
    ua[0, 2] += 5e-4
    e = compare(run(ua, count, np.zeros_like(ub), 1.0), run(ub, count, np.zeros_like(ub), 1.0))
    assert e["uy"] == pytest.approx(1e-2, rel=1e-9)
    assert e["centreline"] == 0.0


def test_compare_rejects_different_sampling():
    c1 = np.ones((2, 2), np.uint64)
    c2 = c1.copy()
    c2[0, 0] = 2
    u = np.ones((2, 2))
    with pytest.raises(ValueError):
        compare(run(u, c1, u, 1.0), run(u, c2, u, 1.0))
    with pytest.raises(ValueError):
        compare(run(u, 0 * c1, u, 1.0), run(u, 0 * c1, u, 1.0))

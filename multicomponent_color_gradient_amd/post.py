"""Mean-profile post-processing of the radial statistics files.

Restates the reference's offline consumers of ``mean_profiles/``:
  read_metadata, read_radial_moments, averaged_radial_fields,
  fit_virtual_origin         (post/mean_profile_io.py:61-239)
  first_half_radius, radial_bin_centers, centerline_summary
                             (post/plot_profiles.py:53-92,120-135,276-308)
Plotting is not reproduced.  Numerics use numpy; results depend on numpy's
least-squares fit (np.polyfit), so they are pinned by hand-computed cases,
not against the reference's own outputs (whose inputs are not shipped).
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

INT_KEYS = ("NX", "NY", "NZ", "NR_BINS", "NradialProfileCells", "NOUTPUT", "NSTATS_SAMPLE", "NSTATS_START_STEP",
            "NPROFILE_OUTPUT", "start_step", "uy_avg_start_step", "uy_avg_samples", "radial_sample_count",
            "uy_avg_first_sample_step", "uy_avg_last_sample_step")
FLOAT_KEYS = ("D", "U_jet", "jet_radius", "jet_velocity")

# Hussein, Capp & George (1994) self-similar round-jet profile, digitised
# (post/plot_profiles.py:31-50): r/(y - y0) and U/U_c.
HUSSEIN_X = np.array([
    0, 0.010599057494345044, 0.019815647572123837, 0.030875562697158625, 0.04009215277493742,
    0.04838709790833872, 0.056221185411050346, 0.06405527291376191, 0.07281104051935197,
    0.08479260058876421, 0.09400919066654304, 0.10737325331102249, 0.11658984338880131,
    0.12949308356109202, 0.14147464363050433, 0.15345620369991658, 0.16866359139165193,
    0.18248847650832015, 0.19723500656936588, 0.21290321673329005, 0.22672810184995826, 0.2447004595333271])
HUSSEIN_U = np.array([
    0.9986970883174464, 0.988273596045439, 0.9622149647712097, 0.9205211944947582, 0.8657980588782979,
    0.8136807963298389, 0.7563517876453765, 0.7042345250969178, 0.6469055164124551, 0.5635178764537632,
    0.514006536676201, 0.4332247200826162, 0.38371338030505414, 0.32899019498569865, 0.2690553629361291,
    0.2247556698887811, 0.16742676061010808, 0.11791512261517773, 0.08143319788051967, 0.05276854442960405,
    0.03452768146806465, 0.024104189196056974])


def read_metadata(mean_dir) -> dict:
    """key value pairs of radial_profile_metadata.txt (post/mean_profile_io.py:61-100)."""
    p = Path(mean_dir) / "radial_profile_metadata.txt"
    md: dict = {}
    if not p.exists():
        return md
    for line in p.read_text().splitlines():
        parts = line.split(maxsplit=1)
        if len(parts) == 2:
            md[parts[0]] = parts[1]
    for k in INT_KEYS:
        if k in md:
            md[k] = int(md[k])
    for k in FLOAT_KEYS:
        if k in md:
            md[k] = float(md[k])
    return md


def read_radial_moments(mean_dir) -> dict:
    """The six radial files reshaped to (NY, NR_BINS) (post/mean_profile_io.py:139-179),
    plus the psi moments when present."""
    mean_dir = Path(mean_dir)
    md = read_metadata(mean_dir)
    ny, nr = int(md.get("NY", 0)), int(md.get("NR_BINS", 0))
    if ny <= 0 or nr <= 0:
        raise ValueError("NY and NR_BINS missing from radial_profile_metadata.txt")

    def rd(name, dtype):
        a = np.fromfile(mean_dir / name, dtype=dtype)
        if a.size != ny * nr:
            raise ValueError(f"{name} has {a.size} values, expected {ny * nr}")
        return a.reshape(ny, nr)

    out = {"sum_uy": rd("radial_sum_uy.bin", np.float64), "sum_uy2": rd("radial_sum_uy2.bin", np.float64),
           "sum_ur": rd("radial_sum_ur.bin", np.float64), "sum_ur2": rd("radial_sum_ur2.bin", np.float64),
           "sum_uruy": rd("radial_sum_uruy.bin", np.float64), "count": rd("radial_count.bin", np.uint64),
           "metadata": md, "mean_dir": mean_dir}
    for extra in ("sum_phi", "sum_phi2"):
        if (mean_dir / f"radial_{extra}.bin").exists():
            out[extra] = rd(f"radial_{extra}.bin", np.float64)
    return out


def averaged_radial_fields(m: dict) -> dict:
    """Means, rms and Reynolds stress per (y, r-bin) (post/mean_profile_io.py:182-210)."""
    count = m["count"].astype(np.float64)
    valid = count > 0

    def avg(name):
        o = np.full(count.shape, np.nan)
        np.divide(m[name], count, out=o, where=valid)
        return o

    uy, uy2, ur, ur2, uruy = (avg(k) for k in ("sum_uy", "sum_uy2", "sum_ur", "sum_ur2", "sum_uruy"))
    out = {"uy": uy, "uy2": uy2, "ur": ur, "ur2": ur2, "uruy": uruy,
           "uy_rms": np.sqrt(np.maximum(uy2 - uy * uy, 0.0)), "ur_rms": np.sqrt(np.maximum(ur2 - ur * ur, 0.0)),
           "uruy_reynolds": uruy - ur * uy, "count": m["count"]}
    if "sum_phi" in m:
        out["phi"] = avg("sum_phi")
        out["phi_rms"] = np.sqrt(np.maximum(avg("sum_phi2") - out["phi"] ** 2, 0.0))
    return out


def fit_virtual_origin(y, uc, diameter, u_jet, yi_over_d=10.0, yf_over_d=15.0):
    """Least-squares line U_jet/U_c = (y/D)/B - y0/(B D) over y/D in [yi, yf]:
    returns (B, y0, x_fit, y_fit, slope, intercept) (post/mean_profile_io.py:221-239)."""
    yd = np.asarray(y, np.float64) / diameter
    uc = np.asarray(uc, np.float64)
    mask = np.isfinite(uc) & (uc > 0.0) & (yd >= yi_over_d) & (yd <= yf_over_d)
    if np.count_nonzero(mask) < 2:
        raise ValueError("not enough valid centerline samples in the fit range")
    x_fit = yd[mask]
    y_fit = u_jet / uc[mask]
    slope, intercept = np.polyfit(x_fit, y_fit, 1)
    B = 1.0 / slope
    y0 = -intercept * B * diameter
    return B, y0, x_fit, y_fit, slope, intercept


def first_half_radius(profile, radius, center_value):
    """First radius where profile/center drops to 1/2, linearly interpolated
    (post/plot_profiles.py:53-70)."""
    if not np.isfinite(center_value) or center_value <= 0.0:
        return np.nan
    ratio = np.asarray(profile, np.float64) / center_value
    below = np.where(np.isfinite(ratio) & (ratio <= 0.5))[0]
    below = below[below > 0]
    if below.size == 0:
        return np.nan
    i = int(below[0])
    r1, r2 = radius[i - 1], radius[i]
    u1, u2 = ratio[i - 1], ratio[i]
    if not np.isfinite(u1) or not np.isfinite(u2) or np.isclose(u1, u2):
        return np.nan
    return r1 + (0.5 - u1) * (r2 - r1) / (u2 - u1)


def radial_bin_centers(nx, nz, nr, x0, z0):
    """Mean radius of the interior columns in each bin; bin 0 is the centreline
    (post/plot_profiles.py:73-92)."""
    xs = np.arange(1, nx - 1, dtype=np.float64) - x0
    zs = np.arange(1, nz - 1, dtype=np.float64) - z0
    r = np.sqrt(xs[:, None] ** 2 + zs[None, :] ** 2).ravel()
    rb = r.astype(np.int64)
    ok = rb < nr
    s = np.bincount(rb[ok], weights=r[ok], minlength=nr)[:nr]
    c = np.bincount(rb[ok], minlength=nr)[:nr]
    radius = np.arange(nr, dtype=np.float64)
    np.divide(s, c, out=radius, where=c > 0)
    radius[0] = 0.0
    return radius


def centerline_summary(mean_dir, diameter=None, u_jet=None, slices=(10, 11, 12, 13, 14, 15), fit_start=10.0,
                       fit_end=15.0) -> dict:
    """Virtual-origin fit and half-radius spreading rates of a run
    (the numbers of profile_summary.txt, post/plot_profiles.py:112-135,276-308)."""
    m = read_radial_moments(mean_dir)
    md = m["metadata"]
    diameter = float(diameter if diameter is not None else md.get("D", 16.0))
    u_jet = float(u_jet if u_jet is not None else md.get("U_jet", 0.05))
    nx, nz = int(md["NX"]), int(md["NZ"])
    x0 = float(md.get("jet_x0", (nx - 1) / 2.0))
    z0 = float(md.get("jet_z0", (nz - 1) / 2.0))
    f = averaged_radial_fields(m)
    uy, count = f["uy"], f["count"]
    ny, nr = uy.shape
    radius = radial_bin_centers(nx, nz, nr, x0, z0)
    uc = uy[:, 0]
    B, y0, *_ = fit_virtual_origin(np.arange(ny, dtype=np.float64), uc, diameter, u_jet, fit_start, fit_end)
    s_h = first_half_radius(HUSSEIN_U, HUSSEIN_X, 1.0)
    rows = []
    for sd in slices:
        yi = int(round(sd * diameter))
        if yi <= 0 or yi >= ny or not np.isfinite(uc[yi]) or uc[yi] <= 0.0:
            continue
        rh = first_half_radius(uy[yi, :], radius, uc[yi])
        S = rh / (yi - y0) if np.isfinite(rh) and yi > y0 else np.nan
        rows.append({"slice_D": sd, "y_index": yi, "r_half": rh, "S": S, "S_over_hussein": S / s_h})
    return {"B": B, "y0": y0, "hussein_S": s_h, "slices": rows, "radial_sample_count": md.get("radial_sample_count", 0)}

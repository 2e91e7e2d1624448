"""Mean-profile post-processing of the radial statistics files.

Restates the reference's offline consumers of ``mean_profiles/``:
  read_metadata, read_radial_moments, averaged_radial_fields,
  fit_virtual_origin         (post/mean_profile_io.py:61-239)
  first_half_radius, radial_bin_centers, centerline_summary
                             (post/plot_profiles.py:53-92,120-135,276-308)
Plotting is not reproduced.  The line fit is solved in closed form here (no
library fit routine), and every function is pinned by hand-computed cases and
the one reference-held number (the Hussein half-radius slope of
resultsJET/profile_summary.txt:7), not against the reference's own outputs,
whose inputs are not shipped.
"""
from __future__ import annotations

import ast
import re
from pathlib import Path
from typing import NamedTuple

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
    return _read_key_values(Path(mean_dir) / "radial_profile_metadata.txt")


def _read_key_values(path) -> dict:
    """'key value' lines of a metadata file, with the known integer and float
    keys converted; {} when the file does not exist."""
    p = Path(path)
    if not p.exists():
        return {}
    md: dict = {}
    for line in p.read_text().splitlines():
        key, _, value = line.strip().partition(" ")
        if key and value.strip():
            md[key] = value.strip()
    for keys, conv in ((INT_KEYS, int), (FLOAT_KEYS, float)):
        for k in keys:
            if k in md:
                md[k] = conv(md[k])
    return md


# ---- case constants when a run left no metadata --------------------------
# The reference's post scripts fall back to its compile-time constants.cuh
# (post/mean_profile_io.py:21-58, post/case_config.py:38-76).  Here the
# fallback chain is: radial_profile_metadata.txt, then case_info.txt of the
# case folder (written by the driver), then a constants.cuh-style file whose
# declarations are evaluated in order (so derived values such as
# jet_x0 = (NX - 1)/2 resolve too).

_DECL = re.compile(r"^\s*(?:inline\s+)?(?:static\s+)?(?:constexpr|const)\s+[\w:<>]+\s+(\w+)\s*=\s*(.+?);", re.M)
_CAST = re.compile(r"static_cast\s*<\s*([\w:]+)\s*>")
_FLOAT_TYPES = ("float", "double", "real_t", "pop_t")


def _eval_expr(expr: str, env: dict):
    """Value of an arithmetic expression over numbers and earlier names
    (+ - * /, parentheses and static_cast to a floating or integer type, with
    C++'s integer division); None when it uses anything else."""
    src = _CAST.sub(lambda mt: "_float_" if mt.group(1) in _FLOAT_TYPES else "_int_", expr)
    try:
        tree = ast.parse(src.replace("std::", ""), mode="eval")
    except SyntaxError:
        return None

    def ev(n):
        if isinstance(n, ast.Expression):
            return ev(n.body)
        if isinstance(n, ast.Constant) and isinstance(n.value, (int, float)) and not isinstance(n.value, bool):
            return n.value
        if isinstance(n, ast.Name):
            if n.id in ("true", "false"):
                return int(n.id == "true")
            return env[n.id]
        if isinstance(n, ast.Call) and isinstance(n.func, ast.Name) and n.func.id in ("_float_", "_int_") \
                and len(n.args) == 1 and not n.keywords:
            v = ev(n.args[0])
            return float(v) if n.func.id == "_float_" else int(v)
        if isinstance(n, ast.UnaryOp) and isinstance(n.op, (ast.USub, ast.UAdd)):
            v = ev(n.operand)
            return -v if isinstance(n.op, ast.USub) else v
        if isinstance(n, ast.BinOp) and isinstance(n.op, (ast.Add, ast.Sub, ast.Mult, ast.Div)):
            a, b = ev(n.left), ev(n.right)
            if isinstance(n.op, ast.Add):
                return a + b
            if isinstance(n.op, ast.Sub):
                return a - b
            if isinstance(n.op, ast.Mult):
                return a * b
            if isinstance(a, int) and isinstance(b, int):
                return int(a / b)  # C++ integer division truncates toward zero
            return a / b
        raise KeyError(type(n).__name__)

    try:
        return ev(tree)
    except (KeyError, ZeroDivisionError, TypeError, ValueError):
        return None


def read_case_constants(path) -> dict:
    """Evaluate the scalar declarations of a constants.cuh-style header
    (constants.cuh:8-76): NX, NY, NZ, NOUTPUT, NSTATS_START_STEP, jet_radius,
    jet_x0, ... in declaration order.  Adds D = 2 jet_radius and U_jet =
    jet_velocity, the metadata names of the same quantities, and NR_BINS =
    NX/2 when the file does not declare it.  {} when the file is missing."""
    p = Path(path)
    if not p.exists():
        return {}
    text = re.sub(r"//[^\n]*", "", p.read_text())
    env: dict = {}
    for name, expr in _DECL.findall(text):
        v = _eval_expr(expr, env)
        if v is not None:
            env[name] = v
    if "jet_radius" in env:
        env.setdefault("D", 2.0 * float(env["jet_radius"]))
    if "jet_velocity" in env:
        env.setdefault("U_jet", float(env["jet_velocity"]))
    if "NX" in env:
        env.setdefault("NR_BINS", int(env["NX"]) // 2)
    return env


def case_metadata(mean_dir, constants_file=None) -> dict:
    """The run's metadata with the fallbacks filled in: keys of
    radial_profile_metadata.txt win, then case_info.txt in the case folder
    above mean_dir, then the constants file (if given)."""
    mean_dir = Path(mean_dir)
    md = read_metadata(mean_dir)
    layers = [_read_key_values(mean_dir.parent / "case_info.txt")]
    if constants_file is not None:
        layers.append(read_case_constants(constants_file))
    for layer in layers:
        if "jet_diameter" in layer:
            layer.setdefault("D", float(layer["jet_diameter"]))
        if "jet_velocity" in layer:
            layer.setdefault("U_jet", float(layer["jet_velocity"]))
        for k, v in layer.items():
            md.setdefault(k, v)
    for k in INT_KEYS:
        if k in md:
            md[k] = int(md[k])
    for k in FLOAT_KEYS:
        if k in md:
            md[k] = float(md[k])
    return md


def read_radial_moments(mean_dir, constants_file=None) -> dict:
    """The six radial files reshaped to (NY, NR_BINS) (post/mean_profile_io.py:139-179),
    plus the psi moments when present.  NY and NR_BINS come from the run's
    metadata, else from case_info.txt or the constants file (case_metadata)."""
    mean_dir = Path(mean_dir)
    md = case_metadata(mean_dir, constants_file)
    ny = int(md.get("NY", 0))
    nr = int(md.get("NR_BINS", int(md.get("NX", 0)) // 2))
    if ny <= 0 or nr <= 0:
        raise ValueError("NY and NR_BINS missing from the run metadata, case_info.txt and the constants file")

    def rd(name, dtype):
        path = mean_dir / name
        if not path.exists():
            raise FileNotFoundError(f"{path} missing: the run wrote no radial statistics yet")
        a = np.fromfile(path, dtype=dtype)
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


def _per_sample(total, count):
    """total / count per bin; NaN in bins without samples."""
    c = np.asarray(count, np.float64)
    has = c > 0
    q = np.asarray(total, np.float64) / np.where(has, c, 1.0)
    return np.where(has, q, np.nan)


def _spread(mean_sq, mean):
    """sqrt(<v^2> - <v>^2), with the rounding-negative variance taken as 0."""
    return np.sqrt(np.clip(mean_sq - np.square(mean), 0.0, None))


def averaged_radial_fields(m: dict) -> dict:
    """Per (y, r-bin) sample means of the accumulated moments and what they
    give: rms of u_y and u_r, the Reynolds stress <u_r u_y> - <u_r><u_y>, and
    with the psi moments the mean and rms of psi.  Bins without samples hold
    NaN.  The quantities of post/mean_profile_io.py:182-210."""
    count = m["count"]
    means = {k[4:]: _per_sample(v, count) for k, v in m.items() if k.startswith("sum_")}
    out = {k: means[k] for k in ("uy", "uy2", "ur", "ur2", "uruy")}
    out["uy_rms"] = _spread(means["uy2"], means["uy"])
    out["ur_rms"] = _spread(means["ur2"], means["ur"])
    out["uruy_reynolds"] = means["uruy"] - means["ur"] * means["uy"]
    out["count"] = count
    if "phi" in means:
        out["phi"] = means["phi"]
        out["phi_rms"] = _spread(means["phi2"], means["phi"])
    return out


class OriginFit(NamedTuple):
    """Straight line U_jet/U_c = slope (y/D) + intercept through the
    centre-line decay, with B = 1/slope and the virtual origin y0 where the
    line crosses zero; x and inv_uc are the points fitted."""
    B: float
    y0: float
    x: np.ndarray
    inv_uc: np.ndarray
    slope: float
    intercept: float


def fit_virtual_origin(y, uc, diameter, u_jet, yi_over_d=10.0, yf_over_d=15.0) -> OriginFit:
    """Least-squares line through U_jet/U_c against y/D over the samples with
    finite, positive U_c and yi <= y/D <= yf (the fit of
    post/mean_profile_io.py:221-239).  The normal equations are solved about
    the centroid of the points: slope = S_xy/S_xx, intercept = mean(v) -
    slope mean(x); then B = 1/slope and y0 = -intercept B D."""
    x = np.asarray(y, np.float64) / diameter
    u = np.asarray(uc, np.float64)
    with np.errstate(invalid="ignore"):
        keep = np.isfinite(u) & (u > 0.0) & (x >= yi_over_d) & (x <= yf_over_d)
    if np.count_nonzero(keep) < 2:
        raise ValueError("not enough valid centerline samples in the fit range")
    xs = x[keep]
    vs = u_jet / u[keep]
    dx = xs - xs.mean()
    sxx = float(np.dot(dx, dx))
    if sxx == 0.0:
        raise ValueError("the fit range holds a single y position")
    slope = float(np.dot(dx, vs - vs.mean())) / sxx
    intercept = float(vs.mean()) - slope * float(xs.mean())
    B = 1.0 / slope
    return OriginFit(B, -intercept * B * diameter, xs, vs, slope, intercept)


def first_half_radius(profile, radius, center_value):
    """Radius where profile/center_value first falls to 1/2 (the half-radius
    of post/plot_profiles.py:53-70): the first sample i >= 1 at or below 1/2,
    linearly interpolated against the sample before it, which must lie above
    1/2 (a genuine crossing).  NaN when the centre value is not positive, no
    sample falls to 1/2, or the sample before the first one at or below 1/2
    is not a finite value above it."""
    if not (np.isfinite(center_value) and center_value > 0.0):
        return np.nan
    ratio = np.asarray(profile, np.float64) / center_value
    with np.errstate(invalid="ignore"):
        at_or_below = np.isfinite(ratio[1:]) & (ratio[1:] <= 0.5)
    hit = np.flatnonzero(at_or_below)
    if hit.size == 0:
        return np.nan
    k = int(hit[0]) + 1
    above, below = ratio[k - 1], ratio[k]
    if not (np.isfinite(above) and above > 0.5):
        return np.nan
    r = np.asarray(radius, np.float64)
    # np.interp wants increasing abscissae: (below, above) -> (r_k, r_{k-1})
    return float(np.interp(0.5, [below, above], [r[k], r[k - 1]]))


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
                       fit_end=15.0, constants_file=None) -> dict:
    """Virtual-origin fit and half-radius spreading rates of a run
    (the numbers of profile_summary.txt, post/plot_profiles.py:112-135,276-308).
    D and U_jet default to the run's metadata, then case_info.txt, then the
    constants file (case_metadata)."""
    m = read_radial_moments(mean_dir, constants_file)
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

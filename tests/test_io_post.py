"""Output formats and post-processing (CPU only; host writers of the C ABI)."""
import base64
import os

import numpy as np
import pytest

import multicomponent_color_gradient_amd as m
from multicomponent_color_gradient_amd import io as mio
from multicomponent_color_gradient_amd import post

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_output_dirs_layout(tmp_path):
    c = m.reference_jet()
    cd, md, vd = mio.output_dirs(c, str(tmp_path))
    assert cd == str(tmp_path / "JET_VTK" / "Re5000_We2000")
    assert md == cd + "/mean_profiles" and vd == cd + "/vti_data"
    assert os.path.isdir(md) and os.path.isdir(vd)


def test_case_info_text(tmp_path):
    c = m.reference_jet()
    mio.write_case_info(c, str(tmp_path))
    expected = ("Re 5000\nWe 2000\nNX 128\nNY 400\nNZ 128\nNSTEP 200000\nNOUTPUT 2000\nNSTATS_START_STEP 6000\n"
                "NR_BINS 64\nwrite_vti_output 1\njet_radius 7\njet_diameter 14\njet_velocity 0.0500000007\n")
    assert (tmp_path / "case_info.txt").read_text() == expected


def test_vti_round_trip_and_header(tmp_path):
    rng = np.random.default_rng(1)
    shape = (5, 7, 6)  # NZ, NY, NX
    f = {k: rng.standard_normal(shape).astype(np.float32) for k in ("rhor", "rhob", "ux", "uy", "uz")}
    p = str(tmp_path / "lbm_00000042.vti")
    mio.write_vti_arrays(p, f)
    text = open(p).read()
    lines = text.split("\n")
    assert lines[0] == '<?xml version="1.0"?>'
    assert lines[1] == '<VTKFile type="ImageData" version="1.0" byte_order="LittleEndian" header_type="UInt64">'
    assert lines[2] == '  <ImageData WholeExtent="0 5 0 6 0 4" Origin="0 0 0" Spacing="1 1 1">'
    assert lines[3] == '    <Piece Extent="0 5 0 6 0 4">'
    assert lines[4] == '      <PointData Scalars="rhor" Vectors="u">'
    assert lines[5] == '        <DataArray type="Float32" Name="rhor" format="binary">'
    assert text.endswith("      </PointData>\n      <CellData>\n      </CellData>\n    </Piece>\n  </ImageData>\n</VTKFile>\n")
    # independent decode of the first block with the standard library
    raw = base64.b64decode(lines[6])
    assert int.from_bytes(raw[:8], "little") == 6 * 7 * 5 * 4
    assert np.array_equal(np.frombuffer(raw[8:], np.float32).reshape(shape), f["rhor"])
    g = mio.read_vti(p)
    assert g["shape"] == shape
    for k in ("rhor", "rhob"):
        assert np.array_equal(g[k], f[k])
    assert np.array_equal(g["rho"], f["rhor"] + f["rhob"])
    assert np.array_equal(g["u"], np.stack([f["ux"], f["uy"], f["uz"]], -1))


@pytest.mark.parametrize("n", [1, 2, 3, 4])
def test_base64_padding_lengths(tmp_path, n):
    # payload sizes 8 + 4*n bytes exercise every padding case of the encoder
    shape = (1, 1, n)
    f = {k: np.full(shape, 1.5, np.float32) for k in ("rhor", "rhob", "ux", "uy", "uz")}
    p = str(tmp_path / "t.vti")
    mio.write_vti_arrays(p, f)
    line = open(p).read().split("\n")[6]
    assert line == base64.b64encode(np.uint64(4 * n).tobytes() + f["rhor"].tobytes()).decode()


def test_radial_files_and_metadata(tmp_path):
    c = m.jet_case(12, 9, 10, 3.0)
    c.NSTATS_START_STEP = 4
    c.NSTATS_SAMPLE = 2
    nb = (9, 6)
    rng = np.random.default_rng(2)
    rad = {k: rng.standard_normal(nb) for k in ("sum_uy", "sum_uy2", "sum_ur", "sum_ur2", "sum_uruy", "sum_phi", "sum_phi2")}
    rad["count"] = rng.integers(0, 100, nb).astype(np.uint64)
    mio.write_radial_files(c, str(tmp_path), rad, started=False, start_step=0, samples=0)
    assert not (tmp_path / "radial_sum_uy.bin").exists()
    assert post.read_metadata(tmp_path)["uy_avg_started"] == "0"
    mio.write_radial_files(c, str(tmp_path), rad, started=True, start_step=4, samples=5)
    got = post.read_radial_moments(tmp_path)
    for k in ("sum_uy", "sum_uy2", "sum_ur", "sum_ur2", "sum_uruy", "sum_phi", "sum_phi2"):
        assert np.array_equal(got[k], rad[k])
        assert os.path.getsize(tmp_path / f"radial_{k}.bin") == 9 * 6 * 8
    assert np.array_equal(got["count"], rad["count"])
    md = got["metadata"]
    assert md["NY"] == 9 and md["NR_BINS"] == 6 and md["NradialProfileCells"] == 54
    assert md["NSTATS_SAMPLE"] == 2 and md["radial_sample_count"] == 5
    assert md["uy_avg_first_sample_step"] == 4 and md["uy_avg_last_sample_step"] == 4 + 4 * 2
    assert md["D"] == 6.0 and md["U_jet"] == pytest.approx(0.05)


def test_tke_series(tmp_path):
    ts = mio.TkeSeries(noutput=100, mean_dir=str(tmp_path))
    totals = [1.0, 3.0, 2.0, 6.0]
    outs = [ts.sample(t, 100 * k) for k, t in enumerate(totals)]
    avgs = np.cumsum(totals) / np.arange(1, 5)
    assert np.allclose([o[1] for o in outs], avgs, rtol=1e-7)
    diffs = [0.0] + list(np.abs(np.diff(avgs)))
    assert np.allclose([o[2] for o in outs], diffs, rtol=1e-6)
    assert np.array_equal(np.fromfile(tmp_path / "tke_total.bin", np.float32), np.float32(totals))
    assert np.fromfile(tmp_path / "tke_avg.bin", np.float32).size == 4


def test_hussein_half_radius_matches_reference_summary():
    # resultsJET/profile_summary.txt:7 of the reference: hussein_r_half_over_y_minus_y0
    assert post.first_half_radius(post.HUSSEIN_U, post.HUSSEIN_X, 1.0) == pytest.approx(0.09632634867890462, rel=1e-14)


def test_spreading_rate_relations_of_reference_summary():
    # rows of resultsJET/profile_summary.txt:9-14: S = r_half/(y - y0), S/S_H
    y0, sh = 50.42832501969275, 0.09632634867890462
    rows = [(140, 6.48003854, 0.0723447289, 0.751037799), (154, 7.36245272, 0.0710855812, 0.737966114),
            (210, 13.627404, 0.0853998932, 0.886568362)]
    for yi, rh, S, ratio in rows:
        assert rh / (yi - y0) == pytest.approx(S, rel=1e-8)
        assert S / sh == pytest.approx(ratio, rel=1e-8)


def test_first_half_radius_linear_case():
    r = np.arange(10, dtype=float)
    prof = 1.0 - 0.08 * r  # crosses 1/2 at r = 6.25
    assert post.first_half_radius(prof, r, 1.0) == pytest.approx(6.25, rel=1e-14)
    assert np.isnan(post.first_half_radius(prof, r, 0.0))
    assert np.isnan(post.first_half_radius(np.ones(5), np.arange(5.0), 1.0))


def test_fit_virtual_origin_recovers_line():
    D, U, B, y0 = 14.0, 0.05, 6.5, 50.0
    y = np.arange(400, dtype=float)
    uc = U * B * D / np.maximum(y - y0, 1e-9)  # U/U_c = (y - y0)/(B D)
    Bf, y0f, *_ = post.fit_virtual_origin(y, uc, D, U)
    assert Bf == pytest.approx(B, rel=1e-10) and y0f == pytest.approx(y0, rel=1e-10)


def test_averaged_fields_and_bin_centers():
    mom = {"sum_uy": np.array([[2.0, 0.0]]), "sum_uy2": np.array([[2.5, 0.0]]), "sum_ur": np.array([[1.0, 0.0]]),
           "sum_ur2": np.array([[1.0, 0.0]]), "sum_uruy": np.array([[1.5, 0.0]]), "count": np.array([[2, 0]], np.uint64)}
    f = post.averaged_radial_fields(mom)
    assert f["uy"][0, 0] == 1.0 and np.isnan(f["uy"][0, 1])
    assert f["uy_rms"][0, 0] == pytest.approx(0.5)
    assert f["uruy_reynolds"][0, 0] == pytest.approx(0.75 - 0.5)
    # 4x4 interior around a half-integer centre: offsets +-0.5, +-1.5 give
    # r = sqrt(0.5) (4 nodes, bin 0 -> centreline), sqrt(2.5) (8 nodes), sqrt(4.5) (4 nodes)
    r = post.radial_bin_centers(6, 6, 3, 2.5, 2.5)
    assert r[0] == 0.0
    assert r[1] == pytest.approx(np.sqrt(2.5), rel=1e-15)
    assert r[2] == pytest.approx(np.sqrt(4.5), rel=1e-15)


def test_fit_virtual_origin_matches_library_least_squares():
    # the closed-form fit against numpy's least-squares polynomial fit on noisy points
    rng = np.random.default_rng(11)
    D, U = 16.0, 0.05
    y = np.arange(300, dtype=float)
    uc = U * 6.0 * D / np.maximum(y - 40.0, 1.0) * (1.0 + 0.02 * rng.standard_normal(y.size))
    uc[170] = np.nan  # skipped: not finite
    uc[200] = -1.0    # skipped: not positive
    fit = post.fit_virtual_origin(y, uc, D, U)
    keep = np.isfinite(uc) & (uc > 0) & (y / D >= 10.0) & (y / D <= 15.0)
    slope, icpt = np.polyfit(y[keep] / D, U / uc[keep], 1)
    assert fit.slope == pytest.approx(slope, rel=1e-12) and fit.intercept == pytest.approx(icpt, rel=1e-10)
    assert fit.B == pytest.approx(1.0 / slope, rel=1e-12)
    assert fit.y0 == pytest.approx(-icpt / slope * D, rel=1e-10)
    assert fit.x.size == np.count_nonzero(keep)
    with pytest.raises(ValueError):
        post.fit_virtual_origin(y, np.full(y.size, np.nan), D, U)


def test_first_half_radius_crossings():
    r = np.array([0.0, 1.0, 2.0, 3.0])
    # a crossing exactly on a sample
    assert post.first_half_radius(np.array([2.0, 1.5, 1.0, 0.2]), r, 2.0) == 2.0
    # the first crossing wins over a later one
    assert post.first_half_radius(np.array([1.0, 0.4, 0.9, 0.1]), r, 1.0) == pytest.approx(5.0 / 6.0, rel=1e-14)
    # a non-finite sample before the first one at or below 1/2: no crossing
    assert np.isnan(post.first_half_radius(np.array([1.0, np.nan, 0.3, 0.1]), r, 1.0))
    # the centre sample itself is ignored as a crossing
    assert np.isnan(post.first_half_radius(np.array([0.4, 0.45, 0.3, 0.1]), r, 1.0))


def test_metadata_fallbacks(tmp_path):
    # no radial_profile_metadata.txt: NY and NR_BINS from case_info.txt, then from a constants header
    case_dir = tmp_path / "JET_VTK" / "Re100_We10"
    mean_dir = case_dir / "mean_profiles"
    mean_dir.mkdir(parents=True)
    ny, nr = 5, 3
    rng = np.random.default_rng(4)
    for k in ("sum_uy", "sum_uy2", "sum_ur", "sum_ur2", "sum_uruy"):
        rng.standard_normal(ny * nr).tofile(mean_dir / f"radial_{k}.bin")
    np.arange(ny * nr, dtype=np.uint64).tofile(mean_dir / "radial_count.bin")
    hdr = tmp_path / "constants.cuh"
    hdr.write_text(
        "// a header in the reference's declaration style\n"
        "inline constexpr int NOUTPUT = 100;\n"
        "inline constexpr label_t NX = static_cast<label_t>(6);\n"
        "inline constexpr label_t NY = static_cast<label_t>(5); // streamwise\n"
        "inline constexpr label_t NZ = static_cast<label_t>(6);\n"
        "inline constexpr label_t NR_BINS = NX / static_cast<label_t>(2);\n"
        "inline constexpr real_t jet_radius = static_cast<real_t>(1.5);\n"
        "inline constexpr real_t jet_x0 = static_cast<real_t>(NX - 1) / static_cast<real_t>(2);\n"
        "inline constexpr real_t jet_velocity = static_cast<real_t>(0.04);\n"
        "inline constexpr real_t nu = (static_cast<real_t>(2) * jet_radius * jet_velocity) / 100.0;\n"
        "inline constexpr std::array<int, 2> skipped = {1, 2};\n")
    k = post.read_case_constants(hdr)
    assert k["NX"] == 6 and k["NY"] == 5 and k["NR_BINS"] == 3 and k["NOUTPUT"] == 100
    assert k["jet_x0"] == 2.5 and k["D"] == 3.0 and k["U_jet"] == pytest.approx(0.04)
    assert k["nu"] == pytest.approx(2 * 1.5 * 0.04 / 100.0)
    assert "skipped" not in k
    with pytest.raises(ValueError):
        post.read_radial_moments(mean_dir)
    got = post.read_radial_moments(mean_dir, constants_file=hdr)
    assert got["count"].shape == (ny, nr) and got["metadata"]["D"] == 3.0
    # case_info.txt of the case folder takes precedence over the constants header
    c = m.jet_case(6, 5, 6, 1.0)
    mio.write_case_info(c, str(case_dir))
    md = post.case_metadata(mean_dir, constants_file=hdr)
    assert md["NY"] == 5 and md["D"] == 2.0 and md["U_jet"] == pytest.approx(0.05)

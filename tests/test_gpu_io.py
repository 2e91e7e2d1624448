"""The run driver and the output files on the GPU, checked against the oracle.

Tolerances are those of test_gpu_parity.py: per radial bin |d sum_u*| <=
1e-6 count (FP32 GPU vs FP64 oracle); TKE relative <= 1e-4.
"""
import os
import subprocess
import sys

import numpy as np
import pytest

import multicomponent_color_gradient_amd as m
from multicomponent_color_gradient_amd import io as mio
from multicomponent_color_gradient_amd import post
from multicomponent_color_gradient_amd.build import DRIVER

pytestmark = pytest.mark.gpu


def test_solver_vti_matches_fields(tmp_path):
    case = m.jet_case(20, 18, 22, jet_radius=4.0, Re=400.0)
    s = m.Solver(case)
    s.init_jet()
    s.step(7)
    p = str(tmp_path / "x.vti")
    s.write_vti(p)
    g = mio.read_vti(p)
    f = s.fields()
    for k in ("rhor", "rhob"):
        assert np.array_equal(g[k], f[k])
    assert np.array_equal(g["u"][..., 1], f["uy"])


def test_driver_run_matches_oracle(tmp_path):
    from oracle.oracle import Oracle

    args = ["--nx", "34", "--ny", "46", "--nz", "30", "--radius", "5", "--Re", "400", "--We", "2000",
            "--steps", "201", "--noutput", "100", "--stats-start", "50", "--out", str(tmp_path)]
    print("driver:", " ".join(args), flush=True)
    res = subprocess.run([DRIVER, *args], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "MLUPS" in res.stdout
    case_dir = tmp_path / "JET_VTK" / "Re400_We2000"
    md = case_dir / "mean_profiles"
    assert (case_dir / "case_info.txt").exists()
    for stp in (0, 100, 200):
        assert (case_dir / "vti_data" / f"lbm_{stp:08d}.vti").exists()
    tke = np.fromfile(md / "tke_total.bin", np.float32)
    assert tke.size == 3
    got = post.read_radial_moments(md)
    assert got["metadata"]["radial_sample_count"] == 151
    # the same case on the oracle
    case = m.jet_case(34, 46, 30, jet_radius=5.0, Re=400.0, We=2000.0)
    case.NSTATS_START_STEP = 50
    o = Oracle(case)
    o.init_jet()
    o.step(201)
    r = o.radial()
    assert np.array_equal(got["count"], r["count"])
    tol = 1e-6 * r["count"].astype(np.float64) + 1e-12
    for k in ("sum_uy", "sum_uy2", "sum_ur", "sum_ur2", "sum_uruy"):
        assert np.all(np.abs(got[k] - r[k]) <= tol), k
    assert abs(float(tke[2]) - o.tke()) <= 1e-4 * o.tke()
    v = mio.read_vti(str(case_dir / "vti_data" / "lbm_00000200.vti"))
    f = o.fields()
    assert np.max(np.abs(v["rhob"] - f["rhob"])) <= 5e-5


def test_driver_with_fp16_storage(tmp_path):
    """mcg_jet --storage fp16: the same short run completes and its radial
    statistics stay within the 16-bit tolerance of the FP32 run's."""
    base = ["--nx", "34", "--ny", "46", "--nz", "30", "--radius", "5", "--Re", "400", "--We", "2000",
            "--steps", "101", "--noutput", "100", "--stats-start", "50", "--no-vti"]
    got = {}
    for st in ("fp32", "fp16"):
        out = tmp_path / st
        print("driver:", st, flush=True)
        res = subprocess.run([DRIVER, *base, "--storage", st, "--out", str(out)], capture_output=True, text=True,
                             timeout=120)
        assert res.returncode == 0, res.stdout + res.stderr
        got[st] = post.read_radial_moments(out / "JET_VTK" / "Re400_We2000" / "mean_profiles")
    assert np.array_equal(got["fp16"]["count"], got["fp32"]["count"])
    mask = got["fp32"]["count"] > 0
    a = got["fp16"]["sum_uy"][mask] / got["fp16"]["count"][mask]
    b = got["fp32"]["sum_uy"][mask] / got["fp32"]["count"][mask]
    assert np.max(np.abs(a - b)) <= 1e-3 * np.max(np.abs(b))
    res = subprocess.run([DRIVER, *base, "--storage", "fp8", "--out", str(tmp_path / "x")], capture_output=True,
                         text=True, timeout=60)
    assert res.returncode == 2

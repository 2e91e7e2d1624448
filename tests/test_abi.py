"""C ABI and host-side checks (no GPU needed: nothing here launches a kernel)."""
import ctypes
import os
import re
import subprocess

import pytest

import multicomponent_color_gradient_amd as m
from multicomponent_color_gradient_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "mcg.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mcg_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = m.load_library()
    out = subprocess.run(["nm", "-D", "--defined-only", _abi.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    declared = declared_functions()
    assert len(declared) >= 28
    missing = [f for f in declared if f not in exported]
    assert not missing, missing
    for f in declared:
        getattr(lib, f)


def test_binding_covers_the_header():
    bound = {name for name, _, _ in _abi.SIGNATURES}
    assert set(declared_functions()) == bound


def test_case_struct_layout_and_reference_defaults():
    lib = m.load_library()
    assert lib.mcg_case_size() == ctypes.sizeof(_abi.McgCase)
    c = _abi.McgCase()
    assert lib.mcg_case_reference_jet(ctypes.byref(c)) == 0
    assert c.as_dict() == m.reference_jet().as_dict()
    # constants.cuh:14-76
    assert (c.NX, c.NY, c.NZ) == (128, 400, 128)
    assert c.jet_x0 == 63.5 and c.jet_radius == 7.0 and c.jet_velocity == 0.05
    assert abs(c.taub - (0.5 + 3 * 2 * 7 * 0.05 / 5000)) < 1e-15
    assert abs(c.sigma - 0.05 ** 2 * 14 / 2000) < 1e-18
    assert c.beta_recolor == 0.8 and c.NSTATS_SAMPLE == 1
    d = m.derived(c)
    assert d["NR_BINS"] == 64 and d["Ncells"] == 128 * 400 * 128
    assert abs(d["A"] - c.sigma / (4 / 9.0) / c.taub) < 1e-18


@pytest.mark.parametrize("field,value,code", [
    ("NX", 3, _abi.MCG_E_INVALID),
    ("NX", 4096, _abi.MCG_E_INVALID),
    ("y_mode", 7, _abi.MCG_E_INVALID),
    ("taub", 0.5, _abi.MCG_E_INVALID),
    ("inlet_profile", 5, _abi.MCG_E_INVALID),
    ("perturbation_amp", -0.1, _abi.MCG_E_INVALID),
])
def test_case_validation_errors(field, value, code):
    lib = m.load_library()
    c = m.reference_jet()
    setattr(c, field, value)
    assert lib.mcg_case_validate(ctypes.byref(c)) == code
    assert lib.mcg_last_error()


def test_case_validation_accepts_baseline_grids():
    lib = m.load_library()
    for (nx, ny, nz) in [(64, 64, 64), (128, 256, 128), (256, 512, 256), (192, 384, 192)]:
        c = m.jet_case(nx, ny, nz, 8.0)
        assert lib.mcg_case_validate(ctypes.byref(c)) == 0


def test_case_validation_grid_limits():
    """Population arrays may exceed 2^32 elements (addresses are a z-plane base
    plus a 32-bit offset within the plane); one z-plane must stay below 4 GiB
    and the padded interior below 2^32 nodes."""
    lib = m.load_library()
    # 512^3: 7.2e9 FP32 population elements per buffer, planes of 56 MB
    c = m.jet_case(512, 512, 512, 16.0)
    assert lib.mcg_case_validate(ctypes.byref(c)) == 0
    # 130 x 130 x 5002, the z-translation test's box: 4.4e9 elements
    c = m.jet_case(130, 130, 5002, 8.0)
    assert lib.mcg_case_validate(ctypes.byref(c)) == 0
    # a plane of 2048 x 10000 nodes: 4.4 GB of FP32 populations
    c = m.jet_case(2048, 10000, 8, 8.0)
    assert lib.mcg_case_validate(ctypes.byref(c)) == _abi.MCG_E_INVALID
    assert b"z-plane" in lib.mcg_last_error()
    # 2048 x 2048 x 1100 padded interior nodes exceed 2^32
    c = m.jet_case(2048, 2048, 1100, 8.0)
    assert lib.mcg_case_validate(ctypes.byref(c)) == _abi.MCG_E_INVALID


def test_errors_without_gpu_are_status_codes():
    """Null handles and pointers are rejected with a status, never a crash."""
    lib = m.load_library()
    assert lib.mcg_step(None, 1) == _abi.MCG_E_INVALID
    assert lib.mcg_destroy(None) == 0
    assert lib.mcg_adjoint_set_checkpoint_limit(None, 3) == _abi.MCG_E_INVALID
    assert lib.mcg_adjoint_recomputed_steps(None, None) == _abi.MCG_E_INVALID
    with pytest.raises(_abi.McgError):
        _abi.check(_abi.MCG_E_INVALID)


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(_abi.McgError):
        _abi.load_library(str(tmp_path / "nope.so"))

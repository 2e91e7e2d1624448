"""Output files in the reference's layout, through the C ABI writers.

  <root>/JET_VTK/Re<Re>_We<We>/case_info.txt
  <root>/JET_VTK/Re<Re>_We<We>/mean_profiles/{tke_*.bin, radial_*.bin, radial_profile_metadata.txt}
  <root>/JET_VTK/Re<Re>_We<We>/vti_data/lbm_%08d.vti

(meanF/meanFieldsRT.cuh:36-375, io/save_data.cuh:22-215).  ``read_vti`` parses
the VTK XML files back into arrays shaped (NZ, NY, NX).
"""
from __future__ import annotations

import base64
import ctypes
import os
import re

import numpy as np

from ._abi import McgCase, McgError, as_ptr, load_library


def _io_check(code: int) -> None:
    if code != 0:
        msg = load_library().mcg_io_last_error()
        raise McgError(code, msg.decode() if msg else "")


def output_dirs(case: McgCase, root: str = ".") -> tuple[str, str, str]:
    """Create and return (case_dir, mean_dir, vti_dir)."""
    lib = load_library()
    bufs = [ctypes.create_string_buffer(4096) for _ in range(3)]
    _io_check(lib.mcg_output_dirs(ctypes.byref(case), os.fsencode(root), *bufs, 4096))
    return tuple(b.value.decode() for b in bufs)


def write_case_info(case: McgCase, case_dir: str) -> None:
    _io_check(load_library().mcg_write_case_info(ctypes.byref(case), os.fsencode(case_dir)))


def write_vti_arrays(path: str, fields: dict) -> None:
    """fields: rhor, rhob, ux, uy, uz arrays shaped (NZ, NY, NX)."""
    arrs = [np.ascontiguousarray(fields[k], np.float32) for k in ("rhor", "rhob", "ux", "uy", "uz")]
    nz, ny, nx = arrs[0].shape
    f = ctypes.c_float
    _io_check(load_library().mcg_write_vti_arrays(os.fsencode(path), nx, ny, nz, *[as_ptr(a, f) for a in arrs]))


def write_radial_files(case: McgCase, mean_dir: str, radial: dict, started: bool, start_step: int, samples: int) -> None:
    d = ctypes.c_double
    names = ["sum_uy", "sum_uy2", "sum_ur", "sum_ur2", "sum_uruy"]
    arrs = [np.ascontiguousarray(radial[n], np.float64) for n in names]
    cnt = np.ascontiguousarray(radial["count"], np.uint64)
    phi = [np.ascontiguousarray(radial[n], np.float64) if n in radial else None for n in ("sum_phi", "sum_phi2")]
    _io_check(load_library().mcg_write_radial_files(
        ctypes.byref(case), os.fsencode(mean_dir), *[as_ptr(a, d) for a in arrs], as_ptr(cnt, ctypes.c_uint64),
        *[as_ptr(p, d) if p is not None else None for p in phi], int(started), int(start_step), int(samples)))


class TkeSeries:
    """tke_total / tke_avg / tke_diff bookkeeping of the reference (FP32),
    appended to the mean_profiles folder when ``mean_dir`` is given."""

    def __init__(self, noutput: int, mean_dir: str | None = None):
        self.noutput = int(noutput)
        self.mean_dir = mean_dir
        self.state = np.array([0.0, 0.0, 1.0], np.float32)
        if mean_dir is not None:
            _io_check(load_library().mcg_tke_files_reset(os.fsencode(mean_dir)))

    def sample(self, tke_total: float, step: int) -> np.ndarray:
        out = np.zeros(3, np.float32)
        f = ctypes.c_float
        _io_check(load_library().mcg_tke_update(as_ptr(self.state, f), float(np.float32(tke_total)), int(step),
                                                self.noutput, as_ptr(out, f)))
        if self.mean_dir is not None:
            _io_check(load_library().mcg_append_tke(os.fsencode(self.mean_dir), as_ptr(out, f)))
        return out


def read_vti(path: str) -> dict:
    """Parse a VTK XML ImageData file written by write_vti (binary, UInt64 headers)."""
    text = open(path, "rb").read().decode()
    ext = re.search(r'WholeExtent="0 (\d+) 0 (\d+) 0 (\d+)"', text)
    nx, ny, nz = (int(v) + 1 for v in ext.groups())
    out = {"shape": (nz, ny, nx)}
    for m in re.finditer(r'<DataArray type="Float32" Name="(\w+)"(?: NumberOfComponents="(\d)")? format="binary">\n(.*?)\n',
                         text, re.S):
        name, ncomp, payload = m.group(1), int(m.group(2) or 1), m.group(3)
        raw = base64.b64decode(payload)
        nbytes = int(np.frombuffer(raw[:8], np.uint64)[0])
        if nbytes != len(raw) - 8:
            raise ValueError(f"{name}: header says {nbytes} bytes, payload has {len(raw) - 8}")
        a = np.frombuffer(raw[8:], np.float32)
        out[name] = a.reshape((nz, ny, nx, ncomp)) if ncomp > 1 else a.reshape((nz, ny, nx))
    return out

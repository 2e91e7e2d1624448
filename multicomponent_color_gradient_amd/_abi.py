"""ctypes mirror of include/mcg.h and the loader of the HIP library.

The library is built in-tree (``__graft_entry__.build()``) as
``multicomponent_color_gradient_amd/libmcg_hip.so``.  There is no fallback:
if the library is missing, every solver call raises ``McgError``.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import (POINTER, c_char_p, c_double, c_float, c_int, c_int32, c_int64, c_size_t, c_uint8, c_uint32, c_uint64,
                    c_void_p)

LIB_NAME = "libmcg_hip.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

MCG_OK = 0
MCG_E_INVALID = -1
MCG_E_HIP = -2
MCG_E_NOMEM = -3
MCG_E_STATE = -4
MCG_E_UNSUPPORTED = -5
MCG_E_NUMERIC = -6
MCG_E_IO = -7

MCG_Y_JET = 0
MCG_Y_PERIODIC = 1
MCG_VISC_SINGLE = 0
MCG_VISC_INTERFACE = 1
MCG_INLET_SHARP = 0
MCG_INLET_TANH = 1
MCG_STORE_FP32 = 0
MCG_STORE_FP16 = 1
MCG_STORE_BF16 = 2
MCG_KERNEL_FUSED = 0
MCG_KERNEL_TWOPASS = 1

STORAGE = {"fp32": MCG_STORE_FP32, "fp16": MCG_STORE_FP16, "bf16": MCG_STORE_BF16}
KERNELS = {"fused": MCG_KERNEL_FUSED, "twopass": MCG_KERNEL_TWOPASS}


class McgError(RuntimeError):
    """A solver call returned a non-zero status."""

    def __init__(self, code: int, message: str):
        super().__init__(f"mcg status {code}: {message}")
        self.code = code


class McgCase(ctypes.Structure):
    """Mirror of ``struct mcg_case`` (run-time form of constants.cuh:8-76)."""

    _fields_ = [
        ("NX", c_int32),
        ("NY", c_int32),
        ("NZ", c_int32),
        ("NSTEP", c_int32),
        ("NOUTPUT", c_int32),
        ("NSTATS_START_STEP", c_int32),
        ("NSTATS_SAMPLE", c_int32),
        ("write_vti_output", c_int32),
        ("sponge_cells", c_int32),
        ("y_mode", c_int32),
        ("wrap_gradient", c_int32),
        ("viscosity_mode", c_int32),
        ("inlet_profile", c_int32),
        ("perturbation_seed", c_int32),
        ("wall_xz", c_int32),
        ("reserved_i", c_int32 * 2),
        ("jet_radius", c_double),
        ("jet_x0", c_double),
        ("jet_z0", c_double),
        ("jet_velocity", c_double),
        ("bubble_radius", c_double),
        ("bubble_x0", c_double),
        ("bubble_y0", c_double),
        ("bubble_z0", c_double),
        ("rhor0", c_double),
        ("rhob0", c_double),
        ("Re", c_double),
        ("We", c_double),
        ("taur", c_double),
        ("taub", c_double),
        ("sigma", c_double),
        ("beta_recolor", c_double),
        ("A_override", c_double),
        ("sponge_K", c_double),
        ("interface_width", c_double),
        ("perturbation_amp", c_double),
        ("gravity", c_double),
        ("reserved_d", c_double * 3),
    ]

    def copy(self) -> "McgCase":
        c = McgCase()
        ctypes.memmove(ctypes.byref(c), ctypes.byref(self), ctypes.sizeof(McgCase))
        return c

    def as_dict(self) -> dict:
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            out[name] = list(v) if hasattr(v, "__len__") else v
        return out

    def __repr__(self) -> str:
        return (f"McgCase(NX={self.NX}, NY={self.NY}, NZ={self.NZ}, y_mode={self.y_mode}, "
                f"taub={self.taub:.6g}, sigma={self.sigma:.6g}, beta={self.beta_recolor:.3g})")


# exported symbols and their signatures: (name, restype, argtypes)
_PSOLVER = c_void_p
_F = POINTER(c_float)
_D = POINTER(c_double)
SIGNATURES = [
    ("mcg_case_reference_jet", c_int, [POINTER(McgCase)]),
    ("mcg_case_validate", c_int, [POINTER(McgCase)]),
    ("mcg_case_size", c_size_t, []),
    ("mcg_create", c_int, [POINTER(McgCase), c_int, c_void_p, POINTER(c_void_p)]),
    ("mcg_destroy", c_int, [_PSOLVER]),
    ("mcg_set_kernel", c_int, [_PSOLVER, c_int]),
    ("mcg_init_jet", c_int, [_PSOLVER]),
    ("mcg_init_bubble", c_int, [_PSOLVER]),
    ("mcg_init_bubble_tanh", c_int, [_PSOLVER]),
    ("mcg_init_spheres", c_int, [_PSOLVER, c_int, c_uint32, c_double, c_double]),
    ("mcg_sphere_list", c_int, [c_uint32, c_int, c_int, c_int, c_int, c_double, c_double, _D]),
    ("mcg_step", c_int, [_PSOLVER, c_int]),
    ("mcg_synchronize", c_int, [_PSOLVER]),
    ("mcg_get_step", c_int, [_PSOLVER, POINTER(c_int64)]),
    ("mcg_stats_info", c_int, [_PSOLVER, POINTER(c_int32), POINTER(c_int64), POINTER(c_int64)]),
    ("mcg_stats_reset", c_int, [_PSOLVER]),
    ("mcg_copy_radial", c_int, [_PSOLVER, _D, _D, _D, _D, _D, POINTER(c_uint64), _D, _D]),
    ("mcg_tke", c_int, [_PSOLVER, _D]),
    ("mcg_global_sums", c_int, [_PSOLVER, _D]),
    ("mcg_global_sums_noalloc", c_int, [_PSOLVER, c_void_p]),
    ("mcg_set_launch_segments", c_int, [_PSOLVER, c_int]),
    ("mcg_launch_info", c_int,
     [_PSOLVER, POINTER(c_int32), POINTER(c_int32), POINTER(c_int32), POINTER(c_int32), POINTER(c_int32)]),
    ("mcg_stats_launch_info", c_int, [_PSOLVER, POINTER(c_int32), POINTER(c_int32), POINTER(c_int32)]),
    ("mcg_hbm_probe", c_int, [c_int, c_int, c_int, c_int, c_uint64, c_int, _D]),
    ("mcg_check_health", c_int, [_PSOLVER]),
    ("mcg_copy_fields", c_int, [_PSOLVER, _F, _F, _F, _F, _F]),
    ("mcg_copy_populations", c_int, [_PSOLVER, _F, _F]),
    ("mcg_set_populations", c_int, [_PSOLVER, _F, _F]),
    ("mcg_checkpoint_size", c_int, [_PSOLVER, POINTER(c_uint64)]),
    ("mcg_checkpoint_save", c_int, [_PSOLVER, c_void_p, c_uint64]),
    ("mcg_checkpoint_load", c_int, [_PSOLVER, c_void_p, c_uint64]),
    ("mcg_adjoint_record", c_int, [_PSOLVER, c_int, c_int]),
    ("mcg_adjoint_backward", c_int, [_PSOLVER, _F, _F, _F, _F, _D]),
    ("mcg_adjoint_last_ms", c_int, [_PSOLVER, _D]),
    ("mcg_adjoint_gradients", c_int, [_PSOLVER, _D, c_int]),
    ("mcg_adjoint_set_checkpoint_limit", c_int, [_PSOLVER, c_int]),
    ("mcg_adjoint_recomputed_steps", c_int, [_PSOLVER, POINTER(c_int64)]),
    ("mcg_adjoint_release", c_int, [_PSOLVER]),
    ("mcg_adjoint_set_uy_seed", c_int, [_PSOLVER, _F, c_int64, c_int64]),
    ("mcg_adjoint_inlet_gradient", c_int, [_PSOLVER, _D, _D]),
    ("mcg_uy_functional", c_int, [_PSOLVER, _D]),
    ("mcg_set_inlet_offset", c_int, [_PSOLVER, _F]),
    ("mcg_cpu_adjoint", c_int, [POINTER(McgCase), c_int, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D]),
    ("mcg_cpu_adjoint_jet", c_int, [POINTER(McgCase), c_int, _D, _D, _D, _D, _D, _D, _D, c_int, c_int, _D, _D, _D, _D, _D,
                                    _D]),
    ("mcg_time_steps", c_int, [_PSOLVER, c_int, _D]),
    ("mcg_device_bytes", c_int, [_PSOLVER, POINTER(c_uint64)]),
    ("mcg_output_dirs", c_int, [POINTER(McgCase), c_char_p, c_char_p, c_char_p, c_char_p, c_size_t]),
    ("mcg_write_case_info", c_int, [POINTER(McgCase), c_char_p]),
    ("mcg_write_vti_arrays", c_int, [c_char_p, c_int32, c_int32, c_int32, _F, _F, _F, _F, _F]),
    ("mcg_write_radial_files", c_int, [POINTER(McgCase), c_char_p, _D, _D, _D, _D, _D, POINTER(c_uint64), _D, _D,
                                       c_int32, c_int64, c_int64]),
    ("mcg_tke_update", c_int, [_F, c_float, c_int64, c_int32, _F]),
    ("mcg_tke_files_reset", c_int, [c_char_p]),
    ("mcg_append_tke", c_int, [c_char_p, _F]),
    ("mcg_write_vti", c_int, [_PSOLVER, c_char_p]),
    ("mcg_write_vti_step", c_int, [_PSOLVER, c_char_p, c_int64]),
    ("mcg_write_radial", c_int, [_PSOLVER, c_char_p]),
    ("mcg_io_last_error", c_char_p, []),
    ("mcg_last_error", c_char_p, []),
    ("mcg_version", c_char_p, []),
]

_lib = None


def library_path() -> str:
    return LIB_PATH


def load_library(path: str | None = None):
    """Load the HIP solver library (once).  Raises McgError if it is missing."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    # MCG_LIB_PATH selects another build of the same library (e.g. a tuning variant)
    p = path or os.environ.get("MCG_LIB_PATH") or LIB_PATH
    if not os.path.exists(p):
        raise McgError(MCG_E_STATE, f"{p} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    lib = ctypes.CDLL(p)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.mcg_case_size() != ctypes.sizeof(McgCase):
        raise McgError(MCG_E_STATE, "McgCase layout does not match the library's struct mcg_case")
    if path is None:
        _lib = lib
    return lib


def check(code: int) -> None:
    if code != MCG_OK:
        msg = load_library().mcg_last_error()
        raise McgError(code, msg.decode() if msg else "")


def as_ptr(arr, ctype):
    return arr.ctypes.data_as(POINTER(ctype))


__all__ = [n for n in dir() if n.startswith("MCG_")] + [
    "McgCase", "McgError", "load_library", "check", "as_ptr", "SIGNATURES", "STORAGE", "KERNELS",
    "c_uint8", "c_uint64", "c_double", "c_float", "c_int64", "c_int32",
]

"""ctypes wrapper of the CPU oracle (oracle/libmcg_oracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, always as the checker, never as the product.
The oracle restates the reference's step in FP64 (mcg_oracle.cpp); arrays use
the reference's layout reshaped to (NZ, NY, NX) / (27, NZ, NY, NX).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_int, c_int32, c_int64, c_uint8, c_uint64, c_void_p

import numpy as np

from multicomponent_color_gradient_amd._abi import McgCase

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmcg_oracle.so")

_D = POINTER(c_double)
_SIGS = [
    ("mcgo_create", c_void_p, [POINTER(McgCase), c_int]),
    ("mcgo_destroy", None, [c_void_p]),
    ("mcgo_init_jet", c_int, [c_void_p]),
    ("mcgo_init_bubble", c_int, [c_void_p]),
    ("mcgo_init_bubble_tanh", c_int, [c_void_p]),
    ("mcgo_step", c_int, [c_void_p, c_int]),
    ("mcgo_get_step", c_int64, [c_void_p]),
    ("mcgo_get_fields", None, [c_void_p, _D, _D, _D, _D, _D]),
    ("mcgo_get_moments", None, [c_void_p, _D]),
    ("mcgo_get_populations", None, [c_void_p, _D, _D]),
    ("mcgo_set_populations", None, [c_void_p, _D, _D]),
    ("mcgo_set_densities", None, [c_void_p, _D, _D]),
    ("mcgo_get_radial", None, [c_void_p, _D, _D, _D, _D, _D, POINTER(c_uint64), _D, _D]),
    ("mcgo_stats_info", None, [c_void_p, POINTER(c_int32), POINTER(c_int64), POINTER(c_int64)]),
    ("mcgo_tke", c_double, [c_void_p]),
    ("mcgo_global_sums", None, [c_void_p, _D]),
    ("mcgo_writeset", c_int, [c_void_p, POINTER(c_uint8), POINTER(c_uint8)]),
    ("mcgo_gradient", None, [c_void_p, c_int, c_int, c_int, _D]),
    ("mcgo_set_inlet_offset", None, [c_void_p, _D]),
    ("mcgo_uy_dot", c_double, [c_void_p, _D]),
    ("mcgo_branch_masks", None, [c_void_p, POINTER(c_uint8)]),
    ("mcgo_lattice", None, [POINTER(c_int), POINTER(c_int), POINTER(c_int), _D, _D, _D]),
    ("mcgo_feq", c_double, [c_int, c_double, c_double, c_double, c_double]),
    ("mcgo_fneqr", c_double, [c_int, _D, c_double, c_double, c_double]),
    ("mcgo_omega2", c_double, [c_int, c_double, c_double, c_double, c_double, c_double]),
    ("mcgo_recolor", c_double, [c_int, c_double, c_double, c_double, c_double, c_double, c_double, c_double]),
    ("mcgo_num_threads", c_int, []),
]

_lib = None


def build() -> str:
    """Compile the oracle with its Makefile (g++ -fopenmp)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_D)


def lattice():
    L = lib()
    cx, cy, cz = (np.zeros(27, np.int32) for _ in range(3))
    w, B, ic = (np.zeros(27) for _ in range(3))
    ip = lambda a: a.ctypes.data_as(POINTER(c_int))  # noqa: E731
    L.mcgo_lattice(ip(cx), ip(cy), ip(cz), _p(w), _p(B), _p(ic))
    return {"cx": cx, "cy": cy, "cz": cz, "w": w, "B": B, "invc": ic}


def feq(i, rho, ux, uy, uz):
    return lib().mcgo_feq(i, rho, ux, uy, uz)


def fneqr(i, P, ux, uy, uz):
    P = np.ascontiguousarray(P, np.float64)
    return lib().mcgo_fneqr(i, _p(P), ux, uy, uz)


def omega2(i, F, absF, A):
    return lib().mcgo_omega2(i, F[0], F[1], F[2], absF, A)


def recolor(i, rr, rb, F, absF, beta):
    return lib().mcgo_recolor(i, rr, rb, F[0], F[1], F[2], absF, beta)


def num_threads() -> int:
    return lib().mcgo_num_threads()


class Oracle:
    """FP64 (precision=64) or FP32 (precision=32) CPU restatement of the reference step."""

    def __init__(self, case: McgCase, precision: int = 64):
        self.case = case.copy()
        self._h = lib().mcgo_create(ctypes.byref(self.case), precision)
        if not self._h:
            raise RuntimeError("oracle: could not create state")
        self.shape = (case.NZ, case.NY, case.NX)
        self.N = case.NX * case.NY * case.NZ

    def close(self):
        if self._h:
            lib().mcgo_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def init_jet(self):
        lib().mcgo_init_jet(self._h)

    def init_bubble(self):
        lib().mcgo_init_bubble(self._h)

    def init_bubble_tanh(self):
        lib().mcgo_init_bubble_tanh(self._h)

    def step(self, n=1):
        rc = lib().mcgo_step(self._h, int(n))
        if rc != 0:
            raise RuntimeError(f"oracle step failed ({rc})")

    @property
    def step_count(self):
        return lib().mcgo_get_step(self._h)

    def fields(self):
        arrs = [np.empty(self.shape) for _ in range(5)]
        lib().mcgo_get_fields(self._h, *[_p(a) for a in arrs])
        return dict(zip(["rhor", "rhob", "ux", "uy", "uz"], arrs))

    def moments(self):
        out = np.empty((11,) + self.shape)
        lib().mcgo_get_moments(self._h, _p(out))
        return dict(zip(["rhor", "rhob", "ux", "uy", "uz", "pxx", "pxy", "pyy", "pyz", "pzz", "pxz"], out))

    def populations(self):
        fr = np.empty((27,) + self.shape)
        fb = np.empty((27,) + self.shape)
        lib().mcgo_get_populations(self._h, _p(fr), _p(fb))
        return fr, fb

    def set_populations(self, fr, fb):
        fr = np.ascontiguousarray(fr, np.float64).reshape((27,) + self.shape)
        fb = np.ascontiguousarray(fb, np.float64).reshape((27,) + self.shape)
        lib().mcgo_set_populations(self._h, _p(fr), _p(fb))

    def set_inlet_offset(self, du):
        """Offset added to the inlet velocity of every jet column, shape (3, NZ, NX); None clears it."""
        if du is None:
            lib().mcgo_set_inlet_offset(self._h, None)
            return
        self._du = np.ascontiguousarray(du, np.float64).reshape((3, self.case.NZ, self.case.NX))
        lib().mcgo_set_inlet_offset(self._h, _p(self._du))

    def set_densities(self, rr, rb):
        rr = np.ascontiguousarray(rr, np.float64).reshape(self.shape)
        rb = np.ascontiguousarray(rb, np.float64).reshape(self.shape)
        lib().mcgo_set_densities(self._h, _p(rr), _p(rb))

    def radial(self):
        nb = (self.case.NY, self.case.NX // 2)
        names = ["sum_uy", "sum_uy2", "sum_ur", "sum_ur2", "sum_uruy", "sum_phi", "sum_phi2"]
        out = {n: np.zeros(nb) for n in names}
        cnt = np.zeros(nb, np.uint64)
        lib().mcgo_get_radial(self._h, *[_p(out[n]) for n in names[:5]], cnt.ctypes.data_as(POINTER(c_uint64)),
                              _p(out["sum_phi"]), _p(out["sum_phi2"]))
        out["count"] = cnt
        return out

    def stats_info(self):
        st, ss, ns = c_int32(), c_int64(), c_int64()
        lib().mcgo_stats_info(self._h, ctypes.byref(st), ctypes.byref(ss), ctypes.byref(ns))
        return {"started": bool(st.value), "start_step": ss.value, "samples": ns.value}

    def tke(self):
        return lib().mcgo_tke(self._h)

    def global_sums(self):
        out = np.zeros(5)
        lib().mcgo_global_sums(self._h, _p(out))
        return out

    def writeset(self):
        cr = np.zeros((27,) + self.shape, np.uint8)
        cb = np.zeros((27,) + self.shape, np.uint8)
        u8 = lambda a: a.ctypes.data_as(POINTER(c_uint8))  # noqa: E731
        rc = lib().mcgo_writeset(self._h, u8(cr), u8(cb))
        if rc != 0:
            raise RuntimeError("writeset failed")
        return cr, cb

    def uy_dot(self, w):
        """sum_x w(x) u_y(x) over the grid (w shaped like the fields), u_y of the
        moments pass of the last step: a u_y functional without copying the fields."""
        w = np.ascontiguousarray(w, np.float64).reshape(self.shape)
        return lib().mcgo_uy_dot(self._h, _p(w))

    def branch_masks(self):
        """Per interior node (shape (NZ-2, NY-2, NX-2)) the branches the last step's
        collision took: bit 0 interface (I > 1e-4), bit 1 |F| > 1e-4 there."""
        out = np.zeros((self.case.NZ - 2, self.case.NY - 2, self.case.NX - 2), np.uint8)
        lib().mcgo_branch_masks(self._h, out.ctypes.data_as(POINTER(c_uint8)))
        return out

    def gradient(self, x, y, z):
        F = np.zeros(3)
        lib().mcgo_gradient(self._h, x, y, z, _p(F))
        return F

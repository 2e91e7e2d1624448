"""Python binding of the C ABI: ``Solver`` owns one ``mcg_solver`` handle.

Arrays returned to Python use the reference's grid layout, reshaped to
``(NZ, NY, NX)`` for cell fields (cell id = x + NX*(y + NY*z),
utilities/indexing.cuh:10-13) and ``(27, NZ, NY, NX)`` for populations
(population id = i*Ncells + id, utilities/indexing.cuh:15-19).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _abi
from ._abi import McgCase, as_ptr, check, load_library


class Solver:
    """One solver on the current HIP device.

    ``storage`` is ``"fp32"``, ``"fp16"`` or ``"bf16"``; ``kernel`` is
    ``"fused"`` (default) or ``"twopass"``.  ``stream`` is a raw hipStream_t
    (int) or None for the null stream.
    """

    def __init__(self, case: McgCase, storage: str = "fp32", kernel: str = "fused", stream: int | None = None):
        self.lib = load_library()
        self.case = case.copy()
        h = ctypes.c_void_p()
        check(self.lib.mcg_create(ctypes.byref(self.case), _abi.STORAGE[storage], ctypes.c_void_p(stream or 0),
                                  ctypes.byref(h)))
        self._h = h
        self.storage = storage
        self.set_kernel(kernel)

    # ---- lifetime
    def close(self) -> None:
        if getattr(self, "_h", None) is not None and self._h.value:
            self.lib.mcg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def shape(self):
        return (self.case.NZ, self.case.NY, self.case.NX)

    @property
    def ncells(self) -> int:
        return self.case.NX * self.case.NY * self.case.NZ

    def set_kernel(self, kernel: str) -> None:
        check(self.lib.mcg_set_kernel(self._h, _abi.KERNELS[kernel]))
        self.kernel = kernel

    # ---- case API (jet/jetIn.cuh, bubble/bubbleIn.cuh)
    def set_launch_segments(self, segments: int = 0) -> None:
        """0: time the candidate launch shapes of the step kernel on the first
        plain steps and keep the fastest (default); n > 0: about n uniform z
        segments, fixed; -k: the k-th banded schedule, fixed."""
        check(self.lib.mcg_set_launch_segments(self._h, int(segments)))

    def launch_info(self) -> dict:
        """The step kernel's launch shape: z segments (or z-bands of a banded
        schedule), workgroups per step, tuned or not, banded or uniform, the
        set_launch_segments argument that fixes it, and the statistics steps' shape."""
        seg, wg, tuned, banded, shape = (ctypes.c_int32() for _ in range(5))
        check(self.lib.mcg_launch_info(self._h, ctypes.byref(seg), ctypes.byref(wg), ctypes.byref(tuned),
                                       ctypes.byref(banded), ctypes.byref(shape)))
        sseg, swg, stuned = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        check(self.lib.mcg_stats_launch_info(self._h, ctypes.byref(sseg), ctypes.byref(swg), ctypes.byref(stuned)))
        return {"segments": seg.value, "workgroups": wg.value, "tuned": bool(tuned.value),
                "schedule": ("uniform", "banded", "banded-paired", "banded-mirrored")[banded.value],
                "shape": shape.value,
                "stats": {"segments": sseg.value, "workgroups": swg.value, "tuned": bool(stuned.value)}}

    def init_jet(self) -> None:
        check(self.lib.mcg_init_jet(self._h))

    def init_bubble(self) -> None:
        check(self.lib.mcg_init_bubble(self._h))

    def init_bubble_tanh(self) -> None:
        check(self.lib.mcg_init_bubble_tanh(self._h))

    def init_spheres(self, n: int = 256, seed: int = 7, rmin: float = 6.0, rmax: float = 14.0) -> None:
        """At rest with a field of n seeded tanh spheres (mcg_init_spheres): the interface-rich workload."""
        check(self.lib.mcg_init_spheres(self._h, int(n), int(seed), float(rmin), float(rmax)))

    # ---- hot path
    def step(self, n: int = 1) -> None:
        check(self.lib.mcg_step(self._h, int(n)))

    def synchronize(self) -> None:
        check(self.lib.mcg_synchronize(self._h))

    @property
    def step_count(self) -> int:
        v = ctypes.c_int64()
        check(self.lib.mcg_get_step(self._h, ctypes.byref(v)))
        return v.value

    def time_steps(self, n: int) -> float:
        ms = ctypes.c_double()
        check(self.lib.mcg_time_steps(self._h, int(n), ctypes.byref(ms)))
        return ms.value

    # ---- statistics (meanF/*)
    def stats_info(self) -> dict:
        st, ss, ns = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int64()
        check(self.lib.mcg_stats_info(self._h, ctypes.byref(st), ctypes.byref(ss), ctypes.byref(ns)))
        return {"started": bool(st.value), "start_step": ss.value, "samples": ns.value}

    def stats_reset(self) -> None:
        check(self.lib.mcg_stats_reset(self._h))

    def radial(self) -> dict:
        nb = (self.case.NY, self.case.NX // 2)
        names = ["sum_uy", "sum_uy2", "sum_ur", "sum_ur2", "sum_uruy", "sum_phi", "sum_phi2"]
        out = {n: np.zeros(nb, np.float64) for n in names}
        cnt = np.zeros(nb, np.uint64)
        d = ctypes.c_double
        check(self.lib.mcg_copy_radial(self._h, as_ptr(out["sum_uy"], d), as_ptr(out["sum_uy2"], d),
                                       as_ptr(out["sum_ur"], d), as_ptr(out["sum_ur2"], d),
                                       as_ptr(out["sum_uruy"], d), as_ptr(cnt, ctypes.c_uint64),
                                       as_ptr(out["sum_phi"], d), as_ptr(out["sum_phi2"], d)))
        out["count"] = cnt
        return out

    def tke(self) -> float:
        v = ctypes.c_double()
        check(self.lib.mcg_tke(self._h, ctypes.byref(v)))
        return v.value

    def global_sums(self) -> np.ndarray:
        out = np.zeros(5, np.float64)
        check(self.lib.mcg_global_sums(self._h, as_ptr(out, ctypes.c_double)))
        return out

    def check_health(self) -> None:
        check(self.lib.mcg_check_health(self._h))

    # ---- state exchange
    def fields(self) -> dict:
        arrs = [np.empty(self.shape, np.float32) for _ in range(5)]
        f = ctypes.c_float
        check(self.lib.mcg_copy_fields(self._h, *[as_ptr(a, f) for a in arrs]))
        return dict(zip(["rhor", "rhob", "ux", "uy", "uz"], arrs))

    def populations(self):
        fr = np.empty((27,) + self.shape, np.float32)
        fb = np.empty((27,) + self.shape, np.float32)
        f = ctypes.c_float
        check(self.lib.mcg_copy_populations(self._h, as_ptr(fr, f), as_ptr(fb, f)))
        return fr, fb

    def set_populations(self, fr: np.ndarray, fb: np.ndarray) -> None:
        fr = np.ascontiguousarray(fr, dtype=np.float32).reshape((27,) + self.shape)
        fb = np.ascontiguousarray(fb, dtype=np.float32).reshape((27,) + self.shape)
        f = ctypes.c_float
        check(self.lib.mcg_set_populations(self._h, as_ptr(fr, f), as_ptr(fb, f)))

    # ---- adjoint (an extension; periodic domain, FP32 populations)
    def adjoint_record(self, nsteps: int, checkpoint_every: int) -> None:
        """Run nsteps steps from the current state, checkpointing every checkpoint_every steps."""
        check(self.lib.mcg_adjoint_record(self._h, int(nsteps), int(checkpoint_every)))

    def adjoint_backward(self, lam_fr: np.ndarray, lam_fb: np.ndarray):
        """dJ/df at the final recorded state -> (dJ/df_r, dJ/df_b at the initial state, (dJ/dA, dJ/dbeta))."""
        lam_fr = np.ascontiguousarray(lam_fr, dtype=np.float32).reshape((27,) + self.shape)
        lam_fb = np.ascontiguousarray(lam_fb, dtype=np.float32).reshape((27,) + self.shape)
        out_r = np.empty((27,) + self.shape, np.float32)
        out_b = np.empty((27,) + self.shape, np.float32)
        dp = np.zeros(2, np.float64)
        f = ctypes.c_float
        check(self.lib.mcg_adjoint_backward(self._h, as_ptr(lam_fr, f), as_ptr(lam_fb, f), as_ptr(out_r, f),
                                            as_ptr(out_b, f), as_ptr(dp, ctypes.c_double)))
        return out_r, out_b, (float(dp[0]), float(dp[1]))

    def adjoint_set_uy_seed(self, w_uy, first: int = 0, last: int = -1) -> None:
        """Jet cases: the u_y functional sum over the moments passes of steps first..last
        (solver step numbers) of sum_x w_uy(x) u_y(x); steps accumulate it (uy_functional)
        and adjoint_backward adds it to the objective (w_uy shape (NZ, NY, NX); None clears it)."""
        if w_uy is None:
            check(self.lib.mcg_adjoint_set_uy_seed(self._h, None, 0, -1))
            return
        w = np.ascontiguousarray(w_uy, dtype=np.float32).reshape(self.shape)
        check(self.lib.mcg_adjoint_set_uy_seed(self._h, as_ptr(w, ctypes.c_float), int(first), int(last)))

    def adjoint_set_checkpoint_limit(self, max_states: int) -> None:
        """Cap the states the next adjoint_record keeps (0: one per segment); the sweep then
        recomputes the missing ones with a binomial (revolve) schedule, same gradients."""
        check(self.lib.mcg_adjoint_set_checkpoint_limit(self._h, int(max_states)))

    def adjoint_recomputed_steps(self) -> int:
        """Forward steps recomputed by the last adjoint_backward."""
        v = ctypes.c_int64()
        check(self.lib.mcg_adjoint_recomputed_steps(self._h, ctypes.byref(v)))
        return v.value

    def adjoint_gradients(self) -> dict:
        """FP64 parameter gradients of the last adjoint_backward sweep."""
        out = np.zeros(5, np.float64)
        check(self.lib.mcg_adjoint_gradients(self._h, as_ptr(out, ctypes.c_double), 5))
        return dict(zip(["dA", "dbeta", "dgravity", "dU", "damp"], out.tolist()))

    def uy_functional(self) -> float:
        """Value of the u_y functional accumulated since adjoint_set_uy_seed (FP64)."""
        v = ctypes.c_double()
        check(self.lib.mcg_uy_functional(self._h, ctypes.byref(v)))
        return v.value

    def set_inlet_offset(self, du) -> None:
        """Jet cases: offset added to the inlet velocity of every jet column, shape (3, NZ, NX)
        (u_x, u_y, u_z planes); None removes it."""
        if du is None:
            check(self.lib.mcg_set_inlet_offset(self._h, None))
            return
        d = np.ascontiguousarray(du, dtype=np.float32).reshape((3, self.case.NZ, self.case.NX))
        check(self.lib.mcg_set_inlet_offset(self._h, as_ptr(d, ctypes.c_float)))

    def adjoint_inlet_gradient(self):
        """Jet cases, after adjoint_backward: ((dJ/djet_velocity, dJ/dperturbation_amp),
        dJ/de per column, shape (3, NZ, NX))."""
        dp = np.zeros(2, np.float64)
        din = np.zeros((3, self.case.NZ, self.case.NX), np.float64)
        check(self.lib.mcg_adjoint_inlet_gradient(self._h, as_ptr(dp, ctypes.c_double), as_ptr(din, ctypes.c_double)))
        return (float(dp[0]), float(dp[1])), din

    def adjoint_last_ms(self) -> float:
        """Device time of the last adjoint_backward sweep (recomputation + adjoint kernels)."""
        ms = ctypes.c_double()
        check(self.lib.mcg_adjoint_last_ms(self._h, ctypes.byref(ms)))
        return ms.value

    def adjoint_release(self) -> None:
        check(self.lib.mcg_adjoint_release(self._h))

    # ---- output files (io/save_data.cuh, meanF/meanFieldsRT.cuh)
    def write_vti(self, path: str) -> None:
        check(self.lib.mcg_write_vti(self._h, os.fsencode(path)))

    def write_vti_step(self, vti_dir: str, step: int) -> None:
        check(self.lib.mcg_write_vti_step(self._h, os.fsencode(vti_dir), int(step)))

    def write_radial(self, mean_dir: str) -> None:
        check(self.lib.mcg_write_radial(self._h, os.fsencode(mean_dir)))

    # ---- checkpoint / restart
    def checkpoint(self) -> np.ndarray:
        n = ctypes.c_uint64()
        check(self.lib.mcg_checkpoint_size(self._h, ctypes.byref(n)))
        buf = np.empty(n.value, np.uint8)
        check(self.lib.mcg_checkpoint_save(self._h, buf.ctypes.data_as(ctypes.c_void_p), n.value))
        return buf

    def restore(self, buf: np.ndarray) -> None:
        buf = np.ascontiguousarray(buf, dtype=np.uint8)
        check(self.lib.mcg_checkpoint_load(self._h, buf.ctypes.data_as(ctypes.c_void_p), buf.size))

    def device_bytes(self) -> int:
        n = ctypes.c_uint64()
        check(self.lib.mcg_device_bytes(self._h, ctypes.byref(n)))
        return n.value


# The reference's case API names (jet/jetIn.cuh, bubble/bubbleIn.cuh).
def jetIn(solver: Solver) -> None:
    """Jet initial state: jetDensity + Injet (kernels.cu:29-56, main.cu:39-40)."""
    solver.init_jet()


def bubbleIn(solver: Solver) -> None:
    """Sharp bubble initial state: bubble (kernels.cu:13-27)."""
    solver.init_bubble()


def sphere_list(seed: int, n: int, NX: int, NY: int, NZ: int, rmin: float, rmax: float) -> np.ndarray:
    """Centres and radii (n, 4) of the spheres mcg_init_spheres places (host computation)."""
    out = np.empty((n, 4), np.float64)
    check(load_library().mcg_sphere_list(int(seed), int(n), int(NX), int(NY), int(NZ), float(rmin), float(rmax),
                                         as_ptr(out, ctypes.c_double)))
    return out

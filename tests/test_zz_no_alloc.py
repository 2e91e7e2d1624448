"""mcg_global_sums_noalloc (the entry point that allocates nothing) against
the allocating entry points it mirrors: the seven FP64 sums it leaves in
device memory equal mcg_tke, mcg_global_sums and the watchdog's count bit for
bit (the same kernels and summation order, only the partials live elsewhere).
Two calls in a row through the C ABI, the second on a smaller grid, into the
same output buffer; then the first solver again, after the smaller call left
its partials in the shared scratch array.
"""
import ctypes

import numpy as np
import pytest

import multicomponent_color_gradient_amd as m
from multicomponent_color_gradient_amd import _abi

pytestmark = pytest.mark.gpu


def hip(lib):
    """The HIP runtime the solver library is linked to, looked up through the
    library's own handle (dlsym searches its dependencies), so that the test's
    buffer lives in the same runtime instance as the solver's."""
    lib.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    lib.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    lib.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    lib.hipFree.argtypes = [ctypes.c_void_p]
    for f in (lib.hipMalloc, lib.hipMemset, lib.hipMemcpy, lib.hipFree):
        f.restype = ctypes.c_int
    return lib


def sums_noalloc(lib, s, out):
    assert lib.hipMemset(out, 0xFF, 7 * 8) == 0  # NaN in every slot
    rc = lib.mcg_global_sums_noalloc(s._h, out)
    assert rc == 0, lib.mcg_last_error().decode()
    s.synchronize()
    host = np.empty(7, np.float64)
    assert lib.hipMemcpy(host.ctypes.data_as(ctypes.c_void_p), out, 7 * 8, 2) == 0  # device to host
    return host


def expected(s):
    return np.concatenate([[s.tke()], s.global_sums(), [0.0]])


def test_noalloc_sums_equal_the_allocating_entry_points():
    lib = hip(m.load_library())
    out = ctypes.c_void_p()
    big = m.perturbed_jet_case(40, 56, 44, jet_radius=6.0)
    small = m.perturbed_jet_case(24, 20, 24, jet_radius=4.0)
    solvers = []
    try:
        for case, steps in ((big, 30), (small, 5)):
            print(f"no-alloc sums: {case.NX}x{case.NY}x{case.NZ} jet, {steps} steps", flush=True)
            s = m.Solver(case)
            solvers.append(s)
            if not out.value:
                assert lib.hipMalloc(ctypes.byref(out), 7 * 8) == 0
            s.init_jet()
            s.step(steps)
            got = sums_noalloc(lib, s, out)
            ref = expected(s)
            assert np.array_equal(got, ref), (got, ref)
            assert got[1] > 0 and got[6] == 0
        # the larger solver again, after the smaller call's partials
        s = solvers[0]
        before = sums_noalloc(lib, s, out)
        s.step(3)
        got = sums_noalloc(lib, s, out)
        assert np.array_equal(got, expected(s))
        assert not np.array_equal(got, before)
    finally:
        if out.value:
            lib.hipFree(out)
        for s in solvers:
            s.close()


def test_noalloc_sums_reject_a_null_output():
    lib = m.load_library()
    s = m.Solver(m.jet_case(20, 16, 20, jet_radius=3.0))
    s.init_jet()
    assert lib.mcg_global_sums_noalloc(s._h, None) == _abi.MCG_E_INVALID
    s.close()

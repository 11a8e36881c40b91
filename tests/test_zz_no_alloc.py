"""The allocation-free FP64 dot product fwi2d_dot_f64_noalloc against fwi2d_engine_dot_f64.

The no-allocation variant keeps its block partials and result in the library's static
device array, resolved per device with hipGetSymbolAddress; it runs the same kernels with
the same block count as the engine's dot product, so the two agree bit for bit.  Two
calls in a row, the second on a smaller input, check that nothing of the first call's
scratch contents leaks into the second.  The test's own device buffers come from the HIP
runtime the library itself uses (hipMalloc through ctypes), so no second runtime is
loaded into the process."""
import ctypes as C

import numpy as np
import pytest

from fwi2d_amd import workloads as W

pytestmark = pytest.mark.gpu


class DeviceArrays:
    """float32 device copies of host arrays, freed on close."""

    def __init__(self):
        # the runtime already loaded by the library (same soname)
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a, np.float32)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(a.nbytes, 4)) == 0
        self.ptrs.append(p)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p.value

    def close(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def test_noalloc_dot_matches_engine_dot():
    import fwi2d_amd
    lib = fwi2d_amd.load()
    prob = W.base_problem(24, 20, 0.01, 0.01, 11, 0.001, [[0.1, 0.02]], [np.array([[0.1, 0.02]])], taper=0)
    eng = fwi2d_amd.Engine(fwi2d_amd.grid_from(prob, 10, 0.001, 0))
    dev = DeviceArrays()
    rng = np.random.default_rng(0)
    for n in (3_000_001, 777):
        host = [rng.standard_normal(n).astype(np.float32) for _ in range(3)]
        a, b, w = [dev.put(x) for x in host]
        ref64 = float(np.sum(host[0].astype(np.float64) * host[1] * host[2]))
        for bp, wp, expect in ((b, w, ref64), (None, None, float(np.sum(host[0].astype(np.float64) ** 2)))):
            out = C.c_double(np.nan)
            rc = lib.fwi2d_dot_f64_noalloc(0, C.c_void_p(a), None if bp is None else C.c_void_p(bp),
                                           None if wp is None else C.c_void_p(wp), n, C.byref(out))
            assert rc == 0, lib.fwi2d_last_error().decode()
            e = eng.dot_f64(a, bp or 0, wp or 0, n)
            print("n %d: noalloc %.17g engine %.17g float64 sum %.17g" % (n, out.value, e, expect))
            assert out.value == e
            assert abs(out.value - expect) <= 1e-12 * np.sum(np.abs(host[0].astype(np.float64)) * 10)
    # an empty input sums to zero; out-of-range sizes, devices and null pointers are errors
    out = C.c_double(np.nan)
    assert lib.fwi2d_dot_f64_noalloc(0, C.c_void_p(a), None, None, 0, C.byref(out)) == 0
    assert out.value == 0.0
    assert lib.fwi2d_dot_f64_noalloc(0, C.c_void_p(a), None, None, -1, C.byref(out)) != 0
    assert "outside" in lib.fwi2d_last_error().decode()
    assert lib.fwi2d_dot_f64_noalloc(0, C.c_void_p(a), None, None, (1 << 40) + 1, C.byref(out)) != 0
    assert lib.fwi2d_dot_f64_noalloc(999, C.c_void_p(a), None, None, 5, C.byref(out)) != 0
    assert lib.fwi2d_dot_f64_noalloc(0, None, None, None, 5, C.byref(out)) != 0
    dev.close()
    eng.close()

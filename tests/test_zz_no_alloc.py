"""The allocation-free boundary current (cicy_boundary_current_f64_noalloc) against the
allocating entry point (cicy_boundary_current_f64) and the oracle, through the C-ABI with
every array in device memory: two calls in a row, the second on a smaller input, then a
refused call (a node outside the voltage vector) and a good call after it."""
import ctypes
import os

import numpy as np
import pytest

import oracle
from cicy_geom_amd import dea, network

pytestmark = pytest.mark.gpu


def hip_runtime():
    """The HIP runtime libcicy_dea.so is linked against (already loaded by it)."""
    dea.lib()
    with open("/proc/self/maps") as f:
        for line in f:
            path = line.split()[-1]
            if "libamdhip64.so" in os.path.basename(path):
                hip = ctypes.CDLL(path)
                hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
                hip.hipFree.argtypes = [ctypes.c_void_p]
                hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                          ctypes.c_int]
                hip.hipDeviceSynchronize.argtypes = []
                return hip
    raise RuntimeError("libamdhip64 is not loaded")


H2D, D2H = 1, 2


class DevArrays:
    """Device copies of host arrays (hipMalloc through the test, not the library)."""

    def __init__(self, hip):
        self.hip, self.ptrs = hip, []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
        if a.nbytes:
            assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        self.ptrs.append(p)
        return p.value

    def get_double(self, ptr):
        out = np.zeros(1)
        assert self.hip.hipMemcpy(out.ctypes.data, ctypes.c_void_p(ptr), 8, D2H) == 0
        return out[0]

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []


def test_noalloc_boundary_current_matches_allocating_entry_point():
    ctx = dea.Context(0)
    hip = hip_runtime()
    dev = DevArrays(hip)
    try:
        s = network.largest_percolating_system(network.generate(1), 0)
        v = np.random.default_rng(5).uniform(0.0, 1.0, s.nodes)
        nb, ni, re = s.boundary_elements[0]
        # the second call: a smaller input (the first 7 elements of the other boundary)
        nb2, ni2, re2 = (a[:7] for a in s.boundary_elements[1])
        out = dev.put(np.zeros(1))
        dv = dev.put(v)
        for (b, i, r) in ((nb, ni, re), (nb2, ni2, re2)):
            want = ctx.boundary_current(b, i, r, v)
            ctx.boundary_current_noalloc(len(b), dev.put(b.astype(np.int64)),
                                         dev.put(i.astype(np.int64)), dev.put(r), dv, len(v),
                                         out)
            got = dev.get_double(out)
            print(f"elements {len(b)}: noalloc {got!r} allocating {want!r}")
            assert got == want           # same kernel, same grid: the same bits
            ref = oracle.boundary_current(b, i, r, v)
            assert abs(got - ref) <= 1e-13 * np.sum(np.abs(v[i] - v[b]) / r)
        # a node outside [0, nv) is refused, and the next call is unaffected
        bad = ni.astype(np.int64).copy()
        bad[3] = len(v) + 5
        with pytest.raises(dea.DEAError, match="outside"):
            ctx.boundary_current_noalloc(len(nb), dev.put(nb.astype(np.int64)), dev.put(bad),
                                         dev.put(re), dv, len(v), out)
        ctx.boundary_current_noalloc(len(nb2), dev.put(nb2.astype(np.int64)),
                                     dev.put(ni2.astype(np.int64)), dev.put(re2), dv, len(v),
                                     out)
        assert dev.get_double(out) == ctx.boundary_current(nb2, ni2, re2, v)
        # sizes beyond the maximum are refused before any launch
        with pytest.raises(dea.DEAError, match="sizes"):
            ctx.boundary_current_noalloc(1 << 41, dv, dv, dv, dv, len(v), out)
        # no elements: the current is zero
        ctx.boundary_current_noalloc(0, None, None, None, None, len(v), out)
        assert dev.get_double(out) == 0.0
    finally:
        hip.hipDeviceSynchronize()
        dev.free()
        ctx.close()

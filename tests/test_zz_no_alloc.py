"""dspt_trace_closest_noalloc (the allocation-free ray query) against
dspt_trace_closest on the same rays, through the C ABI.

Device buffers for the caller's side come from the HIP runtime libdspt.so
links (the entry point itself allocates nothing). Two calls in a row, the
second with fewer rays, must each give exactly the hits, t, u, v and work
counters of the allocating entry point; a grid beyond the scratch is refused.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libdspt.so is linked to."""

    def __init__(self):
        import dspt._lib as L
        L.lib()
        self.h = ctypes.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.h.hipFree.argtypes = [ctypes.c_void_p]
        self.h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.h.hipDeviceSynchronize.argtypes = []
        self.live = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), nbytes) == 0
        self.live.append(p.value)
        return p.value

    def to_device(self, a):
        a = np.ascontiguousarray(a)
        d = self.alloc(a.nbytes)
        assert self.h.hipMemcpy(d, a.ctypes.data, a.nbytes, 1) == 0  # hipMemcpyHostToDevice
        return d

    def to_host(self, d, shape, dtype):
        a = np.empty(shape, dtype)
        assert self.h.hipMemcpy(a.ctypes.data, d, a.nbytes, 2) == 0  # hipMemcpyDeviceToHost
        return a

    def free_all(self):
        for p in self.live:
            self.h.hipFree(p)
        self.live = []


def test_noalloc_matches_trace_closest(sphere_field_small):
    from dspt import Engine, DsptError
    from dspt._lib import check, lib
    from test_gpu_parity import _surface_rays, _upload
    wl = sphere_field_small
    eng = Engine(0)
    _upload(eng, wl)
    hip = _Hip()
    L = lib()
    try:
        for n, seed in ((30000, 41), (7001, 42)):
            print(f"noalloc: {n} rays, seed {seed}", flush=True)
            rays = _surface_rays(wl, n, seed)
            eng.reset_counters()
            tri, t, u, v = eng.trace_closest(rays)
            c_ref = eng.counters()
            d_rays = hip.to_device(rays)
            d_tri, d_t, d_u, d_v = (hip.alloc(n * 4) for _ in range(4))
            eng.reset_counters()
            check(L.dspt_trace_closest_noalloc(eng.h, d_rays, n, d_tri, d_t, d_u, d_v, 0))
            eng.synchronize()
            c_new = eng.counters()
            g_tri = hip.to_host(d_tri, n, np.int32)
            np.testing.assert_array_equal(g_tri, tri)
            for d, ref in ((d_t, t), (d_u, u), (d_v, v)):
                g = hip.to_host(d, n, np.float32)
                hit = tri >= 0
                np.testing.assert_array_equal(g[hit].view(np.uint32), ref[hit].view(np.uint32))
            assert c_new["node_visits"] == c_ref["node_visits"] and c_new["tri_tests"] == c_ref["tri_tests"]
        # a grid beyond the scratch is refused with a clear error
        with pytest.raises(DsptError) as e:
            check(L.dspt_trace_closest_noalloc(eng.h, d_rays, n, d_tri, d_t, d_u, d_v, 513))
        assert e.value.code == -1 and "scratch" in str(e.value)
    finally:
        hip.free_all()
        eng.close()

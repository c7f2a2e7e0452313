"""wd_sort_noalloc, the sort for callers that cannot allocate, against wd_sort (the entry point it is a
variant of) and the oracle: order, count and cov2d bit for bit equal to wd_sort's, on calls in a row
of falling size, the first at its limit (its intermediates share one array per device, which keeps the last call's
contents), and a refusal past its limit. The test's own device buffers come from the HIP runtime the
library itself links (libamdhip64 of /opt/rocm), through ctypes."""
import ctypes as C
import os

import numpy as np
import pytest

import oracle
import winding_amd as wa
from winding_amd import workloads as wl

pytestmark = pytest.mark.gpu


def hip():
    L = C.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipFree.argtypes = [C.c_void_p]
    L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return L


H2D, D2H = 1, 2


class DeviceBuffer:
    def __init__(self, L, nbytes):
        self.L, self.nbytes = L, nbytes
        self.ptr = C.c_void_p()
        assert L.hipMalloc(C.byref(self.ptr), nbytes) == 0
        self.addr = self.ptr.value

    def put(self, array):
        a = np.ascontiguousarray(array)
        assert a.nbytes <= self.nbytes
        assert self.L.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, dtype, count):
        out = np.empty(count, dtype)
        assert self.L.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H) == 0
        return out

    def free(self):
        if self.ptr:
            self.L.hipFree(self.ptr)
            self.ptr = None


def test_noalloc_equals_wd_sort_on_two_calls(ctx):
    L = hip()
    # the largest size it takes (its scratch's bound; tiles of 2048), then smaller ones on the same
    # scratch: tiles of 2048 again and tiles of 1024
    cases = [(0, 11, wa.WD_NOALLOC_MAX_SPLATS, (1.0, 1.0, 11.0), 1920, 1080),
             (1, 5, 1_500_000, (0.0, 2.0, 12.0), 1920, 1080), (0, 9, 70_001, (3.0, -1.0, 8.0), 1280, 720)]
    cap = max(c[2] for c in cases)
    records = DeviceBuffer(L, 32 * cap)
    order = DeviceBuffer(L, 4 * cap)
    count = DeviceBuffer(L, 4)
    cov = DeviceBuffer(L, 16 * cap)
    state = wa.State(ctx, cap)
    try:
        for kind, seed, n, eye, w, h in cases:
            recs = wa.synth_records(kind, seed, n)
            view = wl.view_for(recs, eye, w, h)
            records.put(recs)
            wa.sort_noalloc(ctx, records.addr, n, view, order.addr, count.addr, cov.addr)
            ctx.sync()
            v = int(count.get(np.uint32, 1)[0])
            got_order = order.get(np.uint32, v)
            got_cov = cov.get(np.float32, 4 * n).reshape(n, 4)
            cloud = wa.Cloud(ctx, recs)
            state.sort(cloud, view)
            want_order, want_cov, stats = state.read(n_cov=n)
            cloud.destroy()
            print("n=%d: V %d (wd_sort %d)" % (n, v, stats.count), flush=True)
            assert v == stats.count
            assert np.array_equal(got_order, want_order)
            assert np.array_equal(got_cov.view(np.uint32), want_cov.view(np.uint32))
            o_order, _, _, _ = oracle.run(recs, view)
            assert np.array_equal(got_order, o_order)
        with pytest.raises(wa.WindingError, match="without allocation takes 1 to") as e:
            wa.sort_noalloc(ctx, records.addr, wa.WD_NOALLOC_MAX_SPLATS + 1, view, order.addr, count.addr, cov.addr)
        assert e.value.code == -3
    finally:
        state.destroy()
        for b in (records, order, count, cov):
            b.free()

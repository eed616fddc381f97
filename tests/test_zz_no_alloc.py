"""The allocation-free brute force (nnm_bruteforce_knn_noalloc): device
buffers from the caller, partial lists in the library's fixed device scratch
array.  Its results must equal nnm_bruteforce_knn's bit for bit, on two calls
in a row (the second smaller, so it runs over what the first left in the
scratch array), and a request above the scratch limit is refused.

The caller's device buffers come from the HIP runtime the library itself is
linked against (libamdhip64.so.7), through ctypes."""
import ctypes as C

import numpy as np
import pytest

import nnm
import nnm_data

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class _Hip:
    def __init__(self):
        nnm.lib()  # loads the library and with it its HIP runtime
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.bufs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), nbytes) == 0
        self.bufs.append(p)
        return p.value

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.h.hipMemcpy(C.c_void_p(p), a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0
        return p

    def get(self, p, shape, dtype):
        out = np.empty(shape, dtype)
        assert self.h.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(p), out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def test_noalloc_bruteforce_matches_allocating_entry_point():
    X = nnm_data.sift_like(30000, 96, centers=24, sigma=20.0, seed=1234)
    rng = np.random.default_rng(5)
    hip = _Hip()
    try:
        with nnm.Context(0) as ctx:
            ctx.set_data(X)
            for nq, k in ((5000, 20), (700, 10)):  # the second call is smaller
                rows = np.sort(rng.choice(X.shape[0], nq, replace=False)).astype(np.uint32)
                ref_i, ref_d = ctx.bruteforce(k, rows=rows)
                r = hip.put(rows)
                di, dd = hip.alloc(nq * k * 4), hip.alloc(nq * k * 4)
                ctx.bruteforce_noalloc(k, r, nq, di, dd)
                ids, dist = hip.get(di, (nq, k), np.uint32), hip.get(dd, (nq, k), np.float32)
                assert np.array_equal(ids, ref_i), f"ids differ at nq={nq}"
                assert np.array_equal(dist.view(np.uint32), ref_d.view(np.uint32)), f"distances differ at nq={nq}"
            # above the scratch limit: refused before any launch
            big = hip.put(np.zeros(70000, np.uint32))
            oi, od = hip.alloc(70000 * 4 * 4), hip.alloc(70000 * 4 * 4)
            with pytest.raises(nnm.NNMError):
                ctx.bruteforce_noalloc(4, big, 70000, oi, od)
    finally:
        hip.free()

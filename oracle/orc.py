"""ctypes binding of the CPU oracle (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker.  The product library never
imports it.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")
REF_LIB_PATH = os.path.join(HERE, "_ref", "libref_dist.so")

RAW = 0xFFFFFFFF


class OrcStats(C.Structure):
    _fields_ = [("iters", C.c_uint32), ("converged", C.c_uint32),
                ("dists", C.c_uint64), ("init_dists", C.c_uint64),
                ("secs", C.c_double), ("last_delta", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_lib = None
_ref = None


def build():
    """Compile liborc.so (and _ref/ when the reference sources are present)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        f32p, u32p, u8p, f64p = (C.POINTER(C.c_float), C.POINTER(C.c_uint32),
                                 C.POINTER(C.c_uint8), C.POINTER(C.c_double))
        L.orc_l2_seq.restype = C.c_float
        L.orc_l2_seq.argtypes = [f32p, f32p, C.c_uint32]
        L.orc_l2_simd4.restype = C.c_float
        L.orc_l2_simd4.argtypes = [f32p, f32p, C.c_uint32]
        L.orc_l2_f64.restype = C.c_double
        L.orc_l2_f64.argtypes = [f32p, f32p, C.c_uint32]
        L.orc_bruteforce_f64.restype = None
        L.orc_bruteforce_f64.argtypes = [f32p, C.c_uint64, C.c_uint32, f32p, u32p, C.c_uint64,
                                         C.c_uint32, C.c_int, u32p, f64p]
        L.orc_update_knn_list.restype = C.c_uint32
        L.orc_update_knn_list.argtypes = [u32p, f32p, u8p, C.c_uint32, C.c_uint32, C.c_float]
        L.orc_get_rnd_sds.restype = None
        L.orc_get_rnd_sds.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, u32p]
        L.orc_kgraph.restype = None
        L.orc_kgraph.argtypes = [f32p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, C.c_uint32,
                                 C.c_uint32, C.c_float, C.c_uint32, u32p, f32p, C.POINTER(OrcStats)]
        L.orc_merge.restype = None
        L.orc_merge.argtypes = [f32p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, u32p, u8p,
                                u32p, C.c_uint32, C.c_float, C.c_uint32, C.c_uint32, C.c_float,
                                C.c_uint32, C.c_int, u32p, f32p, C.POINTER(OrcStats)]
        L.orc_recall.restype = C.c_double
        L.orc_recall.argtypes = [u32p, C.c_uint32, u32p, C.c_uint64, u32p, C.c_uint32,
                                 C.c_uint32, C.c_int]
        _lib = L
    return _lib


def ref_lib():
    """The reference's own distance routines, or None when not built."""
    global _ref
    if _ref is None and os.path.exists(REF_LIB_PATH):
        R = C.CDLL(REF_LIB_PATH)
        f32p = C.POINTER(C.c_float)
        R.ref_l2norm.restype = C.c_float
        R.ref_l2norm.argtypes = [f32p, C.c_ulong, f32p, C.c_ulong, C.c_uint]
        R.ref_l2sqr_simd_ext.restype = C.c_float
        R.ref_l2sqr_simd_ext.argtypes = [f32p, C.c_uint, f32p, C.c_uint, C.c_uint]
        _ref = R
    return _ref


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def l2_seq(a, b):
    a, b = _f32(a), _f32(b)
    return lib().orc_l2_seq(_p(a, C.c_float), _p(b, C.c_float), a.size)


def l2_simd4(a, b):
    a, b = _f32(a), _f32(b)
    return lib().orc_l2_simd4(_p(a, C.c_float), _p(b, C.c_float), a.size)


def l2_f64(a, b):
    a, b = _f32(a), _f32(b)
    return lib().orc_l2_f64(_p(a, C.c_float), _p(b, C.c_float), a.size)


def bruteforce(X, k, rows=None, Q=None, exclude_self=True):
    """Exact FP64 k-NN of rows of X (or of query vectors Q) against X."""
    X = _f32(X)
    n, d = X.shape
    if Q is not None:
        Q = _f32(Q)
        nq = Q.shape[0]
        qr = np.zeros(1, np.uint32)
        qp = _p(Q, C.c_float)
    else:
        qr = np.ascontiguousarray(np.arange(n) if rows is None else rows, dtype=np.uint32)
        nq = qr.size
        qp = None
    ids = np.empty((nq, k), np.uint32)
    dist = np.empty((nq, k), np.float64)
    lib().orc_bruteforce_f64(_p(X, C.c_float), n, d, qp, _p(qr, C.c_uint32), nq, k,
                             1 if exclude_self else 0, _p(ids, C.c_uint32), _p(dist, C.c_double))
    return ids, dist


def update_knn_list(ids, dists, flags, new_id, new_dist):
    ids = np.ascontiguousarray(ids, np.uint32).copy()
    dists = np.ascontiguousarray(dists, np.float32).copy()
    flags = np.ascontiguousarray(flags, np.uint8).copy()
    r = lib().orc_update_knn_list(_p(ids, C.c_uint32), _p(dists, C.c_float), _p(flags, C.c_uint8),
                                  ids.size, new_id, new_dist)
    return r, ids, dists, flags


def get_rnd_sds(seed, k, N):
    out = np.empty(k, np.uint32)
    lib().orc_get_rnd_sds(seed, k, N, _p(out, C.c_uint32))
    return out


def kgraph(X, K=20, rho=1.0, rev_cap=100, max_iters=0, delta=0.002, seed=1):
    X = _f32(X)
    n, d = X.shape
    ids = np.empty((n, K), np.uint32)
    dist = np.empty((n, K), np.float32)
    st = OrcStats()
    lib().orc_kgraph(_p(X, C.c_float), n, d, K, rho, rev_cap, max_iters, delta, seed,
                     _p(ids, C.c_uint32), _p(dist, C.c_float), C.byref(st))
    return ids, dist, st.as_dict()


def merge(X, sub_begin, sub_built, init_ids, K=20, cut=None, rho=1.0, rev_cap=100,
          max_iters=0, delta=0.002, seed=1, dedup=True):
    """Partitioned merge (Smerge / Jmerge generalised to P subgraphs)."""
    X = _f32(X)
    n, d = X.shape
    sb = np.ascontiguousarray(sub_begin, np.uint32)
    bu = np.ascontiguousarray(sub_built, np.uint8)
    ii = np.ascontiguousarray(init_ids if init_ids is not None else np.zeros((n, K)), np.uint32)
    cut = K // 2 if cut is None else cut
    ids = np.empty((n, K), np.uint32)
    dist = np.empty((n, K), np.float32)
    st = OrcStats()
    lib().orc_merge(_p(X, C.c_float), n, d, K, bu.size, _p(sb, C.c_uint32), _p(bu, C.c_uint8),
                    _p(ii, C.c_uint32), cut, rho, rev_cap, max_iters, delta, seed,
                    1 if dedup else 0, _p(ids, C.c_uint32), _p(dist, C.c_float), C.byref(st))
    return ids, dist, st.as_dict()


def recall(ids, rows, gt, k=10, unique=True):
    ids = np.ascontiguousarray(ids, np.uint32)
    rows = np.ascontiguousarray(rows, np.uint32)
    gt = np.ascontiguousarray(gt, np.uint32)
    return lib().orc_recall(_p(ids, C.c_uint32), ids.shape[1], _p(rows, C.c_uint32), rows.size,
                            _p(gt, C.c_uint32), gt.shape[1], k, 1 if unique else 0)

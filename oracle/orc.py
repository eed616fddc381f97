"""ctypes binding of the CPU oracle (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, to check results.  The product library never
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
        for nm in ("orc_l1", "orc_chi2", "orc_cosine"):
            getattr(L, nm).restype = C.c_float
            getattr(L, nm).argtypes = [f32p, f32p, C.c_uint32]
        for nm in ("orc_l1_f64", "orc_chi2_f64", "orc_cosine_f64"):
            getattr(L, nm).restype = C.c_double
            getattr(L, nm).argtypes = [f32p, f32p, C.c_uint32]
        L.orc_set_metric.restype = None
        L.orc_set_metric.argtypes = [C.c_int]
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
        L.orc_merge_mt.restype = C.c_int
        L.orc_merge_mt.argtypes = L.orc_merge.argtypes + [C.c_int]
        L.orc_merge2one.restype = None
        L.orc_merge2one.argtypes = [C.c_uint32, C.c_uint32, u32p, f32p, C.c_uint32, u32p, f32p, C.c_int]
        L.orc_nndescent_lists.restype = None
        L.orc_nndescent_lists.argtypes = [f32p, C.c_uint32, C.c_uint32, C.c_uint32, u32p, f32p, u8p, u32p,
                                          C.c_int, C.c_float, C.c_uint32, C.c_uint32, C.c_float, C.c_uint32,
                                          C.POINTER(OrcStats)]
        vp, u64p = C.c_void_p, C.POINTER(C.c_uint64)
        L.orc_search_layered.restype = None
        L.orc_search_layered.argtypes = [f32p, C.c_uint32, C.c_uint32, C.c_uint32, C.POINTER(vp), C.POINTER(vp),
                                         f32p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, u32p, f32p, u64p]
        L.orc_search_flat.restype = None
        L.orc_search_flat.argtypes = [f32p, C.c_uint32, C.c_uint32, u64p, u32p, f32p, C.c_uint32, u32p,
                                      C.c_uint32, C.c_uint32, u32p, f32p, u64p]
        for nm in ("orc_dpg_ref", "orc_dpg_par"):
            f = getattr(L, nm)
            f.restype = None
            f.argtypes = [f32p, C.c_uint32, C.c_uint32, u32p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                          vp, vp, vp, u64p, u32p, f32p]
        L.orc_recall.restype = C.c_double
        L.orc_recall.argtypes = [u32p, C.c_uint32, u32p, C.c_uint64, u32p, C.c_uint32,
                                 C.c_uint32, C.c_int]
        _lib = L
    return _lib


def ref_built():
    """True when the reference's distance routines are built (oracle/_ref);
    checks the file only, without loading it."""
    return os.path.exists(REF_LIB_PATH)


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
        for nm in ("ref_l1norm", "ref_ksquare"):
            getattr(R, nm).restype = C.c_float
            getattr(R, nm).argtypes = [f32p, C.c_ulong, f32p, C.c_ulong, C.c_uint]
        R.ref_cosine.restype = C.c_float
        R.ref_cosine.argtypes = [f32p, C.c_uint, f32p, C.c_uint, C.c_uint]
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


METRIC_L2, METRIC_L1, METRIC_COSINE, METRIC_CHI2 = 0, 1, 2, 3
_METRIC_FN = {METRIC_L2: "orc_l2_seq", METRIC_L1: "orc_l1", METRIC_COSINE: "orc_cosine", METRIC_CHI2: "orc_chi2"}


def dist32(a, b, metric=METRIC_L2):
    """The oracle's FP32 distance of two vectors under a metric."""
    a, b = _f32(a), _f32(b)
    return getattr(lib(), _METRIC_FN[metric])(_p(a, C.c_float), _p(b, C.c_float), a.size)


def dist64(a, b, metric=METRIC_L2):
    a, b = _f32(a), _f32(b)
    fn = {METRIC_L2: "orc_l2_f64", METRIC_L1: "orc_l1_f64", METRIC_COSINE: "orc_cosine_f64",
          METRIC_CHI2: "orc_chi2_f64"}[metric]
    return getattr(lib(), fn)(_p(a, C.c_float), _p(b, C.c_float), a.size)


class metric_scope:
    """Sets the oracle's metric for the graph routines and brute force
    inside a with-block, back to L2 afterwards."""

    def __init__(self, metric):
        self.metric = metric

    def __enter__(self):
        lib().orc_set_metric(int(self.metric))

    def __exit__(self, *exc):
        lib().orc_set_metric(METRIC_L2)


def bruteforce(X, k, rows=None, Q=None, exclude_self=True):
    """Exact FP64 k-NN of rows of X (or of query vectors Q) against X (under
    the metric of an enclosing metric_scope)."""
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


def merge2one(pool_ids, pool_dist, cpn_ids, cpn_dist, dedup=True):
    """merge2one per row: the first K of the two-way merge of the pool rows
    (n x K) and the parked rows (n x cut, an id of RAW ends a row)."""
    ids = np.ascontiguousarray(pool_ids, np.uint32).copy()
    dist = np.ascontiguousarray(pool_dist, np.float32).copy()
    ci = np.ascontiguousarray(cpn_ids, np.uint32)
    cd = np.ascontiguousarray(cpn_dist, np.float32)
    n, K = ids.shape
    lib().orc_merge2one(n, K, _p(ids, C.c_uint32), _p(dist, C.c_float), ci.shape[1],
                        _p(ci, C.c_uint32), _p(cd, C.c_float), 1 if dedup else 0)
    return ids, dist


def nndescent_lists(X, ids, flags, gid_of, rule_all=False, rho=1.0, rev_cap=100, max_iters=0,
                    delta=0.002, seed=1):
    """The NN-Descent loop from given lists (n x K ids and flags) and row
    labels (RAW for raw rows); returns ids, dists, flags, stats."""
    X = _f32(X)
    n, d = X.shape
    ids = np.ascontiguousarray(ids, np.uint32).copy()
    K = ids.shape[1]
    flags = np.ascontiguousarray(flags, np.uint8).copy()
    gid = np.ascontiguousarray(gid_of, np.uint32)
    dist = np.empty((n, K), np.float32)
    st = OrcStats()
    lib().orc_nndescent_lists(_p(X, C.c_float), n, d, K, _p(ids, C.c_uint32), _p(dist, C.c_float),
                              _p(flags, C.c_uint8), _p(gid, C.c_uint32), 1 if rule_all else 0, rho,
                              rev_cap, max_iters, delta, seed, C.byref(st))
    return ids, dist, flags, st.as_dict()


def merge_mt(X, sub_begin, sub_built, init_ids, K=20, cut=None, rho=1.0, rev_cap=100,
             max_iters=0, delta=0.002, seed=1, dedup=True, threads=0):
    """merge on every host core (threads = 0) or on `threads` threads: the
    reference's OpenMP form, one lock per list.  Returns ids, dists, stats
    (with "threads")."""
    X = _f32(X)
    n, d = X.shape
    sb = np.ascontiguousarray(sub_begin, np.uint32)
    bu = np.ascontiguousarray(sub_built, np.uint8)
    ii = np.ascontiguousarray(init_ids, np.uint32)
    cut = K // 2 if cut is None else cut
    ids = np.empty((n, K), np.uint32)
    dist = np.empty((n, K), np.float32)
    st = OrcStats()
    nt = lib().orc_merge_mt(_p(X, C.c_float), n, d, K, bu.size, _p(sb, C.c_uint32), _p(bu, C.c_uint8),
                            _p(ii, C.c_uint32), cut, rho, rev_cap, max_iters, delta, seed,
                            1 if dedup else 0, _p(ids, C.c_uint32), _p(dist, C.c_float), C.byref(st), threads)
    out = st.as_dict()
    out["threads"] = nt
    return ids, dist, out


def recall(ids, rows, gt, k=10, unique=True):
    ids = np.ascontiguousarray(ids, np.uint32)
    rows = np.ascontiguousarray(rows, np.uint32)
    gt = np.ascontiguousarray(gt, np.uint32)
    return lib().orc_recall(_p(ids, C.c_uint32), ids.shape[1], _p(rows, C.c_uint32), rows.size,
                            _p(gt, C.c_uint32), gt.shape[1], k, 1 if unique else 0)


def _csr64(offs):
    return np.ascontiguousarray(offs, np.uint64)


def search_layered(X, layers, Q, entry, beam, k=10):
    """LySearch::knnSearch restated: layers = [(offs, adj), ...], upper first.
    Returns ids, dist (nq x k) and the number of distances computed."""
    X = _f32(X)
    Q = _f32(Q)
    n, d = X.shape
    offs = [_csr64(o) for o, _ in layers]
    adj = [np.ascontiguousarray(a, np.uint32) if len(a) else np.zeros(1, np.uint32) for _, a in layers]
    po = (C.c_void_p * len(offs))(*[o.ctypes.data for o in offs])
    pa = (C.c_void_p * len(adj))(*[a.ctypes.data for a in adj])
    nq = Q.shape[0]
    ids = np.empty((nq, k), np.uint32)
    dist = np.empty((nq, k), np.float32)
    nc = C.c_uint64()
    lib().orc_search_layered(_p(X, C.c_float), n, d, len(layers), po, pa, _p(Q, C.c_float), nq, entry, beam, k,
                             _p(ids, C.c_uint32), _p(dist, C.c_float), C.byref(nc))
    return ids, dist, nc.value


def search_flat(X, offs, adj, Q, seeds, beam, k=10):
    """LySearch::nndFlatSrch restated, seeds: nq x beam."""
    X = _f32(X)
    Q = _f32(Q)
    n, d = X.shape
    o = _csr64(offs)
    a = np.ascontiguousarray(adj, np.uint32)
    sd = np.ascontiguousarray(seeds, np.uint32)
    nq = Q.shape[0]
    ids = np.empty((nq, k), np.uint32)
    dist = np.empty((nq, k), np.float32)
    nc = C.c_uint64()
    lib().orc_search_flat(_p(X, C.c_float), n, d, _p(o, C.c_uint64), _p(a, C.c_uint32), _p(Q, C.c_float), nq,
                          _p(sd, C.c_uint32), beam, k, _p(ids, C.c_uint32), _p(dist, C.c_float), C.byref(nc))
    return ids, dist, nc.value


def dpg(X, knn, M, maxM, prev=None, variant="par"):
    """DelaunayGraph restated on the k-NN lists knn (n x K).  variant "ref":
    the reference's in-place pass; "par": the order-free pass the GPU runs.
    prev: (offs, adj, dist) of the previous layer (carried lists) or None.
    Returns CSR offs (uint32), adj, dist."""
    X = _f32(X)
    n, d = X.shape
    knn = np.ascontiguousarray(knn, np.uint32)
    K = knn.shape[1]
    cap = max(M, maxM, 0 if prev is None else int(np.diff(prev[0]).max(initial=0)))
    offs = np.empty(n + 1, np.uint64)
    adj = np.empty(n * cap + 1, np.uint32)
    dist = np.empty(n * cap + 1, np.float32)
    keep = []
    if prev is not None:
        po = _csr64(prev[0])
        pa = np.ascontiguousarray(prev[1], np.uint32)
        pd = np.ascontiguousarray(prev[2], np.float32)
        keep = [po, pa, pd]
        pargs = (po.size - 1, po.ctypes.data, pa.ctypes.data if pa.size else None, pd.ctypes.data if pd.size else None)
    else:
        pargs = (0, None, None, None)
    f = lib().orc_dpg_ref if variant == "ref" else lib().orc_dpg_par
    f(_p(X, C.c_float), n, d, _p(knn, C.c_uint32), K, M, maxM, *pargs, _p(offs, C.c_uint64),
      _p(adj, C.c_uint32), _p(dist, C.c_float))
    del keep
    m = int(offs[-1])
    return offs.astype(np.uint32), adj[:m].copy(), dist[:m].copy()

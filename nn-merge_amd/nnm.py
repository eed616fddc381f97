"""ctypes binding of libnnmerge.so (the C ABI in include/nnmerge.h).

Loading fails loudly when the library is missing; there is no fallback path.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnnmerge.so")

NNM_OK = 0
NNM_ERR_INVALID = -1
NNM_ERR_HIP = -2
NNM_ERR_NOMEM = -3
NNM_ERR_STATE = -4
NNM_ERR_UNSUPPORTED = -5
RAW = 0xFFFFFFFF
PREC_FP32, PREC_BF16, PREC_LEARNED = 0, 1, 2
RULE_ALL, RULE_MERGE = 0, 1
METRIC_L2, METRIC_L1, METRIC_COSINE, METRIC_CHI2 = 0, 1, 2, 3
RERANK_AUTO, RERANK_ALWAYS, RERANK_NEVER = 0, 1, 2
FLAG_GRAPH, FLAG_AUTOTUNE, FLAG_LEARNED_RENEW, FLAG_LEARNED_FP8, FLAG_LEARNED_MARK_FOUND = 1, 2, 4, 8, 16
SEARCH_LAYERED, SEARCH_FLAT = 0, 1
SEARCH_SRC_KNN, SEARCH_SRC_DPG = 0, 1


class Params(C.Structure):
    _fields_ = [("K", C.c_uint32), ("rho", C.c_float), ("rev_cap", C.c_uint32), ("max_iters", C.c_uint32),
                ("delta", C.c_float), ("rule", C.c_uint32), ("prec", C.c_uint32), ("rerank", C.c_uint32),
                ("offer_cap", C.c_uint32), ("cut", C.c_uint32), ("plain_tail", C.c_uint32), ("flags", C.c_uint32),
                ("learn_tau", C.c_float), ("learn_rho", C.c_float), ("seed", C.c_uint64)]


class Stats(C.Structure):
    _fields_ = [("iters", C.c_uint32), ("converged", C.c_uint32), ("dists", C.c_uint64),
                ("init_dists", C.c_uint64), ("updates", C.c_uint64), ("reranked", C.c_uint64),
                ("list_entries", C.c_uint64), ("offers", C.c_uint64), ("last_delta", C.c_double), ("secs", C.c_double),
                ("join_secs", C.c_double), ("learned_iters", C.c_uint32), ("plain_iters", C.c_uint32),
                ("tile_js", C.c_uint32), ("tile_chunk", C.c_uint32), ("graph_iters", C.c_uint32), ("pad_", C.c_uint32),
                ("long_dists", C.c_uint64), ("long_entries", C.c_uint64), ("long_offers", C.c_uint64),
                ("join_main_secs", C.c_double), ("join_main_bytes", C.c_uint64), ("join_main_flops", C.c_uint64),
                ("near_ties", C.c_uint64), ("tie_swaps", C.c_uint64), ("redo_rows", C.c_uint64),
                ("learned_dists", C.c_uint64), ("redo_rounds", C.c_uint32), ("pad2_", C.c_uint32),
                ("deliver_secs", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class SearchParams(C.Structure):
    _fields_ = [("mode", C.c_uint32), ("beam", C.c_uint32), ("k", C.c_uint32), ("width", C.c_uint32),
                ("prec", C.c_uint32), ("rerank", C.c_uint32), ("entry", C.c_uint32), ("pad_", C.c_uint32),
                ("seed", C.c_uint64)]


class SearchStats(C.Structure):
    _fields_ = [("dists", C.c_uint64), ("reranked", C.c_uint64), ("hash_resets", C.c_uint64),
                ("capped", C.c_uint64), ("steps", C.c_uint64), ("secs", C.c_double), ("wave_secs", C.c_double),
                ("launch_secs", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class DistError(C.Structure):
    _fields_ = [("max_abs", C.c_double), ("max_rel", C.c_double), ("mean_rel", C.c_double),
                ("entries", C.c_uint64), ("order_flips", C.c_uint64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class NNMError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"nnmerge error {code}: {msg}")
        self.code = code


_lib = None

# exported entry points and their signatures (every symbol of include/nnmerge.h)
_vp = C.c_void_p
_SIGS = {
    "nnm_default_params": (None, [C.POINTER(Params)]),
    "nnm_create": (C.c_int, [C.POINTER(_vp), C.c_int]),
    "nnm_destroy": (C.c_int, [_vp]),
    "nnm_last_error": (C.c_char_p, [_vp]),
    "nnm_set_mem_cap": (C.c_int, [_vp, C.c_uint64]),
    "nnm_mem_used": (C.c_uint64, [_vp]),
    "nnm_set_data": (C.c_int, [_vp, _vp, C.c_uint64, C.c_uint32]),
    "nnm_set_metric": (C.c_int, [_vp, C.c_uint32]),
    "nnm_host_hgraph": (C.c_int, [_vp, C.c_uint64, C.c_uint32, C.POINTER(Params), C.c_uint32, C.c_char_p, _vp, _vp,
                                  C.POINTER(Stats), _vp, C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_double)]),
    "nnm_set_partition": (C.c_int, [_vp, _vp, _vp, _vp, C.c_uint32]),
    "nnm_graph_random_init": (C.c_int, [_vp, C.c_uint32, C.c_uint64, C.POINTER(Stats)]),
    "nnm_set_graph": (C.c_int, [_vp, _vp, _vp, _vp, C.c_uint32, C.POINTER(Stats)]),
    "nnm_get_graph": (C.c_int, [_vp, _vp, _vp, _vp]),
    "nnm_nndescent": (C.c_int, [_vp, C.POINTER(Params), C.POINTER(Stats)]),
    "nnm_local_join": (C.c_int, [_vp, C.POINTER(Params), C.c_uint32, C.POINTER(Stats)]),
    "nnm_merge_init": (C.c_int, [_vp, C.POINTER(Params), C.POINTER(Stats)]),
    "nnm_merge_finish": (C.c_int, [_vp, C.c_int]),
    "nnm_bruteforce_knn": (C.c_int, [_vp, _vp, _vp, C.c_uint64, C.c_uint32, C.c_int, _vp, _vp]),
    "nnm_pair_dists": (C.c_int, [_vp, _vp, _vp, C.c_uint64, _vp]),
    "nnm_bruteforce_knn_noalloc": (C.c_int, [_vp, _vp, C.c_uint64, C.c_uint32, C.c_int, _vp, _vp]),
    "nnm_graph_save": (C.c_int, [_vp]),
    "nnm_graph_restore": (C.c_int, [_vp, C.c_int, C.POINTER(Stats)]),
    "nnm_mlp_init": (C.c_int, [_vp, C.c_uint32, C.c_uint32, C.c_uint64]),
    "nnm_mlp_set_weights": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_float]),
    "nnm_mlp_get_weights": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "nnm_mlp_loss_grad": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_uint32, C.c_int, C.POINTER(C.c_float), _vp, _vp,
                                    _vp, _vp]),
    "nnm_mlp_train": (C.c_int, [_vp, C.c_uint32, C.c_uint32, C.c_float, C.c_float, C.c_uint64,
                                C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_double)]),
    "nnm_mlp_project": (C.c_int, [_vp, _vp, C.POINTER(C.c_double)]),
    "nnm_mlp_project_fp8": (C.c_int, [_vp, _vp, C.POINTER(C.c_float)]),
    "nnm_host_hmerge": (C.c_int, [_vp, C.c_uint64, C.c_uint32, C.POINTER(Params), C.c_uint32, _vp, _vp, _vp,
                                  C.POINTER(Stats), _vp, _vp, _vp, _vp, C.c_uint32, C.POINTER(C.c_uint32)]),
    "nnm_host_build": (C.c_int, [C.c_int, _vp, C.c_uint64, C.c_uint32, C.POINTER(Params), C.c_uint32, _vp, _vp,
                                 C.POINTER(Stats), C.POINTER(C.c_double)]),
    "nnm_graph_dist_error": (C.c_int, [_vp, C.POINTER(DistError)]),
    "nnm_io_matrix_shape": (C.c_int, [C.c_char_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)]),
    "nnm_io_read_matrix": (C.c_int, [C.c_char_p, _vp, C.c_uint64, C.c_uint32]),
    "nnm_io_write_matrix": (C.c_int, [C.c_char_p, _vp, C.c_uint64, C.c_uint32]),
    "nnm_io_knn_shape": (C.c_int, [C.c_char_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)]),
    "nnm_io_read_knn": (C.c_int, [C.c_char_p, _vp, C.c_uint64, C.c_uint32]),
    "nnm_io_write_knn": (C.c_int, [C.c_char_p, _vp, C.c_uint64, C.c_uint32]),
    "nnm_io_layers_shape": (C.c_int, [C.c_char_p, C.c_uint32, C.POINTER(C.c_uint32), _vp, _vp]),
    "nnm_io_read_layers": (C.c_int, [C.c_char_p, C.c_uint32, C.POINTER(_vp), C.POINTER(_vp)]),
    "nnm_io_write_layer": (C.c_int, [C.c_char_p, C.c_int, _vp, _vp, C.c_uint32]),
    "nnm_dpg_build": (C.c_int, [_vp, C.c_uint32, C.c_uint32, C.c_int, C.POINTER(Stats)]),
    "nnm_dpg_get": (C.c_int, [_vp, _vp, _vp, _vp]),
    "nnm_search_set_layers": (C.c_int, [_vp, C.c_uint32, _vp, C.POINTER(_vp), C.POINTER(_vp)]),
    "nnm_search_use_graph": (C.c_int, [_vp, C.c_int]),
    "nnm_search": (C.c_int, [_vp, _vp, C.c_uint32, C.POINTER(SearchParams), _vp, _vp, _vp,
                             C.POINTER(SearchStats)]),
    "nnm_search_beams": (C.c_int, [_vp, _vp, C.c_uint32, C.POINTER(SearchParams), _vp, C.c_uint32, _vp, _vp,
                                   _vp]),
    "nnm_host_lysearch": (C.c_int, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p,
                                    C.c_uint32, C.c_uint32, C.c_uint32, _vp, _vp, _vp, _vp, _vp]),
}


def exported_symbols():
    return list(_SIGS)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is not built; run python nn-merge_amd/build.py")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def default_params(**kw):
    p = Params()
    lib().nnm_default_params(C.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def _ptr(a):
    return None if a is None else C.c_void_p(a.ctypes.data)


class Context:
    """One device context (one HIP stream) of the library."""

    def __init__(self, device=0):
        self.h = C.c_void_p()
        rc = lib().nnm_create(C.byref(self.h), device)
        if rc != NNM_OK:
            raise NNMError(rc, "nnm_create failed")
        self.n = 0
        self.d = 0
        self.K = 0

    def close(self):
        if self.h:
            lib().nnm_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != NNM_OK:
            raise NNMError(rc, lib().nnm_last_error(self.h).decode())
        return rc

    def last_error(self):
        return lib().nnm_last_error(self.h).decode()

    def set_mem_cap(self, nbytes):
        self._chk(lib().nnm_set_mem_cap(self.h, nbytes))

    def mem_used(self):
        return lib().nnm_mem_used(self.h)

    def set_data(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        self._chk(lib().nnm_set_data(self.h, _ptr(X), X.shape[0], X.shape[1]))
        self.n, self.d = X.shape
        self.n_act = self.n
        self.K = 0

    def set_metric(self, metric):
        """NNM_METRIC_* for every later distance (L2 by default)."""
        self._chk(lib().nnm_set_metric(self.h, int(metric)))
        self.metric = int(metric)

    def set_partition(self, sub_begin, sub_built, sub_group=None):
        sb = np.ascontiguousarray(sub_begin, dtype=np.uint32)
        bu = np.ascontiguousarray(sub_built, dtype=np.uint8)
        gr = np.ascontiguousarray(np.zeros(bu.size) if sub_group is None else sub_group, dtype=np.uint32)
        self._chk(lib().nnm_set_partition(self.h, _ptr(sb), _ptr(bu), _ptr(gr), bu.size))
        self.n_act = int(sb[-1])

    def random_init(self, K, seed=1):
        st = Stats()
        self._chk(lib().nnm_graph_random_init(self.h, K, seed, C.byref(st)))
        self.K = K
        return st.as_dict()

    def set_graph(self, ids, dist=None, flags=None):
        ids = np.ascontiguousarray(ids, dtype=np.uint32)
        K = ids.shape[1]
        dist = None if dist is None else np.ascontiguousarray(dist, dtype=np.float32)
        flags = None if flags is None else np.ascontiguousarray(flags, dtype=np.uint8)
        st = Stats()
        self._chk(lib().nnm_set_graph(self.h, _ptr(ids), _ptr(dist), _ptr(flags), K, C.byref(st)))
        self.K = K
        return st.as_dict()

    def get_graph(self):
        ids = np.empty((self.n, self.K), np.uint32)
        dist = np.empty((self.n, self.K), np.float32)
        flags = np.empty((self.n, self.K), np.uint8)
        self._chk(lib().nnm_get_graph(self.h, _ptr(ids), _ptr(dist), _ptr(flags)))
        return ids, dist, flags

    def graph_save(self):
        self._chk(lib().nnm_graph_save(self.h))

    def graph_restore(self, recompute=True):
        """recompute: False, True (FP32 rows) or 2 (bf16 rows where exact)."""
        st = Stats()
        self._chk(lib().nnm_graph_restore(self.h, int(recompute), C.byref(st)))
        return st.as_dict()

    def nndescent(self, p):
        st = Stats()
        self._chk(lib().nnm_nndescent(self.h, C.byref(p), C.byref(st)))
        return st.as_dict()

    def local_join(self, p, it=0):
        st = Stats()
        self._chk(lib().nnm_local_join(self.h, C.byref(p), it, C.byref(st)))
        return st.as_dict()

    def merge_init(self, p):
        st = Stats()
        self._chk(lib().nnm_merge_init(self.h, C.byref(p), C.byref(st)))
        return st.as_dict()

    def merge_finish(self, dedup=True):
        self._chk(lib().nnm_merge_finish(self.h, 1 if dedup else 0))

    def bruteforce(self, k, rows=None, Q=None, exclude_self=True):
        if Q is not None:
            Q = np.ascontiguousarray(Q, dtype=np.float32)
            nq = Q.shape[0]
            r = None
        else:
            r = np.ascontiguousarray(np.arange(self.n) if rows is None else rows, dtype=np.uint32)
            nq = r.size
        ids = np.empty((nq, k), np.uint32)
        dist = np.empty((nq, k), np.float32)
        self._chk(lib().nnm_bruteforce_knn(self.h, _ptr(Q), _ptr(r), nq, k, 1 if exclude_self else 0,
                                           _ptr(ids), _ptr(dist)))
        return ids, dist

    def bruteforce_noalloc(self, k, rows_dev, nq, ids_dev, dist_dev, exclude_self=True):
        """nnm_bruteforce_knn_noalloc on device buffers given by address
        (rows_dev: nq uint32 row numbers; ids_dev / dist_dev: nq x k)."""
        self._chk(lib().nnm_bruteforce_knn_noalloc(self.h, C.c_void_p(rows_dev), nq, k, 1 if exclude_self else 0,
                                                   C.c_void_p(ids_dev), C.c_void_p(dist_dev)))

    # ---- learned projection
    def mlp_init(self, hidden=256, d_out=32, seed=1):
        self._chk(lib().nnm_mlp_init(self.h, hidden, d_out, seed))
        self.mlp_shape = (hidden, d_out)

    def mlp_weights(self):
        hid, dout = self.mlp_shape
        W1 = np.empty((hid, (self.d + 31) // 32 * 32), np.float32)
        b1 = np.empty(hid, np.float32)
        W2 = np.empty((dout, hid), np.float32)
        b2 = np.empty(dout, np.float32)
        al, tau = C.c_float(), C.c_float()
        self._chk(lib().nnm_mlp_get_weights(self.h, _ptr(W1), _ptr(b1), _ptr(W2), _ptr(b2), C.byref(al),
                                            C.byref(tau)))
        return W1, b1, W2, b2, al.value, tau.value

    def mlp_set_weights(self, W1, b1, W2, b2, alpha=0.0):
        arrs = [np.ascontiguousarray(a, np.float32) for a in (W1, b1, W2, b2)]
        self._chk(lib().nnm_mlp_set_weights(self.h, *[_ptr(a) for a in arrs], alpha))

    def mlp_loss_grad(self, xa, xb, t, s, bf16=True):
        hid, dout = self.mlp_shape
        xa, xb, t, s = [np.ascontiguousarray(a, np.float32) for a in (xa, xb, t, s)]
        B = t.size
        gW1 = np.empty((hid, self.d), np.float32)
        gb1 = np.empty(hid, np.float32)
        gW2 = np.empty((dout, hid), np.float32)
        gb2 = np.empty(dout, np.float32)
        loss = C.c_float()
        self._chk(lib().nnm_mlp_loss_grad(self.h, _ptr(xa), _ptr(xb), _ptr(t), _ptr(s), B, 1 if bf16 else 0,
                                          C.byref(loss), _ptr(gW1), _ptr(gb1), _ptr(gW2), _ptr(gb2)))
        return loss.value, gW1, gb1, gW2, gb2

    def mlp_train(self, steps=1000, batch=4096, lr=1e-3, keep=0.98, seed=1):
        lf, ll, secs = C.c_float(), C.c_float(), C.c_double()
        self._chk(lib().nnm_mlp_train(self.h, steps, batch, lr, keep, seed, C.byref(lf), C.byref(ll),
                                      C.byref(secs)))
        return {"loss_first": lf.value, "loss_last": ll.value, "secs": secs.value}

    def mlp_project(self, want=False):
        hid, dout = self.mlp_shape
        out = np.empty((self.n, dout), np.float32) if want else None
        secs = C.c_double()
        self._chk(lib().nnm_mlp_project(self.h, _ptr(out), C.byref(secs)))
        return out, secs.value

    def mlp_project_fp8(self, want=True):
        """FP8 rows of the projection (n x d_out uint8, OCP e4m3) and their scale."""
        hid, dout = self.mlp_shape
        out = np.empty((self.n, dout), np.uint8) if want else None
        sc = C.c_float()
        self._chk(lib().nnm_mlp_project_fp8(self.h, _ptr(out), C.byref(sc)))
        return out, sc.value

    def dist_error(self):
        e = DistError()
        self._chk(lib().nnm_graph_dist_error(self.h, C.byref(e)))
        return e.as_dict()

    # -------------------------------------------------------- DPG + search
    def dpg_build(self, M, maxM, carry=False):
        st = Stats()
        self._chk(lib().nnm_dpg_build(self.h, M, maxM, 1 if carry else 0, C.byref(st)))
        return st.as_dict()

    def dpg_get(self, rows=None):
        rows = self.n_act if rows is None else rows
        offs = np.empty(rows + 1, np.uint32)
        self._chk(lib().nnm_dpg_get(self.h, _ptr(offs), None, None))
        adj = np.empty(int(offs[-1]), np.uint32)
        dist = np.empty(int(offs[-1]), np.float32)
        self._chk(lib().nnm_dpg_get(self.h, _ptr(offs), _ptr(adj), _ptr(dist)))
        return offs, adj, dist

    def search_set_layers(self, layers):
        """layers: list of (offs, adj) CSR pairs, upper layers first."""
        offs = [np.ascontiguousarray(o, dtype=np.uint32) for o, _ in layers]
        adj = [np.ascontiguousarray(a, dtype=np.uint32) for _, a in layers]
        rows = np.array([o.size - 1 for o in offs], np.uint32)
        po = (_vp * len(offs))(*[o.ctypes.data for o in offs])
        pa = (_vp * len(adj))(*[a.ctypes.data if a.size else offs[0].ctypes.data for a in adj])
        self._chk(lib().nnm_search_set_layers(self.h, len(offs), _ptr(rows), po, pa))

    def search_use_graph(self, source=SEARCH_SRC_DPG):
        self._chk(lib().nnm_search_use_graph(self.h, source))

    def search(self, Q, beam, k=10, width=1, prec=PREC_FP32, mode=SEARCH_LAYERED, entry=0, seeds=None, seed=1,
               rerank=0):
        Q = np.ascontiguousarray(Q, dtype=np.float32)
        nq = Q.shape[0]
        p = SearchParams(mode=mode, beam=beam, k=k, width=width, prec=prec, rerank=rerank, entry=entry, seed=seed)
        sd = None if seeds is None else np.ascontiguousarray(seeds, dtype=np.uint32)
        ids = np.empty((nq, k), np.uint32)
        dist = np.empty((nq, k), np.float32)
        st = SearchStats()
        self._chk(lib().nnm_search(self.h, _ptr(Q), nq, C.byref(p), _ptr(sd), _ptr(ids), _ptr(dist), C.byref(st)))
        return ids, dist, st.as_dict()

    def search_beams(self, Q, beams, k=10, width=1, prec=PREC_FP32, mode=SEARCH_LAYERED, entry=0, seed=1, rerank=0):
        """Every beam of `beams` in one launch: ids / dists (nbeams x nq x k)
        and one stats dict per beam."""
        Q = np.ascontiguousarray(Q, dtype=np.float32)
        nq = Q.shape[0]
        bl = np.ascontiguousarray(beams, dtype=np.uint32)
        p = SearchParams(mode=mode, beam=0, k=k, width=width, prec=prec, rerank=rerank, entry=entry, seed=seed)
        ids = np.empty((bl.size, nq, k), np.uint32)
        dist = np.empty((bl.size, nq, k), np.float32)
        st = (SearchStats * bl.size)()
        self._chk(lib().nnm_search_beams(self.h, _ptr(Q), nq, C.byref(p), _ptr(bl), bl.size, _ptr(ids), _ptr(dist),
                                         C.cast(st, C.c_void_p)))
        return ids, dist, [s.as_dict() for s in st]

    def pair_dists(self, a, b):
        a = np.ascontiguousarray(a, dtype=np.uint32)
        b = np.ascontiguousarray(b, dtype=np.uint32)
        out = np.empty(a.size, np.float32)
        self._chk(lib().nnm_pair_dists(self.h, _ptr(a), _ptr(b), a.size, _ptr(out)))
        return out


HOST_KGRAPH, HOST_SMERGE, HOST_JMERGE, HOST_PMERGE = 0, 1, 2, 3


def host_lysearch(refer, graph, query, dst, grd, record, entry=0, prec=PREC_FP32):
    """LySearch(refer, graph).knnSearch(query, dst, grd, record) through the
    host class; returns the records [(topk, recall, secs, cmps_per_query)]."""
    n = 32
    topk = np.zeros(n, np.uint32)
    rec = np.zeros(n, np.float64)
    secs = np.zeros(n, np.float64)
    cmps = np.zeros(n, np.uint64)
    nr = C.c_uint32()
    rc = lib().nnm_host_lysearch(*(str(x).encode() for x in (refer, graph, query, dst, grd, record)), entry, prec,
                                 n, C.byref(nr), _ptr(topk), _ptr(rec), _ptr(secs), _ptr(cmps))
    if rc != NNM_OK:
        raise NNMError(rc, "nnm_host_lysearch failed")
    return [(int(topk[i]), float(rec[i]), float(secs[i]), int(cmps[i])) for i in range(min(nr.value, n))]


def host_build(kind, X, p, parts=0):
    """Run a host class (KGraph / Smerge / Jmerge / PMerge) end to end."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    n, d = X.shape
    ids = np.empty((n, p.K), np.uint32)
    dist = np.empty((n, p.K), np.float32)
    st = Stats()
    bs = C.c_double()
    rc = lib().nnm_host_build(kind, _ptr(X), n, d, C.byref(p), parts, _ptr(ids), _ptr(dist), C.byref(st),
                              C.byref(bs))
    if rc != NNM_OK:
        raise NNMError(rc, "nnm_host_build failed")
    return ids, dist, st.as_dict(), bs.value


def host_hmerge(X, p, leaf_begin=None, max_levels=64):
    """LyNNDescent: doubling schedule (leaf_begin None) or leaf hierarchy."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    n, d = X.shape
    ids = np.empty((n, p.K), np.uint32)
    dist = np.empty((n, p.K), np.float32)
    st = Stats()
    lb = None if leaf_begin is None else np.ascontiguousarray(leaf_begin, np.uint32)
    nl = 0 if lb is None else lb.size - 1
    rows = np.zeros(max_levels, np.uint32)
    groups = np.zeros(max_levels, np.uint32)
    secs = np.zeros(max_levels, np.float64)
    dists = np.zeros(max_levels, np.uint64)
    L = C.c_uint32()
    rc = lib().nnm_host_hmerge(_ptr(X), n, d, C.byref(p), nl, _ptr(lb), _ptr(ids), _ptr(dist), C.byref(st),
                               _ptr(rows), _ptr(groups), _ptr(secs), _ptr(dists), max_levels, C.byref(L))
    if rc != NNM_OK:
        raise NNMError(rc, "nnm_host_hmerge failed")
    m = min(L.value, max_levels)
    levels = [{"rows": int(rows[i]), "groups": int(groups[i]), "secs": float(secs[i]), "dists": int(dists[i])}
              for i in range(m)]
    return ids, dist, st.as_dict(), levels


def host_hgraph(X, p, M, dst_prefix=None, max_layers=16):
    """LyNNDescent::buildhgraph with its DPG search layers; returns ids,
    dist, stats, the layers' row counts and the layers' device seconds."""
    X = np.ascontiguousarray(X, dtype=np.float32)
    n, d = X.shape
    ids = np.empty((n, p.K), np.uint32)
    dist = np.empty((n, p.K), np.float32)
    st = Stats()
    rows = np.zeros(max_layers, np.uint32)
    L = C.c_uint32()
    ds = C.c_double()
    pre = None if dst_prefix is None else str(dst_prefix).encode()
    rc = lib().nnm_host_hgraph(_ptr(X), n, d, C.byref(p), M, pre, _ptr(ids), _ptr(dist), C.byref(st), _ptr(rows),
                               max_layers, C.byref(L), C.byref(ds))
    if rc != NNM_OK:
        raise NNMError(rc, "nnm_host_hgraph failed")
    return ids, dist, st.as_dict(), [int(r) for r in rows[:min(L.value, max_layers)]], ds.value


def recall(ids, gt, rows, k=10, unique=True):
    """recall@k of lists `ids` (n x K) on `rows` against ground truth `gt`
    (len(rows) x >=k): Evaluator::getKGraphRecall semantics
    (src/evaluator.cpp:258-295); unique=False counts repeated ids each time."""
    L = ids[np.asarray(rows)][:, :k]
    G = gt[:, :k]
    hits = 0
    for i in range(L.shape[0]):
        row = L[i]
        if unique:
            row = np.unique(row)
        hits += np.isin(row, G[i]).sum()
    return hits / (L.shape[0] * k)


# ------------------------------------------------------------ file formats

def _io_chk(rc, what):
    if rc != NNM_OK:
        raise NNMError(rc, what)


def read_matrix(path, max_rows=None):
    """.txt / .fvecs / .bvecs matrix as float32 (rows x dim)."""
    r, d = C.c_uint64(), C.c_uint32()
    _io_chk(lib().nnm_io_matrix_shape(path.encode(), C.byref(r), C.byref(d)), f"cannot read {path}")
    rows = r.value if max_rows is None else min(r.value, max_rows)
    X = np.empty((rows, d.value), np.float32)
    _io_chk(lib().nnm_io_read_matrix(path.encode(), _ptr(X), rows, d.value), f"malformed {path}")
    return X


def write_matrix(path, X):
    X = np.ascontiguousarray(X, dtype=np.float32)
    _io_chk(lib().nnm_io_write_matrix(path.encode(), _ptr(X), X.shape[0], X.shape[1]), f"cannot write {path}")


def read_knn(path, topk=None):
    r, k = C.c_uint64(), C.c_uint32()
    _io_chk(lib().nnm_io_knn_shape(path.encode(), C.byref(r), C.byref(k)), f"cannot read {path}")
    topk = k.value if topk is None else topk
    ids = np.empty((r.value, topk), np.uint32)
    _io_chk(lib().nnm_io_read_knn(path.encode(), _ptr(ids), r.value, topk), f"malformed {path}")
    return ids


def write_knn(path, ids):
    ids = np.ascontiguousarray(ids, dtype=np.uint32)
    _io_chk(lib().nnm_io_write_knn(path.encode(), _ptr(ids), ids.shape[0], ids.shape[1]), f"cannot write {path}")


def read_layers(path, max_layers=8):
    """DPG blocks of a (layered) graph file as [(offs, adj), ...]."""
    nl = C.c_uint32()
    rows = np.zeros(max_layers, np.uint32)
    edges = np.zeros(max_layers, np.uint64)
    _io_chk(lib().nnm_io_layers_shape(path.encode(), max_layers, C.byref(nl), _ptr(rows), _ptr(edges)),
            f"cannot read {path}")
    offs = [np.empty(int(rows[g]) + 1, np.uint32) for g in range(nl.value)]
    adj = [np.empty(max(int(edges[g]), 1), np.uint32) for g in range(nl.value)]
    po = (_vp * nl.value)(*[o.ctypes.data for o in offs])
    pa = (_vp * nl.value)(*[a.ctypes.data for a in adj])
    _io_chk(lib().nnm_io_read_layers(path.encode(), nl.value, po, pa), f"malformed {path}")
    return [(offs[g], adj[g][:int(edges[g])]) for g in range(nl.value)]


def write_layers(path, layers):
    for g, (offs, adj) in enumerate(layers):
        o = np.ascontiguousarray(offs, dtype=np.uint32)
        a = np.ascontiguousarray(adj, dtype=np.uint32)
        _io_chk(lib().nnm_io_write_layer(path.encode(), 1 if g else 0, _ptr(o), _ptr(a) if a.size else _ptr(o),
                                         o.size - 1), f"cannot write {path}")

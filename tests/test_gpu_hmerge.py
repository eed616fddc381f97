"""H-Merge (LyNNDescent): the reference's doubling schedule and the batched
hierarchy over uneven leaves, against the CPU restatement run level by
level (oracle KGraph for the seed graph and the leaves, oracle merge for
every level)."""
import numpy as np
import pytest

import nnm
import nnm_data
import orc

pytestmark = pytest.mark.gpu

TOL = 0.005


def _oracle_doubling(X, K, seed):
    """LyNNDescent::buildhgraph restated (src/lynndescent.cpp:735-831)."""
    n = X.shape[0]
    N0 = 64
    g, _, _ = orc.kgraph(X[:N0], K=K, seed=seed)
    cur, smpl = N0, N0
    while cur < n:
        b = min(smpl, n - cur)
        m = cur + b
        init = np.zeros((m, K), np.uint32)
        init[:cur] = g
        g, _, _ = orc.merge(X[:m], [0, cur, m], [1, 0], init, K=K, cut=K - K // 2, seed=seed + cur)
        cur += b
        smpl = cur
    return g


def _oracle_leaves(X, K, begin, seed):
    n = X.shape[0]
    g = np.zeros((n, K), np.uint32)
    for i in range(len(begin) - 1):
        a, b = begin[i], begin[i + 1]
        li, _, _ = orc.kgraph(X[a:b], K=K, seed=seed + i)
        g[a:b] = li + a
    begin = list(begin)
    while len(begin) > 2:
        nb = []
        for s in range(0, len(begin) - 1, 2):
            if s + 2 < len(begin):
                a, m, b = begin[s], begin[s + 1], begin[s + 2]
                sub = g[a:b] - a
                mi, _, _ = orc.merge(X[a:b], [0, m - a, b - a], [1, 1], sub, K=K, seed=seed + a)
                g[a:b] = mi + a
            nb.append(begin[s])
        nb.append(n)
        begin = nb
    return g


@pytest.fixture(scope="module")
def data():
    X = nnm_data.sift_like(12000, 128, centers=48, seed=81, rank=12)
    rows = np.arange(0, 12000, 6, dtype=np.uint32)
    gt, _ = orc.bruteforce(X, 10, rows=rows)
    return X, rows, gt


def test_hmerge_doubling_matches_restatement(data):
    X, rows, gt = data
    K = 20
    p = nnm.default_params(K=K, seed=3)
    ids, dist, st, levels = nnm.host_hmerge(X, p)
    assert [l["rows"] for l in levels] == [64, 128, 256, 512, 1024, 2048, 4096, 8192, 12000]
    r_gpu = nnm.recall(ids, gt, rows)
    r_cpu = orc.recall(_oracle_doubling(X, K, 3), rows, gt, 10)
    print("doubling gpu", r_gpu, "cpu", r_cpu, st)
    assert r_gpu >= r_cpu - TOL


def test_hmerge_uneven_leaves_matches_restatement(data):
    X, rows, gt = data
    K = 20
    sizes = [900, 3100, 1500, 2200, 800, 3500]  # uneven leaves, one left over at the first level
    begin = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint32)
    p = nnm.default_params(K=K, seed=4)
    ids, dist, st, levels = nnm.host_hmerge(X, p, leaf_begin=begin)
    # level 0 builds 6 leaves together, then 3 merges together, then 2, then 1
    assert [l["groups"] for l in levels] == [6, 3, 2, 1]
    r_gpu = nnm.recall(ids, gt, rows)
    r_cpu = orc.recall(_oracle_leaves(X, K, begin, 4), rows, gt, 10)
    print("leaves gpu", r_gpu, "cpu", r_cpu, levels)
    assert r_gpu >= r_cpu - TOL
    assert r_gpu > 0.95


def _layer_dists(X, offs, adj):
    """FP32 distances of a layer's edges (integer rows: exact in any order)."""
    src = np.repeat(np.arange(offs.size - 1), np.diff(offs.astype(np.int64)))
    d = X[src].astype(np.float64) - X[adj].astype(np.float64)
    return (d * d).sum(1).astype(np.float32)


def test_hgraph_layer_emission(tmp_path):
    """buildhgraph's level[] emission (src/lynndescent.cpp:755-800): layers
    after the seed graph (64 rows), after levels 3 and 6 of the doubling
    schedule (512 and 4096 rows) and after the last level (all rows), in that
    order in the layers file, base k-NN graph in the _baselayer file.  The
    base layer must equal the oracle's DelaunayGraph of the final lists
    carried from the layer before it; search over the emitted layers
    (LySearch::knnSearch) must return what the oracle's search returns on the
    same layers."""
    X = nnm_data.sift_like(5000, 64, centers=16, seed=41)
    K, M = 16, 16
    p = nnm.default_params(K=K, seed=7, max_iters=10)
    pre = tmp_path / "h"
    ids, dist, st, rows, dpg_secs = nnm.host_hgraph(X, p, M, dst_prefix=pre)
    assert rows == [64, 512, 4096, 5000]
    assert dpg_secs > 0 and st["secs"] > 0
    layers = nnm.read_layers(str(pre) + ".txt")
    assert [o.size - 1 for o, _ in layers] == rows
    assert np.array_equal(nnm.read_knn(str(pre) + "_baselayer.txt"), ids)
    # base layer = oracle DelaunayGraph(final lists, K/2, M) carried from layer 3
    po, pa = layers[2]
    prev = (po, pa, _layer_dists(X, po, pa))
    ro, ra, _ = orc.dpg(X, ids, K // 2, M, prev=prev, variant="par")
    bo, ba = layers[3]
    assert np.array_equal(bo, ro) and np.array_equal(ba, ra)
    # upper layers bound their reverse lists by M / 2 beyond the occlusion lists
    for (o, a), r in zip(layers[:3], rows[:3]):
        assert a.size == 0 or (a < r).all()
    # the layers serve search: layered descent then best-first on the base
    Q = nnm_data.sift_like(5000, 64, centers=16, seed=41, nq=300)[1]
    with nnm.Context(0) as ctx:
        ctx.set_data(X)
        ctx.search_set_layers(layers)
        got, gd, _ = ctx.search(Q, 64, k=10, entry=0)
        gt, _ = ctx.bruteforce(10, Q=Q)
    oi, od, _ = orc.search_layered(X, [(o.astype(np.uint64), a) for o, a in layers], Q, 0, 64, 10)
    assert np.array_equal(got, oi)
    assert np.array_equal(gd.view(np.uint32), od.view(np.uint32))
    print("search recall@10 over the emitted layers (clustered rows, beam 64):",
          nnm.recall(got, gt, np.arange(300, dtype=np.uint32), 10))

"""The C++ host classes (Smerge, Jmerge, KGraph, PMerge) end to end through
the C ABI, against the CPU restatement."""
import numpy as np
import pytest

import nnm
import nnm_data
import orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def data():
    X = nnm_data.sift_like(12000, 128, centers=48, sigma=20.0, seed=31)
    rows = np.arange(0, 12000, 8, dtype=np.uint32)
    gt, _ = orc.bruteforce(X, 10, rows=rows)
    return X, rows, gt


def test_host_kgraph(data):
    X, rows, gt = data
    p = nnm.default_params(K=20, seed=3)
    ids, dist, st, bs = nnm.host_build(nnm.HOST_KGRAPH, X, p)
    oi, _, _ = orc.kgraph(X, K=20, seed=3)
    assert nnm.recall(ids, gt, rows) >= orc.recall(oi, rows, gt, 10) - 0.005


@pytest.mark.parametrize("kind,built", [(nnm.HOST_SMERGE, [1, 1]), (nnm.HOST_JMERGE, [1, 0])])
def test_host_two_way_merges(data, kind, built):
    X, rows, gt = data
    n = X.shape[0]
    p = nnm.default_params(K=20, seed=5)
    ids, dist, st, bs = nnm.host_build(kind, X, p)
    assert st["dists"] > 0 and st["secs"] > 0 and bs > 0
    r_gpu = nnm.recall(ids, gt, rows)
    half = n * 5 // 10
    g0, _, _ = orc.kgraph(X[:half], K=20, seed=7)
    g1, _, _ = orc.kgraph(X[half:], K=20, seed=8)
    init = np.vstack([g0, g1 + half]).astype(np.uint32)
    oi, _, _ = orc.merge(X, [0, half, n], built, init, K=20, seed=9)
    r_cpu = orc.recall(oi, rows, gt, 10)
    print(kind, r_gpu, r_cpu, st)
    assert r_gpu >= r_cpu - 0.005


def test_host_pmerge_four_parts(data):
    X, rows, gt = data
    p = nnm.default_params(K=20, seed=6)
    ids, dist, st, bs = nnm.host_build(nnm.HOST_PMERGE, X, p, parts=4)
    assert nnm.recall(ids, gt, rows) > 0.98

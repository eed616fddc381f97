"""GPU parity of the other metrics (L1, cosine, chi-square) against the
oracle (whose metric routines equal the reference's DistOpt routines bit
for bit, tests/test_metrics.py).

Distances are evaluated with the same IEEE operations in the same order on
both sides, so FP32 distances are compared for bit equality; neighbour ids
against the FP64 ground truth wherever no FP32 near-tie can reorder them.
NN-Descent and J-Merge recall@10 must be within 0.005 of the oracle's
restatement on the same inputs.
"""
import numpy as np
import pytest

import nnm
import nnm_data
import orc

pytestmark = pytest.mark.gpu

TOL = 0.005
METRICS = [("l1", nnm.METRIC_L1), ("cosine", nnm.METRIC_COSINE), ("chi2", nnm.METRIC_CHI2)]


@pytest.fixture(scope="module")
def ctx():
    with nnm.Context(0) as c:
        yield c
        c.set_metric(nnm.METRIC_L2)


def _data(metric, n, seed):
    if metric == nnm.METRIC_COSINE:
        return nnm_data.unit_mixture(n, 64, centers=32, sigma=0.3, seed=seed)
    # histogram-like non-negative integer rows (SIFT-like), the usual domain
    # of L1 and chi-square
    return nnm_data.sift_like(n, 64, centers=32, sigma=20.0, seed=seed)


@pytest.mark.parametrize("name,metric", METRICS)
def test_bruteforce_and_pair_distances_bit_exact(ctx, name, metric):
    X = _data(metric, 3000, 31)
    ctx.set_metric(metric)
    ctx.set_data(X)
    rows = np.arange(0, 3000, 13, dtype=np.uint32)
    gi, gd = ctx.bruteforce(10, rows=rows)
    with orc.metric_scope(metric):
        oi, od = orc.bruteforce(X, 10, rows=rows)
    # every returned distance is the oracle's FP32 value of that pair
    for q, r in enumerate(rows):
        for j in range(10):
            assert np.float32(gd[q, j]).tobytes() == np.float32(orc.dist32(X[r], X[gi[q, j]], metric)).tobytes()
    # FP32 error bound against FP64: relative 1e-5, plus for cosine the
    # absolute cancellation of 2 - 2 cos near 0 (a few ulps of 2)
    tol = 1e-5 * np.abs(od) + (2e-6 if metric == nnm.METRIC_COSINE else 0.0)
    assert (np.abs(gd.astype(np.float64) - od) <= tol).all()
    # ids equal the FP64 ground truth away from FP32 near-ties
    gap_ok = np.ones_like(od, dtype=bool)
    sep = np.diff(od, axis=1) > tol[:, 1:] + tol[:, :-1]
    gap_ok[:, 1:] &= sep
    gap_ok[:, :-1] &= sep
    assert np.array_equal(gi[gap_ok], oi[gap_ok])
    a = np.arange(0, 3000, 3, dtype=np.uint32)
    b = (a * 17 + 11) % 3000
    pd = ctx.pair_dists(a, b)
    ref = np.array([orc.dist32(X[i], X[j], metric) for i, j in zip(a, b)], np.float32)
    assert pd.tobytes() == ref.tobytes()


@pytest.mark.parametrize("name,metric", METRICS)
def test_kgraph_and_jmerge_recall_parity(ctx, name, metric):
    n, K = 6000, 16
    X = _data(metric, n, 32)
    ctx.set_metric(metric)
    ctx.set_data(X)
    rows = np.arange(0, n, 5, dtype=np.uint32)
    gt, _ = ctx.bruteforce(10, rows=rows)
    with orc.metric_scope(metric):
        ot, _ = orc.bruteforce(X, 10, rows=rows)
    assert (gt == ot).mean() > 0.99
    # NN-Descent from random lists
    ctx.random_init(K, seed=3)
    st = ctx.nndescent(nnm.default_params(K=K, rule=nnm.RULE_ALL, seed=3, max_iters=20))
    ids, dist, _ = ctx.get_graph()
    assert (np.diff(dist.astype(np.float64), axis=1) >= 0).all()
    for v in range(0, n, 97):
        for j in range(K):
            assert np.float32(dist[v, j]).tobytes() == np.float32(orc.dist32(X[v], X[ids[v, j]], metric)).tobytes()
    r_gpu = nnm.recall(ids, ot, rows, 10)
    with orc.metric_scope(metric):
        oi, _, _ = orc.kgraph(X, K=K, seed=3, max_iters=20)
    r_cpu = orc.recall(oi, rows, ot, 10)
    print(name, "kgraph recall gpu", r_gpu, "cpu", r_cpu, "iters", st["iters"])
    assert r_gpu >= r_cpu - TOL
    # J-Merge: first half built, second half raw
    half = n // 2
    ctx.set_data(X[:half])
    ctx.random_init(K, seed=4)
    ctx.nndescent(nnm.default_params(K=K, rule=nnm.RULE_ALL, seed=4, max_iters=20))
    g0, _, _ = ctx.get_graph()
    init = np.zeros((n, K), np.uint32)
    init[:half] = g0
    init[half:] = ((np.arange(half, n)[:, None] + 1 + np.arange(K)[None, :]) % n).astype(np.uint32)
    ctx.set_data(X)
    ctx.set_partition([0, half, n], [1, 0], [0, 0])
    ctx.set_graph(init)
    p = nnm.default_params(K=K, rule=nnm.RULE_MERGE, seed=6, max_iters=20)
    ctx.merge_init(p)
    ctx.nndescent(p)
    ctx.merge_finish(True)
    mi, _, _ = ctx.get_graph()
    r_gpu = nnm.recall(mi, ot, rows, 10)
    with orc.metric_scope(metric):
        om, _, _ = orc.merge(X, [0, half, n], [1, 0], init, K=K, seed=6, max_iters=20)
    r_cpu = orc.recall(om, rows, ot, 10)
    print(name, "jmerge recall gpu", r_gpu, "cpu", r_cpu)
    assert r_gpu >= r_cpu - TOL


def test_l2_only_paths_refuse_other_metrics(ctx):
    X = _data(nnm.METRIC_L1, 2000, 33)
    ctx.set_metric(nnm.METRIC_L1)
    ctx.set_data(X)
    ctx.random_init(10, seed=1)
    with pytest.raises(nnm.NNMError):
        ctx.nndescent(nnm.default_params(K=10, prec=nnm.PREC_BF16, max_iters=2))
    with pytest.raises(nnm.NNMError):
        ctx.dpg_build(5, 10)
    with pytest.raises(nnm.NNMError):
        ctx.dist_error()
    ctx.set_metric(nnm.METRIC_L2)
    st = ctx.nndescent(nnm.default_params(K=10, prec=nnm.PREC_BF16, max_iters=2))
    assert st["iters"] >= 1 and st["dists"] > 0

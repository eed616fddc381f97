"""Captured iterations and the join autotuner change nothing in the result.

NNM_FLAG_GRAPH replays each NN-Descent iteration from one HIP graph per join
precision (the iteration number and the salts derive from a device counter,
so every replay is a fresh iteration); NNM_FLAG_AUTOTUNE times the join launch
shapes on the first iteration.  Both must reproduce the eager default run bit
for bit: same lists, distances, flags and counted distances, after every
replay.
"""
import numpy as np
import pytest

import nnm
import nnm_data

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    with nnm.Context(0) as c:
        yield c


@pytest.fixture(scope="module")
def X():
    return nnm_data.sift_like(30000, 128, centers=64, seed=71, rank=12)


def _kgraph(ctx, X, flags, prec, iters=6, seed=9):
    ctx.set_data(X)
    ctx.set_partition([0, X.shape[0]], [0], [0])
    ctx.random_init(20, seed=seed)
    st = ctx.nndescent(nnm.default_params(K=20, prec=prec, seed=seed, max_iters=iters, flags=flags))
    ids, dist, fl = ctx.get_graph()
    return ids, dist, fl, st


def _jmerge(ctx, X, flags, prec, iters=6):
    n, K = X.shape[0], 20
    half = n // 2
    ctx.set_data(X[:half])
    ctx.random_init(K, seed=3)
    ctx.nndescent(nnm.default_params(K=K, seed=3, max_iters=8))
    g0, _, _ = ctx.get_graph()
    init = ((np.arange(n)[:, None] + 1 + np.arange(K)[None, :]) % n).astype(np.uint32)
    init[:half] = g0
    ctx.set_data(X)
    ctx.set_partition([0, half, n], [1, 0], [0, 0])
    ctx.set_graph(init)
    p = nnm.default_params(K=K, rule=nnm.RULE_MERGE, prec=prec, seed=11, max_iters=iters, flags=flags)
    ctx.merge_init(p)
    st = ctx.nndescent(p)
    ctx.merge_finish(True)
    ids, dist, fl = ctx.get_graph()
    return ids, dist, fl, st


def _same(a, b):
    ia, da, fa, sa = a
    ib, db, fb, sb = b
    assert np.array_equal(ia, ib)
    assert np.array_equal(da.view(np.uint32), db.view(np.uint32))
    assert np.array_equal(fa, fb)
    for k in ("iters", "dists", "updates", "offers", "list_entries", "reranked"):
        assert sa[k] == sb[k], k


@pytest.mark.parametrize("prec", [nnm.PREC_FP32, nnm.PREC_BF16])
@pytest.mark.parametrize("build", ["kgraph", "jmerge"])
def test_graph_replay_matches_eager(ctx, X, prec, build):
    run = _kgraph if build == "kgraph" else _jmerge
    eager = run(ctx, X, 0, prec)
    replay = run(ctx, X, nnm.FLAG_GRAPH, prec)
    print(build, prec, "eager", eager[3], "graph", replay[3])
    assert eager[3]["graph_iters"] == 0
    assert replay[3]["graph_iters"] == replay[3]["iters"] >= 5
    _same(eager, replay)


@pytest.mark.parametrize("prec", [nnm.PREC_FP32, nnm.PREC_BF16])
def test_autotune_matches_default(ctx, X, prec):
    base = _kgraph(ctx, X, 0, prec)
    tuned = _kgraph(ctx, X, nnm.FLAG_AUTOTUNE, prec)
    print(prec, "default", base[3], "tuned", tuned[3])
    _same(base, tuned)
    # the tuned shape is kept for later calls, also under graph replay
    again = _kgraph(ctx, X, nnm.FLAG_AUTOTUNE | nnm.FLAG_GRAPH, prec)
    assert again[3]["tile_js"] == tuned[3]["tile_js"]
    assert again[3]["tile_chunk"] == tuned[3]["tile_chunk"]
    assert again[3]["graph_iters"] == again[3]["iters"]
    _same(base, again)

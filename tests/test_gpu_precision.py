"""Chosen precision: the BF16 MFMA join against the plain FP32 join.

On integer rows of magnitude <= 256 (d <= 128) the bf16 norm form is exact,
so the BF16 run must reproduce the FP32 run bit for bit (same distances
counted, same lists).  On real-valued rows the join offers lower bounds and
every candidate that reaches a list is recomputed in exact FP32: stored
distances must equal the canonical FP32 pair distance, and recall must stay
within 0.005 of the FP32 run.
"""
import numpy as np
import pytest

import nnm
import nnm_data
import orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    with nnm.Context(0) as c:
        yield c


def _run(ctx, X, prec, rule=nnm.RULE_ALL, seed=5, iters=8):
    ctx.set_data(X)
    ctx.random_init(20, seed=seed)
    st = ctx.nndescent(nnm.default_params(K=20, rule=rule, prec=prec, seed=seed, max_iters=iters))
    ids, dist, flags = ctx.get_graph()
    return ids, dist, flags, st


@pytest.mark.parametrize("d", [128, 64])
def test_bf16_equals_fp32_on_integer_rows(ctx, d):
    X = nnm_data.sift_like(20000, d, centers=64, seed=50 + d)
    i32, d32, f32, s32 = _run(ctx, X, nnm.PREC_FP32)
    i16, d16, f16, s16 = _run(ctx, X, nnm.PREC_BF16)
    assert s16["dists"] == s32["dists"] and s16["iters"] == s32["iters"]
    assert s16["reranked"] == 0
    assert np.array_equal(i16, i32) and np.array_equal(d16, d32) and np.array_equal(f16, f32)


def test_bf16_merge_equals_fp32_on_integer_rows(ctx):
    X = nnm_data.sift_like(16000, 128, centers=64, seed=61)
    n, h, K = 16000, 8000, 20
    ctx.set_data(X)
    ctx.set_partition([0, h, n], [1, 1], [0, 1])
    ctx.random_init(K, seed=2)
    ctx.nndescent(nnm.default_params(K=K, seed=2))
    sub, _, _ = ctx.get_graph()
    out = {}
    for prec in (nnm.PREC_FP32, nnm.PREC_BF16):
        ctx.set_partition([0, h, n], [1, 0], [0, 0])
        ctx.set_graph(sub)
        p = nnm.default_params(K=K, rule=nnm.RULE_MERGE, prec=prec, rho=0.5, max_iters=10, seed=3)
        ctx.merge_init(p)  # bf16: the fresh candidates' distances from rows staged in LDS
        init = ctx.get_graph()
        st = ctx.nndescent(p)
        ctx.merge_finish(True)
        out[prec] = (init, ctx.get_graph(), st)
    (i32, g32, s32), (i16, g16, s16) = out[nnm.PREC_FP32], out[nnm.PREC_BF16]
    for a, b in zip(i32, i16):
        assert np.array_equal(a, b)
    assert s32["dists"] == s16["dists"]
    for a, b in zip(g32, g16):
        assert np.array_equal(a, b)


def test_bf16_real_rows_rerank_exact(ctx):
    X = nnm_data.unit_mixture(20000, 96, centers=64, sigma=0.35, seed=62)
    rows = np.arange(0, 20000, 10, dtype=np.uint32)
    ctx.set_data(X)
    gt, _ = ctx.bruteforce(10, rows=rows)
    i32, d32, _, s32 = _run(ctx, X, nnm.PREC_FP32, iters=12)
    i16, d16, _, s16 = _run(ctx, X, nnm.PREC_BF16, iters=12)
    assert s16["reranked"] > 0
    v = np.repeat(np.arange(20000, dtype=np.uint32), 20)
    ref = ctx.pair_dists(v, i16.ravel())
    assert np.array_equal(ref, d16.ravel()), "a stored distance is not the exact FP32 value"
    # FP32 values within 1e-5 relative of FP64
    sel = np.arange(0, v.size, 37)
    ex = ((X[v[sel]].astype(np.float64) - X[i16.ravel()[sel]].astype(np.float64)) ** 2).sum(1)
    assert (np.abs(d16.ravel()[sel] - ex) <= 1e-5 * np.maximum(ex, 1e-6)).all()
    r32, r16 = nnm.recall(i32, gt, rows), nnm.recall(i16, gt, rows)
    print("fp32", r32, s32, "bf16", r16, s16)
    assert r16 >= r32 - 0.005


def test_fp64_error_report_and_near_ties(ctx):
    """nnm_graph_dist_error against an FP64 numpy recomputation; integer rows
    have no error and no near-tie decisions, real rows stay within the FP32
    accumulation bound (d + 2) * 2^-24 and resolve near-ties in FP64."""
    Xi = nnm_data.sift_like(20000, 128, centers=64, seed=61)
    _, _, _, st = _run(ctx, Xi, nnm.PREC_FP32)
    e = ctx.dist_error()
    assert e["entries"] == 20000 * 20
    assert e["max_abs"] == 0.0 and e["order_flips"] == 0
    assert st["near_ties"] == 0
    X = nnm_data.unit_mixture(20000, 96, centers=256, seed=62)
    for prec in (nnm.PREC_FP32, nnm.PREC_BF16):
        ids, dist, _, st = _run(ctx, X, prec)
        e = ctx.dist_error()
        rows = np.arange(0, 20000, 37)
        d64 = ((X[ids[rows]].astype(np.float64) - X[rows][:, None, :].astype(np.float64)) ** 2).sum(-1)
        rel = np.abs(dist[rows].astype(np.float64) - d64) / d64
        print("prec", prec, e, "sample max rel", rel.max(), "ties", st["near_ties"], st["tie_swaps"])
        assert rel.max() <= e["max_rel"] * (1 + 1e-9)
        assert e["max_rel"] <= (96 + 2) * 2.0 ** -24
        assert st["tie_swaps"] <= st["near_ties"]


def test_graph_restore_recompute_int8_rows_equal_fp32(ctx):
    """The saved lists' distances recomputed from the exact low-precision
    rows (recompute=2: the int8 rows for integer data spanning at most 255
    over 128 values) equal the FP32 recompute bit for bit, and the J-Merge
    set-up from them too."""
    X = nnm_data.sift_like(6000, 128, centers=16, seed=63)
    n, h, K = 6000, 3000, 20
    ctx.set_data(X)
    ctx.set_partition([0, h, n], [1, 1], [0, 1])
    ctx.random_init(K, seed=2)
    ctx.nndescent(nnm.default_params(K=K, seed=2, max_iters=4))
    ids, _, _ = ctx.get_graph()
    ctx.set_partition([0, h, n], [1, 0], [0, 0])
    ctx.set_graph(ids)
    ctx.graph_save()
    out = {}
    for rc in (1, 2):
        ctx.graph_restore(recompute=rc)
        restored = ctx.get_graph()
        ctx.merge_init(nnm.default_params(K=K, rule=nnm.RULE_MERGE, prec=nnm.PREC_BF16, seed=3))
        out[rc] = (restored, ctx.get_graph())
    for a, b in zip(out[1][0] + out[1][1], out[2][0] + out[2][1]):
        assert np.array_equal(a, b)
    v = np.repeat(np.arange(0, h, 37), K)
    e = out[2][0][0][:h:37].ravel()
    exact = np.array([orc.l2_seq(X[i], X[j]) for i, j in zip(v, e)], np.float32)
    assert np.array_equal(out[2][0][1][:h:37].ravel(), exact)


def test_int8_rows_padded_dimension_equal_fp32(ctx):
    """Integer rows of 100 values (row pitch 128: 28 zero-padded values in the
    int8 rows): NN-Descent on the BF16 precision (int8 joins and set-up
    distances) gives the FP32 lists bit for bit."""
    X = nnm_data.sift_like(8000, 100, centers=32, seed=64)
    out = {}
    for prec in (nnm.PREC_FP32, nnm.PREC_BF16):
        ctx.set_data(X)
        ctx.random_init(20, seed=4)
        st = ctx.nndescent(nnm.default_params(K=20, prec=prec, seed=4, max_iters=6))
        out[prec] = (ctx.get_graph(), st)
    (g32, s32), (g16, s16) = out[nnm.PREC_FP32], out[nnm.PREC_BF16]
    assert s32["dists"] == s16["dists"]
    for a, b in zip(g32, g16):
        assert np.array_equal(a, b)

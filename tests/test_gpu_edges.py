"""Edge cases of the GPU path: the smallest groups, odd row lengths,
duplicate rows, extreme reverse caps, many ragged partitions, and the
inputs the library must refuse (checked on the host before any launch)."""
import numpy as np
import pytest

import nnm
import nnm_data
import orc

pytestmark = pytest.mark.gpu

TOL = 0.005


@pytest.fixture(scope="module")
def ctx():
    with nnm.Context(0) as c:
        yield c


def _invariants(ids, dist, n_rows=None):
    n = ids.shape[0]
    s = np.sort(ids, axis=1)
    assert (np.diff(s.astype(np.int64), axis=1) != 0).all(), "duplicate ids"
    assert (ids != np.arange(n)[:, None]).all(), "self loop"
    if n_rows is not None:
        assert (ids < n_rows).all()
    d64 = dist.astype(np.float64)
    ok = (np.diff(d64, axis=1) > 0) | ((np.diff(d64, axis=1) == 0) & (np.diff(ids.astype(np.int64), axis=1) > 0))
    assert ok.all(), "rows not sorted by (dist, id)"


@pytest.mark.parametrize("n,K", [(2, 1), (9, 8), (33, 32)])
def test_smallest_groups_give_exact_lists(ctx, n, K):
    """n = K + 1: every list holds all other rows, in (dist, id) order."""
    rng = np.random.default_rng(n)
    X = rng.integers(0, 50, size=(n, 16)).astype(np.float32)
    ctx.set_data(X)
    ctx.random_init(K, seed=1)
    ctx.nndescent(nnm.default_params(K=K, seed=1, max_iters=3))
    ids, dist, _ = ctx.get_graph()
    gi, gd = ctx.bruteforce(K)
    assert np.array_equal(ids, gi) and np.array_equal(dist, gd)
    _invariants(ids, dist, n)


@pytest.mark.parametrize("d", [1, 3, 100, 200])
def test_odd_row_lengths(ctx, d):
    """Rows padded to a multiple of 32 floats: distances exact on integer
    rows; mean recall over three seeds within 0.005 of the oracle's (the
    reference seeds from the clock, so parity is statistical)."""
    X = nnm_data.sift_like(4000, d, centers=16, sigma=20.0, seed=50 + d)
    K = 10
    rows = np.arange(0, 4000, 4, dtype=np.uint32)
    gt, _ = orc.bruteforce(X, 10, rows=rows)
    ctx.set_data(X)
    r_gpu, r_cpu = [], []
    for seed in (2, 3, 4):
        ctx.random_init(K, seed=seed)
        ctx.nndescent(nnm.default_params(K=K, seed=seed))  # until the delta test, as the reference runs
        ids, dist, _ = ctx.get_graph()
        _invariants(ids, dist, 4000)
        for v in range(0, 4000, 211):
            for j in range(K):
                assert dist[v, j] == np.float32(orc.l2_f64(X[v], X[ids[v, j]]))
        r_gpu.append(nnm.recall(ids, gt, rows, 10))
        oi, _, _ = orc.kgraph(X, K=K, seed=seed)
        r_cpu.append(orc.recall(oi, rows, gt, 10))
    print("d", d, "gpu", r_gpu, "cpu", r_cpu)
    assert np.mean(r_gpu) >= np.mean(r_cpu) - TOL


def test_duplicate_rows(ctx):
    """Every row five times over: distance ties everywhere, broken by id; the
    lists never repeat an id or hold the row itself."""
    base = nnm_data.sift_like(600, 32, centers=8, seed=61)
    X = np.repeat(base, 5, axis=0)
    K = 8
    ctx.set_data(X)
    rows = np.arange(0, X.shape[0], 7, dtype=np.uint32)
    gi, gd = ctx.bruteforce(K, rows=rows)
    oi, od = orc.bruteforce(X, K, rows=rows)
    assert np.array_equal(gi, oi) and np.array_equal(gd.astype(np.float64), od)
    ctx.random_init(K, seed=3)
    for prec in (nnm.PREC_FP32, nnm.PREC_BF16):
        ctx.random_init(K, seed=3)
        ctx.nndescent(nnm.default_params(K=K, seed=3, max_iters=12, prec=prec))
        ids, dist, _ = ctx.get_graph()
        _invariants(ids, dist, X.shape[0])
        # the four copies of a row are at distance 0: they lead every list
        assert (dist[:, :4] == 0).mean() > 0.9


@pytest.mark.parametrize("rev_cap", [1, 300])
def test_reverse_cap_extremes(ctx, rev_cap):
    """rev_cap 1 (almost no reverse sampling) and 300 (lists longer than the
    long-list window, joined by the windows kernel in FP32).  The reference
    always samples 100."""
    X = nnm_data.sift_like(8000, 64, centers=16, sigma=20.0, seed=71)
    K = 16
    tol = TOL
    rows = np.arange(0, 8000, 8, dtype=np.uint32)
    gt, _ = orc.bruteforce(X, 10, rows=rows)
    oi, _, _ = orc.kgraph(X, K=K, rev_cap=rev_cap, seed=4)
    r_cpu = orc.recall(oi, rows, gt, 10)
    ctx.set_data(X)
    for prec in (nnm.PREC_FP32, nnm.PREC_BF16):
        ctx.random_init(K, seed=4)
        st = ctx.nndescent(nnm.default_params(K=K, rev_cap=rev_cap, seed=4, prec=prec))
        ids, dist, _ = ctx.get_graph()
        _invariants(ids, dist, 8000)
        r_gpu = nnm.recall(ids, gt, rows, 10)
        print("rev_cap", rev_cap, "prec", prec, "gpu", r_gpu, "iters", st["iters"], "cpu", r_cpu, "long",
              st["long_entries"])
        assert r_gpu >= r_cpu - tol


def test_many_ragged_partitions(ctx):
    """A P-way merge of 40 built subgraphs of uneven sizes (17 .. 400 rows)."""
    rng = np.random.default_rng(81)
    sizes = rng.integers(17, 400, size=40)
    n = int(sizes.sum())
    X = nnm_data.sift_like(n, 48, centers=16, sigma=20.0, seed=82)
    K = 16
    b = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint32)
    P = sizes.size
    ctx.set_data(X)
    ctx.set_partition(b, [1] * P, np.arange(P))
    ctx.random_init(K, seed=5)
    ctx.nndescent(nnm.default_params(K=K, rule=nnm.RULE_ALL, seed=5, max_iters=12))
    sub, _, _ = ctx.get_graph()
    for s in range(P):
        blk = sub[b[s]:b[s + 1]]
        assert (blk >= b[s]).all() and (blk < b[s + 1]).all()
    ctx.set_partition(b, [1] * P, [0] * P)
    ctx.set_graph(sub)
    p = nnm.default_params(K=K, rule=nnm.RULE_MERGE, seed=6, max_iters=15)
    ctx.merge_init(p)
    ctx.nndescent(p)
    ctx.merge_finish(True)
    ids, dist, _ = ctx.get_graph()
    _invariants(ids, dist, n)
    rows = np.arange(0, n, 5, dtype=np.uint32)
    gt, _ = orc.bruteforce(X, 10, rows=rows)
    oi, _, _ = orc.merge(X, b, [1] * P, sub, K=K, seed=6, max_iters=15)
    r_gpu, r_cpu = nnm.recall(ids, gt, rows, 10), orc.recall(oi, rows, gt, 10)
    print("ragged P-merge gpu", r_gpu, "cpu", r_cpu)
    assert r_gpu >= r_cpu - TOL


@pytest.mark.parametrize("K", [31, 32])
def test_merge_init_rows_both_widths(ctx, K):
    """J-Merge set-up at K = 31 (two rows per wave, K + 1 draws fill a half)
    and K = 32 (one row per wave).  Built rows keep their K - cut nearest
    entries and take the rest from the raw half when fresh ones exist; raw
    rows draw K distinct others; flags of new entries set; distances exact.
    The merge then reaches the oracle's recall."""
    n, half = 3000, 1500
    X = nnm_data.sift_like(n, 64, centers=16, sigma=20.0, seed=95 + K)
    ctx.set_data(X[:half])
    ctx.random_init(K, seed=3)
    ctx.nndescent(nnm.default_params(K=K, seed=3, max_iters=12))
    g0, d0, _ = ctx.get_graph()
    init = ((np.arange(n)[:, None] + 1 + np.arange(K)[None, :]) % n).astype(np.uint32)
    init[:half] = g0
    ctx.set_data(X)
    ctx.set_partition([0, half, n], [1, 0], [0, 0])
    ctx.set_graph(init)
    p = nnm.default_params(K=K, rule=nnm.RULE_MERGE, seed=6, max_iters=12)
    ctx.merge_init(p)
    ids, dist, flags = ctx.get_graph()
    _invariants(ids, dist, n)
    cut = K // 2
    remain = K - cut
    for v in range(0, half, 7):
        kept = set(g0[v, :remain].tolist())
        row = set(ids[v].tolist())
        assert kept <= row, f"built row {v} lost a near entry"
        extra = row - kept
        far = set(g0[v, remain:].tolist())
        assert all(u >= half or u in far for u in extra), f"built row {v}: entry from nowhere"
        assert any(u >= half for u in extra), f"built row {v}: no fresh entry"
    assert flags[half:].all()
    v = np.repeat(np.arange(0, n, 11), K)
    e = ids[::11].ravel()
    exact = np.array([orc.l2_seq(X[i], X[j]) for i, j in zip(v, e)], np.float32)
    assert np.array_equal(dist[::11].ravel(), exact)
    ctx.nndescent(p)
    ctx.merge_finish(True)
    ids, dist, _ = ctx.get_graph()
    _invariants(ids, dist, n)
    rows = np.arange(0, n, 3, dtype=np.uint32)
    gt, _ = orc.bruteforce(X, 10, rows=rows)
    oi, _, _ = orc.merge(X, [0, half, n], [1, 0], init, K=K, seed=6, max_iters=12)
    r_gpu, r_cpu = nnm.recall(ids, gt, rows, 10), orc.recall(oi, rows, gt, 10)
    print("K", K, "merge gpu", r_gpu, "cpu", r_cpu)
    assert r_gpu >= r_cpu - TOL


def test_refused_inputs(ctx):
    """Inputs the host must refuse before launching anything."""
    X = nnm_data.sift_like(200, 16, centers=4, seed=91)
    ctx.set_data(X)
    with pytest.raises(nnm.NNMError):
        ctx.random_init(33, seed=1)  # K above 32
    with pytest.raises(nnm.NNMError):
        ctx.random_init(0, seed=1)
    ctx.set_partition([0, 10, 200], [1, 1], [0, 1])
    with pytest.raises(nnm.NNMError):
        ctx.random_init(10, seed=1)  # a group of 10 rows cannot hold K + 1 = 11
    ctx.set_partition([0, 200], [1], [0])
    ctx.random_init(8, seed=1)
    for bad in (dict(rev_cap=0), dict(offer_cap=2), dict(rho=0.0), dict(rho=1.5), dict(rule=7), dict(prec=9)):
        with pytest.raises(nnm.NNMError):
            ctx.nndescent(nnm.default_params(K=8, max_iters=2, **bad))
    with pytest.raises(nnm.NNMError):
        ctx.nndescent(nnm.default_params(K=9, max_iters=2))  # K differs from the graph's
    with pytest.raises(nnm.NNMError):
        ctx.bruteforce(33)

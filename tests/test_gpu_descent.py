"""GPU NN-Descent and merges against the CPU restatement of the reference.

Parity is statistical, as the reference's own runs are (it seeds with
time(NULL) and shuffles with rand()): merged-graph recall@10 against exact
brute force must be within 0.005 of the CPU restatement's on the same seeded
input and the same starting subgraphs.  Graph invariants are exact.
"""
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


def _invariants(ids, dist, K):
    n = ids.shape[0]
    assert ids.shape == (n, K)
    s = np.sort(ids, axis=1)
    assert (np.diff(s.astype(np.int64), axis=1) != 0).all(), "duplicate ids"
    assert (ids != np.arange(n)[:, None]).all(), "self loop"
    d64 = dist.astype(np.float64)
    ok = (np.diff(d64, axis=1) > 0) | ((np.diff(d64, axis=1) == 0) & (np.diff(ids.astype(np.int64), axis=1) > 0))
    assert ok.all(), "rows not sorted by (dist, id)"


@pytest.fixture(scope="module")
def data20k():
    X = nnm_data.sift_like(20000, 128, centers=64, sigma=20.0, seed=21)
    rows = np.arange(0, 20000, 10, dtype=np.uint32)
    gt, _ = orc.bruteforce(X, 10, rows=rows)
    return X, rows, gt


def test_kgraph_recall_parity(ctx, data20k):
    X, rows, gt = data20k
    K = 20
    ctx.set_data(X)
    ctx.random_init(K, seed=7)
    p = nnm.default_params(K=K, rule=nnm.RULE_ALL, seed=7, max_iters=30)
    st = ctx.nndescent(p)
    ids, dist, _ = ctx.get_graph()
    _invariants(ids, dist, K)
    r_gpu = nnm.recall(ids, gt, rows, 10)
    oi, _, ost = orc.kgraph(X, K=K, seed=7, max_iters=30)
    r_cpu = orc.recall(oi, rows, gt, 10)
    print("kgraph", st, "gpu", r_gpu, "cpu", r_cpu, ost)
    assert st["dists"] > 0 and st["iters"] >= 2
    assert r_gpu >= r_cpu - TOL


def _build_halves(ctx, X, K, seed, parts):
    n = X.shape[0]
    b = np.linspace(0, n, parts + 1).astype(np.int64)
    ctx.set_data(X)
    ctx.set_partition(b, [1] * parts, np.arange(parts))
    ctx.random_init(K, seed=seed)
    st = ctx.nndescent(nnm.default_params(K=K, rule=nnm.RULE_ALL, seed=seed, max_iters=30))
    ids, dist, _ = ctx.get_graph()
    for i in range(parts):
        blk = ids[b[i]:b[i + 1]]
        assert (blk >= b[i]).all() and (blk < b[i + 1]).all(), "subgraph build crossed its group"
    return b, ids, st


def test_smerge_recall_parity(ctx, data20k):
    X, rows, gt = data20k
    K = 20
    b, sub_ids, _ = _build_halves(ctx, X, K, 3, 2)
    r_sub = nnm.recall(sub_ids, gt, rows, 10)
    # merge on the GPU from the GPU subgraphs
    ctx.set_partition(b, [1, 1], [0, 0])
    ctx.set_graph(sub_ids)
    p = nnm.default_params(K=K, rule=nnm.RULE_MERGE, seed=9, max_iters=30)
    ist = ctx.merge_init(p)
    st = ctx.nndescent(p)
    ctx.merge_finish(True)
    ids, dist, _ = ctx.get_graph()
    _invariants(ids, dist, K)
    r_gpu = nnm.recall(ids, gt, rows, 10)
    # CPU restatement from the same subgraphs
    oi, _, ost = orc.merge(X, b, [1, 1], sub_ids, K=K, seed=9, max_iters=30)
    r_cpu = orc.recall(oi, rows, gt, 10)
    print("smerge sub", r_sub, "gpu", r_gpu, st, ist, "cpu", r_cpu, ost)
    assert r_gpu >= r_cpu - TOL


def _jmerge(ctx, X, K, offer_cap=0, prec=nnm.PREC_FP32):
    n = X.shape[0]
    half = n // 2
    ctx.set_data(X[:half])
    ctx.random_init(K, seed=4)
    ctx.nndescent(nnm.default_params(K=K, seed=4, max_iters=30))
    g0, _, _ = ctx.get_graph()
    init = np.zeros((n, K), np.uint32)
    init[:half] = g0
    ctx.set_data(X)
    ctx.set_partition([0, half, n], [1, 0], [0, 0])
    dummy = ((np.arange(n)[:, None] + 1 + np.arange(K)[None, :]) % n).astype(np.uint32)
    init[half:] = dummy[half:]
    ctx.set_graph(init)
    p = nnm.default_params(K=K, rule=nnm.RULE_MERGE, seed=6, max_iters=30, offer_cap=offer_cap, prec=prec)
    ctx.merge_init(p)
    st = ctx.nndescent(p)
    ctx.merge_finish(True)
    ids, dist, _ = ctx.get_graph()
    return ids, dist, st, init


def test_jmerge_recall_parity(ctx, data20k):
    X, rows, gt = data20k
    K = 20
    n = X.shape[0]
    half = n // 2
    ids, dist, st, init = _jmerge(ctx, X, K)
    _invariants(ids, dist, K)
    r_gpu = nnm.recall(ids, gt, rows, 10)
    oi, _, ost = orc.merge(X, [0, half, n], [1, 0], init, K=K, seed=6, max_iters=30)
    r_cpu = orc.recall(oi, rows, gt, 10)
    print("jmerge gpu", r_gpu, st, "cpu", r_cpu, ost)
    assert st["redo_rows"] == 0
    assert r_gpu >= r_cpu - TOL


@pytest.mark.parametrize("prec", [nnm.PREC_FP32, nnm.PREC_BF16])
def test_offer_log_overflow_is_exact(ctx, data20k, prec):
    """Every offer reaches its owner however small the offer log: with room
    for 4 offers per row the join runs out of log chunks every iteration and
    the left vertices are joined again in redo rounds.  The merged lists
    (ids and distance bits) and the distance count equal the run whose log
    never fills, and both meet the oracle's recall to 0.005."""
    X, rows, gt = data20k
    K = 20
    n = X.shape[0]
    half = n // 2
    ids0, dist0, st0, init = _jmerge(ctx, X, K, prec=prec)
    ids1, dist1, st1, _ = _jmerge(ctx, X, K, offer_cap=4, prec=prec)
    print("full log", st0["redo_rows"], st0["dists"], "small log", st1["redo_rows"], st1["redo_rounds"], st1["dists"])
    assert st0["redo_rows"] == 0
    assert st1["redo_rows"] > 0 and st1["redo_rounds"] > 0
    assert np.array_equal(ids0, ids1)
    assert np.array_equal(dist0.view(np.uint32), dist1.view(np.uint32))
    assert st0["dists"] == st1["dists"] and st0["iters"] == st1["iters"]
    assert st0["updates"] == st1["updates"] and st0["last_delta"] == st1["last_delta"]
    oi, _, _ = orc.merge(X, [0, half, n], [1, 0], init, K=K, seed=6, max_iters=30)
    r_cpu = orc.recall(oi, rows, gt, 10)
    assert nnm.recall(ids1, gt, rows, 10) >= r_cpu - TOL


def test_batched_builds_match_single(ctx):
    """Two subgraphs built in one launch give the same recall as separate builds."""
    X = nnm_data.sift_like(8000, 128, centers=32, seed=22)
    K = 16
    b, ids, st = _build_halves(ctx, X, K, 5, 4)
    for i in range(4):
        Xi = X[b[i]:b[i + 1]]
        rows = np.arange(0, Xi.shape[0], 7, dtype=np.uint32)
        gt, _ = orc.bruteforce(Xi, 10, rows=rows)
        r = nnm.recall(ids[b[i]:b[i + 1]] - b[i], gt, rows, 10)
        assert r > 0.97


def test_single_iteration_counts(ctx):
    """nCmps of one iteration from random lists equals the closed-form pair
    count C(n_new, 2) + n_new * n_old summed over the rows, when every list
    entry is new (rho = 1) and every pair passes the rule."""
    X = nnm_data.sift_like(600, 128, centers=4, seed=23)
    ctx.set_data(X)
    ctx.random_init(10, seed=2)
    p = nnm.default_params(K=10, rule=nnm.RULE_ALL, seed=2, rev_cap=1000)
    ids, dist, _ = ctx.get_graph()
    st = ctx.local_join(p, 0)
    # expected lists: all K entries are new (rho = 1); reverse entries v -> u
    # when u's k-th distance is below d(v, u); no old entries yet
    n = 600
    kth = dist[:, -1]
    n_new = np.full(n, 10, np.int64)
    for v in range(n):
        for e in range(10):
            u = ids[v, e]
            if kth[u] < dist[v, e]:
                n_new[u] += 1
    assert st["list_entries"] == int(n_new.sum())
    assert st["dists"] == int((n_new * (n_new - 1) // 2).sum())
    assert st["iters"] == 1


@pytest.mark.parametrize("cap", [1000, 100])
def test_hub_bucket_overflow_counts(ctx, cap):
    """A hub: every row lists row 0, whose reverse bucket then holds far more
    entries than the 64 a bucket keeps (the rest go through the overflow
    pass).  The joined list sizes and pair counts equal the closed form: all
    reverse entries under a large cap, exactly cap of them under cap 100."""
    n, K = 600, 10
    X = nnm_data.sift_like(n, 32, centers=4, seed=61)
    gi, _ = orc.bruteforce(X, K)
    ids = gi.astype(np.uint32).copy()
    for v in range(1, n):
        if 0 not in ids[v]:
            ids[v, -1] = 0
    ctx.set_data(X)
    ctx.set_graph(ids)
    ids, dist, _ = ctx.get_graph()
    kth = dist[:, -1]
    rev = np.zeros(n, np.int64)
    for v in range(n):
        for e in range(K):
            u = ids[v, e]
            if kth[u] < dist[v, e]:
                rev[u] += 1
    assert rev[0] > 64 + 100
    p = nnm.default_params(K=K, rule=nnm.RULE_ALL, seed=2, rev_cap=cap)
    st = ctx.local_join(p, 0)
    n_new = K + np.minimum(rev, cap)
    assert st["list_entries"] == int(n_new.sum())
    assert st["dists"] == int((n_new * (n_new - 1) // 2).sum())


@pytest.mark.parametrize("d", [128, 96, 64, 200])
def test_join_distances_exact(ctx, d):
    """Every distance the join writes into a list equals the library's
    canonical pair distance bit for bit (checks the row staging of every
    join path: long and short lists, several row lengths)."""
    X = nnm_data.sift_like(3000, d, centers=8, seed=40 + d)
    ctx.set_data(X)
    ctx.random_init(20, seed=3)
    p = nnm.default_params(K=20, rule=nnm.RULE_ALL, seed=3, rev_cap=100)
    for it in range(3):
        ctx.local_join(p, it)
    ids, dist, _ = ctx.get_graph()
    v = np.repeat(np.arange(3000, dtype=np.uint32), 20)
    ref = ctx.pair_dists(v, ids.ravel())
    assert np.array_equal(ref, dist.ravel())
    # integer rows: the canonical order is exact
    ex = ((X[v].astype(np.float64) - X[ids.ravel()].astype(np.float64)) ** 2).sum(1)
    assert np.array_equal(ex, dist.ravel().astype(np.float64))


def _same_up_to_ties(gi, gd, oi, od):
    """Equal distance rows, and equal ids wherever the distance is below the
    row's last one (entries tied at the cut may differ: the library orders
    ties by id, merge2one's two-way merge puts the pool first)."""
    assert np.array_equal(gd, od)
    for v in range(gi.shape[0]):
        m = od[v] < od[v, -1]
        assert sorted(gi[v][m].tolist()) == sorted(oi[v][m].tolist()), f"row {v}"


@pytest.mark.parametrize("dedup", [True, False])
def test_merge_finish_matches_merge2one(ctx, dedup):
    """merge2one (src/abstractmerge.cpp:232-277), with and without the
    documented de-dup: the parked far entries of each built list (cut = K/2
    of them) are merged back with the final pool by (dist, id), parked ids
    already in the pool dropped under de-dup.  The expected rows come from
    the oracle's merge2one on the pool read before the finish and the parked
    entries read at set-up, and from a restatement here.  The descent joins
    every pair, so pools re-find many parked ids and the de-dup path runs."""
    X = nnm_data.sift_like(6000, 64, centers=8, seed=21)
    K, cut = 16, 8
    b, sub, _ = _build_halves(ctx, X, K, 5, 2)
    ctx.set_partition(b, [1, 1], [0, 0])
    ctx.set_graph(sub)
    li, ld, _ = ctx.get_graph()  # sorted, exact distances
    p = nnm.default_params(K=K, rule=nnm.RULE_MERGE, seed=8, max_iters=6, cut=cut)
    ctx.merge_init(p)
    # pairs inside a built subgraph are joined too, so pools can re-find
    # parked ids (under the merge rule they never can)
    ctx.nndescent(nnm.default_params(K=K, rule=nnm.RULE_ALL, seed=8, max_iters=6, cut=cut))
    pi, pd, _ = ctx.get_graph()
    ctx.merge_finish(dedup)
    gi, gd, _ = ctx.get_graph()
    park_i, park_d = li[:, K - cut:], ld[:, K - cut:]
    oi, od = orc.merge2one(pi, pd, park_i, park_d, dedup)
    _same_up_to_ties(gi, gd, oi, od)
    if dedup:
        _invariants(gi, gd, K)
    overlap = 0
    for v in range(X.shape[0]):
        keep = ~np.isin(park_i[v], pi[v]) if dedup else np.ones(cut, bool)
        overlap += int(np.isin(park_i[v], pi[v]).sum())
        ids = np.concatenate([pi[v], park_i[v][keep]])
        ds = np.concatenate([pd[v], park_d[v][keep]])
        o = np.lexsort((ids, ds))[:K]
        assert np.array_equal(gi[v], ids[o]) and np.array_equal(gd[v], ds[o]), f"row {v}"
    print("parked ids re-found by the pool:", overlap)
    assert overlap > 1000
    if not dedup:
        assert any(len(set(r.tolist())) < K for r in gi), "no repeated id without de-dup"


def _complete(K, d=16, seed=0):
    """n = K + 1 random points, every list holding all other rows."""
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((K + 1, d)).astype(np.float32)
    ids = np.array([[u for u in range(K + 1) if u != v] for v in range(K + 1)], np.uint32)
    return X, ids


@pytest.mark.parametrize("K,a", [(20, 9), (16, 1), (12, 12)])
def test_complete_graph_pair_counts_match_oracle(ctx, K, a):
    """Complete graph on K + 1 rows, all entries new, rho = 1: no reverse
    entries, so one iteration joins C(K, 2) pairs per row under KGraph's
    rule, a*b*(a+b-2) under the merge rule with two built subgraphs of a and
    b rows (S-Merge), and a(C(K,2)-C(a-1,2)) + b(C(K,2)-C(a,2)) with one
    built subgraph and raw rows (J-Merge); the oracle gives the same counts."""
    X, ids = _complete(K, seed=K + a)
    n, b = K + 1, K + 1 - a
    c2 = lambda m: m * (m - 1) // 2
    ones = np.ones_like(ids, np.uint8)
    RAW = orc.RAW
    cases = [(nnm.RULE_ALL, [0, n], [0], [RAW] * n, n * c2(K)),
             (nnm.RULE_MERGE, [0, a, n], [1, 1], [0] * a + [1] * b, a * b * (a + b - 2)),
             (nnm.RULE_MERGE, [0, a, n], [1, 0], [0] * a + [RAW] * b,
              a * (c2(K) - c2(a - 1)) + b * (c2(K) - c2(a)))]
    ctx.set_data(X)
    for rule, sb, built, gid, want in cases:
        ctx.set_partition(sb, built, [0] * len(built))
        ctx.set_graph(ids, flags=ones)
        st = ctx.local_join(nnm.default_params(K=K, rule=rule, seed=1), 0)
        _, _, _, ost = orc.nndescent_lists(X, ids, ones, gid, rule_all=rule == nnm.RULE_ALL, max_iters=1)
        assert st["dists"] == want == ost["dists"], (rule, built, st["dists"], want, ost["dists"])


@pytest.mark.parametrize("K,rho", [(20, 0.5), (16, 0.25)])
def test_delta_closed_form_matches_oracle(ctx, K, rho):
    """Forward sampling takes round(rho K) new entries a row; on the complete
    graph nothing changes, so after one iteration delta = (K - S) / K, the
    reported delta equals the flags left in the lists over N K, and the run
    stops after ceil(K / S) iterations with C(K, 2) pairs a row in all, as
    in the oracle."""
    X, ids = _complete(K, seed=3)
    n = K + 1
    S = int(round(rho * K))
    ones = np.ones_like(ids, np.uint8)
    ctx.set_data(X)
    ctx.set_partition([0, n], [0], [0])
    ctx.set_graph(ids, flags=ones)
    st = ctx.nndescent(nnm.default_params(K=K, rule=nnm.RULE_ALL, seed=1, rho=rho, max_iters=1))
    _, _, f = ctx.get_graph()
    _, _, of, ost = orc.nndescent_lists(X, ids, ones, [orc.RAW] * n, rule_all=True, rho=rho, max_iters=1)
    assert st["iters"] == 1 and st["converged"] == 0
    assert st["last_delta"] == f.sum() / (n * K)
    assert abs(st["last_delta"] - (K - S) / K) < 1e-12
    assert abs(st["last_delta"] - ost["last_delta"]) < 1e-7
    assert st["dists"] == ost["dists"] == n * S * (S - 1) // 2
    assert (f.sum(1) == of.sum(1)).all()
    ctx.set_graph(ids, flags=ones)
    st = ctx.nndescent(nnm.default_params(K=K, rule=nnm.RULE_ALL, seed=1, rho=rho))
    _, _, _, ost = orc.nndescent_lists(X, ids, ones, [orc.RAW] * n, rule_all=True, rho=rho)
    assert st["iters"] == ost["iters"] == -(-K // S)
    assert st["converged"] == ost["converged"] == 1
    assert st["dists"] == ost["dists"] == n * K * (K - 1) // 2
    assert st["last_delta"] == 0.0


def test_last_delta_is_flag_fraction(ctx, data20k):
    """On real data the reported delta after a capped run is the fraction of
    entries still flagged new in the lists the run leaves."""
    X, _, _ = data20k
    K = 20
    ctx.set_data(X)
    ctx.random_init(K, seed=12)
    for iters in (2, 1):
        st = ctx.nndescent(nnm.default_params(K=K, rule=nnm.RULE_ALL, seed=12, max_iters=iters))
        _, _, f = ctx.get_graph()
        assert st["last_delta"] == f.astype(np.int64).sum() / (X.shape[0] * K)
        assert 0.0 < st["last_delta"] < 1.0

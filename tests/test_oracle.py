"""Pins the CPU oracle (oracle/oracle.cpp) before the tests trust it:
against the reference's own distance routines compiled from its
sources (oracle/_ref), the DistOpt::test vectors (src/distopt.cpp:208-215),
FP64 closed forms, brute force on tiny inputs and list invariants."""
import ctypes as C

import numpy as np
import pytest

import orc
import nnm_data


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def test_distopt_known_answer():
    # DistOpt::test prints L2^2 of these two 6-d vectors (src/distopt.cpp:210-213)
    v1 = np.array([1.2, 4.1, 6.1, 2.2, 4.5, 7.1], np.float32)
    v2 = np.array([2.2, 3.1, 5.1, 3.2, 5.5, 3.3], np.float32)
    exact = float(np.sum((v1.astype(np.float64) - v2.astype(np.float64)) ** 2))
    assert abs(orc.l2_f64(v1, v2) - exact) < 1e-12
    assert abs(orc.l2_simd4(v1, v2) - exact) / exact < 1e-6
    assert abs(orc.l2_seq(v1, v2) - exact) / exact < 1e-6
    assert abs(exact - 19.44) < 1e-5  # 1+1+1+1+1+14.44


@pytest.mark.skipif(not orc.ref_built(), reason="reference sources not built (oracle/_ref)")
def test_restatement_bit_equal_to_reference_distances():
    R = orc.ref_lib()
    rng = np.random.default_rng(0)
    for d in (1, 3, 4, 5, 6, 17, 96, 128, 131):
        for _ in range(40):
            a = rng.standard_normal(d).astype(np.float32) * rng.uniform(0.01, 100)
            b = rng.standard_normal(d).astype(np.float32) * rng.uniform(0.01, 100)
            assert np.float32(orc.l2_seq(a, b)) == np.float32(R.ref_l2norm(_fp(a), 0, _fp(b), 0, d))
            assert np.float32(orc.l2_simd4(a, b)) == np.float32(R.ref_l2sqr_simd_ext(_fp(a), 0, _fp(b), 0, d))


@pytest.mark.skipif(not orc.ref_built(), reason="reference sources not built (oracle/_ref)")
def test_reference_row_indexing():
    R = orc.ref_lib()
    X = nnm_data.sift_like(10, 16, centers=2, seed=1)
    for i, j in [(0, 1), (3, 7), (9, 2)]:
        assert R.ref_l2norm(_fp(X), i, _fp(X), j, 16) == orc.l2_seq(X[i], X[j])


def test_integer_rows_exact_in_any_order():
    X = nnm_data.sift_like(50, 128, centers=4, seed=2)
    for i in range(0, 50, 7):
        for j in range(1, 50, 11):
            e = orc.l2_f64(X[i], X[j])
            assert orc.l2_seq(X[i], X[j]) == e and orc.l2_simd4(X[i], X[j]) == e


def test_bruteforce_tiny_against_itertools():
    rng = np.random.default_rng(1)
    X = rng.integers(0, 4, size=(12, 3)).astype(np.float32)  # many ties
    ids, dist = orc.bruteforce(X, 5)
    for q in range(12):
        cand = sorted((float(np.sum((X[q].astype(np.float64) - X[r]) ** 2)), r) for r in range(12) if r != q)
        assert [c[1] for c in cand[:5]] == ids[q].tolist()
        assert [c[0] for c in cand[:5]] == dist[q].tolist()


def _brute_update(ids, dists, new_id, new_dist):
    # the reference rule restated as a set operation: reject when the k-th
    # distance is <= d or an equal-distance entry has the same id; otherwise
    # insert after every entry with dist <= d and drop the last
    K = len(ids)
    if dists[-1] <= new_dist:
        return 0, ids, dists
    if any(i == new_id and d == new_dist for i, d in zip(ids, dists)):
        return 0, ids, dists
    pos = sum(1 for d in dists if d <= new_dist)
    ids = ids[:pos] + [new_id] + ids[pos:]
    dists = dists[:pos] + [new_dist] + dists[pos:]
    return 1, ids[:K], dists[:K]


def test_update_knn_list_against_set_rule():
    rng = np.random.default_rng(2)
    for trial in range(300):
        K = int(rng.integers(1, 8))
        ids = list(map(int, rng.choice(50, K, replace=False)))
        dists = sorted(map(float, rng.integers(0, 10, K).astype(np.float32)))
        flags = [0] * K
        for _ in range(10):
            nid = int(rng.integers(0, 50))
            nd = float(np.float32(rng.integers(0, 12)))
            r, gi, gd, gf = orc.update_knn_list(ids, dists, flags, nid, nd)
            er, ei, ed = _brute_update(ids, dists, nid, nd)
            assert r == er and gi.tolist() == ei and gd.tolist() == ed
            ids, dists, flags = gi.tolist(), gd.tolist(), gf.tolist()
            assert dists == sorted(dists)


def test_get_rnd_sds_distinct_and_in_range():
    for seed in range(50):
        for (k, N) in [(21, 50000), (20, 20), (41, 45), (1, 2), (10, 11)]:
            s = orc.get_rnd_sds(seed, k, N)
            assert len(set(s.tolist())) == k
            assert (s < N).all()


def _mt_stream(seed):
    # numpy's MT19937 with the legacy integer seeding is std::mt19937(seed)
    # (init_genrand); random_raw yields its 32-bit outputs in order
    bg = np.random.MT19937()
    bg._legacy_seeding(seed)
    while True:
        for x in bg.random_raw(64):
            yield int(x)


def _rnd_sds_numpy(seed, k, N):
    # getRndSds as written at src/abstractgraph.cpp:7-35, on numpy's stream
    if N == k:
        return list(range(k))
    g = _mt_stream(seed)
    a = sorted(next(g) % (N - k) for _ in range(k))
    for i in range(1, k):
        if a[i] <= a[i - 1]:
            a[i] = a[i - 1] + 1
    off = next(g) % N
    return [(x + off) % N for x in a]


def test_get_rnd_sds_against_numpy_mt19937():
    # std::mt19937 with seed 5489 outputs 3499211612 first (C++ standard)
    assert next(_mt_stream(5489)) == 3499211612
    rng = np.random.default_rng(5)
    for _ in range(400):
        seed = int(rng.integers(0, 2**32))
        N = int(rng.integers(2, 100000))
        k = int(rng.integers(1, min(N, 70) + 1))
        assert orc.get_rnd_sds(seed, k, N).tolist() == _rnd_sds_numpy(seed, k, N), (seed, k, N)


def test_kgraph_small_exact():
    # a handful of points: NN-Descent reaches the exact graph
    X = nnm_data.sift_like(60, 16, centers=3, seed=3)
    ids, dist, st = orc.kgraph(X, K=5, seed=1)
    gt, gd = orc.bruteforce(X, 5)
    assert np.allclose(dist, gd.astype(np.float32))
    assert st["converged"] == 1


def test_kgraph_and_merges_reach_high_recall():
    X = nnm_data.sift_like(4000, 128, centers=16, seed=4)
    rows = np.arange(0, 4000, 4, dtype=np.uint32)
    gt, _ = orc.bruteforce(X, 10, rows=rows)
    ids, _, st = orc.kgraph(X, K=20, seed=2)
    assert orc.recall(ids, rows, gt, 10) > 0.98
    h = 2000
    g0, _, _ = orc.kgraph(X[:h], K=20, seed=3)
    g1, _, _ = orc.kgraph(X[h:], K=20, seed=4)
    init = np.vstack([g0, g1 + h])
    sub = orc.recall(init, rows, gt, 10)
    s_ids, _, sst = orc.merge(X, [0, h, 4000], [1, 1], init, K=20, seed=5)
    j_ids, _, jst = orc.merge(X, [0, h, 4000], [1, 0], init, K=20, seed=6)
    assert sub < 0.8
    assert orc.recall(s_ids, rows, gt, 10) > 0.98
    assert orc.recall(j_ids, rows, gt, 10) > 0.98
    # the S-merge join only evaluates cross-subgraph pairs
    assert sst["dists"] < st["dists"]


def _complete(K, d=8, seed=0):
    """n = K + 1 random points, every list holding all other rows."""
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((K + 1, d)).astype(np.float32)
    ids = np.array([[u for u in range(K + 1) if u != v] for v in range(K + 1)], np.uint32)
    return X, ids


def _c2(m):
    return m * (m - 1) // 2


@pytest.mark.parametrize("K,a", [(20, 9), (16, 1), (12, 12)])
def test_one_iteration_pair_counts_closed_form(K, a):
    # complete graph on n = K + 1 rows, every entry new, rho = 1: no reverse
    # entry exists (each row's k-th distance reaches every row), so row v
    # joins its K entries pairwise.  KGraph's rule keeps all C(K, 2) pairs;
    # the merge rule drops the pairs inside one built subgraph.  Every
    # candidate is already listed, so the iteration changes nothing and
    # the run stops after it with delta 0.
    X, ids = _complete(K, seed=K + a)
    n, b = K + 1, K + 1 - a
    flags = np.ones_like(ids, np.uint8)
    RAW = orc.RAW
    _, _, f, st = orc.nndescent_lists(X, ids, flags, [RAW] * n, rule_all=True)
    assert st["dists"] == n * _c2(K) and st["iters"] == 1 and st["converged"] == 1
    assert st["last_delta"] == 0.0 and not f.any()
    # S-Merge: two built subgraphs [0, a) and [a, n): a*b*(a+b-2) pairs
    gid = [0] * a + [1] * b
    _, _, _, st = orc.nndescent_lists(X, ids, flags, gid, rule_all=False)
    assert st["dists"] == a * b * (a + b - 2)
    # J-Merge: one built subgraph [0, a), raw rows after it
    gid = [0] * a + [RAW] * b
    _, _, _, st = orc.nndescent_lists(X, ids, flags, gid, rule_all=False)
    assert st["dists"] == a * (_c2(K) - _c2(a - 1)) + b * (_c2(K) - _c2(a))
    # no built rows: the merge rule equals KGraph's
    _, _, _, st = orc.nndescent_lists(X, ids, flags, [RAW] * n, rule_all=False)
    assert st["dists"] == n * _c2(K)


@pytest.mark.parametrize("K,rho", [(20, 0.5), (16, 0.25), (10, 0.7)])
def test_delta_and_stop_closed_form(K, rho):
    # forward sampling takes round(rho K) new entries per row; the others
    # keep their flag and sit out the iteration, so after one iteration
    # delta = (K - S) / K exactly; the run goes on until every entry has been
    # sampled, and the pair count over all iterations adds up to the one
    # of a single full iteration, C(K, 2) per row
    X, ids = _complete(K, seed=3)
    n = K + 1
    S = int(round(rho * K))
    flags = np.ones_like(ids, np.uint8)
    gid = [orc.RAW] * n
    _, _, f, st = orc.nndescent_lists(X, ids, flags, gid, rule_all=True, rho=rho, max_iters=1)
    assert st["iters"] == 1 and st["converged"] == 0
    assert st["dists"] == n * _c2(S)
    assert abs(st["last_delta"] - (K - S) / K) < 1e-7
    assert (f.sum(1) == K - S).all()
    assert st["last_delta"] == pytest.approx(f.sum() / (n * K), abs=1e-7)
    # the stop rule: delta below the threshold ends the loop, not at it
    _, _, _, st = orc.nndescent_lists(X, ids, flags, gid, rule_all=True, rho=rho,
                                      delta=(K - S) / K + 1e-6)
    assert st["iters"] == 1 and st["converged"] == 1
    _, _, _, st = orc.nndescent_lists(X, ids, flags, gid, rule_all=True, rho=rho)
    iters = -(-K // S)
    assert st["iters"] == iters and st["converged"] == 1 and st["last_delta"] == 0.0
    assert st["dists"] == n * _c2(K)


def test_reverse_entries_closed_form():
    # four points on a line at 0, 1, 3, 7 with K = 1: rows 2 and 3 list a
    # nearer row whose own k-th distance is smaller, so they enter that
    # row's reverse list (maxDst < nb.dist, src/abstractmerge.cpp:94-104);
    # rows 1 and 2 then each join one pair: (0, 2) and (1, 3)
    X = np.array([[0.0], [1.0], [3.0], [7.0]], np.float32)
    ids = np.array([[1], [0], [1], [2]], np.uint32)
    flags = np.ones_like(ids, np.uint8)
    gid = [orc.RAW] * 4
    _, _, _, st = orc.nndescent_lists(X, ids, flags, gid, rule_all=True)
    assert st["dists"] == 2 and st["converged"] == 1
    _, _, _, st = orc.nndescent_lists(X, ids, flags, gid, rule_all=True, rev_cap=0)
    assert st["dists"] == 0
    # old entries join new ones: with row 3's entry already old, row 2's new
    # list is [1] and its old list [3] (through reverse), one pair again
    f2 = flags.copy()
    f2[3, 0] = 0
    _, _, _, st = orc.nndescent_lists(X, ids, f2, gid, rule_all=True)
    assert st["dists"] == 2


def _merge2one_brute(pi, pd, ci, cd, dedup):
    # a global stable order of all entries: distance, then the pool before
    # the parked entries, then position; the first K (distinct ids when
    # de-duplicating) are the merge's output
    K = len(pi)
    allv = [(float(pd[j]), 0, j, int(pi[j])) for j in range(K)]
    allv += [(float(cd[j]), 1, j, int(ci[j])) for j in range(len(ci))]
    allv.sort()
    out, seen = [], set()
    for dd, _, _, i in allv:
        if dedup and i in seen:
            continue
        seen.add(i)
        out.append((i, dd))
        if len(out) == K:
            break
    return [o[0] for o in out], [o[1] for o in out]


@pytest.mark.parametrize("dedup", [True, False])
def test_merge2one_against_brute_force(dedup):
    rng = np.random.default_rng(11 + dedup)
    n, K, cut = 400, 12, 6
    pi = np.empty((n, K), np.uint32)
    pd = np.empty((n, K), np.float32)
    ci = np.full((n, cut), orc.RAW, np.uint32)
    cd = np.zeros((n, cut), np.float32)
    dist_of = rng.integers(0, 20, 60).astype(np.float32)  # a fixed distance per id: ties, repeats
    for v in range(n):
        ids = rng.choice(60, K, replace=False)
        o = np.argsort(dist_of[ids], kind="stable")
        pi[v], pd[v] = ids[o], dist_of[ids[o]]
        m = int(rng.integers(0, cut + 1))
        c = rng.choice(60, m, replace=False)  # may repeat pool ids
        o = np.argsort(dist_of[c], kind="stable")
        ci[v, :m], cd[v, :m] = c[o], dist_of[c[o]]
    gi, gd = orc.merge2one(pi, pd, ci, cd, dedup)
    dup_rows = 0
    for v in range(n):
        m = int((ci[v] != orc.RAW).sum())
        ei, ed = _merge2one_brute(pi[v], pd[v], ci[v, :m], cd[v, :m], dedup)
        assert gi[v].tolist() == ei and gd[v].tolist() == ed, f"row {v}"
        dup_rows += len(set(gi[v].tolist())) < K
    # without de-dup a parked id already in the pool can come back twice
    assert (dup_rows > 0) == (not dedup)


def test_recall_semantics():
    ids = np.array([[1, 1, 2, 3]], np.uint32)
    gt = np.array([[1, 2, 9, 8]], np.uint32)
    rows = np.array([0], np.uint32)
    # reference counts the repeated id twice (src/evaluator.cpp:568-579)
    assert orc.recall(ids, rows, gt, 4, unique=False) == 3 / 4
    assert orc.recall(ids, rows, gt, 4, unique=True) == 2 / 4


def test_mlp_oracle_gradients_match_finite_differences():
    import mlp
    rng = np.random.default_rng(9)
    d, H, O, B = 12, 16, 4, 10
    W1 = rng.standard_normal((H, d)) * 0.4
    b1 = rng.standard_normal(H) * 0.1
    W2 = rng.standard_normal((O, H)) * 0.3
    b2 = rng.standard_normal(O) * 0.1
    xa = rng.standard_normal((B, d))
    xb = rng.standard_normal((B, d))
    t = rng.uniform(0.5, 3.0, B)
    s = rng.uniform(0.5, 2.0, B)
    L, gW1, gb1, gW2, gb2 = mlp.loss_grad(W1, b1, W2, b2, xa, xb, t, s)
    params = [W1, b1, W2, b2]
    grads = [gW1, gb1, gW2, gb2]
    h = 1e-6
    for P, G in zip(params, grads):
        flat = P.reshape(-1)
        for idx in rng.choice(flat.size, min(12, flat.size), replace=False):
            old = flat[idx]
            flat[idx] = old + h
            lp = mlp.loss(*params, xa, xb, t, s)
            flat[idx] = old - h
            lm = mlp.loss(*params, xa, xb, t, s)
            flat[idx] = old
            fd = (lp - lm) / (2 * h)
            assert abs(fd - G.reshape(-1)[idx]) <= 1e-6 + 1e-5 * abs(fd)


def test_mlp_oracle_loss_closed_form():
    import mlp
    # identity-like network on 2-d points: f(x) = x (W1 = I, W2 = I) on the
    # positive quadrant, so p = |xa - xb|^2 and L = 0 when t = p
    W1 = np.eye(2); b1 = np.zeros(2); W2 = np.eye(2); b2 = np.zeros(2)
    xa = np.array([[1.0, 2.0], [3.0, 1.0]]); xb = np.array([[2.0, 4.0], [1.0, 1.0]])
    t = ((xa - xb) ** 2).sum(1)
    assert mlp.loss(W1, b1, W2, b2, xa, xb, t, np.ones(2)) == 0.0
    L = mlp.loss(W1, b1, W2, b2, xa, xb, 2 * t, np.zeros(2))  # p = t/2 of target: r = -1/2
    assert abs(L - 0.25) < 1e-15


def test_merge_mt_agrees_with_sequential_merge():
    # the all-core form (the CPU baseline) runs the same algorithm with other
    # random draws: equal list invariants, recall within 0.005, pair counts
    # within 2%, the same set-up distance count
    X = nnm_data.sift_like(6000, 64, centers=16, seed=12)
    rows = np.arange(0, 6000, 5, dtype=np.uint32)
    gt, _ = orc.bruteforce(X, 10, rows=rows)
    h = 3000
    g0, _, _ = orc.kgraph(X[:h], K=16, seed=3)
    init = np.vstack([g0, np.zeros((h, 16), np.uint32)])
    si, sd, sst = orc.merge(X, [0, h, 6000], [1, 0], init, K=16, seed=5)
    mi, md, mst = orc.merge_mt(X, [0, h, 6000], [1, 0], init, K=16, seed=5, threads=4)
    assert mst["threads"] == 4
    assert (np.diff(md.astype(np.float64), axis=1) >= 0).all()
    assert all(len(set(r.tolist())) == 16 for r in mi)
    rs, rm = orc.recall(si, rows, gt, 10), orc.recall(mi, rows, gt, 10)
    assert abs(rs - rm) < 0.005 and rm > 0.97
    assert abs(mst["dists"] / sst["dists"] - 1) < 0.02
    assert mst["init_dists"] == sst["init_dists"]

"""Pins of the CPU oracle for graph search and the DPG layer builder.

Closed forms and invariants that a plausible mistake in the restatement
(comparison direction, list bound, reverse-rule order, a dropped distance
count) would break:
  * collinear points: the occlusion prune keeps exactly the two adjacent
    points, so the DPG of points on a line is the path graph;
  * greedy descent on a path counts 1 + deg(v) distances per visited node;
  * best-first search on a complete graph with beam >= n is exact brute
    force and computes every distance once;
  * reverse edges: a target with room receives every source that kept it.
"""
import numpy as np

import nnm_data
import orc


def _line(n, d=8, seed=0):
    rng = np.random.default_rng(seed)
    t = np.sort(rng.choice(np.arange(1, 20 * n), n, replace=False)).astype(np.float32)
    X = np.zeros((n, d), np.float32)
    X[:, 0] = t
    return X


def _knn(X, K):
    ids, _ = orc.bruteforce(X, K)
    return ids


def test_dpg_of_collinear_points_is_the_path():
    n = 60
    X = _line(n)
    for variant in ("ref", "par"):
        offs, adj, dist = orc.dpg(X, _knn(X, 8), M=4, maxM=6, variant=variant)
        for i in range(n):
            got = sorted(adj[offs[i]:offs[i + 1]].tolist())
            want = [j for j in (i - 1, i + 1) if 0 <= j < n]
            assert got == want, (variant, i, got)
            dd = dist[offs[i]:offs[i + 1]]
            assert (np.diff(dd) >= 0).all()
            for j, dj in zip(adj[offs[i]:offs[i + 1]], dd):
                assert dj == np.float32((X[i, 0] - X[j, 0]) ** 2)


def test_dpg_star_keeps_the_nearest_axes():
    # a centre row and axis rows r_j e_j: an axis row keeps only the centre
    # (every other axis row is farther from it than the centre is), the
    # centre keeps its M nearest axis rows (no axis row occludes another:
    # r_i^2 + r_j^2 > r_j^2), and the reverse pass offers every axis row to
    # the centre, which ends with its maxM nearest, replacing its farthest
    # entry whenever a nearer offer comes in (src/lynndescent.cpp:685-705)
    m, M, maxM = 9, 3, 5
    rng = np.random.default_rng(2)
    r = rng.permutation(np.arange(3, 3 + m)).astype(np.float32)
    order = rng.permutation(m + 1)  # the centre's row index is order[0]
    X = np.zeros((m + 1, m), np.float32)
    for j in range(m):
        X[order[j + 1], j] = r[j]
    c = int(order[0])
    axis_by_r = [int(order[j + 1]) for j in np.argsort(r)]
    for variant in ("ref", "par"):
        offs, adj, dist = orc.dpg(X, _knn(X, m), M, maxM, variant=variant)
        assert adj[offs[c]:offs[c + 1]].tolist() == axis_by_r[:maxM], variant
        assert dist[offs[c]:offs[c + 1]].tolist() == sorted(float(x * x) for x in r)[:maxM]
        for a in axis_by_r:
            assert adj[offs[a]:offs[a + 1]].tolist() == [c], (variant, a)


def test_dpg_in_place_pass_closed_form():
    # four rows where the reference's in-place reverse pass and the
    # order-free pass differ (ties included).  Squared distances:
    # d01 = 261, d02 = 185, d03 = 722, d12 = 20, d13 = 185, d23 = 185.
    # Phase 1 (M = 2): row 0 keeps [2] (1 and 3 are occluded by 2); row 1
    # keeps [2, 3] (d23 = d13 is not "closer"); row 2 keeps 1, then 0, then
    # 3 (equal to the last distance) replaces 0: [1, 3]; row 3 keeps [1].
    # Reverse pass (maxM = 2): row 0 offers itself to row 2, whose full list
    # has no entry closer to row 0 than d02, so 0 replaces 3: [1, 0].  In the
    # reference row 2 then walks its current list [1, 0] and never offers
    # itself to row 3; the order-free pass walks row 2's phase-1 list [1, 3]
    # and row 3 (with room) takes 2 at the tied distance 185: [1, 2].
    X = np.array([[23, 28], [8, 22], [12, 20], [4, 9]], np.float32)
    knn = _knn(X, 3)
    want = {"ref": [[2], [2, 3], [1, 0], [1]], "par": [[2], [2, 3], [1, 0], [1, 2]]}
    for variant, rows in want.items():
        offs, adj, _ = orc.dpg(X, knn, 2, 2, variant=variant)
        assert [adj[offs[i]:offs[i + 1]].tolist() for i in range(4)] == rows, variant


def test_dpg_bounds_and_reverse_rule():
    X = nnm_data.sift_like(1500, 16, centers=6, seed=8)
    knn = _knn(X, 10)
    M, maxM = 5, 8
    for variant in ("ref", "par"):
        offs, adj, dist = orc.dpg(X, knn, M, maxM, variant=variant)
        deg = np.diff(offs)
        assert deg.max() <= max(M, maxM)
        for i in range(X.shape[0]):
            row = adj[offs[i]:offs[i + 1]]
            assert len(set(row.tolist())) == row.size
            assert (np.diff(dist[offs[i]:offs[i + 1]]) >= 0).all()
    # with room for every offer (maxM = 64) the reverse pass mirrors every
    # edge: i -> u kept in phase 1 makes u receive i, and i -> u received as
    # an offer means u kept i; nothing is ever evicted
    big_o, big_a, _ = orc.dpg(X, knn, M, 64, variant="par")
    for i in range(X.shape[0]):
        for u in big_a[big_o[i]:big_o[i + 1]]:
            assert i in set(big_a[big_o[u]:big_o[u + 1]].tolist()), (i, int(u))
    # a tight reverse bound never grows a list past maxM unless its own
    # occlusion list was longer
    offs_t, _, _ = orc.dpg(X, knn, M, M, variant="par")
    assert np.diff(offs_t).max() <= M


def test_dpg_par_agrees_with_reference_pass():
    X = nnm_data.sift_like(4000, 32, centers=12, seed=9)
    knn = _knn(X, 16)
    o1, a1, _ = orc.dpg(X, knn, 8, 16, variant="ref")
    o2, a2, _ = orc.dpg(X, knn, 8, 16, variant="par")
    e1 = {(i, int(j)) for i in range(X.shape[0]) for j in a1[o1[i]:o1[i + 1]]}
    e2 = {(i, int(j)) for i in range(X.shape[0]) for j in a2[o2[i]:o2[i + 1]]}
    assert len(e1 & e2) / len(e1 | e2) >= 0.99


def test_greedy_descent_on_a_path_counts():
    n = 50
    X = _line(n)
    offs = np.zeros(n + 1, np.uint64)
    adj = []
    for i in range(n):
        nb = [j for j in (i - 1, i + 1) if 0 <= j < n]
        adj += nb
        offs[i + 1] = offs[i] + len(nb)
    adj = np.array(adj, np.uint32)
    p = 23
    Q = X[p:p + 1] + np.float32(0.25)
    # two layers (the path above, then a base layer with no edges): greedy from
    # row 0 walks to p, scanning every neighbour of rows 0..p
    base_offs = np.zeros(n + 1, np.uint64)
    ids, dist, nc = orc.search_layered(X, [(offs, adj), (base_offs, np.zeros(0, np.uint32))], Q, 0, 1, 1)
    assert ids[0, 0] == p
    assert nc == 1 + 1 + 2 * p  # entry + deg(0) = 1 + deg(v) = 2 for v = 1..p


def test_best_first_on_complete_graph_is_exact():
    X = nnm_data.sift_like(300, 16, centers=4, seed=10)
    n = X.shape[0]
    offs = (np.arange(n + 1, dtype=np.uint64) * (n - 1))
    adj = np.array([j for i in range(n) for j in range(n) if j != i], np.uint32)
    Q = nnm_data.sift_like(20, 16, centers=4, seed=11)
    gt, gd = orc.bruteforce(X, 10, Q=Q)
    ids, dist, nc = orc.search_layered(X, [(offs, adj)], Q, 5, n, 10)
    assert (ids == gt).all()
    assert np.allclose(dist, gd, rtol=1e-6)
    assert nc == Q.shape[0] * n  # entry once, then every other row once
    seeds = np.tile(np.arange(n, dtype=np.uint32), (Q.shape[0], 1))
    fi, fd, fnc = orc.search_flat(X, offs, adj, Q, seeds, n, 10)
    assert (fi == gt).all() and fnc == Q.shape[0] * n

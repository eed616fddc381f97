"""GPU parity of the basic entry points against the CPU oracle.

Distances: the library's FP32 order (0..d-1, fused multiply-add) is exact on
integer-valued rows (every partial sum is an integer below 2^24), so on
sift-like data the GPU must equal the FP64 oracle exactly and neighbour ids
must match bit for bit with ties broken by id.  On real-valued rows the
tolerance is 1e-5 relative to FP64.
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


def test_bruteforce_integer_rows_bit_exact(ctx):
    X = nnm_data.sift_like(3000, 128, centers=32, seed=11)
    ctx.set_data(X)
    rows = np.arange(0, 3000, 15, dtype=np.uint32)
    gi, gd = ctx.bruteforce(20, rows=rows)
    oi, od = orc.bruteforce(X, 20, rows=rows)
    assert np.array_equal(gi, oi)
    assert np.array_equal(gd.astype(np.float64), od)


def test_bruteforce_real_rows_within_1e5(ctx):
    X = nnm_data.unit_mixture(2500, 96, centers=16, sigma=0.3, seed=12)
    ctx.set_data(X)
    rows = np.arange(0, 2500, 25, dtype=np.uint32)
    gi, gd = ctx.bruteforce(10, rows=rows)
    oi, od = orc.bruteforce(X, 10, rows=rows)
    rel = np.abs(gd.astype(np.float64) - od) / np.maximum(od, 1e-30)
    assert rel.max() < 1e-5
    # ids agree wherever the FP64 distances are not within 1e-5 of a neighbour
    gap_ok = np.ones_like(od, dtype=bool)
    gap = np.diff(od, axis=1) / np.maximum(od[:, 1:], 1e-30)
    gap_ok[:, 1:] &= gap > 1e-5
    gap_ok[:, :-1] &= gap > 1e-5
    assert np.array_equal(gi[gap_ok], oi[gap_ok])


def test_bruteforce_query_vectors(ctx):
    X = nnm_data.sift_like(2000, 128, centers=16, seed=13)
    Q = nnm_data.sift_like(64, 128, centers=16, seed=14)
    ctx.set_data(X)
    gi, gd = ctx.bruteforce(8, Q=Q)
    oi, od = orc.bruteforce(X, 8, Q=Q)
    assert np.array_equal(gi, oi)
    assert np.array_equal(gd.astype(np.float64), od)


@pytest.mark.parametrize("d,vmax", [(20, 255), (64, 3), (96, 255), (128, 255), (128, 1)])
def test_bruteforce_mfma_form_bit_exact(ctx, d, vmax):
    """The MFMA form (integer rows: bf16 operands, FP32 norm form, exact)
    against the FP64 oracle, ids and distances bit for bit with ties broken
    by id: row widths of one to four 32-element steps, tiny value ranges
    (mostly ties), lists of 16 and 32, query counts around the 256-query
    block, query rows (self excluded) and query vectors."""
    rng = np.random.default_rng(d * 7 + vmax)
    n = 2100
    X = rng.integers(0, vmax + 1, size=(n, d)).astype(np.float32)
    ctx.set_data(X)
    for k in (1, 10, 16, 17, 32):
        for nq in (1, 255, 257):
            rows = rng.choice(n, nq, replace=False).astype(np.uint32)
            gi, gd = ctx.bruteforce(k, rows=rows)
            oi, od = orc.bruteforce(X, k, rows=rows)
            assert np.array_equal(gi, oi), (k, nq)
            assert np.array_equal(gd.astype(np.float64), od), (k, nq)
    Q = rng.integers(0, vmax + 1, size=(300, d)).astype(np.float32)
    for k in (10, 32):
        gi, gd = ctx.bruteforce(k, Q=Q)
        oi, od = orc.bruteforce(X, k, Q=Q)
        assert np.array_equal(gi, oi), k
        assert np.array_equal(gd.astype(np.float64), od), k


def test_pair_dists_match_reference_order_on_integers(ctx):
    X = nnm_data.sift_like(500, 128, centers=8, seed=15)
    ctx.set_data(X)
    rng = np.random.default_rng(0)
    a = rng.integers(0, 500, 300).astype(np.uint32)
    b = rng.integers(0, 500, 300).astype(np.uint32)
    g = ctx.pair_dists(a, b)
    ref = np.array([orc.l2_seq(X[i], X[j]) for i, j in zip(a, b)], np.float32)
    assert np.array_equal(g, ref)


def _check_rows(ids, dist, X, K, lo=None, hi=None):
    n = ids.shape[0]
    for v in range(n):
        row = ids[v]
        assert len(set(row.tolist())) == K, f"duplicate ids in row {v}"
        assert v not in row
        if lo is not None:
            assert (row >= lo[v]).all() and (row < hi[v]).all()
        keys = list(zip(dist[v].tolist(), row.tolist()))
        assert keys == sorted(keys), f"row {v} not sorted by (dist, id)"


def test_random_init_invariants(ctx):
    X = nnm_data.sift_like(1200, 128, centers=8, seed=16)
    ctx.set_data(X)
    ctx.set_partition([0, 500, 1200], [1, 1], [0, 1])
    ctx.random_init(20, seed=5)
    ids, dist, flags = ctx.get_graph()
    lo = np.where(np.arange(1200) < 500, 0, 500)
    hi = np.where(np.arange(1200) < 500, 500, 1200)
    _check_rows(ids, dist, X, 20, lo, hi)
    assert flags.all()
    v = np.repeat(np.arange(1200), 20)
    exact = np.array([orc.l2_seq(X[i], X[j]) for i, j in zip(v[:4000], ids.ravel()[:4000])], np.float32)
    assert np.array_equal(dist.ravel()[:4000], exact)


def test_set_graph_recomputes_and_sorts(ctx):
    X = nnm_data.sift_like(800, 128, centers=8, seed=17)
    ctx.set_data(X)
    rng = np.random.default_rng(3)
    ids = np.stack([rng.choice(np.delete(np.arange(800), v), 12, replace=False) for v in range(800)]).astype(np.uint32)
    ctx.set_graph(ids)
    gi, gd, gf = ctx.get_graph()
    _check_rows(gi, gd, X, 12)
    for v in range(0, 800, 50):
        assert set(gi[v].tolist()) == set(ids[v].tolist())


def test_bad_graph_rejected(ctx):
    X = nnm_data.sift_like(100, 128, centers=4, seed=18)
    ctx.set_data(X)
    ids = np.tile(np.arange(10, dtype=np.uint32), (100, 1))
    with pytest.raises(nnm.NNMError) as e:
        ctx.set_graph(ids)  # row v < 10 contains itself
    assert e.value.code == nnm.NNM_ERR_INVALID


def test_real_sift_rows_parity(ctx):
    """Real SIFT descriptors (the reference's data/sift_query.txt, first 1000
    rows): brute force bit-exact against the FP64 oracle, the reference's own
    distance code bit-exact on pairs, NN-Descent recall within 0.005 of the
    CPU restatement."""
    import os
    p = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sift_query_head1000.txt")
    X = nnm.read_matrix(p)
    ctx.set_data(X)
    ids, dist = ctx.bruteforce(10)
    oi, od = orc.bruteforce(X, 10)
    assert np.array_equal(ids, oi)
    assert np.array_equal(dist.astype(np.float64), od)  # integer rows: exact
    a = np.arange(0, 1000, 7, dtype=np.uint32)
    b = (a * 31 + 5) % 1000
    pd = ctx.pair_dists(a, b)
    ref = np.array([orc.l2_seq(X[i], X[j]) for i, j in zip(a, b)], np.float32)
    assert np.array_equal(pd, ref)
    ctx.random_init(10, seed=2)
    ctx.nndescent(nnm.default_params(K=10, rule=nnm.RULE_ALL, seed=2, max_iters=20))
    g, _, _ = ctx.get_graph()
    rows = np.arange(1000, dtype=np.uint32)
    r_gpu = nnm.recall(g, oi, rows, 10)
    gi, _, _ = orc.kgraph(X, K=10, seed=2, max_iters=20)
    r_cpu = orc.recall(gi, rows, oi, 10)
    print("sift head recall gpu", r_gpu, "cpu", r_cpu)
    assert r_gpu >= r_cpu - 0.005

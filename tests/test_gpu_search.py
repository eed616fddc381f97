"""GPU DPG builder and batched graph search against the CPU restatements.

Integer rows make every distance exact in FP32 and in bf16, so the GPU must
reproduce the oracle bit for bit: the DPG lists (ids and distances), and the
search results and distance counts of the reference's best-first order
(width 1).  Wider expansion and real-valued rows are checked by recall.

Queries come from the data's own mixture.  The mixtures are well-separated
clusters, so their k-NN and DPG graphs have one component per cluster: a
search from one entry point stays in its cluster (here and in the
reference), and recall is measured with the seeded flat search, whose seeds
cover every cluster.
"""
import numpy as np
import pytest

import nnm
import nnm_data
import orc

pytestmark = pytest.mark.gpu

N, D, K = 20000, 64, 20


@pytest.fixture(scope="module")
def ctx():
    with nnm.Context(0) as c:
        yield c


@pytest.fixture(scope="module")
def data():
    return nnm_data.sift_like(N, D, centers=32, seed=101, nq=400)


def _knn(X, k=K):
    with nnm.Context(0) as c:
        c.set_data(X)
        ids, _ = c.bruteforce(k)
    return ids


def _build_layers(X):
    out = []
    for rows, maxM in ((1000, 10), (5000, 10), (X.shape[0], 20)):
        knn = _knn(X[:rows])
        with nnm.Context(0) as c:
            c.set_data(X[:rows])
            c.set_graph(knn)
            st = c.dpg_build(10, maxM)
            o, a, dd = c.dpg_get()
        out.append((rows, knn, (o, a, dd), st))
    return out


@pytest.fixture(scope="module")
def layers(data):
    """Three DPG layers on prefixes of 1000, 5000 and all rows, built on the
    GPU from exact k-NN lists."""
    return _build_layers(data[0])


def test_dpg_matches_oracle(data, layers):
    X, _ = data
    for rows, knn, (o, a, dd), st in layers:
        maxM = 10 if rows < N else 20
        ro, ra, rd = orc.dpg(X[:rows], knn, 10, maxM, variant="par")
        assert np.array_equal(o, ro), rows
        assert np.array_equal(a, ra), rows
        assert np.array_equal(dd.view(np.uint32), rd.view(np.uint32)), rows
        # the order-free pass against the reference's in-place pass
        fo, fa, _ = orc.dpg(X[:rows], knn, 10, maxM, variant="ref")
        e1 = {(i, int(j)) for i in range(rows) for j in a[o[i]:o[i + 1]]}
        e2 = {(i, int(j)) for i in range(rows) for j in fa[fo[i]:fo[i + 1]]}
        print("dpg rows", rows, "edges", len(e1), "only in the GPU layer", len(e1 - e2),
              "only in the reference pass", len(e2 - e1), st["secs"])
        assert len(e1 & e2) / len(e1 | e2) >= 0.99


def test_dpg_in_place_tie_case():
    """Four rows on which the reference's in-place reverse pass and the
    order-free pass differ (derived by hand in the oracle's tests): the GPU
    layer is the order-free one, one edge (3 -> 2) more than the reference."""
    X = np.array([[23, 28], [8, 22], [12, 20], [4, 9]], np.float32)
    knn, _ = orc.bruteforce(X, 3)
    with nnm.Context(0) as c:
        c.set_data(X)
        c.set_graph(knn.astype(np.uint32))
        c.dpg_build(2, 2)
        o, a, _ = c.dpg_get()
    assert [a[o[i]:o[i + 1]].tolist() for i in range(4)] == [[2], [2, 3], [1, 0], [1, 2]]


def test_dpg_carry_matches_oracle(ctx, data):
    X, _ = data
    pre = 5000
    knn_pre = _knn(X[:pre])
    knn_all = _knn(X)
    ctx.set_data(X)
    g = knn_all.copy()
    g[:pre] = knn_pre
    ctx.set_graph(g)
    ctx.set_partition([0, pre], [0], [0])
    ctx.dpg_build(10, 10)
    ctx.set_graph(knn_all)
    ctx.set_partition([0, N], [0], [0])
    ctx.dpg_build(10, 20, carry=True)
    o, a, dd = ctx.dpg_get()
    prev = orc.dpg(X[:pre], knn_pre, 10, 10, variant="par")
    ro, ra, rd = orc.dpg(X, knn_all, 10, 20, prev=prev, variant="par")
    assert np.array_equal(o, ro) and np.array_equal(a, ra)
    assert np.array_equal(dd.view(np.uint32), rd.view(np.uint32))


def _set_layers(ctx, X, layers, upto=3):
    ctx.set_data(X)
    ctx.search_set_layers([(o, a) for _, _, (o, a, _), _ in layers[:upto]])


@pytest.mark.parametrize("beam", [16, 64])
@pytest.mark.parametrize("prec", [nnm.PREC_FP32, nnm.PREC_BF16])
def test_layered_search_matches_oracle(ctx, data, layers, beam, prec):
    X, Q = data
    _set_layers(ctx, X, layers)
    ids, dist, st = ctx.search(Q, beam, k=10, width=1, prec=prec, entry=7)
    oi, od, onc = orc.search_layered(X, [(o.astype(np.uint64), a) for _, _, (o, a, _), _ in layers], Q, 7, beam, 10)
    print("beam", beam, "prec", prec, st, "oracle dists", onc)
    assert np.array_equal(ids, oi)
    assert np.array_equal(dist.view(np.uint32), od.view(np.uint32))
    if st["hash_resets"] == 0:
        assert st["dists"] == onc
    else:
        assert st["dists"] >= onc
    assert st["capped"] == 0


def test_flat_search_with_seeds_matches_oracle(ctx, data, layers):
    X, Q = data
    _set_layers(ctx, X, layers[2:])
    beam = 32
    rng = np.random.default_rng(5)
    seeds = np.stack([rng.choice(N, beam, replace=False) for _ in range(Q.shape[0])]).astype(np.uint32)
    ids, dist, st = ctx.search(Q, beam, k=10, mode=nnm.SEARCH_FLAT, seeds=seeds)
    o, a = layers[2][2][0], layers[2][2][1]
    oi, od, onc = orc.search_flat(X, o, a, Q, seeds, beam, 10)
    assert np.array_equal(ids, oi)
    assert np.array_equal(dist.view(np.uint32), od.view(np.uint32))
    if st["hash_resets"] == 0:
        assert st["dists"] == onc


def test_search_recall_rises_with_beam_and_width(ctx, data, layers):
    X, Q = data
    ctx.set_data(X)
    gt, _ = ctx.bruteforce(10, Q=Q)
    _set_layers(ctx, X, layers[2:])
    qrows = np.arange(Q.shape[0], dtype=np.uint32)
    rec = {}
    for beam in (16, 32, 64, 128, 256):
        for width in (1, 2, 4):
            ids, _, st = ctx.search(Q, beam, k=10, width=width, prec=nnm.PREC_BF16, mode=nnm.SEARCH_FLAT, seed=9)
            rec[beam, width] = nnm.recall(ids, gt, qrows, 10)
    print(rec)
    assert rec[256, 1] >= 0.99
    assert rec[256, 1] >= rec[32, 1] >= rec[16, 1] - 0.005
    for beam in (32, 64, 128, 256):
        assert rec[beam, 4] >= rec[beam, 1] - 0.01


def test_search_real_rows_recall_parity(ctx):
    X, Q = nnm_data.unit_mixture(8000, 96, centers=64, seed=303, nq=300)
    knn = _knn(X)
    ctx.set_data(X)
    ctx.set_graph(knn)
    ctx.dpg_build(10, 20)
    ctx.search_use_graph(nnm.SEARCH_SRC_DPG)
    o, a, _ = ctx.dpg_get()
    gt, _ = orc.bruteforce(X, 10, Q=Q)
    qrows = np.arange(Q.shape[0], dtype=np.uint32)
    beam = 64
    rng = np.random.default_rng(6)
    seeds = np.stack([rng.choice(X.shape[0], beam, replace=False) for _ in range(Q.shape[0])]).astype(np.uint32)
    oi, _, _ = orc.search_flat(X, o, a, Q, seeds, beam, 10)
    r_cpu = orc.recall(oi, qrows, gt, 10)
    for prec in (nnm.PREC_FP32, nnm.PREC_BF16):
        ids, dist, st = ctx.search(Q, beam, k=10, prec=prec, mode=nnm.SEARCH_FLAT, seeds=seeds)
        r_gpu = orc.recall(ids, qrows, gt, 10)
        print("real rows prec", prec, r_gpu, r_cpu, st)
        assert r_gpu >= r_cpu - 0.005
        # returned distances are exact FP32 (within FMA rounding of the oracle)
        ref = ((X[ids].astype(np.float64) - Q[:, None, :]) ** 2).sum(-1)
        assert np.allclose(dist, ref, rtol=2e-5, atol=1e-6)


def test_search_rejects_bad_input(ctx, data, layers):
    X, Q = data
    _set_layers(ctx, X, layers)
    with pytest.raises(nnm.NNMError):
        ctx.search(Q, 300, k=10)
    with pytest.raises(nnm.NNMError):
        ctx.search(Q, 64, k=10, entry=5000)  # not a row of layer 0
    with pytest.raises(nnm.NNMError):
        ctx.search(Q, 64, k=10, width=9)
    bad = [(np.array([0, 1], np.uint32), np.array([5], np.uint32))]
    with pytest.raises(nnm.NNMError):
        ctx.search_set_layers(bad)  # neighbour outside the layer


@pytest.mark.parametrize("prec", [nnm.PREC_FP32, nnm.PREC_BF16])
def test_beams_in_one_launch_equal_single_beam_searches(ctx, data, layers, prec):
    """nnm_search_beams: every beam of the list in one launch gives, beam for
    beam, the results and counts of a launch of that beam alone (layered
    mode from one entry row, and the flat mode with drawn seeds); each beam
    gets its share of the launch time."""
    X, Q = data
    beams = [64, 16, 256, 32, 128, 48]
    _set_layers(ctx, X, layers)
    for mode in (nnm.SEARCH_LAYERED, nnm.SEARCH_FLAT):
        bi, bd, bst = ctx.search_beams(Q, beams, k=10, prec=prec, mode=mode, entry=7, seed=3)
        for j, beam in enumerate(beams):
            ids, dist, st = ctx.search(Q, beam, k=10, prec=prec, mode=mode, entry=7, seed=3)
            assert np.array_equal(bi[j], ids), (mode, beam)
            assert np.array_equal(bd[j].view(np.uint32), dist.view(np.uint32)), (mode, beam)
            for f in ("dists", "reranked", "hash_resets", "steps", "capped"):
                assert bst[j][f] == st[f], (mode, beam, f)
        assert abs(sum(s["secs"] for s in bst) - bst[0]["launch_secs"]) < 1e-6
        assert all(s["wave_secs"] > 0 for s in bst)
        print("mode", mode, [(b, round(s["secs"] * 1e3, 3)) for b, s in zip(beams, bst)])


def test_lysearch_host_class_files(ctx, data, layers, tmp_path):
    """LySearch(refer, graph).knnSearch(query, dst, grd, record) through the
    host class: the reference's beam list in one launch, one record line per
    beam with recall@1 against the "query nn" ground truth and distances per
    query, stopping after the first beam of recall 1; dst holds the last
    beam's top-1.  Every line equals a single-beam search of that beam."""
    X, Q = data
    refer, graph, query = tmp_path / "refer.txt", tmp_path / "graph.txt", tmp_path / "query.txt"
    dst, grd, record = tmp_path / "dst.txt", tmp_path / "grd.txt", tmp_path / "record.txt"
    nnm.write_matrix(str(refer), X)
    nnm.write_matrix(str(query), Q)
    lay = [(o, a) for _, _, (o, a, _), _ in layers]
    nnm.write_layers(str(graph), lay)
    gt, _ = orc.bruteforce(X, 1, Q=Q)
    with open(grd, "w") as f:
        for q in range(Q.shape[0]):
            f.write(f"{q} {gt[q, 0]}\n")
    recs = nnm.host_lysearch(refer, graph, query, dst, grd, record, entry=7)
    ks1 = [60, 50, 40, 36, 32, 28, 24, 20, 18, 15, 14, 12, 10, 8, 6, 4, 2]
    assert [r[0] for r in recs] == ks1[:len(recs)]
    assert len(recs) == len(ks1) or recs[-1][1] == 1.0
    assert all(r[1] < 1.0 for r in recs[:-1])
    lines = [ln.split("\t") for ln in open(record).read().splitlines()]
    assert [int(ln[0]) for ln in lines] == [r[0] for r in recs]
    ctx.set_data(X)
    ctx.search_set_layers(lay)
    for topk, rec, secs, cmps in recs:
        ids, dist, st = ctx.search(Q, topk, k=1, entry=7)
        assert rec == float(np.mean(ids[:, 0] == gt[:, 0]))
        assert cmps == st["dists"] // Q.shape[0]
        assert secs > 0
    out = [ln.split("\t") for ln in open(dst).read().splitlines()]
    assert [int(o[1]) for o in out] == ids[:, 0].tolist()
    # distances as the reference prints them (default stream precision)
    assert np.allclose([float(o[2]) for o in out], dist[:, 0], rtol=1e-5, atol=0)
    print("records", recs)

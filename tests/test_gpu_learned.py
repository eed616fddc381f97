"""Learned candidate pruning: GPU training kernels against the FP64 oracle,
the projection kernel, and the ablation of the learned prefilter against
the plain join (which it falls back to)."""
import time

import numpy as np
import pytest

import mlp
import nnm
import nnm_data
import orc

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    with nnm.Context(0) as c:
        yield c


def _batch(X, B, rng):
    a = rng.integers(0, X.shape[0], B)
    b = rng.integers(0, X.shape[0], B)
    xa, xb = X[a] / 128.0, X[b] / 128.0
    t = ((xa.astype(np.float64) - xb) ** 2).sum(1)
    s = rng.uniform(0.5, 1.5, B) * np.median(t)
    return xa.astype(np.float32), xb.astype(np.float32), t.astype(np.float32), s.astype(np.float32)


def _rel(a, b):
    return np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30)


@pytest.mark.parametrize("bf16,tol", [(False, 1e-4), (True, 4e-2)])
def test_mlp_gradients_match_oracle(ctx, bf16, tol):
    X = nnm_data.sift_like(2000, 128, centers=8, seed=71)
    ctx.set_data(X)
    ctx.mlp_init(256, 32, seed=3)
    W1, b1, W2, b2, alpha, _ = ctx.mlp_weights()
    rng = np.random.default_rng(0)
    b1r = rng.standard_normal(256).astype(np.float32) * 0.05
    b2r = rng.standard_normal(32).astype(np.float32) * 0.05
    ctx.mlp_set_weights(W1, b1r, W2, b2r)
    xa, xb, t, s = _batch(X, 64, rng)
    L, gW1, gb1, gW2, gb2 = ctx.mlp_loss_grad(xa, xb, t, s, bf16=bf16)
    oL, ogW1, ogb1, ogW2, ogb2 = mlp.loss_grad(W1[:, :128], b1r, W2, b2r, xa, xb, t, s)
    assert abs(L - oL) <= tol * abs(oL)
    for g, o in [(gW1, ogW1), (gb1, ogb1), (gW2, ogW2), (gb2, ogb2)]:
        assert _rel(g, o) < tol


def test_mlp_projection_matches_oracle(ctx):
    X = nnm_data.sift_like(3000, 128, centers=8, seed=72)
    ctx.set_data(X)
    ctx.mlp_init(256, 32, seed=4)
    W1, b1, W2, b2, alpha, _ = ctx.mlp_weights()
    P, secs = ctx.mlp_project(want=True)
    _, _, z = mlp.forward(W1[:, :128], b1, W2, b2, X.astype(np.float64) * alpha)
    assert _rel(P, z) < 2e-2


def test_mlp_training_reduces_loss(ctx):
    X = nnm_data.sift_like(30000, 128, centers=64, seed=73)
    ctx.set_data(X)
    ctx.random_init(20, seed=1)
    ctx.nndescent(nnm.default_params(K=20, seed=1, max_iters=8))
    ctx.mlp_init(256, 32, seed=5)
    st = ctx.mlp_train(steps=300, batch=2048, lr=2e-3, keep=0.98, seed=2)
    print(st)
    assert st["loss_last"] < 0.5 * st["loss_first"]
    _, _, _, _, _, tau = ctx.mlp_weights()
    assert 1.0 <= tau < 10.0


@pytest.fixture(scope="module")
def jm100k(ctx):
    """J-Merge input of 100k low-rank sift-like rows: the built half's
    NN-Descent lists plus K placeholder neighbours for the raw half."""
    n, K = 100000, 20
    half = n // 2
    X = nnm_data.sift_like(n, 128, centers=256, seed=nnm_data.SEED_BASE + 4, rank=12)
    rows = np.arange(0, n, 25, dtype=np.uint32)
    ctx.set_data(X)
    gt, _ = ctx.bruteforce(10, rows=rows)
    ctx.set_data(X[:half])
    ctx.random_init(K, seed=3)
    ctx.nndescent(nnm.default_params(K=K, seed=3))
    g0, _, _ = ctx.get_graph()
    init = np.vstack([g0, ((np.arange(half, n)[:, None] + 1 + np.arange(K)[None, :]) % n).astype(np.uint32)])
    return X, rows, gt, init


def _jmerge(ctx, jm, prec, flags):
    X, rows, gt, init = jm
    n, K = init.shape
    half = n // 2
    ctx.set_partition([0, half, n], [1, 0], [0, 0])
    ctx.set_graph(init)
    # learned runs: 5 learned iterations, then exact lists and 3 plain iterations
    p = nnm.default_params(K=K, rule=nnm.RULE_MERGE, rho=0.5, max_iters=8 if prec == nnm.PREC_LEARNED else 10,
                           prec=prec, seed=5, plain_tail=3, flags=flags)
    ctx.merge_init(p)
    st = ctx.nndescent(p)
    ctx.merge_finish(True)
    ids, _, _ = ctx.get_graph()
    return nnm.recall(ids, gt, rows), st


def test_learned_join_ablation(ctx, jm100k):
    """J-Merge of 100k low-rank sift-like rows: plain BF16 join vs the learned
    phase with a trained network vs with the untrained (random) network,
    each list entry renewed at the switch.  The learned run keeps recall
    within 0.005 of the plain join and computes fewer exact distances (one
    per list entry at the switch plus the plain tail's joins); the trained
    network beats the untrained one by 0.01 of recall.  Device time at this
    size is launch-bound; the time gain is asserted on the 5M merge."""
    X, rows, gt, init = jm100k
    n, K = init.shape
    ctx.set_data(X)
    res = {}

    def merge(prec):
        return _jmerge(ctx, jm100k, prec, nnm.FLAG_LEARNED_RENEW)

    res["plain"] = merge(nnm.PREC_BF16)
    # network trained on pairs of the built half (the plain NN-Descent's own lists)
    ctx.set_partition([0, n // 2, n], [1, 0], [0, 0])
    ctx.set_graph(init)
    ctx.mlp_init(256, 32, seed=7)
    ctx.mlp_train(steps=1, batch=1024, lr=0.0, keep=0.98, seed=8)  # calibrates tau only
    res["untrained"] = merge(nnm.PREC_LEARNED)
    ctx.set_partition([0, n // 2, n], [1, 0], [0, 0])
    ctx.set_graph(init)
    ctx.mlp_init(256, 32, seed=7)
    tst = ctx.mlp_train(steps=600, batch=4096, lr=2e-3, keep=0.98, seed=9)
    res["trained"] = merge(nnm.PREC_LEARNED)

    def exact(st):
        return st["dists"] - st["learned_dists"] + st["reranked"]

    for k in ("plain", "untrained", "trained"):
        r, st = res[k]
        print(k, round(r, 4), "iters", st["iters"], "learned", st["learned_iters"], "exact dists", exact(st),
              "switch", st["reranked"], "pairs", st["dists"], "secs", round(st["secs"], 4))
    print("train", tst)
    r_plain, s_plain = res["plain"]
    r_tr, s_tr = res["trained"]
    r_un, s_un = res["untrained"]
    assert s_tr["learned_iters"] > 0 and s_tr["plain_iters"] >= 1
    assert s_tr["reranked"] == n * K  # one exact recompute per list entry at the switch
    assert r_tr >= r_plain - 0.005
    assert exact(s_tr) < 0.7 * exact(s_plain)
    assert r_tr >= r_un + 0.01


def test_learned_merge_matches_oracle(ctx, jm100k):
    """The learned J-Merge (trained network, 5 learned iterations, then 3
    plain BF16 ones) against the oracle's J-Merge of the same lists and the
    same iteration cap: recall@10 within 0.005."""
    X, rows, gt, init = jm100k
    n, K = init.shape
    ctx.set_data(X)
    ctx.set_partition([0, n // 2, n], [1, 0], [0, 0])
    ctx.set_graph(init)
    ctx.mlp_init(256, 32, seed=7)
    ctx.mlp_train(steps=600, batch=4096, lr=2e-3, keep=0.98, seed=9)
    r_l, st = _jmerge(ctx, jm100k, nnm.PREC_LEARNED, nnm.FLAG_LEARNED_MARK_FOUND)
    oi, _, ost = orc.merge(X, [0, n // 2, n], [1, 0], init, K=K, rho=0.5, max_iters=8, seed=5)
    r_o = orc.recall(oi, rows, gt, 10)
    print(f"\nlearned {r_l:.4f} (iters {st['iters']}, learned {st['learned_iters']}) oracle {r_o:.4f} "
          f"(iters {ost['iters']}, {ost['secs']:.1f}s)")
    assert st["learned_iters"] == 5 and st["plain_iters"] == 3
    assert r_l >= r_o - 0.005


def test_learned_refuses_long_candidate_lists(ctx):
    """The learned join handles candidate lists of up to 256 entries
    (round(rho K) + K + 2 rev_cap): K = 32 at rho = 1 with rev_cap 100 asks
    for 264 and is refused with NNM_ERR_UNSUPPORTED before any launch;
    rev_cap 96 (256 entries) runs."""
    X = nnm_data.sift_like(6000, 128, centers=16, seed=77, rank=12)
    n, K = X.shape[0], 32
    ctx.set_data(X)
    ctx.set_partition([0, n // 2, n], [1, 0], [0, 0])
    ctx.random_init(K, seed=2)
    ctx.mlp_init(256, 32, seed=3)
    ctx.mlp_train(steps=20, batch=1024, lr=2e-3, keep=0.98, seed=4)
    p = nnm.default_params(K=K, rule=nnm.RULE_MERGE, rho=1.0, rev_cap=100, max_iters=3, prec=nnm.PREC_LEARNED,
                           plain_tail=1)
    with pytest.raises(nnm.NNMError) as e:
        ctx.nndescent(p)
    assert e.value.code == nnm.NNM_ERR_UNSUPPORTED
    p.rev_cap = 96
    st = ctx.nndescent(p)
    assert st["learned_iters"] == 2 and st["plain_iters"] == 1


_E4M3 = np.array([(2.0 ** ((c >> 3) - 7)) * (1 + (c & 7) / 8.0) if (c >> 3) else (c & 7) * 2.0 ** -9
                  for c in range(127)])  # positive OCP e4m3 values, codes 0..126 ascending


def _e4m3_value(b):
    b = b.astype(np.int64)
    v = _E4M3[np.minimum(b & 0x7F, 126)]
    return np.where(b & 0x80, -v, v)


def _e4m3_encode(x):
    """Round to nearest e4m3 (ties to the even code), |x| <= 448."""
    ax = np.abs(x.astype(np.float64))
    hi = np.clip(np.searchsorted(_E4M3, ax), 1, 126)
    lo = hi - 1
    dl, dh = ax - _E4M3[lo], _E4M3[hi] - ax
    code = np.where(dl < dh, lo, np.where(dh < dl, hi, np.where(lo % 2 == 0, lo, hi)))
    return (code | np.where(x < 0, 0x80, 0)).astype(np.uint8)


def test_fp8_rows_encode_the_projection(ctx):
    """FP8 rows of the learned phase: OCP e4m3 of clamp(s p, +-448), rounded
    to nearest even, s = 448 / max |p|, against a table encoder."""
    X = nnm_data.sift_like(6000, 128, centers=32, seed=73, rank=12)
    ctx.set_data(X)
    ctx.random_init(16, seed=2)
    ctx.nndescent(nnm.default_params(K=16, seed=2, max_iters=5))
    ctx.mlp_init(256, 32, seed=3)
    ctx.mlp_train(steps=50, batch=2048, lr=2e-3, keep=0.98, seed=4)
    P, _ = ctx.mlp_project(want=True)  # the bf16 projection, as floats
    P8, s = ctx.mlp_project_fp8()
    assert s == np.float32(448.0) / np.float32(np.abs(P).max())
    x = np.clip(P.astype(np.float32) * np.float32(s), np.float32(-448), np.float32(448))
    ref = _e4m3_encode(x)
    assert np.array_equal(_e4m3_value(P8), _e4m3_value(ref)), "FP8 rows differ from round-to-nearest-even e4m3"
    # a second projection (the model trained on) re-encodes
    ctx.mlp_train(steps=5, batch=2048, lr=2e-3, keep=0.98, seed=5)
    P2, _ = ctx.mlp_project(want=True)
    P82, s2 = ctx.mlp_project_fp8()
    x2 = np.clip(P2.astype(np.float32) * np.float32(s2), np.float32(-448), np.float32(448))
    assert np.array_equal(_e4m3_value(P82), _e4m3_value(_e4m3_encode(x2)))


def test_learned_fp8_prefilter(ctx, jm100k):
    """The learned phase on FP8 rows (16x16x32 fp8 MFMA) against the same
    phase on bf16 rows, both with a trained network and the plain BF16
    tail: the FP8 prefilter may lose a little recall, the tail judges it."""
    X, rows, gt, init = jm100k
    n, K = init.shape
    ctx.set_data(X)
    ctx.set_partition([0, n // 2, n], [1, 0], [0, 0])
    ctx.set_graph(init)
    ctx.mlp_init(256, 32, seed=7)
    ctx.mlp_train(steps=600, batch=4096, lr=2e-3, keep=0.98, seed=9)
    r_bf, s_bf = _jmerge(ctx, jm100k, nnm.PREC_LEARNED, 0)
    r_f8, s_f8 = _jmerge(ctx, jm100k, nnm.PREC_LEARNED, nnm.FLAG_LEARNED_FP8)
    r_pl, _ = _jmerge(ctx, jm100k, nnm.PREC_BF16, 0)
    print(f"\nlearned bf16 rows {r_bf:.4f} ({s_bf['secs']:.4f} s), fp8 rows {r_f8:.4f} ({s_f8['secs']:.4f} s), "
          f"plain {r_pl:.4f}; offers bf16 {s_bf['offers']} fp8 {s_f8['offers']}")
    assert s_f8["learned_iters"] == 5 and s_f8["plain_iters"] == 3
    assert r_f8 >= r_bf - 0.03


def test_learned_fp8_rejects_other_widths(ctx):
    """FP8 rows are built for d_out = 32: another width is refused with
    NNM_ERR_UNSUPPORTED before any launch, and the context stays usable."""
    X = nnm_data.sift_like(4000, 128, centers=16, seed=75, rank=12)
    n, K = X.shape[0], 16
    ctx.set_data(X)
    ctx.set_partition([0, n // 2, n], [1, 0], [0, 0])
    ctx.random_init(K, seed=2)
    ctx.mlp_init(256, 16, seed=3)
    ctx.mlp_train(steps=5, batch=1024, lr=2e-3, keep=0.98, seed=4)
    p = nnm.default_params(K=K, rule=nnm.RULE_MERGE, max_iters=4, prec=nnm.PREC_LEARNED, plain_tail=1,
                           flags=nnm.FLAG_LEARNED_FP8)
    with pytest.raises(nnm.NNMError) as e:
        ctx.nndescent(p)
    assert e.value.code == nnm.NNM_ERR_UNSUPPORTED
    p.flags = 0
    st = ctx.nndescent(p)  # the bf16 learned phase still runs
    assert st["learned_iters"] >= 1


def test_learned_switch_marks_found_entries(ctx, jm100k):
    """NNM_FLAG_LEARNED_MARK_FOUND: when the lists leave the learned space,
    every entry whose id was not in its row when the learned phase started
    is flagged new; with kept flags some of those are not (so the check
    is not vacuous).  A run of learned iterations only ends with that
    switch."""
    X, rows, gt, init = jm100k
    n, K = init.shape
    ctx.set_data(X)
    ctx.set_partition([0, n // 2, n], [1, 0], [0, 0])
    ctx.set_graph(init)
    ctx.mlp_init(256, 32, seed=7)
    ctx.mlp_train(steps=200, batch=4096, lr=2e-3, keep=0.98, seed=9)
    out = {}
    for name, flags in (("kept", 0), ("found", nnm.FLAG_LEARNED_MARK_FOUND)):
        ctx.set_graph(init)
        p = nnm.default_params(K=K, rule=nnm.RULE_MERGE, rho=0.5, max_iters=5, prec=nnm.PREC_LEARNED, seed=5,
                               plain_tail=0, flags=flags)
        ctx.merge_init(p)
        prior, _, _ = ctx.get_graph()
        st = ctx.nndescent(p)
        ids, dist, fl = ctx.get_graph()
        assert st["learned_iters"] >= 1 and st["plain_iters"] == 0
        in_prior = (ids[:, :, None] == prior[:, None, :]).any(-1)
        out[name] = (in_prior, fl)
    in_prior, fl = out["found"]
    assert (fl[~in_prior] == 1).all(), "an entry the learned phase found is not flagged new"
    in_prior_k, fl_k = out["kept"]
    assert (fl_k[~in_prior_k] == 0).any(), "kept flags already mark every found entry: the check is vacuous"
    print(f"\nfound entries: {(~in_prior).mean():.3f} of all; with kept flags {(fl_k[~in_prior_k] == 0).mean():.3f} "
          f"of them were not new")


def test_projection_follows_weight_changes(ctx):
    """New weights (training or nnm_mlp_set_weights) invalidate the cached
    projection and its FP8 rows: the next use projects again, without an
    explicit nnm_mlp_project."""
    X = nnm_data.sift_like(3000, 128, centers=16, seed=77, rank=12)
    ctx.set_data(X)
    ctx.random_init(16, seed=2)
    ctx.mlp_init(256, 32, seed=3)
    P8a, sa = ctx.mlp_project_fp8()  # projects on first use
    W1, b1, W2, b2, alpha, _ = ctx.mlp_weights()
    ctx.mlp_set_weights(W1, b1, -W2, -b2)
    P8b, sb = ctx.mlp_project_fp8()  # no explicit mlp_project in between
    assert sb == sa
    assert np.array_equal(_e4m3_value(P8b), -_e4m3_value(P8a)), "FP8 rows not rebuilt after new weights"
    ctx.mlp_train(steps=5, batch=1024, lr=2e-3, keep=0.98, seed=4)
    P8c, sc = ctx.mlp_project_fp8()
    P, _ = ctx.mlp_project(want=True)
    x = np.clip(P.astype(np.float32) * np.float32(sc), np.float32(-448), np.float32(448))
    assert np.array_equal(_e4m3_value(P8c), _e4m3_value(_e4m3_encode(x))), "FP8 rows not rebuilt after training"

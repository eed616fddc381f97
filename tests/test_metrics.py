"""The other metrics of the reference (L1, cosine, chi-square): the oracle's
restatements against the reference's own routines (when they are built
from its sources), closed forms and FP64 forms.  Host code only."""
import ctypes as C

import numpy as np
import pytest

import nnm
import orc

L1, COS, CHI2 = orc.METRIC_L1, orc.METRIC_COSINE, orc.METRIC_CHI2


def _cases(seed, count):
    rng = np.random.default_rng(seed)
    for t in range(count):
        d = int(rng.integers(1, 140))
        a = (rng.standard_normal(d) * rng.choice([1e-3, 1.0, 100.0])).astype(np.float32)
        b = (rng.standard_normal(d) * rng.choice([1e-3, 1.0, 100.0])).astype(np.float32)
        if t % 3 == 0:
            a, b = np.abs(a), np.abs(b)
        if t % 7 == 0:
            a[rng.random(d) < 0.5] = 0
            b[rng.random(d) < 0.5] = 0
        if t % 11 == 0:
            a[:] = 0
        yield a, b


def _same(x, y):
    x, y = np.float32(x), np.float32(y)
    return x.tobytes() == y.tobytes() or (np.isnan(x) and np.isnan(y))


def test_oracle_metrics_equal_reference_routines_bit_for_bit():
    R = orc.ref_lib()
    if R is None:
        pytest.skip("the reference's distance routines are not built here")
    fp = C.POINTER(C.c_float)
    for a, b in _cases(5, 2000):
        pa, pb, d = a.ctypes.data_as(fp), b.ctypes.data_as(fp), a.size
        assert _same(R.ref_l1norm(pa, 0, pb, 0, d), orc.dist32(a, b, L1))
        assert _same(R.ref_ksquare(pa, 0, pb, 0, d), orc.dist32(a, b, CHI2))
        assert _same(R.ref_cosine(pa, 0, pb, 0, d), orc.dist32(a, b, COS))


def test_metric_closed_forms():
    a = np.array([1, 2, 3, 4], np.float32)
    b = np.array([4, 0, 3, 1], np.float32)
    assert orc.dist32(a, b, L1) == 3 + 2 + 0 + 3
    # chi-square: 9/5 + 4/2 + 0/6 + 9/5
    assert abs(orc.dist64(a, b, CHI2) - (9 / 5 + 4 / 2 + 0 + 9 / 5)) < 1e-12
    z = np.zeros(4, np.float32)
    assert orc.dist32(z, z, CHI2) == 0.0  # both zero: every element skipped
    e1, e2 = np.eye(4, dtype=np.float32)[:2]
    assert orc.dist32(e1, e2, COS) == 2.0      # orthogonal
    assert orc.dist32(e1, 3 * e1, COS) == 0.0  # parallel
    assert orc.dist32(e1, -e1, COS) == 4.0     # opposite
    assert orc.dist32(z, e1, COS) == 2.0       # a zero norm gives res = 0


def test_fp64_forms_against_numpy():
    for a, b in _cases(9, 300):
        x, y = a.astype(np.float64), b.astype(np.float64)
        assert abs(orc.dist64(a, b, L1) - np.abs(x - y).sum()) <= 1e-9 * max(1.0, np.abs(x - y).sum())
        m = (x != 0) | (y != 0)
        with np.errstate(divide="ignore", invalid="ignore"):
            c2 = ((x - y)[m] ** 2 / (x + y)[m]).sum()
        got = orc.dist64(a, b, CHI2)
        assert (np.isnan(c2) and np.isnan(got)) or (np.isinf(c2) and got == c2) or abs(got - c2) <= 1e-9 * max(1.0, abs(c2))
        na, nb = np.sqrt((x * x).sum()), np.sqrt((y * y).sum())
        cs = 2 - 2 * ((x * y).sum() / (na * nb)) if na and nb else 2.0
        assert abs(orc.dist64(a, b, COS) - cs) <= 1e-12


def test_oracle_brute_force_follows_the_metric_scope():
    rng = np.random.default_rng(4)
    X = rng.integers(0, 20, size=(300, 16)).astype(np.float32)
    rows = np.arange(0, 300, 7, dtype=np.uint32)
    for m in (L1, COS, CHI2):
        with orc.metric_scope(m):
            ids, dist = orc.bruteforce(X, 5, rows=rows)
        for qi, r in enumerate(rows):
            full = np.array([orc.dist64(X[r], X[j], m) if j != r else np.inf for j in range(300)])
            order = np.lexsort((np.arange(300), full))[:5]
            assert np.array_equal(ids[qi], order)
            assert np.allclose(dist[qi], full[order], rtol=0, atol=0)
    ids2, _ = orc.bruteforce(X, 5, rows=rows)  # back to L2 outside the scope
    full = ((X[rows][:, None, :].astype(np.float64) - X[None, :, :]) ** 2).sum(-1)
    full[np.arange(rows.size), rows] = np.inf
    assert np.array_equal(ids2, np.lexsort((np.broadcast_to(np.arange(300), full.shape), full), axis=1)[:, :5])


def test_metric_constants_match_the_library():
    assert (nnm.METRIC_L2, nnm.METRIC_L1, nnm.METRIC_COSINE, nnm.METRIC_CHI2) == (0, 1, 2, 3)
    assert (orc.METRIC_L2, orc.METRIC_L1, orc.METRIC_COSINE, orc.METRIC_CHI2) == (0, 1, 2, 3)

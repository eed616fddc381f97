"""Seeded synthetic inputs for the merge workloads.

Every generator is deterministic in its seed (numpy PCG64).  The workloads:

* ``sift_like(n, d, centers, sigma)``: integer-valued float32 rows in
  [0, 255], drawn from a Gaussian mixture whose centres are uniform in
  [0, 255]^d; x = round(clip(c + N(0, sigma^2), 0, 255)).  Every value is an
  integer <= 255, so squared L2 is exact in float32 for d <= 256 and the
  rows are exactly representable in bfloat16.
* ``sift_like(..., nq=m)``: also m queries from the same mixture (same
  centres and factors, independent draws).
* ``sift_like(..., rank=r)``: the same mixture with SIFT-like low intrinsic
  dimension: x = round(clip(c + A_c z + e, 0, 255)) with z ~ N(0, I_r), a
  random d x r factor A_c per centre scaled so the total per-coordinate
  spread stays sigma, and isotropic noise e ~ N(0, noise^2).
* ``unit_mixture(n, d, centers, sigma)``: Gaussian mixture with centres drawn
  from N(0, I), x = c + N(0, sigma^2 I), then L2-normalised.
* ``log_uniform_leaves(total, lo, hi)``: leaf sizes drawn log-uniformly in
  [lo, hi] until they sum to ``total`` (the last leaf takes the remainder,
  merged into its neighbour if it falls under ``lo``).
"""
import numpy as np

SEED_BASE = 20261203


def sift_like(n, d=128, centers=64, sigma=20.0, seed=SEED_BASE + 1, chunk=1 << 20, rank=None, noise=2.0, nq=0):
    """n rows of the mixture; with nq > 0 also nq query rows drawn from the
    same mixture (same centres and factors) by an independent stream, and
    (X, Q) is returned."""
    rng = np.random.default_rng(seed)
    C = rng.uniform(0.0, 255.0, size=(centers, d)).astype(np.float32)
    lab = rng.integers(0, centers, size=n)
    A = None
    if rank is not None:
        scale = np.sqrt(max(sigma * sigma - noise * noise, 0.0) / rank)
        A = (rng.standard_normal(size=(centers, d, rank)) * scale).astype(np.float32)
    X = _mixture_rows(rng, C, A, lab, d, sigma, noise, rank, chunk)
    if nq:
        qrng = np.random.default_rng([seed, 1])
        qlab = qrng.integers(0, centers, size=nq)
        return X, _mixture_rows(qrng, C, A, qlab, d, sigma, noise, rank, chunk)
    return X


def _mixture_rows(rng, C, A, lab, d, sigma, noise, rank, chunk):
    n, centers = lab.size, C.shape[0]
    X = np.empty((n, d), np.float32)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        lb = lab[s:e]
        if rank is None:
            blk = rng.standard_normal(size=(e - s, d), dtype=np.float32)
            blk *= np.float32(sigma)
        else:
            z = rng.standard_normal(size=(e - s, rank), dtype=np.float32)
            blk = rng.standard_normal(size=(e - s, d), dtype=np.float32)
            blk *= np.float32(noise)
            order = np.argsort(lb, kind="stable")
            bounds = np.searchsorted(lb[order], np.arange(centers + 1))
            for k in range(centers):
                idx = order[bounds[k]:bounds[k + 1]]
                if idx.size:
                    blk[idx] += z[idx] @ A[k].T
        blk += C[lb]
        np.clip(blk, 0.0, 255.0, out=blk)
        np.rint(blk, out=blk)
        X[s:e] = blk
    return X


def unit_mixture(n, d=96, centers=512, sigma=0.35, seed=SEED_BASE + 3, chunk=1 << 20, nq=0):
    """With nq > 0, also nq queries from the same mixture; returns (X, Q)."""
    rng = np.random.default_rng(seed)
    C = rng.standard_normal(size=(centers, d)).astype(np.float32)
    lab = rng.integers(0, centers, size=n)

    def rows(r, lb):
        out = np.empty((lb.size, d), np.float32)
        for s in range(0, lb.size, chunk):
            e = min(lb.size, s + chunk)
            blk = r.standard_normal(size=(e - s, d), dtype=np.float32)
            blk *= np.float32(sigma)
            blk += C[lb[s:e]]
            blk /= np.linalg.norm(blk, axis=1, keepdims=True)
            out[s:e] = blk
        return out

    X = rows(rng, lab)
    if nq:
        qrng = np.random.default_rng([seed, 1])
        return X, rows(qrng, qrng.integers(0, centers, size=nq))
    return X


def log_uniform_leaves(total, lo=5000, hi=80000, seed=SEED_BASE + 33):
    rng = np.random.default_rng(seed)
    sizes = []
    left = total
    while left > 0:
        s = int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
        s = min(s, left)
        sizes.append(s)
        left -= s
    if len(sizes) > 1 and sizes[-1] < lo:
        sizes[-2] += sizes[-1]
        sizes.pop()
    return np.array(sizes, dtype=np.int64)


def random_partition_perm(n, parts, seed):
    """A random assignment of n rows to `parts` equal-size partitions, as a
    permutation that makes each partition a contiguous block."""
    rng = np.random.default_rng(seed)
    return rng.permutation(n)

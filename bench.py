#!/usr/bin/env python3
"""Benchmark of the NN-Descent local-join path on one MI355X.

Workload (one step): a J-Merge of n = 5,000,000 x 128 integer-valued
sift-like rows (256-centre Gaussian mixture, sigma 20, clipped to [0, 255]):
the first half arrives as an NN-Descent k-NN graph (built beforehand, not
timed), the second half is raw.  One step is the reference's timed region,
Jmerge::buildKGraph src/jmerge.cpp:63-70: InitSGraph (distance recompute of
the built half), InitialGraph (cross samples + random lists of the raw half),
the NN-Descent loop (K = 20, rho = 0.5, reverse cap 100, T = 10,
delta = 0.002) and merge2one.  All inputs sit in HBM when a step starts (the
built graph is restored from a device-side checkpoint).

Printed: one JSON line.  value = distances evaluated per second (Gdist/s,
reference nCmps semantics plus the set-up distances of the timed region),
with recall@10 of the merged graph against exact brute force on a seeded
sample of rows, the join kernel's roofline (HIP events on the library's
stream) and a CPU baseline (the oracle's restatement of the reference merge,
single thread, on a bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "nn-merge_amd"))

PEAK_HBM_GBS = 8000.0
PEAK = {"fp32": 157.3, "bf16": 2500.0, "fp8": 5000.0}  # TFLOP/s dense


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_first_half(nnm, ctx, X, half, K, rho, seed):
    ctx.set_data(X[:half])
    ctx.random_init(K, seed=seed)
    p = nnm.default_params(K=K, rule=nnm.RULE_ALL, rho=rho, max_iters=10, seed=seed)
    st = ctx.nndescent(p)
    g0, _, _ = ctx.get_graph()
    return g0, st


def cpu_baseline(nnm, X, half, K, rho, T, sample, seed):
    """The oracle restatement of the reference J-Merge on `sample` rows (half
    from each side of the workload), one thread, timed by the oracle."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    s2 = sample // 2
    Xs = np.ascontiguousarray(np.vstack([X[:s2], X[half:half + s2]]))
    with nnm.Context(0) as c2:
        g0, _ = build_first_half(nnm, c2, Xs, s2, K, rho, seed)
    init = np.zeros((2 * s2, K), np.uint32)
    init[:s2] = g0
    init[s2:] = ((np.arange(2 * s2)[s2:, None] + 1 + np.arange(K)[None, :]) % (2 * s2)).astype(np.uint32)
    ids, _, st = orc.merge(Xs, [0, s2, 2 * s2], [1, 0], init, K=K, rho=rho, max_iters=T, seed=seed)
    dists = st["dists"] + st["init_dists"]
    return {"value": dists / st["secs"] / 1e9, "unit": "Gdist/s", "cores": 1, "kind": "port",
            "sample": f"J-Merge of {2 * s2} rows ({s2} built + {s2} raw) of the same mixture, "
                      f"K={K}, rho={rho}, T={T}, 1 thread; {dists} distances in {st['secs']:.2f} s",
            "secs": st["secs"], "dists": dists}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=5_000_000)
    ap.add_argument("--d", type=int, default=128)
    ap.add_argument("--centers", type=int, default=256)
    ap.add_argument("--K", type=int, default=20)
    ap.add_argument("--rho", type=float, default=0.5)
    ap.add_argument("--iters", type=int, default=10)
    ap.add_argument("--prec", choices=["fp32", "bf16", "learned"], default="fp32")
    ap.add_argument("--recall-rows", type=int, default=20000)
    ap.add_argument("--cpu-sample", type=int, default=100_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-recall", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (the multi-GPU sharding "
                                   "of the local join is a design, not built)"}))
        sys.exit(2)

    import nnm
    import nnm_data

    K, n, d = args.K, args.n, args.d
    half = n // 2
    seed = 1
    t0 = time.time()
    X = nnm_data.sift_like(n, d, centers=args.centers, sigma=20.0, seed=nnm_data.SEED_BASE + 4)
    log(f"data {n}x{d} in {time.time() - t0:.1f}s")

    ctx = nnm.Context(0)
    t0 = time.time()
    g0, bst = build_first_half(nnm, ctx, X, half, K, args.rho, seed + 10)
    log(f"first-half build {time.time() - t0:.1f}s: {bst}")
    ctx.set_data(X)
    ctx.set_partition([0, half, n], [1, 0], [0, 0])
    init = np.empty((n, K), np.uint32)
    init[:half] = g0
    init[half:] = ((np.arange(half, n)[:, None] + 1 + np.arange(K)[None, :]) % n).astype(np.uint32)
    ctx.set_graph(init)
    ctx.graph_save()
    del init

    prec = {"fp32": nnm.PREC_FP32, "bf16": nnm.PREC_BF16, "learned": nnm.PREC_LEARNED}[args.prec]
    p = nnm.default_params(K=K, rule=nnm.RULE_MERGE, rho=args.rho, max_iters=args.iters, prec=prec, seed=seed)

    def step():
        r = ctx.graph_restore(recompute=True)
        i = ctx.merge_init(p)
        s = ctx.nndescent(p)
        ctx.merge_finish(True)
        return r, i, s

    for w in range(args.warmup):
        r, i, s = step()
        log(f"warmup {w}: restore {r['secs']:.3f}s init {i['secs']:.3f}s descent {s['secs']:.3f}s "
            f"iters {s['iters']} dists {s['dists']}")
    recs = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        recs.append(step())
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) / args.steps * 1e3
    for k, (r, i, s) in enumerate(recs):
        log(f"step {k}: restore {r['secs']:.3f}s init {i['secs']:.3f}s descent {s['secs']:.3f}s "
            f"join {s['join_secs']:.3f}s iters {s['iters']} dists {s['dists']} offers {s['offers']}")
    r, i, s = recs[-1]
    dists_step = np.mean([a["init_dists"] + b["init_dists"] + c["dists"] for a, b, c in recs])
    value = dists_step / (ms_per_step * 1e-3) / 1e9

    # join kernel roofline from the library's HIP events (summed over the
    # iterations of a step, so average launch = join_secs / iters)
    jd = np.mean([c["dists"] for _, _, c in recs])
    je = np.mean([c["list_entries"] for _, _, c in recs])
    jo = np.mean([c["offers"] for _, _, c in recs])
    jt = np.mean([c["join_secs"] for _, _, c in recs])
    jit = np.mean([c["iters"] for _, _, c in recs])
    ld = (d + 31) // 32 * 32
    row_bytes = ld * (4 if args.prec == "fp32" else 2)
    bytes_per_iter = (je * (row_bytes + 12) + jo * 8) / jit
    flops_per_iter = jd * 3 * d / jit
    launch_s = jt / jit
    gbs = bytes_per_iter / launch_s / 1e9
    tfs = flops_per_iter / launch_s / 1e12
    peak_c = PEAK["fp32"] if args.prec == "fp32" else PEAK["bf16"]
    frac_m, frac_c = gbs / PEAK_HBM_GBS, tfs / peak_c
    if frac_m >= frac_c:
        roof = {"bound": "hbm", "achieved": round(gbs, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(frac_m, 4), "traffic": None}
    else:
        # compute-bound; the FP32 join runs on the vector ALUs, priced
        # against the published FP32 peak (compute_unit says which units)
        roof = {"bound": "mfma", "achieved": round(tfs, 2), "peak": peak_c, "unit": "TFLOP/s",
                "frac": round(frac_c, 4), "traffic": None}
    roof["compute_unit"] = "fp32 vector ALU" if args.prec == "fp32" else "bf16 MFMA"
    roof.update({"kernel": "k_join_fp32" if args.prec == "fp32" else "k_join_" + args.prec,
                 "launch_ms": round(launch_s * 1e3, 3), "alg_bytes_per_launch": int(bytes_per_iter),
                 "alg_flops_per_launch": int(flops_per_iter), "hbm_frac": round(frac_m, 4),
                 "compute_frac": round(frac_c, 4)})

    out = {"metric": "merge secs at recall@10>=0.95 (learned+BF16 vs plain FP32 GPU), Gdist/s",
           "value": round(value, 3), "unit": "Gdist/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms_per_step, 2), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32" if args.prec == "fp32" else args.prec,
           "data": "synthetic",
           "config": {"workload": f"J-Merge {n}x{d} sift-like ({half} built + {n - half} raw), K={K}, "
                                  f"rho={args.rho}, T={args.iters}, prec={args.prec}",
                      "n": n, "d": d, "K": K, "rho": args.rho, "max_iters": args.iters, "prec": args.prec,
                      "parallelism": "single GPU"},
           "merge_secs": round(ms_per_step * 1e-3, 4), "iters": s["iters"], "dists_per_step": int(dists_step),
           "scan_rate": dists_step / (n / 2.0 * (n - 1.0)), "first_half_build": bst,
           "roofline": roof}
    if not args.no_recall:
        t0 = time.time()
        rng = np.random.default_rng(7)
        rows = np.sort(rng.choice(n, args.recall_rows, replace=False)).astype(np.uint32)
        gt, _ = ctx.bruteforce(10, rows=rows)
        ids, _, _ = ctx.get_graph()
        out["recall@10"] = round(float(nnm.recall(ids, gt, rows, 10)), 5)
        out["recall@10_refcount"] = round(float(nnm.recall(ids, gt, rows, 10, unique=False)), 5)
        out["recall_rows"] = int(args.recall_rows)
        log(f"recall {out['recall@10']} in {time.time() - t0:.1f}s")
    ctx.close()
    if not args.no_cpu:
        t0 = time.time()
        out["cpu_baseline"] = cpu_baseline(nnm, X, half, K, args.rho, args.iters, args.cpu_sample, seed)
        log(f"cpu baseline {time.time() - t0:.1f}s")
    print(json.dumps(out))


if __name__ == "__main__":
    main()

"""Benchmark of yael's exhaustive-kNN / k-means hot path on one MI355X.

Default workload (the headline): exact kNN, nb = 1,000,000 base vectors and
nq = 10,000 queries, d = 128, k = 100, squared L2, vectors drawn from a
1,024-component Gaussian mixture (centres U[0,255]^128, sigma = 20, clipped to
[0,255], rounded to integers), float32.  One step = one full search of the
10,000 queries against the resident base (norm kernels + fused distance /
top-k kernel + split merge), inputs already in HBM.  value = queries/s.

Also reported ("extras"): k-means iterations/s for n = 100,000, d = 128,
k = 1,024, niter = 10 (same mixture, seed 0), device-resident.

The CPU baseline is yael's own C path (oracle/_ref: the reference's sources
compiled with OpenBLAS + OpenMP), timed on the host cores on a bounded sample
of the same workload.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-extras]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3  # MI355X dense FP32 (matrix)
PEAK_BF16_TFLOPS = 2500.0
PEAK_HBM_GBS = 8000.0


def mixture(seed, n, d, ncomp, sigma, integer=True, centres_seed=None):
    """Gaussian-mixture vectors in [0,255]^d (float32)."""
    crng = np.random.default_rng(centres_seed if centres_seed is not None else seed)
    cen = crng.uniform(0, 255, (ncomp, d)).astype(np.float32)
    rng = np.random.default_rng(seed)
    out = np.empty((n, d), np.float32)
    step = 1 << 18
    for i in range(0, n, step):
        m = min(step, n - i)
        x = cen[rng.integers(0, ncomp, m)] + np.float32(sigma) * rng.standard_normal((m, d), np.float32)
        np.clip(x, 0, 255, out=x)
        if integer:
            np.rint(x, out=x)
        out[i:i + m] = x
    return out


class Dev:
    """A padded device copy of an (n, d) float32 array."""

    def __init__(self, ya, x):
        self.ya, self.lib = ya, ya.lib()
        self.n, self.d = x.shape
        self.ld = self.lib.yk_ld(self.d)
        self.ptr = ctypes.c_void_p()
        ya.check(self.lib.yk_malloc(ctypes.byref(self.ptr), self.n * self.ld * 4))
        ya.check(self.lib.yk_upload_vectors(self.ptr, x.ctypes.data_as(ctypes.c_void_p), self.n, self.d))


def dalloc(ya, nbytes):
    p = ctypes.c_void_p()
    ya.check(ya.lib().yk_malloc(ctypes.byref(p), nbytes))
    return p


def prof_avg_ms(ya, slot):
    tot, cnt = ctypes.c_double(), ctypes.c_long()
    ya.check(ya.lib().yk_prof_get(slot, ctypes.byref(tot), ctypes.byref(cnt)))
    return (tot.value / cnt.value if cnt.value else float("nan")), cnt.value


def bench_knn(ya, args, base, queries, k):
    lib = ya.lib()
    db, dq = Dev(ya, base), Dev(ya, queries)
    nb, d = base.shape
    nq = queries.shape[0]
    bn = dalloc(ya, ((nb + 31) // 32 * 32) * 4)
    qn = dalloc(ya, nq * 8)
    oi = dalloc(ya, nq * k * 4)
    od = dalloc(ya, nq * k * 4)

    def step():
        ya.check(lib.yk_norms(db.ptr, nb, d, bn, None))
        ya.check(lib.yk_norms(dq.ptr, nq, d, None, qn))
        ya.check(lib.yk_knn_f32(2, nq, nb, d, k, dq.ptr, qn, db.ptr, bn, None, oi, od))

    for _ in range(args.warmup):
        step()
    ya.check(lib.yk_sync())
    ya.check(lib.yk_prof_reset())
    ya.check(lib.yk_prof_enable(1))
    ya.check(lib.yk_sync())
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ya.check(lib.yk_sync())
    t1 = time.perf_counter()
    ya.check(lib.yk_prof_enable(0))
    fused_ms, nlaunch = prof_avg_ms(ya, 0)
    norm_ms, _ = prof_avg_ms(ya, 1)
    merge_ms, _ = prof_avg_ms(ya, 3)
    idx = np.empty((nq, k), np.int32)
    dis = np.empty((nq, k), np.float32)
    ya.check(lib.yk_memcpy_d2h(idx.ctypes.data_as(ctypes.c_void_p), oi, idx.nbytes))
    ya.check(lib.yk_memcpy_d2h(dis.ctypes.data_as(ctypes.c_void_p), od, dis.nbytes))
    # FP64 error report of the returned distances (GPU FP64 kernel)
    o64 = dalloc(ya, nq * k * 8)
    ya.check(lib.yk_pair_dist_f64(2, nq, d, k, dq.ptr, db.ptr, oi, o64))
    d64 = np.empty((nq, k), np.float64)
    ya.check(lib.yk_memcpy_d2h(d64.ctypes.data_as(ctypes.c_void_p), o64, d64.nbytes))
    rel = np.abs(dis.astype(np.float64) - d64) / np.maximum(np.abs(d64), 1e-30)
    for p in (db.ptr, dq.ptr, bn, qn, oi, od, o64):
        lib.yk_free(p)
    sec = (t1 - t0) / args.steps
    return dict(sec=sec, fused_ms=fused_ms, norm_ms=norm_ms, merge_ms=merge_ms, nlaunch=nlaunch,
                idx=idx, dis=dis, max_rel_err_f64=float(np.nanmax(rel)),
                sorted_ok=bool(np.all(np.diff(dis, axis=1) >= 0)))


def bench_kmeans(ya, v, k, niter, seed):
    lib = ya.lib()
    dv = Dev(ya, v)
    n, d = v.shape
    cent = np.empty((k, d), np.float32)
    ret = ctypes.c_float()
    flags = 0x10000  # quiet
    # warm-up run (allocations, code objects)
    ya.check(lib.yael_kmeans_device(d, n, k, 2, dv.ptr, flags, seed, 1,
                                    cent.ctypes.data_as(ctypes.c_void_p), None, None, None,
                                    ctypes.byref(ret)))
    ya.check(lib.yk_sync())
    t0 = time.perf_counter()
    ya.check(lib.yael_kmeans_device(d, n, k, niter, dv.ptr, flags, seed, 1,
                                    cent.ctypes.data_as(ctypes.c_void_p), None, None, None,
                                    ctypes.byref(ret)))
    ya.check(lib.yk_sync())
    t1 = time.perf_counter()
    lib.yk_free(dv.ptr)
    return dict(sec=t1 - t0, iters=niter, obj=float(ret.value))


def cpu_baseline_knn(base, queries, k, budget_s=12.0):
    """yael's own knn_full_thread (compiled reference, OpenBLAS + OpenMP) on a
    bounded sample of the queries against the full base."""
    os.environ["OPENBLAS_NUM_THREADS"] = "1"
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_bind import Reference, Oracle, have_ref
    cores = len(os.sched_getaffinity(0))
    cores = min(cores, 16)
    if have_ref():
        api, kind = Reference(), "reference"
    else:
        api, kind = Oracle(), "port"
    nq_cal = cores * 32
    t0 = time.perf_counter()
    api.knn_full(2, base, queries[:nq_cal], k, nt=cores)
    cal = time.perf_counter() - t0
    nq_s = int(min(len(queries), max(nq_cal, nq_cal * budget_s / max(cal, 1e-3))))
    t0 = time.perf_counter()
    idx, dis = api.knn_full(2, base, queries[:nq_s], k, nt=cores)
    sec = time.perf_counter() - t0
    return dict(value=nq_s / sec, unit="queries/s", cores=cores, kind=kind,
                sample="%d of the %d queries against all %d base vectors, k=%d, %d OpenMP threads"
                       % (nq_s, len(queries), len(base), k, cores),
                idx=idx, dis=dis, nq=nq_s)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--nb", type=int, default=1_000_000)
    ap.add_argument("--nq", type=int, default=10_000)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (got --gpus %d); the multi-GPU "
              "sharding is described in DESIGN.md" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import yael_amd as ya
    ya.init()
    info = ya.device_info()
    print("# device:", info, file=sys.stderr, flush=True)

    d, k = 128, 100
    t = time.perf_counter()
    base = mixture(1, args.nb, d, 1024, 20, centres_seed=0)
    queries = mixture(2, args.nq, d, 1024, 20, centres_seed=0)
    print("# data generated in %.1fs" % (time.perf_counter() - t), file=sys.stderr, flush=True)

    r = bench_knn(ya, args, base, queries, k)
    nq, nb = args.nq, args.nb
    qps = nq / r["sec"]
    flops = 2.0 * nq * nb * d
    achieved = flops / (r["fused_ms"] * 1e-3) / 1e12
    print("# knn fp32: %.3f ms/step, fused kernel %.3f ms (%d launches), norms %.3f ms, merge %.3f ms"
          % (r["sec"] * 1e3, r["fused_ms"], r["nlaunch"], r["norm_ms"], r["merge_ms"]),
          file=sys.stderr, flush=True)

    out = {
        "metric": "exact kNN queries/s, d=128, k=100, nb=1e6, nq=1e4 (FP32 MFMA form)",
        "value": qps,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["sec"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic: 1,024-component Gaussian mixture, integer-valued float32",
        "config": {"workload": "knn_full type 2 (squared L2), nb=%d, nq=%d, d=%d, k=%d" % (nb, nq, d, k),
                   "nb": nb, "nq": nq, "d": d, "k": k},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None,
                     "kernel": "yk_fused_f32 (distance GEMM + top-k)",
                     "avg_launch_ms": r["fused_ms"]},
        "accuracy": {"max_rel_err_vs_fp64": r["max_rel_err_f64"], "sorted": r["sorted_ok"]},
    }

    if not args.no_extras:
        v = mixture(3, 100_000, 128, 1024, 20, centres_seed=0)
        km = bench_kmeans(ya, v, 1024, 10, 1)
        out["extras"] = {"kmeans_iter_per_s": km["iters"] / km["sec"],
                         "kmeans_config": "n=100000 d=128 k=1024 niter=10, device-resident",
                         "kmeans_obj": km["obj"]}
        print("# kmeans: %.2f iter/s" % (km["iters"] / km["sec"]), file=sys.stderr, flush=True)

    if not args.no_cpu:
        cb = cpu_baseline_knn(base, queries, k)
        g_idx = r["idx"][:cb["nq"]]
        agree = float((np.sort(g_idx, 1) == np.sort(cb["idx"], 1)).mean())
        out["cpu_baseline"] = {"value": cb["value"], "unit": cb["unit"], "cores": cb["cores"],
                               "kind": cb["kind"], "sample": cb["sample"],
                               "index_agreement_with_gpu": agree}
        print("# cpu baseline: %.1f queries/s on %d cores" % (cb["value"], cb["cores"]),
              file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

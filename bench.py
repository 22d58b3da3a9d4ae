"""Benchmark of yael's exhaustive-kNN / k-means hot path on one MI355X.

Workload (the headline): exact kNN, nb = 1,000,000 base vectors and
nq = 10,000 queries, d = 128, k = 100, squared L2, vectors drawn from a
1,024-component Gaussian mixture (centres U[0,255]^128, sigma = 20, clipped to
[0,255], rounded to integers), float32.  One step = one full search of the
10,000 queries against the resident base, inputs already in HBM:
  * FP32 form: norm kernels + fused FP32-MFMA distance / top-k kernels;
  * precision form (the headline value): norm + BF16-conversion kernels,
    BF16-MFMA first pass keeping k' = 200 candidates, FP32 re-rank
    (bit-identical to the FP32 form), certificate, FP32 fallback.
The line reports the precision form's queries/s; the FP32 form's rate, the
ratio and the index/distance agreement between the two are in "forms".

Also reported ("extras"): k-means iterations/s for n = 100,000, d = 128,
k = 1,024, niter = 10 (same mixture), device-resident.

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

from yael_amd.workloads import mixture, zipf_sizes  # noqa: E402

PEAK_FP32_TFLOPS = 157.3  # MI355X dense FP32 (matrix)
PEAK_BF16_TFLOPS = 2500.0  # MI355X dense BF16
PEAK_HBM_GBS = 8000.0

(PROF_FUSED, PROF_NORM, PROF_UPDATE, PROF_MERGE, PROF_RERANK, PROF_BATCH, PROF_BF16,
 PROF_BF16_SKETCH, PROF_F32_SKETCH) = range(9)
NSLOTS = 9


def traffic_from_profile(name):
    """HBM bytes per launch of the dominant kernel, as measured by the
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs summarised in profiles/<name>
    (a line 'hbm_bytes_per_launch: <fetch+write>'); None when absent."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("hbm_bytes_per_launch:"):
                    return float(line.split(":", 1)[1].split()[0])
    except OSError:
        pass
    return None


def prof(lib, slot):
    tot, cnt = ctypes.c_double(), ctypes.c_long()
    lib.yk_prof_get(slot, ctypes.byref(tot), ctypes.byref(cnt))
    return tot.value, cnt.value


def timed(lib, fn, steps, warmup):
    """Run fn warmup + steps times; wall time per step and per-slot kernel
    times (HIP events on the library's stream, summed over the steps)."""
    for _ in range(warmup):
        fn()
    lib.yk_sync()
    lib.yk_prof_reset()
    lib.yk_prof_enable(1)
    lib.yk_sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    lib.yk_sync()
    t1 = time.perf_counter()
    lib.yk_prof_enable(0)
    slots = {s: prof(lib, s) for s in range(NSLOTS)}
    return (t1 - t0) / steps, slots


def bench_knn(ya, dev, args, base, queries, k, kp):
    lib = ya.lib()
    B, Q = dev.DeviceSet(base), dev.DeviceSet(queries)
    nq, nb, d = Q.n, B.n, B.d
    oi, od = dev.DeviceBuffer(nq * k * 4), dev.DeviceBuffer(nq * k * 4)
    oi16, od16 = dev.DeviceBuffer(nq * k * 4), dev.DeviceBuffer(nq * k * 4)
    nfb = ctypes.c_int(0)

    def step_fp32():
        Q.compute_norms()
        B.compute_norms()
        ya.check(lib.yk_knn_f32(2, nq, nb, d, k, Q.vec.ptr, Q.dnorm.ptr, B.vec.ptr, B.fnorm.ptr,
                                None, oi.ptr, od.ptr))

    def step_bf16():
        # the base's float norms come out of its BF16 conversion pass
        Q.compute_norms()
        qh, qst, _ = Q.compute_bf16(2, True)
        bh, _, bmx = B.compute_bf16(2, False, with_norms=True)
        ya.check(lib.yk_knn_bf16(2, nq, nb, d, k, kp, Q.vec.ptr, Q.dnorm.ptr, qh.ptr, qst.ptr,
                                 B.vec.ptr, B.fnorm.ptr, bh.ptr, bmx.ptr, oi16.ptr, od16.ptr,
                                 ctypes.byref(nfb)))

    sec32, sl32 = (0.0, None) if args.only_bf16 else timed(lib, step_fp32, args.steps, args.warmup)
    if not args.only_bf16:
        print("# fp32 form: %.3f ms/step" % (sec32 * 1e3), file=sys.stderr, flush=True)
    sec16, sl16 = timed(lib, step_bf16, args.steps, args.warmup)
    print("# bf16+rerank form: %.3f ms/step, fallback queries %d" % (sec16 * 1e3, nfb.value),
          file=sys.stderr, flush=True)
    i32, d32 = oi.to_host(np.int32, (nq, k)), od.to_host(np.float32, (nq, k))
    i16, d16 = oi16.to_host(np.int32, (nq, k)), od16.to_host(np.float32, (nq, k))
    # FP64 error report of the returned distances (GPU FP64 kernel)
    o64 = dev.DeviceBuffer(nq * k * 8)
    ya.check(lib.yk_pair_dist_f64(2, nq, d, k, Q.vec.ptr, B.vec.ptr, oi16.ptr, o64.ptr))
    d64 = o64.to_host(np.float64, (nq, k))
    rel = np.abs(d16.astype(np.float64) - d64) / np.maximum(np.abs(d64), 1e-30)
    B.free()
    Q.free()
    return dict(sec32=sec32, sl32=sl32, sec16=sec16, sl16=sl16, i32=i32, d32=d32, i16=i16,
                d16=d16, nfb=nfb.value, max_rel_err_f64=float(np.nanmax(rel)),
                sorted_ok=bool(np.all(np.diff(d16, axis=1) >= 0)))


def bench_kmeans(ya, dev, v, k, niter, env=None):
    """k-means on device-resident points, seed 0 (kmeans() then draws its
    seed from lrand48; srand48(0) before every run makes that draw the same
    each time): iterations counted by the library, median of 3 runs; with
    env, the reference's kmeans() on the same points and seed."""
    lib = ya.lib()
    libc = ctypes.CDLL(None)
    V = dev.DeviceSet(v)
    n, d = v.shape
    cent = np.empty((k, d), np.float32)
    asg = np.empty(n, np.int32)
    ret = ctypes.c_float()
    flags = 0x10000  # quiet

    def run(it):
        libc.srand48(ctypes.c_long(0))
        ya.check(lib.yael_kmeans_device(d, n, k, it, V.vec.ptr, flags, 0, 1,
                                        cent.ctypes.data_as(ctypes.c_void_p), None,
                                        asg.ctypes.data_as(ctypes.c_void_p), None,
                                        ctypes.byref(ret)))
        lib.yk_sync()
        st = [ctypes.c_long() for _ in range(3)]
        lib.yael_kmeans_last_stats(*[ctypes.byref(x) for x in st])
        return st[0].value
    run(2)
    recs = []
    for _ in range(3):
        t0 = time.perf_counter()
        iters = run(niter)
        recs.append((time.perf_counter() - t0, iters))
    V.free()
    sec, iters = sorted(recs)[1]
    rec = dict(sec=sec, iters=iters, obj=float(ret.value))
    if env:
        cb = cpu_baseline_kmeans(v, k, niter, env)
        rec["cpu"] = {key: val for key, val in cb.items() if key != "assign"}
        rec["cpu"]["assignment_agreement_with_gpu"] = float(np.mean(cb["assign"] == asg))
        rec["cpu"]["objective_rel_diff_gpu_vs_cpu"] = float(ret.value) / cb["objective"] - 1.0
    return rec


def bench_kmeanspp(ya, dev, v, k, seed, cpu):
    """k-means++ seeding: one k-means step after k-means++ seeding minus one
    step after random seeding, device-resident; with cpu, the same two runs
    of the compiled reference on the host (OpenMP distance steps)."""
    lib = ya.lib()
    V = dev.DeviceSet(v)
    n, d = v.shape
    cent = np.empty((k, d), np.float32)
    asg = np.empty(n, np.int32)
    ret = ctypes.c_float()

    def run(flags):
        lib.yk_sync()
        t0 = time.perf_counter()
        ya.check(lib.yael_kmeans_device(d, n, k, 1, V.vec.ptr, flags, seed, 1,
                                        cent.ctypes.data_as(ctypes.c_void_p), None,
                                        asg.ctypes.data_as(ctypes.c_void_p), None,
                                        ctypes.byref(ret)))
        lib.yk_sync()
        return time.perf_counter() - t0
    run(0x10000 | 0x20000)
    t_rand = min(run(0x10000) for _ in range(2))
    t_kpp = min(run(0x10000 | 0x20000) for _ in range(2))
    V.free()
    rec = dict(init_s=t_kpp - t_rand, obj=float(ret.value))
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_bind import Reference, have_ref
        if have_ref():
            cores = cpu_env()["threads"]
            ref = Reference()
            t0 = time.perf_counter()
            ref.kmeans(v, k, 1, flags=0x10000 | cores, seed=seed)
            t1 = time.perf_counter()
            r, _, _, a_ref, _ = ref.kmeans(v, k, 1, flags=0x10000 | 0x20000 | cores, seed=seed)
            t2 = time.perf_counter()
            rec.update(cpu_reference_init_s=(t2 - t1) - (t1 - t0), cpu_threads=cores,
                       same_result_as_reference=bool(np.float32(r) == np.float32(rec["obj"])
                                                     and np.array_equal(a_ref, asg)))
    return rec


def bench_hkm(ya, base, cpu, n=200_000, nlevel=2, bf=32, niter=10, nq_cpu=20_000):
    """Hierarchical k-means: hkm_learn on n mixture points (every level one
    batched k-means) and hkm_quantize of the whole base; with cpu, the
    reference's own hkm_learn on the same points and hkm_quantize on a
    sample, on the host (YAEL_COUNT_CPU threads)."""
    from yael_amd import hkm
    libc = ctypes.CDLL(None)
    v = np.ascontiguousarray(base[:n])
    libc.srand48(ctypes.c_long(1))
    t, _ = hkm.hkm_learn(v[:20000], 2, 4, niter=2)  # warm-up (builds, caches)
    libc.srand48(ctypes.c_long(1))
    t0 = time.perf_counter()
    t, labels = hkm.hkm_learn(v, nlevel, bf, niter=niter)
    t1 = time.perf_counter()
    idx = t.quantize(base)
    t2 = time.perf_counter()
    idx = t.quantize(base)
    t3 = time.perf_counter()
    rec = dict(config="hkm_learn n=%d d=%d nlevel=%d bf=%d (%d leaves) niter=%d, mixture data; "
                      "hkm_quantize of %d points" % (n, v.shape[1], nlevel, bf, bf ** nlevel, niter,
                                                     len(base)),
               learn_s=t1 - t0, quantize_points_per_s=len(base) / (t3 - t2))
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_bind import Reference, have_ref, make_hkm
        if have_ref():
            cores = cpu_env()["threads"]
            os.environ["YAEL_COUNT_CPU"] = str(cores)
            ref = Reference()
            libc.srand48(ctypes.c_long(1))
            t0 = time.perf_counter()
            lv, lab_ref, h = ref.hkm_learn(v, nlevel, bf, niter)
            t1 = time.perf_counter()
            q = np.ascontiguousarray(base[:nq_cpu])
            t2 = time.perf_counter()
            qi_ref = ref.hkm_quantize(h, q)
            t3 = time.perf_counter()
            def qerr(leaves, lab):
                # mean squared distance of the points to their leaf centroid (FP64)
                c = leaves[lab].astype(np.float64)
                return float(((v.astype(np.float64) - c) ** 2).sum(1).mean())
            # leaf-level agreement is low once a level-0 near-tie moves one
            # point: the children's random seeding is by position in the node
            rec.update(cpu_reference_learn_s=t1 - t0, cpu_threads=cores,
                       cpu_reference_quantize_points_per_s=nq_cpu / (t3 - t2),
                       cpu_quantize_sample=nq_cpu,
                       level0_agreement_with_reference=float(
                           (lab_ref // bf ** (nlevel - 1) == labels // bf ** (nlevel - 1)).mean()),
                       leaf_agreement_with_reference=float((lab_ref == labels).mean()),
                       mean_sq_dist_to_leaf=qerr(t.centroids(nlevel - 1), labels),
                       reference_mean_sq_dist_to_leaf=qerr(lv[nlevel - 1], lab_ref))
    return rec


def bench_dropin(ya, base, queries, k):
    """knn_full through the C API on host arrays (the drop-in call: uploads,
    norms, search, downloads), automatic form vs the FP32 form only."""
    lib = ya.lib()

    def rate(mode, reps):
        old = lib.yael_knn_precision(mode)
        try:
            ya.knn(queries, base, k)
            t0 = time.perf_counter()
            for _ in range(reps):
                idx, dis = ya.knn(queries, base, k)
            sec = (time.perf_counter() - t0) / reps
        finally:
            lib.yael_knn_precision(old)
        return len(queries) / sec, idx, dis
    r0, i0, d0 = rate(0, 3)
    r1, i1, d1 = rate(1, 2)
    return dict(config="knn_full(2, nq=%d, nb=%d, d=%d, k=%d) on host arrays, PCIe copies "
                       "included" % (len(queries), len(base), base.shape[1], k),
                queries_per_s=r0, fp32_form_queries_per_s=r1,
                same_bits_as_fp32_form=bool(np.array_equal(i0, i1) and
                                            np.array_equal(d0.view(np.int32), d1.view(np.int32))))


def bench_kmeans_large(ya, dev, n=10_000_000, d=128, k=65536, niter=20, seed=1, nsub=8,
                       ksub=256, env=None, cpu_points=8192):
    """n=1e7, d=128, k=65,536 k-means with the BF16 assign form (BF16 first
    pass, FP32 re-rank of each point's 8 best centroids, certificate, FP32
    search of the uncertified points), FP64 sums and objective; data from a
    65,536-component mixture (sigma 15, clipped to [0,255]) generated in HBM.
    Then the assign step alone on the final centroids in both forms, the
    FP64 objective of that assignment, and 8 PQ sub-spaces (d=16, k=256)
    trained in one batched run."""
    lib = ya.lib()
    t = time.perf_counter()
    C0 = dev.DeviceSet.synth(65536, d, 0, 41, 0.0, 255.0)
    V = dev.DeviceSet.synth(n, d, 1, 42, 0.0, 255.0, centres=C0, sigma=15.0)
    C0.free()
    lib.yk_sync()
    out = {"generate_s": time.perf_counter() - t}
    cent = np.empty((k, d), np.float32)
    ret = ctypes.c_float()
    # one untimed iteration first: the assign step's launch plan is tuned on
    # the first call of its shape
    ya.check(lib.yael_kmeans_device(d, n, k, 1, V.vec.ptr, 0x10000 | 0x1000000, seed, 1,
                                    cent.ctypes.data_as(ctypes.c_void_p), None, None, None,
                                    ctypes.byref(ret)), "kmeans bf16 warm-up")
    lib.yk_prof_reset()
    lib.yk_prof_enable(1)
    lib.yk_sync()
    t0 = time.perf_counter()
    ya.check(lib.yael_kmeans_device(d, n, k, niter, V.vec.ptr, 0x10000 | 0x1000000, seed, 1,
                                    cent.ctypes.data_as(ctypes.c_void_p), None, None, None,
                                    ctypes.byref(ret)), "kmeans bf16")
    lib.yk_sync()
    sec = time.perf_counter() - t0
    lib.yk_prof_enable(0)
    sl = {s_: prof(lib, s_) for s_ in range(NSLOTS)}
    st = [ctypes.c_long() for _ in range(3)]
    lib.yael_kmeans_last_stats(*[ctypes.byref(x) for x in st])
    iters, assigned, fallback = [x.value for x in st]
    out.update(sec=sec, iters=iters, obj=float(ret.value), assigned=assigned, fallback=fallback,
               bf16_ms=sl[PROF_BF16][0], rerank_ms=sl[PROF_RERANK][0], fp32_ms=sl[PROF_FUSED][0],
               update_ms=sl[PROF_UPDATE][0])
    print("# kmeans n=%d k=%d: %d iterations in %.2f s (%.3f iter/s), bf16 scan %.1f ms, "
          "rerank %.1f ms, update %.1f ms, fp32 fallback %d of %d assignments"
          % (n, k, iters, sec, iters / sec, sl[PROF_BF16][0], sl[PROF_RERANK][0],
             sl[PROF_UPDATE][0], fallback, assigned), file=sys.stderr, flush=True)

    # the assign step alone, both forms, on the final centroids
    C = dev.DeviceSet(cent)
    V.compute_norms()
    C.compute_norms()
    qh, qst, _ = V.compute_bf16(2, True)
    bh, _, bmx = C.compute_bf16(2, False)
    oi16, od16 = dev.DeviceBuffer(n * 4), dev.DeviceBuffer(n * 4)
    oi32, od32 = dev.DeviceBuffer(n * 4), dev.DeviceBuffer(n * 4)
    nfb = ctypes.c_int(0)

    def a16():
        ya.check(lib.yk_knn_bf16(2, n, k, d, 1, 8, V.vec.ptr, V.dnorm.ptr, qh.ptr, qst.ptr,
                                 C.vec.ptr, C.fnorm.ptr, bh.ptr, bmx.ptr, oi16.ptr, od16.ptr,
                                 ctypes.byref(nfb)))

    def a32():
        ya.check(lib.yk_knn_f32(2, n, k, d, 1, V.vec.ptr, V.dnorm.ptr, C.vec.ptr, C.fnorm.ptr,
                                None, oi32.ptr, od32.ptr))
    s16, sl16 = timed(lib, a16, 1, 0)
    s32, sl32 = timed(lib, a32, 1, 0)
    i16, d16 = oi16.to_host(np.int32, n), od16.to_host(np.float32, n)
    i32, d32 = oi32.to_host(np.int32, n), od32.to_host(np.float32, n)
    o64 = dev.DeviceBuffer(n * 8)
    ya.check(lib.yk_pair_dist_f64(2, n, d, 1, V.vec.ptr, C.vec.ptr, oi16.ptr, o64.ptr))
    d64 = o64.to_host(np.float64, n)
    obj32, obj64 = float(np.sum(d16.astype(np.float64))), float(np.sum(d64))
    flops = 2.0 * n * k * d
    out["assign"] = {
        "bf16_rerank_s": s16, "fp32_s": s32, "speedup": s32 / s16,
        "bf16_scan_ms": sl16[PROF_BF16][0], "rerank_ms": sl16[PROF_RERANK][0],
        "fp32_scan_ms": sl32[PROF_FUSED][0],
        "bf16_scan_tflops": flops / (sl16[PROF_BF16][0] * 1e-3) / 1e12,
        "fp32_scan_tflops": flops / (sl32[PROF_FUSED][0] * 1e-3) / 1e12,
        "fallback_points": nfb.value,
        "assign_identical": bool(np.array_equal(i16, i32)),
        "dis_bit_identical": bool(np.array_equal(d16.view(np.int32), d32.view(np.int32))),
        "objective_fp32": obj32, "objective_fp64": obj64,
        "objective_rel_err_vs_fp64": abs(obj32 - obj64) / abs(obj64),
    }
    print("# assign alone: bf16+rerank %.3f s, fp32 %.3f s (x%.2f), identical %s, "
          "objective rel err vs fp64 %.2e" % (s16, s32, s32 / s16, out["assign"]["assign_identical"],
                                              out["assign"]["objective_rel_err_vs_fp64"]),
          file=sys.stderr, flush=True)
    if env:
        # the assign step of the reference on a sample of the points against
        # the same final centroids; its per-iteration time at n follows
        vs = np.empty((cpu_points, d), np.float32)
        ya.check(lib.yk_download_vectors(vs.ctypes.data_as(ctypes.c_void_p), V.vec.ptr,
                                         cpu_points, d))
        cb = cpu_baseline_assign(vs, cent, env, "the first %d points" % cpu_points)
        cb["index_agreement_with_gpu"] = float(np.mean(cb.pop("idx")[:, 0] == i16[:cpu_points]))
        cb.pop("dis")
        cb["assign_step_s_at_full_n"] = n / cb["value"]
        cb["gpu_assign_step_s"] = s16
        out["assign"]["cpu_baseline"] = cb
    for b in (oi16, od16, oi32, od32, o64):
        b.free()
    C.free()

    # product-quantiser sub-spaces, one batched run
    cq = np.empty((nsub, ksub, d // nsub), np.float32)
    rq = np.empty(nsub, np.float32)
    itq = np.empty(nsub, np.int32)
    chunk = ctypes.c_int(0)
    seeds = np.arange(1, nsub + 1, dtype=np.int64)

    def pq(it):
        ya.check(lib.yael_kmeans_subspaces_device(d, n, nsub, ksub, it, V.vec.ptr, 0x10000,
                                                  seeds.ctypes.data, cq.ctypes.data, None, None,
                                                  None, rq.ctypes.data, itq.ctypes.data,
                                                  ctypes.byref(chunk)), "subspaces")
    pq(1)  # the update chunk is autotuned on the first call of a shape
    lib.yk_sync()
    t0 = time.perf_counter()
    pq(niter)
    lib.yk_sync()
    out["pq"] = {"sec": time.perf_counter() - t0, "iters": itq.tolist(), "chunk": chunk.value,
                 "objective": rq.tolist()}
    print("# pq %d x (d=%d, k=%d): %.3f s for %s iterations" % (nsub, d // nsub, ksub,
                                                               out["pq"]["sec"], itq.tolist()),
          file=sys.stderr, flush=True)
    V.free()
    return out


def bench_precision_d960(ya, dev, nb=1_000_000, nq=10_000, d=960, k=10, kp=64, steps=3,
                         env=None, fp32=True):
    """Precision form at d=960: |N(0,1)| vectors scaled to unit norm,
    generated in HBM; k=10 from k'=64 BF16 candidates.  With env, the
    reference's knn_full on a sample of the same queries against the whole
    base (copied to the host)."""
    lib = ya.lib()
    B = dev.DeviceSet.synth(nb, d, 2, 51, unit_rows=True)
    Q = dev.DeviceSet.synth(nq, d, 2, 52, unit_rows=True)
    oi, od = dev.DeviceBuffer(nq * k * 4), dev.DeviceBuffer(nq * k * 4)
    oi16, od16 = dev.DeviceBuffer(nq * k * 4), dev.DeviceBuffer(nq * k * 4)
    nfb = ctypes.c_int(0)

    def f32():
        Q.compute_norms()
        B.compute_norms()
        ya.check(lib.yk_knn_f32(2, nq, nb, d, k, Q.vec.ptr, Q.dnorm.ptr, B.vec.ptr, B.fnorm.ptr,
                                None, oi.ptr, od.ptr))

    def b16():
        Q.compute_norms()
        qh, qst, _ = Q.compute_bf16(2, True)
        bh, _, bmx = B.compute_bf16(2, False, with_norms=True)
        ya.check(lib.yk_knn_bf16(2, nq, nb, d, k, kp, Q.vec.ptr, Q.dnorm.ptr, qh.ptr, qst.ptr,
                                 B.vec.ptr, B.fnorm.ptr, bh.ptr, bmx.ptr, oi16.ptr, od16.ptr,
                                 ctypes.byref(nfb)))
    s16, sl16 = timed(lib, b16, steps, 1)
    if not fp32:  # profile runs: the precision form alone
        for b in (oi, od, oi16, od16):
            b.free()
        B.free()
        Q.free()
        return {"bf16_rerank_queries_per_s": nq / s16,
                "bf16_scan_ms_per_step": sl16[PROF_BF16][0] / steps,
                "bf16_sketch_ms_per_step": sl16[PROF_BF16_SKETCH][0] / steps,
                "rerank_ms_per_step": sl16[PROF_RERANK][0] / steps, "fallback_queries": nfb.value}
    s32, sl32 = timed(lib, f32, 2, 1)
    i32, d32 = oi.to_host(np.int32, (nq, k)), od.to_host(np.float32, (nq, k))
    i16, d16 = oi16.to_host(np.int32, (nq, k)), od16.to_host(np.float32, (nq, k))
    o64 = dev.DeviceBuffer(nq * k * 8)
    ya.check(lib.yk_pair_dist_f64(2, nq, d, k, Q.vec.ptr, B.vec.ptr, oi16.ptr, o64.ptr))
    d64 = o64.to_host(np.float64, (nq, k))
    rel = np.abs(d16.astype(np.float64) - d64) / np.maximum(np.abs(d64), 1e-30)
    flops = 2.0 * nq * nb * d
    cpu = None
    if env:
        bh_ = np.empty((nb, d), np.float32)
        ya.check(lib.yk_download_vectors(bh_.ctypes.data_as(ctypes.c_void_p), B.vec.ptr, nb, d))
        qh_ = np.empty((nq, d), np.float32)
        ya.check(lib.yk_download_vectors(qh_.ctypes.data_as(ctypes.c_void_p), Q.vec.ptr, nq, d))
        cb = cpu_baseline_knn(bh_, qh_, k, env)
        cpu = {key: val for key, val in cb.items() if key not in ("idx", "dis", "nq")}
        cpu["index_agreement_with_gpu"] = float(np.mean(np.all(cb["idx"] == i16[:cb["nq"]],
                                                               axis=1)))
        del bh_
    for b in (oi, od, oi16, od16, o64):
        b.free()
    B.free()
    Q.free()
    r = {"config": "nb=%d nq=%d d=%d k=%d k'=%d, |N(0,1)| rows scaled to unit L2 norm"
                   % (nb, nq, d, k, kp),
         "bf16_rerank_queries_per_s": nq / s16, "fp32_queries_per_s": nq / s32,
         "speedup": s32 / s16,
         "bf16_scan_tflops": flops / (sl16[PROF_BF16][0] / steps * 1e-3) / 1e12,
         "fp32_scan_tflops": flops / (sl32[PROF_FUSED][0] / 2 * 1e-3) / 1e12,
         "index_agreement_with_fp32": float(np.mean(np.all(i16 == i32, axis=1))),
         "distances_bit_identical_to_fp32": bool(np.array_equal(d16.view(np.int32),
                                                                d32.view(np.int32))),
         "max_rel_err_vs_fp64": float(np.nanmax(rel)), "fallback_queries": nfb.value}
    if cpu:
        r["cpu_baseline"] = cpu
    print("# d=960: bf16+rerank %.1f q/s, fp32 %.1f q/s (x%.2f), agreement %.4f, fallback %d"
          % (nq / s16, nq / s32, s32 / s16, r["index_agreement_with_fp32"], nfb.value),
          file=sys.stderr, flush=True)
    return r


def tuned_plans(tag=None):
    """The launch plans the library's autotuner recorded (shape class -> plan
    index), read from its tuning cache file; with tag, only the shape
    classes whose key contains it."""
    path = os.environ.get("YAEL_MI355X_TUNE_CACHE") or os.path.join(
        os.path.expanduser("~"), ".cache", "yael_mi355x_tune.txt")
    out = {}
    try:
        with open(path) as f:
            for line in f:
                kv = line.split()
                if len(kv) == 2 and (tag is None or tag in kv[0]):
                    out[kv[0]] = int(kv[1])
    except OSError:
        pass
    return out


def bench_batched(ya, dev, ninst=256, total=2_000_000, d=128, k=64, niter=15, env=None):
    """hkm level-2 pattern: 256 independent k-means instances of Zipf sizes,
    one batched run (grouped launches, graph-replayed iteration); with env,
    the reference's loop of kmeans() calls over the same instances."""
    lib = ya.lib()
    sizes = zipf_sizes(ninst, total, 256, 120_000)
    v = mixture(4, int(sizes.sum()), d, 1024, 20, centres_seed=0)
    V = dev.DeviceSet(v)
    cent = np.empty((ninst, k, d), np.float32)
    ret = np.empty(ninst, np.float32)
    iters = np.empty(ninst, np.int32)
    seeds = np.arange(1, ninst + 1, dtype=np.int64)
    chunk = ctypes.c_int(0)

    def run(graph):
        ya.check(lib.yael_kmeans_batched_device(d, ninst, sizes.ctypes.data, k, niter, V.vec.ptr,
                                                0x10000, seeds.ctypes.data, cent.ctypes.data,
                                                None, None, None, ret.ctypes.data,
                                                iters.ctypes.data, graph, ctypes.byref(chunk)))
    run(1)  # warm-up (and autotuning on a fresh cache)
    lib.yk_sync()
    lib.yk_prof_reset()
    lib.yk_prof_enable(1)
    t0 = time.perf_counter()
    run(1)
    lib.yk_sync()
    t1 = time.perf_counter()
    lib.yk_prof_enable(0)
    loop_ms, _ = prof(lib, PROF_BATCH)
    # the same instances one kmeans() call at a time (the reference's hkm loop)
    off = np.concatenate([[0], np.cumsum(sizes)])
    c1 = np.empty((k, d), np.float32)
    r1 = ctypes.c_float()
    t2 = time.perf_counter()
    for i in range(ninst):
        ya.check(lib.yael_kmeans_device(d, int(sizes[i]), k, niter,
                                        ctypes.c_void_p(V.vec.ptr.value + int(off[i]) * V.ld * 4),
                                        0x10000, int(seeds[i]), 1, c1.ctypes.data, None, None, None,
                                        ctypes.byref(r1)))
    lib.yk_sync()
    t3 = time.perf_counter()
    V.free()
    out = dict(sizes=sizes, call_s=t1 - t0, loop_ms=loop_ms, seq_s=t3 - t2, chunk=chunk.value,
               iters_max=int(iters.max()), iters_min=int(iters.min()))
    if env:
        cb = cpu_baseline_batched(v, sizes, k, niter, seeds, env)
        rel = np.asarray(ret, np.float64) / np.asarray(cb.pop("objectives")) - 1.0
        cb["objective_rel_diff_gpu_vs_cpu"] = {"mean": float(rel.mean()),
                                               "max_abs": float(np.abs(rel).max())}
        out["cpu"] = cb
    return out


# ---------------------------------------------------------------------------
# CPU baselines: yael's own C path (oracle/_ref, the reference's sources
# compiled here against OpenBLAS, OpenMP threads), or the oracle's
# restatement of it where the compiled reference is absent.


def cpu_env():
    """Facts of the host the CPU baselines run on.  threads = the OpenMP
    thread count the run uses: OMP_NUM_THREADS when set (a shared GPU host
    gives each GPU's jobs that share of its cores), else every CPU this
    process may run on."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    logical = len(os.sched_getaffinity(0))
    omp = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(logical, int(omp)) if omp.isdigit() and int(omp) > 0 else logical
    blas = "unknown"
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_bind import openblas_path
    path = openblas_path()
    if path:
        try:
            ob = ctypes.CDLL(path)
            ob.scipy_openblas_get_config.restype = ctypes.c_char_p
            blas = "%s (%s)" % (ob.scipy_openblas_get_config().decode(), os.path.basename(path))
        except (OSError, AttributeError):
            blas = os.path.basename(path)
    return dict(cpu_model=model, logical_cpus_visible=logical, threads=threads,
                omp_num_threads=omp or None, blas=blas, blas_threads_inside_yael=1)


def cpu_api():
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_bind import Reference, Oracle, have_ref
    os.environ["OPENBLAS_NUM_THREADS"] = "1"  # BLAS single-threaded inside yael's threads
    if have_ref():
        return Reference(), "reference"
    o = Oracle()
    o.set_blas(0, 0)
    return o, "port"


def median_runs(fn, reps=3, budget_s=40.0):
    """Median wall time of up to reps runs of fn (fewer when one run alone
    exceeds budget_s / reps); returns (median seconds, runs, last result)."""
    times, out = [], None
    for _ in range(reps):
        t0 = time.perf_counter()
        out = fn()
        times.append(time.perf_counter() - t0)
        if sum(times) > budget_s or times[0] > budget_s / reps:
            break
    return float(np.median(times)), len(times), out


def cpu_baseline_knn(base, queries, k, env, budget_s=4.0):
    """knn_full_thread on a bounded sample of the queries against the full
    base (about budget_s per run), median of 3 runs."""
    api, kind = cpu_api()
    nt = env["threads"]
    nq_cal = nt * 8
    t0 = time.perf_counter()
    api.knn_full(2, base, queries[:nq_cal], k, nt=nt)
    cal = time.perf_counter() - t0
    nq_s = int(min(len(queries), max(nq_cal, nq_cal * budget_s / max(cal, 1e-3))))
    sec, runs, (idx, dis) = median_runs(lambda: api.knn_full(2, base, queries[:nq_s], k, nt=nt))
    return dict(value=nq_s / sec, unit="queries/s", cores=nt, kind=kind,
                sample="%d of the %d queries against all %d base vectors, d=%d, k=%d; %d "
                       "OpenMP threads; median of %d runs"
                       % (nq_s, len(queries), len(base), base.shape[1], k, nt, runs),
                idx=idx, dis=dis, nq=nq_s)


def cpu_baseline_kmeans(v, k, niter, env):
    """kmeans() of the reference on the same points, seed 0 (its seed drawn
    from lrand48 after srand48(0), as the GPU run draws it), median of 3."""
    api, kind = cpu_api()
    libc = ctypes.CDLL(None)
    nt = env["threads"]

    def run():
        libc.srand48(ctypes.c_long(0))
        return api.kmeans(v, k, niter, flags=0x10000 | nt, seed=0)
    sec, runs, (obj, cent, dis, asg, nas) = median_runs(run)
    return dict(value=niter / sec, unit="iterations/s", cores=nt, kind=kind, objective=obj,
                sample="the whole workload: n=%d d=%d k=%d niter=%d (a fixed-iteration run: "
                       "the reference stops only on a repeated objective), %d OpenMP threads; "
                       "median of %d runs" % (len(v), v.shape[1], k, niter, nt, runs),
                assign=asg)


def cpu_baseline_batched(v, sizes, k, niter, seeds, env):
    """The reference's hkm loop: one kmeans() call per instance (hkm.c:66-103),
    each with its own seed; grouped iterations/s = niter / the loop's time."""
    api, kind = cpu_api()
    nt = env["threads"]
    off = np.concatenate([[0], np.cumsum(sizes)])

    def run():
        objs = []
        for i in range(len(sizes)):
            objs.append(api.kmeans(v[off[i]:off[i + 1]], k, niter, flags=0x10000 | nt,
                                   seed=int(seeds[i]))[0])
        return objs
    sec, runs, objs = median_runs(run)
    return dict(value=niter / sec, unit="grouped iterations/s", cores=nt, kind=kind,
                seconds=sec, objectives=objs,
                sample="all %d instances (sum n=%d, d=%d, k=%d, niter=%d), one kmeans() call "
                       "each, %d OpenMP threads; median of %d runs"
                       % (len(sizes), int(sizes.sum()), v.shape[1], k, niter, nt, runs))


def cpu_baseline_assign(points, cent, env, what):
    """One k-means assign step (nn(): the nearest centroid of each point,
    knn_full_thread with k=1) of a sample of the points against all the
    centroids; the unit is points/s, from which the per-iteration time of
    the whole set follows."""
    api, kind = cpu_api()
    nt = env["threads"]
    sec, runs, (idx, dis) = median_runs(lambda: api.knn_full(2, cent, points, 1, nt=nt))
    return dict(value=len(points) / sec, unit="points/s (assign step)", cores=nt, kind=kind,
                sample="%s: %d points against all %d centroids (d=%d), k=1; %d OpenMP "
                       "threads; median of %d runs" % (what, len(points), len(cent),
                                                       cent.shape[1], nt, runs),
                idx=idx, dis=dis)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--no-large", action="store_true",
                    help="skip the d=960 precision-form and n=1e7, k=65536 k-means extras")
    ap.add_argument("--only-bf16", action="store_true", help="profile runs: the precision form only")
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="profile runs: first copy this many MiB three times with a plain "
                         "copy kernel (the reference point of the HBM traffic counters)")
    ap.add_argument("--only", choices=["kmeans", "kmeans-large", "d960", "d960-bf16", "batched"],
                    default=None,
                    help="profile runs: run one extra workload alone and print its record")
    ap.add_argument("--nb", type=int, default=1_000_000)
    ap.add_argument("--nq", type=int, default=10_000)
    ap.add_argument("--kprime", type=int, default=200)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this library runs on one GPU only (got --gpus %d): the base "
              "(kNN) and point (k-means) sharding over several GPUs is not built"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    import yael_amd as ya
    from yael_amd import device as dev
    env = None if args.no_cpu else cpu_env()
    ya.init()
    print("# device:", ya.device_info(), file=sys.stderr, flush=True)

    if args.calib_copy_mb:
        nbytes = args.calib_copy_mb << 20
        src, dst = dev.DeviceBuffer(nbytes), dev.DeviceBuffer(nbytes)
        ya.check(ya.lib().yk_memset(src.ptr, 0, nbytes))
        for _ in range(3):
            ya.check(ya.lib().yk_copy_kernel(dst.ptr, src.ptr, nbytes))
        ya.lib().yk_sync()
        src.free()
        dst.free()

    if args.only:
        only = {"kmeans": lambda ya_, dev_: bench_kmeans(
                    ya_, dev_, mixture(3, 100_000, 128, 1024, 20, centres_seed=0), 1024, 10),
                "kmeans-large": bench_kmeans_large, "d960": bench_precision_d960,
                "d960-bf16": lambda ya_, dev_: bench_precision_d960(ya_, dev_, fp32=False),
                "batched": bench_batched}[args.only]
        rec = only(ya, dev)
        print(json.dumps(rec, default=lambda x: x.tolist() if hasattr(x, "tolist") else float(x)),
              flush=True)
        return

    d, k = 128, 100
    t = time.perf_counter()
    base = mixture(1, args.nb, d, 1024, 20, centres_seed=0)
    queries = mixture(2, args.nq, d, 1024, 20, centres_seed=0)
    print("# data generated in %.1fs" % (time.perf_counter() - t), file=sys.stderr, flush=True)

    r = bench_knn(ya, dev, args, base, queries, k, args.kprime)
    nq, nb = args.nq, args.nb
    flops = 2.0 * nq * nb * d
    qps16 = nq / r["sec16"]
    # dominant kernel: the BF16 filtered pass (one launch per step); its
    # algorithmic work is the whole distance matrix, 2*nq*nb*d flop.  The
    # sketch pass (a 1-in-stride sample) is reported beside it.
    bf_ms, bf_n = r["sl16"][PROF_BF16]
    sk_ms, sk_n = r["sl16"][PROF_BF16_SKETCH]
    bf_ms_launch = bf_ms / max(bf_n, 1)
    ach16 = flops / (bf_ms_launch * 1e-3) / 1e12
    traffic = traffic_from_profile("knn_nb1e6_nq1e4_d128_k100_traffic.md")
    out = {
        "metric": "exact kNN queries/s at d=128 (BF16+FP32-rerank form, exact-NN match with FP32)",
        "value": qps16,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["sec16"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16 first pass + fp32 re-rank",
        "data": "synthetic: 1,024-component Gaussian mixture, integer-valued float32",
        "config": {"workload": "knn_full type 2 (squared L2), nb=%d, nq=%d, d=%d, k=%d, k'=%d"
                               % (nb, nq, d, k, args.kprime),
                   "nb": nb, "nq": nq, "d": d, "k": k, "kprime": args.kprime},
        "roofline": {"bound": "mfma", "achieved": ach16, "peak": PEAK_BF16_TFLOPS,
                     "unit": "TFLOP/s", "frac": ach16 / PEAK_BF16_TFLOPS,
                     "traffic": traffic,
                     "kernel": "yk_scan_bf16<3,9,3,1> (BF16 filtered pass, MFMA 32x32x16)",
                     "flop_per_launch": flops, "kernel_ms_per_launch": bf_ms_launch,
                     "launches_per_step": bf_n / args.steps,
                     "sketch_kernel_ms_per_launch": sk_ms / max(sk_n, 1)},
        "accuracy": {"max_rel_err_vs_fp64": r["max_rel_err_f64"], "sorted": r["sorted_ok"],
                     "fallback_queries": r["nfb"]},
        # the plans of this search's shape class (queries and base sizes as
        # floor(log2)) in the library's tuning cache
        "tuned_plans": tuned_plans(":q%d:b%d" % (int(np.log2(args.nq)), int(np.log2(args.nb)))),
    }
    print("# bf16 filtered pass: %.3f ms/launch -> %.1f TFLOP/s; sketch %.3f ms/launch"
          % (bf_ms_launch, ach16, sk_ms / max(sk_n, 1)), file=sys.stderr, flush=True)
    if not args.only_bf16:
        qps32 = nq / r["sec32"]
        f32_ms, f32_n = r["sl32"][PROF_FUSED]
        f32_ms_launch = f32_ms / max(f32_n, 1)
        ach32 = flops / (f32_ms_launch * 1e-3) / 1e12
        f32_sk_ms, f32_sk_n = r["sl32"][PROF_F32_SKETCH]
        out["forms"] = {
            "fp32": {"queries_per_s": qps32, "ms_per_step": r["sec32"] * 1e3,
                     "roofline": {"bound": "mfma", "achieved": ach32, "peak": PEAK_FP32_TFLOPS,
                                  "unit": "TFLOP/s", "frac": ach32 / PEAK_FP32_TFLOPS,
                                  "kernel": "yk_scan_f32<1,8,3,1> (FP32 filtered pass)",
                                  "kernel_ms_per_launch": f32_ms_launch,
                                  "launches_per_step": f32_n / args.steps,
                                  "sketch_kernel_ms_per_launch":
                                      f32_sk_ms / max(f32_sk_n, 1)}},
            "bf16_rerank_over_fp32": qps16 / qps32,
            "index_agreement_with_fp32": float(np.mean(np.all(r["i16"] == r["i32"], axis=1))),
            "distances_bit_identical_to_fp32": bool(
                np.array_equal(r["d16"].view(np.int32), r["d32"].view(np.int32))),
        }
        print("# fp32 filtered pass: %.3f ms/launch -> %.1f TFLOP/s" % (f32_ms_launch, ach32),
              file=sys.stderr, flush=True)

    if not args.no_extras:
        v = mixture(3, 100_000, 128, 1024, 20, centres_seed=0)
        km = bench_kmeans(ya, dev, v, 1024, 10, env)
        out["extras"] = {"kmeans_iter_per_s": km["iters"] / km["sec"],
                         "kmeans_config": "n=100000 d=128 k=1024 niter=10 seed 0, FP32 assign, "
                                          "device-resident; iterations counted by the library, "
                                          "median of 3 runs",
                         "kmeans_iterations": km["iters"], "kmeans_obj": km["obj"]}
        if "cpu" in km:
            out["extras"]["kmeans_cpu_baseline"] = km["cpu"]
        print("# kmeans: %.2f iter/s (%d iterations)%s" % (
            km["iters"] / km["sec"], km["iters"],
            ", cpu reference %.2f iter/s" % km["cpu"]["value"] if "cpu" in km else ""),
            file=sys.stderr, flush=True)
        kpp = bench_kmeanspp(ya, dev, v, 1024, 1, cpu=not args.no_cpu)
        out["extras"]["kmeanspp"] = dict(
            config="k-means++ seeding (KMEANS_INIT_BERKELEY), n=100000 d=128 k=1024: draws "
                   "among the first 8k = 8192 points; one k-means step with it minus one with "
                   "random seeding, device-resident", **kpp)
        print("# kmeans++ seeding: %.1f ms (%s)" % (kpp["init_s"] * 1e3, "cpu reference %.1f ms"
              % (kpp["cpu_reference_init_s"] * 1e3) if "cpu_reference_init_s" in kpp else "no cpu"),
              file=sys.stderr, flush=True)
        di = bench_dropin(ya, base, queries, k)
        out["extras"]["drop_in_knn_full"] = di
        print("# drop-in knn_full: %.0f queries/s (FP32 form %.0f), same bits %s" % (
            di["queries_per_s"], di["fp32_form_queries_per_s"], di["same_bits_as_fp32_form"]),
            file=sys.stderr, flush=True)
        hk = bench_hkm(ya, base, cpu=not args.no_cpu)
        out["extras"]["hkm"] = hk
        print("# hkm: learn %.3f s, quantize %.3g points/s%s" % (
            hk["learn_s"], hk["quantize_points_per_s"],
            (" (cpu reference: learn %.2f s, quantize %.3g points/s)"
             % (hk["cpu_reference_learn_s"], hk["cpu_reference_quantize_points_per_s"]))
            if "cpu_reference_learn_s" in hk else ""), file=sys.stderr, flush=True)
        bt = bench_batched(ya, dev, env=env)
        out["extras"]["batched"] = {
            "config": "256 k-means instances (hkm level 2), Zipf(1.1) sizes sum 2e6 in [256, "
                      "120000], d=128, k=64, niter=15, FP32, one batched run, graph-replayed "
                      "iteration, autotuned update chunk %d" % bt["chunk"],
            "grouped_iter_per_s": 15.0 / (bt["loop_ms"] * 1e-3),
            "whole_call_s": bt["call_s"],
            "sequential_kmeans_calls_s": bt["seq_s"],
            "speedup_vs_sequential_calls": bt["seq_s"] / bt["call_s"],
            "iterations_run": [bt["iters_min"], bt["iters_max"]],
        }
        if "cpu" in bt:
            out["extras"]["batched"]["cpu_baseline"] = bt["cpu"]
        print("# batched: %.1f grouped iter/s (loop %.2f ms, call %.3f s, sequential %.3f s)"
              % (15.0 / (bt["loop_ms"] * 1e-3), bt["loop_ms"], bt["call_s"], bt["seq_s"]),
              file=sys.stderr, flush=True)

    if not args.no_large:
        out.setdefault("extras", {})
        out["extras"]["precision_d960"] = bench_precision_d960(ya, dev, env=env)
        kl = bench_kmeans_large(ya, dev, env=env)
        out["extras"]["kmeans_large"] = {
            "config": "n=1e7 d=128 k=65536 niter=20 seed 1, 65,536-component mixture sigma 15 in "
                      "[0,255] generated in HBM; BF16 first pass + FP32 re-rank of the 8 best "
                      "centroids + certificate, FP64 centroid sums and objective",
            "iter_per_s": kl["iters"] / kl["sec"], "iterations": kl["iters"], "sec": kl["sec"],
            "objective": kl["obj"], "fp32_fallback_assignments": kl["fallback"],
            "assignments": kl["assigned"], "bf16_scan_ms": kl["bf16_ms"],
            "rerank_ms": kl["rerank_ms"], "update_ms": kl["update_ms"],
            "assign_step": kl["assign"],
            "pq_subspaces": dict(kl["pq"], config="8 sub-spaces d=16 k=256 niter=20 on the same "
                                                  "points, one batched run"),
        }

    if not args.no_cpu:
        cb = cpu_baseline_knn(base, queries, k, env)
        g_idx = r["i16"][:cb["nq"]]
        agree = float(np.mean(np.sort(g_idx, 1) == np.sort(cb["idx"], 1)))
        # integer-valued data: every distance is exact in both, so the
        # distance lists must be identical; indices differ only where equal
        # distances tie at the k-th place (the reference's heap breaks such
        # ties in its own order)
        dagree = float(np.mean(np.all(r["d16"][:cb["nq"]].view(np.int32)
                                      == cb["dis"].view(np.int32), axis=1)))
        out["cpu_baseline"] = {"value": cb["value"], "unit": cb["unit"], "cores": cb["cores"],
                               "kind": cb["kind"], "sample": cb["sample"],
                               "index_agreement_with_gpu": agree,
                               "distance_lists_identical_to_gpu": dagree, "host": env}
        print("# cpu baseline: %.1f queries/s on %d cores" % (cb["value"], cb["cores"]),
              file=sys.stderr, flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

"""GPU tests of the measured workloads' shapes, checked against the oracle
(the C restatement of yael's CPU path) or against FP64, not only against
the library's own FP32 form:

* the precision form reached through knn_full's automatic choice (past its
  BF16 cut-off) and forced onto the reference's golden k=100 case, against
  the oracle / golden: bit-exact on integer data;
* the d = 960 precision workload's distribution (|N(0,1)| rows scaled to
  unit norm, k = 10, k' = 64) against the FP64 oracle at a size the oracle
  finishes in seconds, and at the full 1e6 x 1e4 size through properties:
  identical to the FP32 form, every distance within the FP32 rounding bound
  of its FP64 value;
* the batched form on the 256-instance Zipf(1.1) workload at full size
  against single-instance runs, and on 12 instances against the oracle;
* k-means with k = 65,536 (the large workload's shape at n = 1e6, 2
  iterations): the precision-form run equals the FP32 run, the objective
  is within 1e-5 of FP64; the 8 product-quantiser sub-spaces against the
  oracle per sub-space.
"""
import ctypes

import numpy as np
import pytest

import parity
from oracle_bind import Oracle

pytestmark = pytest.mark.gpu

QUIET, BF16, FP32 = 0x10000, 0x1000000, 0x2000000
PROF_BF16 = 6
vp = ctypes.c_void_p


@pytest.fixture(scope="module")
def ya():
    import yael_amd
    yael_amd.init()
    return yael_amd


@pytest.fixture(scope="module")
def oracle():
    o = Oracle()
    o.set_blas(0, 0)
    return o


def bf16_launches(ya, fn):
    """Runs fn with the per-kernel timers on; returns (fn's result, the
    number of BF16 first-pass launches it made)."""
    L = ya.lib()
    L.yk_sync()
    L.yk_prof_reset()
    L.yk_prof_enable(1)
    try:
        out = fn()
    finally:
        L.yk_sync()
        L.yk_prof_enable(0)
    tot, cnt = ctypes.c_double(), ctypes.c_long()
    L.yk_prof_get(PROF_BF16, ctypes.byref(tot), ctypes.byref(cnt))
    return out, cnt.value


# ---------------------------------------------------------------------------
# precision form through knn_full, against the oracle and the golden


def test_knn_full_automatic_precision_form_vs_oracle(ya, oracle):
    """300 queries x 300,000 base vectors, d=128, k=100 (integer mixture):
    past the automatic BF16 cut-off (>= 256 queries, >= 2^26 pairs), so
    knn_full runs the precision form; every distance is exact in FP32, so
    the result is the oracle's bit for bit (exact ties in heap order
    aside)."""
    from yael_amd.workloads import mixture
    b = mixture(11, 300_000, 128, 1024, 20, centres_seed=0)
    q = mixture(12, 300, 128, 1024, 20, centres_seed=0)
    assert ya.lib().yael_knn_precision(0) == 0
    (gi, gd), nbf = bf16_launches(ya, lambda: ya.knn(q, b, 100))
    assert nbf >= 1, "the automatic choice did not take the precision form"
    print("bf16 first-pass launches", nbf, flush=True)
    oi, od = oracle.knn_full(2, b, q, 100, nt=16)
    parity.check_knn(2, q, b, 100, gi, gd, oi, od, exact=True, label="auto precision")


def test_golden_mix_k100_forced_precision_form(ya):
    """The reference's golden k=100 case with the precision form forced."""
    import os
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mix_k100.npz")
    with np.load(gold, allow_pickle=False) as z:
        g = {k: z[k] for k in z.files}
    old = ya.lib().yael_knn_precision(2)
    try:
        (gi, gd), nbf = bf16_launches(ya, lambda: ya.knn(g["q"], g["b"], int(g["k"])))
    finally:
        ya.lib().yael_knn_precision(old)
    assert nbf >= 1
    parity.check_knn(2, g["q"], g["b"], int(g["k"]), gi, gd, g["assign"], g["dis"], exact=True,
                     label="mix_k100 precision")


# ---------------------------------------------------------------------------
# the d = 960 precision workload


def _fp32_bound(q, b, idx):
    """The FP32 expanded-form rounding bound 2 * gamma_d * (|q|^2 + |b|^2)
    of each returned (query, neighbour) distance."""
    d = q.shape[1]
    qn = (q.astype(np.float64) ** 2).sum(1)
    bn = (b.astype(np.float64) ** 2).sum(1)
    return 2 * parity.gamma(d) * (qn[:, None] + bn[idx])


def test_d960_unit_rows_vs_fp64_oracle(ya, oracle):
    """60,000 x 300 at d=960, k=10, k'=64: indices equal the FP64 oracle's
    up to near-ties, every distance within the FP32 bound of its FP64
    value."""
    from yael_amd import device as dev
    from yael_amd.workloads import unit_abs_normal
    b = unit_abs_normal(51, 60_000, 960)
    q = unit_abs_normal(52, 300, 960)
    Q, B = dev.DeviceSet(q), dev.DeviceSet(b)
    gi, gd, nfb = dev.knn_bf16(Q, B, 10, 64)
    Q.free()
    B.free()
    print("fallback queries", nfb, flush=True)
    oi, od = oracle.knn_f64(2, b, q, 10)
    r = parity.check_knn(2, q, b, 10, gi, gd, oi, od, label="d960 vs fp64")
    assert r["near_ties"] <= 3, r
    d64 = parity.f64_dist(2, q, b, gi)
    assert np.all(np.abs(gd.astype(np.float64) - d64) <= _fp32_bound(q, b, gi))


def test_d960_full_size_properties(ya):
    """nb=1e6, nq=1e4, d=960, k=10, k'=64, generated in HBM: the precision
    form's indices and distance bits equal the FP32 form's, lists are
    sorted, and every distance is within the FP32 bound of its FP64 value
    (the FP64 distances from the GPU's FP64 pair kernel)."""
    from yael_amd import device as dev
    L = ya.lib()
    nb, nq, d, k = 1_000_000, 10_000, 960, 10
    B = dev.DeviceSet.synth(nb, d, 2, 51, unit_rows=True)
    Q = dev.DeviceSet.synth(nq, d, 2, 52, unit_rows=True)
    print("d960 full: fp32 form", flush=True)
    i32, d32 = dev.knn_fp32(Q, B, k)
    print("d960 full: precision form", flush=True)
    i16, d16, nfb = dev.knn_bf16(Q, B, k, 64)
    assert np.array_equal(i16, i32)
    assert np.array_equal(d16.view(np.int32), d32.view(np.int32))
    assert np.all(np.diff(d16, axis=1) >= 0)
    oi = dev.DeviceBuffer(nq * k * 4)
    o64 = dev.DeviceBuffer(nq * k * 8)
    L.yk_memcpy_h2d(oi.ptr, i16.ctypes.data_as(vp), nq * k * 4)
    ya.check(L.yk_pair_dist_f64(2, nq, d, k, Q.vec.ptr, B.vec.ptr, oi.ptr, o64.ptr))
    d64 = o64.to_host(np.float64, (nq, k))
    # unit rows: |q|^2 + |b|^2 = 2 up to rounding
    bound = 2 * parity.gamma(d) * 2.0 * (1 + 1e-6)
    err = np.abs(d16.astype(np.float64) - d64)
    print("d960 full: fallback %d, max |fp32 - fp64| %.3g (bound %.3g), max rel %.3g"
          % (nfb, err.max(), bound, (err / d64).max()), flush=True)
    assert err.max() <= bound
    for x in (oi, o64, Q, B):
        x.free()


# ---------------------------------------------------------------------------
# the batched form


def _batched_device(ya, V, sizes, k, niter, seeds):
    L = ya.lib()
    ninst, n, d = len(sizes), int(sizes.sum()), V.d
    cent = np.empty((ninst, k, d), np.float32)
    dis = np.empty(n, np.float32)
    asg = np.empty(n, np.int32)
    nas = np.empty((ninst, k), np.int32)
    ret = np.empty(ninst, np.float32)
    iters = np.empty(ninst, np.int32)
    chunk = ctypes.c_int(0)
    ya.check(L.yael_kmeans_batched_device(d, ninst, sizes.ctypes.data, k, niter, V.vec.ptr,
                                          QUIET, seeds.ctypes.data, cent.ctypes.data,
                                          dis.ctypes.data, asg.ctypes.data, nas.ctypes.data,
                                          ret.ctypes.data, iters.ctypes.data, 1,
                                          ctypes.byref(chunk)), "batched")
    return cent, dis, asg, nas, ret, iters


def test_batched_zipf_256_full_size_vs_single_runs(ya):
    """256 instances of Zipf(1.1) sizes (sum 2e6, in [256, 120000]), d=128,
    k=64, niter=15, mixture data generated in HBM: every instance's
    centroids, distances, assignment, sizes and objective equal its
    single-instance kmeans() run with the same seed."""
    from yael_amd import device as dev
    from yael_amd.workloads import zipf_sizes
    L = ya.lib()
    sizes = zipf_sizes(256, 2_000_000, 256, 120_000)
    assert sizes.sum() == 2_000_000 and sizes.min() >= 256 and sizes.max() <= 120_000
    C0 = dev.DeviceSet.synth(1024, 128, 0, 3, 0.0, 255.0)
    V = dev.DeviceSet.synth(int(sizes.sum()), 128, 1, 4, 0.0, 255.0, centres=C0, sigma=20.0,
                            round_values=True)
    C0.free()
    seeds = np.arange(1, 257, dtype=np.int64)
    cent, dis, asg, nas, ret, iters = _batched_device(ya, V, sizes, 64, 15, seeds)
    print("batched: iterations %d..%d" % (iters.min(), iters.max()), flush=True)
    off = np.concatenate([[0], np.cumsum(sizes)])
    c1 = np.empty((64, 128), np.float32)
    n1 = np.empty(64, np.int32)
    r1 = ctypes.c_float()
    for i in range(256):
        n = int(sizes[i])
        d1 = np.empty(n, np.float32)
        a1 = np.empty(n, np.int32)
        ya.check(L.yael_kmeans_device(128, n, 64, 15,
                                      ctypes.c_void_p(V.vec.ptr.value + int(off[i]) * V.ld * 4),
                                      QUIET, int(seeds[i]), 1, c1.ctypes.data, d1.ctypes.data,
                                      a1.ctypes.data, n1.ctypes.data, ctypes.byref(r1)))
        s = slice(off[i], off[i + 1])
        assert np.array_equal(cent[i].view(np.int32), c1.view(np.int32)), i
        assert np.array_equal(asg[s], a1) and np.array_equal(nas[i], n1), i
        assert np.array_equal(dis[s].view(np.int32), d1.view(np.int32)), i
        assert np.float32(ret[i]) == np.float32(r1.value), i
    V.free()


def test_batched_instances_vs_oracle(ya, oracle):
    """12 instances of uneven size on integer mixture data, k=64, one Lloyd
    step from each instance's seeded random initialisation: the first step
    is exact arithmetic, so every instance equals the oracle's kmeans()
    with the same seed bit for bit."""
    from yael_amd.workloads import mixture, zipf_sizes
    sizes = zipf_sizes(12, 30_000, 300, 12_000)
    v = mixture(21, int(sizes.sum()), 128, 512, 20, centres_seed=0)
    seeds = np.arange(101, 113, dtype=np.int64)
    cent, ret, dis, asg, nas = ya.kmeans_batched(v, sizes, 64, niter=1, seeds=seeds)
    off = np.concatenate([[0], np.cumsum(sizes)])
    for i in range(12):
        s = slice(off[i], off[i + 1])
        r, c, d1, a1, n1 = oracle.kmeans(v[s], 64, 1, flags=QUIET, seed=int(seeds[i]))
        assert np.array_equal(cent[i].view(np.int32), c.view(np.int32)), i
        assert np.array_equal(asg[s], a1) and np.array_equal(nas[i], n1), i
        assert np.array_equal(dis[s].view(np.int32), d1.view(np.int32)), i
        assert np.float32(ret[i]) == np.float32(r), i


def test_batched_many_iterations_vs_oracle_statistics(ya, oracle):
    """256 Zipf instances at a tenth of the workload's size (sum 2e5), k=64,
    niter=15, against the oracle's kmeans() per instance.  Past the first
    step, FP32 rounding differences (MFMA vs the oracle's GEMM order) move
    near-tie points, and a moved point changes the later trajectory, so the
    runs are compared by their objectives: no systematic bias over the 256
    instances, and no instance far off."""
    from yael_amd.workloads import mixture, zipf_sizes
    sizes = zipf_sizes(256, 200_000, 256, 12_000)
    v = mixture(22, int(sizes.sum()), 128, 1024, 20, centres_seed=0)
    seeds = np.arange(1, 257, dtype=np.int64)
    cent, ret, dis, asg, nas = ya.kmeans_batched(v, sizes, 64, niter=15, seeds=seeds)
    off = np.concatenate([[0], np.cumsum(sizes)])
    rel = np.empty(256)
    same = 0
    for i in range(256):
        s = slice(off[i], off[i + 1])
        r, c, d1, a1, n1 = oracle.kmeans(v[s], 64, 15, flags=QUIET, seed=int(seeds[i]))
        rel[i] = float(ret[i]) / r - 1.0
        same += int(np.array_equal(asg[s], a1))
    print("objective gpu/oracle - 1: mean %.2e, median |.| %.2e, max |.| %.2e; identical "
          "assignments in %d of 256 instances" % (rel.mean(), np.median(np.abs(rel)),
                                                 np.abs(rel).max(), same), flush=True)
    assert abs(rel.mean()) <= 5e-4
    assert np.abs(rel).max() <= 2e-2
    assert same >= 230


# ---------------------------------------------------------------------------
# k-means at k = 65,536 and the product-quantiser sub-spaces


def test_kmeans_k65536_precision_run_equals_fp32_run(ya):
    """n=1e6, d=128, k=65,536, niter=2, seed 1, data from a 65,536-component
    mixture (sigma 15) generated in HBM: the run with the precision-form
    assign step (BF16 first pass, 8 candidates re-ranked in FP32,
    certificate) gives the FP32 run's centroids, distances, assignment,
    sizes and objective; on the final centroids, the assign step's FP32
    objective is within 1e-5 (relative) of the FP64 objective of the same
    assignment."""
    from yael_amd import device as dev
    L = ya.lib()
    n, d, k = 1_000_000, 128, 65536
    C0 = dev.DeviceSet.synth(65536, d, 0, 41, 0.0, 255.0)
    V = dev.DeviceSet.synth(n, d, 1, 42, 0.0, 255.0, centres=C0, sigma=15.0)
    C0.free()
    runs = {}
    for name, flag in (("bf16", BF16), ("fp32", FP32)):
        print("kmeans k=65536 run", name, flush=True)
        cent = np.empty((k, d), np.float32)
        dis = np.empty(n, np.float32)
        asg = np.empty(n, np.int32)
        nas = np.empty(k, np.int32)
        ret = ctypes.c_float()
        ya.check(L.yael_kmeans_device(d, n, k, 2, V.vec.ptr, QUIET | flag, 1, 1,
                                      cent.ctypes.data, dis.ctypes.data, asg.ctypes.data,
                                      nas.ctypes.data, ctypes.byref(ret)), name)
        runs[name] = (cent, dis, asg, nas, ret.value)
    a, b = runs["bf16"], runs["fp32"]
    assert np.array_equal(a[0].view(np.int32), b[0].view(np.int32))
    assert np.array_equal(a[1].view(np.int32), b[1].view(np.int32))
    assert np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3])
    assert a[4] == b[4]
    # the assign step alone on the final centroids, FP32 objective vs FP64
    C = dev.DeviceSet(a[0])
    qi, qd, nfb = dev.knn_bf16(V, C, 1, 8)
    oi = dev.DeviceBuffer(n * 4)
    o64 = dev.DeviceBuffer(n * 8)
    L.yk_memcpy_h2d(oi.ptr, qi.ctypes.data_as(vp), n * 4)
    ya.check(L.yk_pair_dist_f64(2, n, d, 1, V.vec.ptr, C.vec.ptr, oi.ptr, o64.ptr))
    obj64 = float(o64.to_host(np.float64, n).sum())
    obj32 = float(qd.astype(np.float64).sum())
    print("assign step: fallback %d, objective rel err vs fp64 %.2e"
          % (nfb, abs(obj32 - obj64) / obj64), flush=True)
    assert abs(obj32 - obj64) <= 1e-5 * obj64
    for x in (oi, o64, C, V):
        x.free()


def test_pq_subspaces_vs_oracle(ya, oracle):
    """8 product-quantiser sub-spaces (d=16 each, k=256) of 20,000 integer
    mixture points in one batched run, one Lloyd step from each sub-space's
    seeded random initialisation: each sub-space equals the oracle's
    kmeans() on its columns bit for bit."""
    from yael_amd import device as dev
    from yael_amd.workloads import mixture
    L = ya.lib()
    n, d, nsub, ksub = 20_000, 128, 8, 256
    v = mixture(31, n, d, 1024, 20, centres_seed=0)
    V = dev.DeviceSet(v)
    ds = d // nsub
    cq = np.empty((nsub, ksub, ds), np.float32)
    dis = np.empty((nsub, n), np.float32)
    asg = np.empty((nsub, n), np.int32)
    nas = np.empty((nsub, ksub), np.int32)
    rq = np.empty(nsub, np.float32)
    itq = np.empty(nsub, np.int32)
    chunk = ctypes.c_int(0)
    seeds = np.arange(1, nsub + 1, dtype=np.int64)
    ya.check(L.yael_kmeans_subspaces_device(d, n, nsub, ksub, 1, V.vec.ptr, QUIET,
                                            seeds.ctypes.data, cq.ctypes.data, dis.ctypes.data,
                                            asg.ctypes.data, nas.ctypes.data, rq.ctypes.data,
                                            itq.ctypes.data, ctypes.byref(chunk)), "subspaces")
    V.free()
    for j in range(nsub):
        sub = np.ascontiguousarray(v[:, j * ds:(j + 1) * ds])
        r, c, d1, a1, n1 = oracle.kmeans(sub, ksub, 1, flags=QUIET, seed=int(seeds[j]))
        assert np.array_equal(cq[j].view(np.int32), c.view(np.int32)), j
        assert np.array_equal(asg[j], a1) and np.array_equal(nas[j], n1), j
        assert np.array_equal(dis[j].view(np.int32), d1.view(np.int32)), j
        assert np.float32(rq[j]) == np.float32(r), j

"""GPU parity of kmeans() (device Lloyd loop + host split/seed logic).

* Integer-valued data: the first iteration is exact arithmetic, so one
  user-initialised step is bit-exact against the oracle (assign, dis, counts,
  centroids: FP64 sums of integers equal the reference's FP32 sums while a
  cluster stays below 2^24 / 255 points).
* Float data: every GPU iteration is checked against the oracle's step from
  the GPU's own previous centroids (step-wise parity, near-tie rule).
* Full seeded runs against the reference's golden runs: same seed chain,
  same splits, objective within tolerance.
"""
import os

import numpy as np
import pytest

import parity
from oracle_bind import Oracle

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
QUIET, INIT_USER, NORMALIZE = 0x10000, 0x100000, 0x40000


def load(name):
    with np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


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


def mixture(rng, n, d, ncomp, sigma, integer=True):
    cen = rng.uniform(0, 255, (ncomp, d))
    x = np.clip(cen[rng.integers(0, ncomp, n)] + sigma * rng.standard_normal((n, d)), 0, 255)
    if integer:
        x = np.rint(x)
    return x.astype(np.float32)


def gpu_kmeans(ya, v, k, niter, seed=1, redo=1, init=None, normalize=False):
    c, q, dis, asg, nas = ya.kmeans(v, k, niter=niter, seed=seed, redo=redo, verbose=False,
                                    init_centroids=init, normalize=normalize, output="all")
    return q, c, dis, asg, nas


def test_first_step_bitexact_integer_data(ya, oracle):
    rng = np.random.default_rng(1)
    v = mixture(rng, 100000, 128, 1024, 20)
    init = v[rng.choice(len(v), 1024, replace=False)].copy()
    print("step n=1e5 k=1024", flush=True)
    g = gpu_kmeans(ya, v, 1024, 1, init=init)
    o = oracle.kmeans(v, 1024, 1, flags=QUIET | INIT_USER, seed=1, init=init)
    assert np.array_equal(g[3], o[3]), "assign"
    assert np.array_equal(g[4], o[4]), "nassign"
    assert np.array_equal(g[2].view(np.int32), o[2].view(np.int32)), "dis"
    assert np.array_equal(g[1].view(np.int32), o[1].view(np.int32)), "centroids"
    assert g[0] == o[0]


@pytest.mark.parametrize("n,d,k", [(5000, 16, 2048), (5000, 16, 2049), (140000, 8, 900),
                                   (140000, 8, 1000)])
def test_update_permutation_paths_bitexact(ya, oracle, n, d, k):
    """The update orders each cluster's points by a counting sort for k <=
    2048 and k * ceil(n / 1024) <= 2^17, else by a radix sort: one
    user-initialised step on integer data on each side of both limits is
    bit-exact against the oracle (counts, centroids, empty-cluster splits)."""
    rng = np.random.default_rng(n + k)
    v = mixture(rng, n, d, 64, 20)
    init = v[rng.choice(n, k, replace=False)].copy()
    print("step n=%d d=%d k=%d" % (n, d, k), flush=True)
    g = gpu_kmeans(ya, v, k, 1, init=init)
    o = oracle.kmeans(v, k, 1, flags=QUIET | INIT_USER, seed=1, init=init)
    assert np.array_equal(g[3], o[3]), "assign"
    assert np.array_equal(g[4], o[4]), "nassign"
    assert np.array_equal(g[1].view(np.int32), o[1].view(np.int32)), "centroids"
    assert g[0] == o[0]


def test_stepwise_parity_float_data(ya, oracle):
    """Each GPU iteration vs the oracle's iteration from the same centroids."""
    rng = np.random.default_rng(2)
    v = mixture(rng, 20000, 32, 50, 30, integer=False)
    cent = v[:64].copy()
    for it in range(6):
        print("stepwise iteration", it, flush=True)
        g = gpu_kmeans(ya, v, 64, 1, init=cent, seed=3 + it)
        o = oracle.kmeans(v, 64, 1, flags=QUIET | INIT_USER, seed=3 + it, init=cent)
        # assignment: equal up to near-ties (against the same centroids `cent`);
        # a split leaves two centroids 2e-7 |c| apart, whose points are all
        # near-ties, so only their share is bounded
        r = parity.check_knn(2, v, cent, 1, g[3], g[2], o[3], o[2], label="iter%d" % it)
        assert r["near_ties"] <= 0.02 * len(v)
        if r["near_ties"] == 0:
            assert np.array_equal(g[4], o[4])
            # FP64 vs FP32 sums: centroids within a few ulps
            np.testing.assert_allclose(g[1], o[1], rtol=2e-6, atol=1e-4)
        cent = g[1]


def test_stepwise_parity_config0_size(ya, oracle):
    """The n=1e5, d=128, k=1024 workload (integer mixture, 1,024 components,
    sigma 20): four GPU iterations, each against the oracle's iteration from
    the same centroids (assignment equal up to near-ties, counts equal when
    there are none, centroids within FP32-vs-FP64 rounding), then a whole
    seeded 10-iteration run against the oracle's run from the same seed
    (seed 1: seed 0 would draw each side's seed from its own lrand48
    state).  Whole runs may still diverge: a split's twins are near-ties,
    one assignment that flips changes a count, and a later split then draws
    another cluster.  The bound: the objectives agree within 1%."""
    rng = np.random.default_rng(5)
    v = mixture(rng, 100000, 128, 1024, 20)
    cent = v[rng.choice(len(v), 1024, replace=False)].copy()
    for it in range(4):
        print("config-0 stepwise iteration", it, flush=True)
        g = gpu_kmeans(ya, v, 1024, 1, init=cent, seed=7 + it)
        o = oracle.kmeans(v, 1024, 1, flags=QUIET | INIT_USER, seed=7 + it, init=cent)
        r = parity.check_knn(2, v, cent, 1, g[3], g[2], o[3], o[2], label="cfg0 iter%d" % it)
        assert r["near_ties"] <= 0.02 * len(v)
        if r["near_ties"] == 0:
            assert np.array_equal(g[4], o[4])
            np.testing.assert_allclose(g[1], o[1], rtol=2e-6, atol=1e-4)
        cent = g[1]
    print("config-0 whole run", flush=True)
    g = gpu_kmeans(ya, v, 1024, 10, seed=1)
    o = oracle.kmeans(v, 1024, 10, flags=QUIET, seed=1)
    print("objective: GPU %.1f, oracle %.1f" % (g[0], o[0]), flush=True)
    assert abs(g[0] - o[0]) <= 0.01 * abs(o[0]), (g[0], o[0])


@pytest.mark.parametrize("case", ["km_plain", "km_redo", "km_norm"])
def test_full_runs_vs_reference_golden(ya, case):
    gl = load(case)
    k, niter, seed, redo = int(gl["k"]), int(gl["niter"]), int(gl["seed"]), int(gl["redo"])
    print("golden", case, flush=True)
    g = gpu_kmeans(ya, gl["v"], k, niter, seed=seed, redo=redo,
                   normalize=bool(int(gl["flags"]) & NORMALIZE))
    # same seed chain -> same initial centroids; trajectories may separate at
    # near-ties, so the objective is compared with a tolerance
    assert abs(g[0] - float(gl["ret"])) <= 2e-3 * abs(float(gl["ret"])), (g[0], gl["ret"])
    agree = (g[3] == gl["assign"]).mean()
    assert agree > 0.98, agree


@pytest.mark.parametrize("d", [1000, 1100, 2053])
def test_normalized_step_long_vectors(ya, oracle, d):
    """KMEANS_NORMALIZE_CENTS on rows longer than the update kernel's
    1024-column LDS row (a separate normalisation pass): one user-initialised
    step on integer data in [0, 63] (every distance below 2^24, so exact in
    FP32) gives the oracle's bits, normalised centroids included."""
    rng = np.random.default_rng(d)
    cen = rng.uniform(0, 63, (30, d))
    v = np.clip(np.rint(cen[rng.integers(0, 30, 3000)] + 8 * rng.standard_normal((3000, d))),
                0, 63).astype(np.float32)
    init = v[rng.choice(len(v), 40, replace=False)].copy()
    print("normalized step d=%d" % d, flush=True)
    g = gpu_kmeans(ya, v, 40, 1, init=init, normalize=True)
    o = oracle.kmeans(v, 40, 1, flags=QUIET | INIT_USER | NORMALIZE, seed=1, init=init)
    assert np.array_equal(g[3], o[3]) and np.array_equal(g[4], o[4])
    assert np.array_equal(g[2].view(np.int32), o[2].view(np.int32))
    assert np.array_equal(g[1].view(np.int32), o[1].view(np.int32))
    np.testing.assert_allclose(np.linalg.norm(g[1].astype(np.float64), axis=1), 1.0, rtol=1e-6)


def test_split_step_bitexact_golden(ya):
    """Ten empty clusters split in one step on integer data: the assignment,
    the update and the host-side split (rand_r / gaussrand_r streams) give the
    reference's bits."""
    gl = load("km_split_step")
    g = gpu_kmeans(ya, gl["v"], int(gl["k"]), 1, seed=int(gl["seed"]), init=gl["init"])
    assert (gl["nassign"] == 0).sum() == 10
    assert np.array_equal(g[3], gl["assign"]) and np.array_equal(g[4], gl["nassign"])
    assert np.array_equal(g[2].view(np.int32), gl["dis"].view(np.int32))
    assert np.array_equal(g[1].view(np.int32), gl["cent"].view(np.int32))
    assert np.float32(g[0]) == gl["ret"]


@pytest.mark.parametrize("case", ["km_overflow", "km_step"])
def test_overflow_and_step_golden(ya, case):
    gl = load(case)
    print("golden", case, flush=True)
    g = gpu_kmeans(ya, gl["v"], int(gl["k"]), int(gl["niter"]), seed=int(gl["seed"]), init=gl["init"])
    assert g[4].sum() == len(gl["v"])
    assert abs(g[0] - float(gl["ret"])) <= 1e-5 * max(1.0, abs(float(gl["ret"]))), (case, g[0], gl["ret"])
    if case == "km_overflow":
        # the split draws its source cluster with the wrapped nassign^2 - 1
        assert gl["nassign"].max() > 46340
        assert np.array_equal(g[4], gl["nassign"])


def test_duplicate_points_split_every_iteration(ya):
    """Many duplicates: clusters empty out and split in every iteration.  The
    split pairs are 2e-7 |c| apart, so which twin a point joins is decided at
    rounding level and trajectories may part from the reference's; the run
    must still keep every invariant and reach a comparable objective."""
    gl = load("km_split")
    g = gpu_kmeans(ya, gl["v"], int(gl["k"]), int(gl["niter"]), seed=int(gl["seed"]))
    assert g[4].sum() == len(gl["v"])
    assert np.array_equal(np.bincount(g[3], minlength=int(gl["k"])), g[4])
    assert np.isclose(g[0] * len(gl["v"]), g[2].astype(np.float64).sum(), rtol=1e-6)
    assert g[0] <= 2.0 * float(gl["ret"]) + 1e-3


def test_abandon_returns_error(ya):
    gl = load("km_abandon")
    with pytest.raises(RuntimeError):
        ya.kmeans(gl["v"], int(gl["k"]), niter=int(gl["niter"]), seed=int(gl["seed"]), verbose=False)


def test_kmeans_invariants(ya):
    rng = np.random.default_rng(4)
    v = mixture(rng, 30000, 64, 40, 10, integer=False)
    c, q, dis, asg, nas = ya.kmeans(v, 40, niter=10, seed=2, verbose=False, output="all")
    assert nas.sum() == 30000
    assert np.array_equal(np.bincount(asg, minlength=40), nas)
    assert np.isclose(q * 30000, dis.astype(np.float64).sum(), rtol=1e-6)


def test_kmeans_k1_mean_and_k_equals_n(ya):
    rng = np.random.default_rng(5)
    v = rng.integers(0, 100, (5000, 12)).astype(np.float32)
    c = ya.kmeans(v, 1, niter=2, seed=1, verbose=False)
    assert np.array_equal(c[0], (v.astype(np.float64).sum(0) / 5000).astype(np.float32))
    w = rng.permutation(np.arange(64 * 3, dtype=np.float32)).reshape(64, 3)
    c, q, dis, asg, nas = ya.kmeans(w, 64, niter=4, seed=3, verbose=False, output="all")
    assert q == 0.0 and np.all(nas == 1)


def test_kmeans_deterministic(ya):
    rng = np.random.default_rng(6)
    v = mixture(rng, 50000, 48, 100, 25, integer=False)
    r1 = ya.kmeans(v, 100, niter=8, seed=5, verbose=False, output="all")
    r2 = ya.kmeans(v, 100, niter=8, seed=5, verbose=False, output="all")
    assert r1[1] == r2[1]
    assert np.array_equal(r1[0].view(np.int32), r2[0].view(np.int32))
    assert np.array_equal(r1[3], r2[3])


def _kmeans_flags(ya, v, k, niter, seed, flags):
    import ctypes
    n, d = v.shape
    c = np.empty((k, d), np.float32)
    dis = np.empty(n, np.float32)
    asg = np.empty(n, np.int32)
    nas = np.empty(k, np.int32)
    q = ctypes.c_float()
    rc = ya.lib().yael_kmeans_checked(d, n, k, niter, v.ctypes.data, flags, seed, 1,
                                     c.ctypes.data, dis.ctypes.data, asg.ctypes.data,
                                     nas.ctypes.data, ctypes.byref(q))
    ya.check(rc, "kmeans")
    st = [ctypes.c_long() for _ in range(3)]
    ya.lib().yael_kmeans_last_stats(*[ctypes.byref(x) for x in st])
    return (c, q.value, dis, asg, nas), [x.value for x in st]


@pytest.mark.parametrize("integer", [True, False])
def test_bf16_assign_same_run_as_fp32(ya, integer):
    """The precision form of the assign step (BF16 first pass, FP32 re-rank
    of 8 candidates, certificate, FP32 search of the uncertified points)
    gives the FP32 form's run bit for bit."""
    rng = np.random.default_rng(21)
    n, d, k = 40000, 64, 600
    cen = rng.uniform(0, 255, (300, d))
    v = np.clip(cen[rng.integers(0, 300, n)] + 15 * rng.standard_normal((n, d)), 0, 255)
    v = (np.rint(v) if integer else v).astype(np.float32)
    base = 0x10000 | 0x80000
    a, st32 = _kmeans_flags(ya, v, k, 8, 5, base | 0x2000000)
    b, st16 = _kmeans_flags(ya, v, k, 8, 5, base | 0x1000000)
    print("iters", st32[0], st16[0], "assigned", st16[1], "fallback", st16[2], flush=True)
    assert st32[2] == 0 and st16[0] == st32[0]
    assert np.array_equal(a[3], b[3]) and np.array_equal(a[4], b[4])
    assert np.array_equal(a[0].view(np.int32), b[0].view(np.int32))
    assert np.array_equal(a[2].view(np.int32), b[2].view(np.int32))
    assert a[1] == b[1]
    assert st16[2] < st16[1]


@pytest.mark.parametrize("case", ["km_kpp_step", "km_kpp_sub_step"])
def test_kmeanspp_step_bitexact_golden(ya, case):
    """k-means++ seeding (kmeans.c:27-82; the second case seeds from the
    first 8k points) then one step, on integer data: the reference's bits."""
    gl = load(case)
    print("golden", case, flush=True)
    c, q, dis, asg, nas = ya.kmeans(gl["v"], int(gl["k"]), niter=1, seed=int(gl["seed"]),
                                    verbose=False, init="kmeans++", output="all")
    assert np.array_equal(asg, gl["assign"]) and np.array_equal(nas, gl["nassign"])
    assert np.array_equal(dis.view(np.int32), gl["dis"].view(np.int32))
    assert np.array_equal(c.view(np.int32), gl["cent"].view(np.int32))
    assert np.float32(q) == gl["ret"]


def test_kmeanspp_float_runs_golden(ya):
    """Float data, two runs: the same seeds are drawn (the draws sit far from
    the rounding of the distances), the trajectories may part at near-ties."""
    gl = load("km_kpp")
    c, q, dis, asg, nas = ya.kmeans(gl["v"], int(gl["k"]), niter=int(gl["niter"]),
                                    seed=int(gl["seed"]), redo=int(gl["redo"]), verbose=False,
                                    init="kmeans++", output="all")
    assert abs(q - float(gl["ret"])) <= 2e-3 * abs(float(gl["ret"])), (q, gl["ret"])
    assert (asg == gl["assign"]).mean() > 0.98

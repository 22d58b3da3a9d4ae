"""Pins the oracle (oracle/yael_oracle.c, the C restatement of yael's CPU path).

1. Golden vectors written by the compiled reference (tests/golden/*.npz):
   with the same OpenBLAS plugged in, the restatement must reproduce them
   bit for bit -- kNN (k = 1, k > 1, dot, weights, ties, NaN padding), cross
   distances, full k-means runs (redo, normalisation, empty-cluster splits,
   the 32-bit wrap of nassign^2 - 1, abandon), the RNG streams, the shortlist
   re-rank.
2. With its own plain-loop GEMM (the configuration the GPU tests use), the
   restatement must agree exactly on integer data and up to near-ties
   elsewhere.
3. Closed forms, invariants and brute force on tiny inputs, chosen so that a
   dropped term, a wrong sign or a transposed operand fails.
"""
import glob
import math
import os

import numpy as np
import pytest

from oracle_bind import Oracle, Reference, have_ref, openblas_entry_points
import parity

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


# Both fixtures drive the one loaded liboracle.so, whose GEMM is a process-wide
# setting: each test sets the mode it needs on entry.
@pytest.fixture
def oracle_blas():
    o = Oracle()
    ep = openblas_entry_points()
    if ep is None:
        pytest.skip("OpenBLAS of the compiled reference not present")
    o.set_blas(*ep)
    yield o
    o.set_blas(0, 0)


@pytest.fixture
def oracle():
    o = Oracle()
    o.set_blas(0, 0)
    return o


KNN_CASES = ["l2_k1", "l2_k10", "dot_k5", "ties_k1", "ties_k12", "w_k1", "w_k7", "nan_k1", "nan_k39",
             "mix_k1", "mix_k100"]


def bits(x):
    return np.asarray(x, np.float32).view(np.int32)


@pytest.mark.parametrize("case", KNN_CASES)
def test_knn_golden_bitexact_with_reference_blas(oracle_blas, case):
    g = load(case)
    w = g.get("w")
    a, d = oracle_blas.knn_full(int(g["typ"]), g["b"], g["q"], int(g["k"]), w=w)
    assert np.array_equal(a, g["assign"])
    assert np.array_equal(bits(d), bits(g["dis"]))


@pytest.mark.parametrize("case", KNN_CASES)
def test_knn_golden_plain_gemm(oracle, case):
    g = load(case)
    w = g.get("w")
    typ, k = int(g["typ"]), int(g["k"])
    a, d = oracle.knn_full(typ, g["b"], g["q"], k, w=w)
    exact = case.startswith(("mix", "ties"))
    parity.check_knn(typ, g["q"], g["b"], k, a, d, g["assign"], g["dis"], w=w, exact=exact, label=case)


@pytest.mark.parametrize("case", ["cross_l2", "cross_dot"])
def test_cross_golden(oracle_blas, case):
    g = load(case)
    out = oracle_blas.cross_distances(g["a"], g["b"], int(g["typ"]))
    assert np.array_equal(bits(out), bits(g["dist"]))


@pytest.mark.parametrize("case", ["cross_l2", "cross_dot"])
def test_cross_golden_plain_gemm(oracle, case):
    g = load(case)
    a, b = g["a"], g["b"]
    out = oracle.cross_distances(a, b, int(g["typ"]))
    # the expanded-form rounding bound: 2 gamma_d (|a_i|^2 + |b_j|^2)
    na = (a.astype(np.float64) ** 2).sum(1)
    nb = (b.astype(np.float64) ** 2).sum(1)
    bound = 2 * parity.gamma(a.shape[1]) * (na[None, :] + nb[:, None])
    assert np.all(np.abs(out.astype(np.float64) - g["dist"]) <= bound)


KM_CASES = ["km_plain", "km_redo", "km_norm", "km_split", "km_step", "km_overflow", "km_split_step",
            "km_kpp", "km_kpp_step", "km_kpp_sub_step"]


@pytest.mark.parametrize("case", KM_CASES)
def test_kmeans_golden_bitexact_with_reference_blas(oracle_blas, case):
    g = load(case)
    init = g.get("init")
    r, c, dis, asg, nas = oracle_blas.kmeans(g["v"], int(g["k"]), int(g["niter"]), flags=int(g["flags"]),
                                             seed=int(g["seed"]), redo=int(g["redo"]), init=init)
    assert np.float32(r) == g["ret"]
    assert np.array_equal(asg, g["assign"])
    assert np.array_equal(nas, g["nassign"])
    assert np.array_equal(bits(c), bits(g["cent"]))
    assert np.array_equal(bits(dis), bits(g["dis"]))


@pytest.mark.parametrize("case", ["km_kpp_step", "km_kpp_sub_step"])
def test_kmeanspp_integer_data_plain_gemm(oracle, case):
    """On integer-valued data every k-means++ distance is exact, so the
    restatement's own GEMV loop gives the reference's seeds and first step."""
    g = load(case)
    r, c, dis, asg, nas = oracle.kmeans(g["v"], int(g["k"]), int(g["niter"]), flags=int(g["flags"]),
                                        seed=int(g["seed"]))
    assert np.float32(r) == g["ret"]
    assert np.array_equal(asg, g["assign"]) and np.array_equal(bits(c), bits(g["cent"]))


def test_kmeans_golden_abandon(oracle_blas):
    g = load("km_abandon")
    r, *_ = oracle_blas.kmeans(g["v"], int(g["k"]), int(g["niter"]), flags=int(g["flags"]),
                               seed=int(g["seed"]))
    assert r == -1 and g["ret"] == -1


def test_overflow_case_really_wraps():
    g = load("km_overflow")
    assert g["nassign"].max() > 46340  # nassign^2 overflows int32


def test_rng_streams(oracle):
    g = load("rng")
    assert np.array_equal(bits(oracle.randn_r(64, int(g["randn_seed"]))), bits(g["randn"]))
    sel = oracle.random_init(1000, 1000, int(g["perm_seed"]))
    assert np.array_equal(sel, g["perm"])


def test_shortlist_golden(oracle_blas):
    g = load("shortlist")
    a, d = oracle_blas.knn_reorder_shortlist(g["b"], g["q"], g["assign_in"], np.zeros_like(g["dis"]))
    assert np.array_equal(a, g["assign"])
    valid = g["assign"] >= 0
    assert np.array_equal(bits(d)[valid], bits(g["dis"])[valid])


def test_recompute_exact_dists_golden(oracle):
    """Three windows of the base in turn (nn.c:583-600): after each, the
    advanced kp and the distances equal the reference's bit for bit."""
    g = load("recompute")
    kp = np.zeros(len(g["q"]), np.int32)
    dis = np.full(g["idx"].shape, -1.0, np.float32)
    for w, (label0, nbw) in enumerate(g["windows"]):
        kp, dis = oracle.knn_recompute_exact_dists(g["b"][label0:label0 + nbw], g["q"], int(label0),
                                                   kp, g["idx"], dis)
        assert np.array_equal(kp, g["after"][w][:, 0].astype(np.int32)), w
        assert np.array_equal(bits(dis), bits(g["after"][w][:, 1:])), w


def test_recompute_exact_dists_closed_form(oracle):
    """The recomputed values are the FP64 difference-form distances rounded
    to float; entries outside the window keep their value; an entry below
    the window (the reference's assert) is reported."""
    rng = np.random.default_rng(4)
    b = rng.standard_normal((100, 7)).astype(np.float32)
    q = rng.standard_normal((5, 7)).astype(np.float32)
    idx = np.sort(np.stack([rng.choice(100, 6, replace=False) for _ in range(5)]), 1)
    idx = idx.astype(np.int32)
    dis = np.full((5, 6), 7.0, np.float32)
    # kp at 0: an entry below the window comes first (the reference asserts)
    assert (idx[:, 0] < 40).any()
    with pytest.raises(ValueError):
        oracle.knn_recompute_exact_dists(b[40:70], q, 40, np.zeros(5, np.int32), idx, dis)
    start = np.array([np.searchsorted(r, 40) for r in idx], np.int32)
    kp, out = oracle.knn_recompute_exact_dists(b[40:70], q, 40, start, idx, dis)
    for i in range(5):
        stop = np.searchsorted(idx[i], 70)
        assert kp[i] == stop
        for j in range(6):
            if start[i] <= j < stop:
                x = q[i].astype(np.float64) - b[idx[i, j]].astype(np.float64)
                exact = math.fsum((x * x).tolist())  # correctly rounded sum
                assert abs(float(out[i, j]) - exact) <= 6e-8 * exact
            else:
                assert out[i, j] == 7.0


def test_reference_build_matches_fixtures():
    """The fixtures are reproducible from the compiled reference (when built)."""
    if not have_ref():
        pytest.skip("oracle/_ref not built")
    ref = Reference()
    g = load("mix_k100")
    a, d = ref.knn_full(2, g["b"], g["q"], 100)
    assert np.array_equal(a, g["assign"]) and np.array_equal(bits(d), bits(g["dis"]))


# ---------------------------------------------------------------------------
# closed forms, invariants, brute force


def test_knn_matches_fp64_bruteforce_on_integer_data(oracle):
    rng = np.random.default_rng(1)
    b = rng.integers(0, 256, (500, 64)).astype(np.float32)
    q = rng.integers(0, 256, (40, 64)).astype(np.float32)
    for typ in (2, 16):
        a, d = oracle.knn_full(typ, b, q, 7)
        a64, d64 = oracle.knn_f64(typ, b, q, 7)
        # exact arithmetic: same values; order of exact ties by index in the f64 oracle
        assert np.array_equal(d.astype(np.float64), d64)
        assert np.array_equal(np.sort(a, 1), np.sort(a64, 1)) or np.array_equal(d.astype(np.float64), d64)


def test_cross_distance_closed_form(oracle):
    rng = np.random.default_rng(2)
    a = rng.integers(-20, 20, (13, 9)).astype(np.float32)
    b = rng.integers(-20, 20, (11, 9)).astype(np.float32)
    out = oracle.cross_distances(a, b, 2)
    ref = ((b[:, None, :].astype(np.float64) - a[None, :, :]) ** 2).sum(-1)  # [j, i]
    assert np.array_equal(out.astype(np.float64), ref)
    dot = oracle.cross_distances(a, b, 16)
    assert np.array_equal(dot.astype(np.float64), b.astype(np.float64) @ a.T.astype(np.float64))


def test_pair_dist_f64_matches_numpy(oracle):
    rng = np.random.default_rng(3)
    b = rng.standard_normal((100, 17)).astype(np.float32)
    q = rng.standard_normal((9, 17)).astype(np.float32)
    idx = rng.integers(-1, 100, (9, 4)).astype(np.int32)
    out = oracle.pair_dist_f64(2, b, q, idx)
    ref = parity.f64_dist(2, q, b, idx)
    np.testing.assert_allclose(out, ref, rtol=1e-14, equal_nan=True)


def test_kmeans_k_equals_n_is_exact(oracle):
    rng = np.random.default_rng(4)
    v = rng.permutation(np.arange(50 * 3, dtype=np.float32)).reshape(50, 3)
    r, c, dis, asg, nas = oracle.kmeans(v, 50, 5, flags=0x10000, seed=7)
    assert r == 0.0 and np.all(nas == 1) and np.all(dis == 0)
    assert np.array_equal(np.sort(c, 0), np.sort(v, 0))


def test_kmeans_k1_is_the_mean(oracle):
    rng = np.random.default_rng(5)
    v = rng.integers(0, 100, (1000, 6)).astype(np.float32)
    r, c, dis, asg, nas = oracle.kmeans(v, 1, 2, flags=0x10000, seed=3)
    assert np.array_equal(c[0], (v.astype(np.float64).sum(0) / 1000).astype(np.float32))
    assert nas[0] == 1000


def test_kmeans_invariants(oracle):
    rng = np.random.default_rng(6)
    v = rng.standard_normal((3000, 8)).astype(np.float32)
    r, c, dis, asg, nas, tr = oracle.kmeans_traced(v, 30, 12, flags=0x10000, seed=8)
    assert nas.sum() == 3000
    assert np.array_equal(np.bincount(asg, minlength=30), nas)
    # the objective is the sum of the returned distances (pre-update centroids)
    assert np.isclose(r * 3000, dis.astype(np.float64).sum(), rtol=1e-6)
    # Lloyd iterations never increase the objective (no splits here)
    assert np.all(np.diff(tr["qerr"]) <= 1e-6 * tr["qerr"][0])


def test_planted_mistake_is_caught():
    """A restatement with a sign error in the cross term must fail the golden check."""
    g = load("mix_k1")
    b, q = g["b"].astype(np.float64), g["q"].astype(np.float64)
    wrong = (q ** 2).sum(1)[:, None] + (b ** 2).sum(1)[None, :] + 2 * q @ b.T
    assert not np.array_equal(wrong.argmin(1).astype(np.int32), g["assign"][:, 0])
    right = (q ** 2).sum(1)[:, None] + (b ** 2).sum(1)[None, :] - 2 * q @ b.T
    assert np.array_equal(right.argmin(1).astype(np.int32), g["assign"][:, 0])


def test_golden_files_present():
    assert len(glob.glob(os.path.join(GOLD, "*.npz"))) >= 20


@pytest.mark.parametrize("case", ["hkm_step", "hkm_float"])
def test_hkm_learn_golden_bitexact_with_reference_blas(oracle_blas, case):
    """The restated tree driver (label sort, kmeans(seed 0) per node with
    lrand48 seeds, label update) gives the reference's tree and labels."""
    from oracle_bind import libc_srand48
    g = load(case)
    nlevel, bf = int(g["nlevel"]), int(g["bf"])
    libc_srand48(int(g["srand48"]))
    levels, labels = oracle_blas.hkm_learn(g["v"], nlevel, bf, int(g["niter"]))
    for l in range(nlevel):
        assert np.array_equal(bits(levels[l]), bits(g["c%d" % l])), l
    assert np.array_equal(labels, g["labels"])


def test_hkm_quantize_golden(oracle, oracle_blas):
    """Descent through a tree of integer centroids (exact distances, any
    GEMM) and through the learned tree (reference BLAS)."""
    g = load("hkm_quant")
    levels = [g["c0"], g["c1"], g["c2"]]
    assert np.array_equal(oracle.hkm_quantize(levels, int(g["bf"]), g["q"]), g["qidx"])
    s = load("hkm_step")
    idx = oracle_blas.hkm_quantize([s["c0"], s["c1"], s["c2"]], int(s["bf"]), s["q"])
    assert np.array_equal(idx, s["qidx"])


def test_hkm_invariants(oracle):
    """Labels are leaves (< bf^nlevel); the level-0 centroids are the k-means
    of the whole set; each leaf label's parent is the point's level-1 node."""
    g = load("hkm_step")
    from oracle_bind import libc_srand48
    libc_srand48(int(g["srand48"]))
    levels, labels = oracle.hkm_learn(g["v"], 3, 4, 1)
    assert labels.min() >= 0 and labels.max() < 4 ** 3
    # one step on integer data from the same seeds: the restatement's own GEMM
    # loop gives the reference's tree too (exact arithmetic)
    assert np.array_equal(labels, g["labels"])

"""GPU parity of the kNN / cross-distance path (through the C-ABI library).

Against the compiled reference's golden vectors and against the oracle on
seeded inputs: bit-exact on integer-valued data (S11: all arithmetic exact),
near-tie rule elsewhere (tests/parity.py).
"""
import ctypes
import os

import numpy as np
import pytest

import parity
from oracle_bind import Oracle

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with np.load(os.path.join(GOLD, name + ".npz"), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="module")
def ya():
    import yael_amd
    yael_amd.init()
    print("device:", yael_amd.device_info(), flush=True)
    return yael_amd


@pytest.fixture(scope="module")
def oracle():
    o = Oracle()
    o.set_blas(0, 0)
    return o


def mixture(rng, n, d, ncomp, sigma, integer=True):
    cen = rng.uniform(0, 255, (ncomp, d))
    x = cen[rng.integers(0, ncomp, n)] + sigma * rng.standard_normal((n, d))
    x = np.clip(x, 0, 255)
    if integer:
        x = np.rint(x)
    return x.astype(np.float32)


KNN_CASES = ["l2_k1", "l2_k10", "dot_k5", "ties_k1", "ties_k12", "w_k1", "w_k7", "nan_k1", "nan_k39",
             "mix_k1", "mix_k100"]


@pytest.mark.parametrize("case", KNN_CASES)
def test_knn_vs_reference_golden(ya, case):
    g = load(case)
    typ, k = int(g["typ"]), int(g["k"])
    w = g.get("w")
    print("case", case, flush=True)
    idx, dis = ya.knn(g["q"], g["b"], k, distance_type=typ, weights=w)
    exact = case.startswith(("mix", "ties"))
    parity.check_knn(typ, g["q"], g["b"], k, idx, dis, g["assign"], g["dis"], w=w, exact=exact, label=case)


@pytest.mark.parametrize("k", [1, 10, 100])
@pytest.mark.parametrize("typ", [2, 16])
def test_knn_integer_mixture_bitexact(ya, oracle, k, typ):
    rng = np.random.default_rng(100 + k + typ)
    b = mixture(rng, 20000, 128, 64, 20)
    q = mixture(rng, 700, 128, 64, 20)
    print("knn integer mixture", typ, k, flush=True)
    gi, gd = ya.knn(q, b, k, distance_type=typ)
    oi, od = oracle.knn_full(typ, b, q, k)
    parity.check_knn(typ, q, b, k, gi, gd, oi, od, exact=True, label="mix%d" % k)


@pytest.mark.parametrize("d", [8, 20, 32])
@pytest.mark.parametrize("k", [1, 40])
@pytest.mark.parametrize("typ", [2, 16])
def test_knn_short_vectors_bitexact(ya, oracle, d, k, typ):
    """d <= 32 (one or two 16-column chunks: base norms loaded a block ahead;
    k = 1 and the k > 16 sketch take the branch-free exact pass on full
    blocks); nb = 20011 leaves a partial last block."""
    rng = np.random.default_rng(300 + 3 * d + k + typ)
    b = mixture(rng, 20011, d, 64, 20)
    q = mixture(rng, 555, d, 64, 20)
    print("knn short vectors", typ, d, k, flush=True)
    gi, gd = ya.knn(q, b, k, distance_type=typ)
    oi, od = oracle.knn_full(typ, b, q, k)
    parity.check_knn(typ, q, b, k, gi, gd, oi, od, exact=True, label="short%d_%d_%d" % (typ, d, k))


@pytest.mark.parametrize("d", [16, 20, 128, 200, 960])
@pytest.mark.parametrize("k", [1, 8])
def test_knn_float_dims(ya, oracle, d, k):
    rng = np.random.default_rng(d * 7 + k)
    b = rng.standard_normal((3000, d)).astype(np.float32)
    q = rng.standard_normal((333, d)).astype(np.float32)
    print("knn float d", d, "k", k, flush=True)
    gi, gd = ya.knn(q, b, k)
    oi, od = oracle.knn_full(2, b, q, k)
    r = parity.check_knn(2, q, b, k, gi, gd, oi, od, label="d%d" % d)
    assert r["near_ties"] <= 2


def test_knn_split_merge_large_k(ya, oracle):
    """Few queries, many base rows: the base range is split across workgroups
    and the partial lists merged; results must not depend on the split."""
    rng = np.random.default_rng(7)
    b = mixture(rng, 100000, 64, 128, 15)
    q = mixture(rng, 37, 64, 128, 15)
    for k in (1, 64, 250):
        print("split/merge k", k, flush=True)
        gi, gd = ya.knn(q, b, k)
        oi, od = oracle.knn_full(2, b, q, k)
        parity.check_knn(2, q, b, k, gi, gd, oi, od, exact=True, label="split k%d" % k)


def test_fp32_starting_bound_underflow(ya, oracle):
    """The FP32 form (k > 16) starts its filtered pass from an estimated
    bound (an m-th smallest sketch minimum, m < k); a query left with fewer
    than k keys within it is scanned again without a bound, which overflows
    its list, then with the overflow's bound.  At this size the sketch
    samples one block in 8 and splits the base into 37 ranges of 1,792 rows;
    the only rows near the queries are the first rows of those ranges, one
    per sketch subset, so the m-th smallest minimum (m = 35 for k = 40) has
    35 rows at or below it.  Integer data: the lists must equal the oracle's
    bit for bit, after three filtered passes over one sketch."""
    import yael_amd
    from yael_amd import device as dev
    rng = np.random.default_rng(31)
    nb, d, k = 65536, 32, 40
    b = np.rint(40.0 + 3.0 * rng.standard_normal((nb, d))).astype(np.float32)
    near = np.arange(0, nb, 1792)
    b[near] = np.rint(3.0 * rng.standard_normal((len(near), d))).astype(np.float32)
    q = np.rint(3.0 * rng.standard_normal((100, d))).astype(np.float32)
    lib = yael_amd.lib()
    Q, B = dev.DeviceSet(q), dev.DeviceSet(b)
    lib.yk_prof_enable(1)
    lib.yk_prof_reset()
    try:
        gi, gd = dev.knn_fp32(Q, B, k)
        counts = []
        for slot in (0, 8):  # filtered-pass scans, sketch scans
            tot, cnt = ctypes.c_double(), ctypes.c_long()
            lib.yk_prof_get(slot, ctypes.byref(tot), ctypes.byref(cnt))
            counts.append(cnt.value)
    finally:
        lib.yk_prof_enable(0)
        Q.free()
        B.free()
    print("filtered-pass and sketch launches", counts, flush=True)
    oi, od = oracle.knn_full(2, b, q, k)
    parity.check_knn(2, q, b, k, gi, gd, oi, od, exact=True, label="fp32 underflow")
    assert counts == [3, 1]


def test_knn_edge_shapes(ya, oracle):
    rng = np.random.default_rng(8)
    for nb, nq, d, k in [(1, 1, 1, 1), (5, 3, 7, 5), (33, 65, 3, 33), (257, 1, 128, 256), (64, 129, 17, 1)]:
        b = rng.integers(0, 9, (nb, d)).astype(np.float32)
        q = rng.integers(0, 9, (nq, d)).astype(np.float32)
        print("edge", nb, nq, d, k, flush=True)
        gi, gd = ya.knn(q, b, k)
        oi, od = oracle.knn_full(2, b, q, k)
        parity.check_knn(2, q, b, k, gi, gd, oi, od, exact=True, label="edge%d_%d_%d_%d" % (nb, nq, d, k))


def test_knn_errors(ya):
    b = np.zeros((10, 4), np.float32)
    q = np.zeros((3, 4), np.float32)
    with pytest.raises(ya.YaelGPUError):
        ya.knn(q, b, 11)  # k > nb (nn.c:456)
    with pytest.raises(ya.YaelGPUError):
        ya.knn(q, b, 2, distance_type=3)  # chi2: no GPU path


def test_c_api_knn_full(ya, oracle):
    """The reference's own symbol, called the way a C caller would."""
    lib = ya.lib()
    rng = np.random.default_rng(9)
    b = rng.integers(0, 256, (5000, 32)).astype(np.float32)
    q = rng.integers(0, 256, (100, 32)).astype(np.float32)
    k = 5
    assign = np.empty((100, k), np.int32)
    dis = np.empty((100, k), np.float32)
    vp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
    lib.knn_full_thread(2, 100, 5000, 32, k, vp(b), vp(q), None, vp(assign), vp(dis), 8)
    oi, od = oracle.knn_full(2, b, q, k)
    parity.check_knn(2, q, b, k, assign, dis, oi, od, exact=True, label="c_api")
    # nn(): sum of the nearest distances, in double
    vw = np.empty(100, np.int32)
    s = lib.nn(100, 5000, 32, vp(b), vp(q), vp(vw))
    assert np.array_equal(vw, oi[:, 0]) and s == od[:, 0].astype(np.float64).sum()


@pytest.mark.parametrize("typ", [2, 12, 16])
def test_cross_distances(ya, oracle, typ):
    rng = np.random.default_rng(10 + typ)
    a = rng.integers(0, 256, (301, 40)).astype(np.float32)
    b = rng.integers(0, 256, (129, 40)).astype(np.float32)
    g = ya.cross_distances(a, b, typ)
    o = oracle.cross_distances(a, b, 2 if typ == 12 else typ)
    assert np.array_equal(g.view(np.int32), o.view(np.int32))


def test_cross_distances_golden(ya):
    for case in ("cross_l2", "cross_dot"):
        gl = load(case)
        a, b = gl["a"], gl["b"]
        out = ya.cross_distances(a, b, int(gl["typ"]))
        na = (a.astype(np.float64) ** 2).sum(1)
        nb = (b.astype(np.float64) ** 2).sum(1)
        bound = 2 * parity.gamma(a.shape[1]) * (na[None, :] + nb[:, None])
        assert np.all(np.abs(out.astype(np.float64) - gl["dist"]) <= bound), case


def test_cross_distances_nonpacked(ya, oracle):
    lib = ya.lib()
    rng = np.random.default_rng(11)
    d, na, nb, lda, ldb, ldd = 24, 50, 30, 31, 40, 64
    A = rng.integers(0, 50, (na, lda)).astype(np.float32)
    B = rng.integers(0, 50, (nb, ldb)).astype(np.float32)
    out = np.full((nb, ldd), -7.0, np.float32)
    vp = lambda x: x.ctypes.data_as(ctypes.c_void_p)
    lib.compute_cross_distances_nonpacked(d, na, nb, vp(A), lda, vp(B), ldb, vp(out), ldd)
    ref = oracle.cross_distances(np.ascontiguousarray(A[:, :d]), np.ascontiguousarray(B[:, :d]), 2)
    assert np.array_equal(out[:, :na], ref)
    assert np.all(out[:, na:] == -7.0)  # untouched beyond na


def test_pair_dist_f64_bitexact(ya, oracle):
    lib = ya.lib()
    rng = np.random.default_rng(12)
    b = rng.standard_normal((400, 33)).astype(np.float32)
    q = rng.standard_normal((50, 33)).astype(np.float32)
    idx = rng.integers(-1, 400, (50, 6)).astype(np.int32)
    ld = lib.yk_ld(33)
    out = np.empty((50, 6), np.float64)
    ptrs = {}
    vp = ctypes.c_void_p
    for name, arr in (("b", b), ("q", q)):
        p = vp()
        ya.check(lib.yk_malloc(ctypes.byref(p), arr.shape[0] * ld * 4))
        ya.check(lib.yk_upload_vectors(p, arr.ctypes.data_as(vp), arr.shape[0], 33))
        ptrs[name] = p
    for name, nbytes in (("i", idx.nbytes), ("o", out.nbytes)):
        p = vp()
        ya.check(lib.yk_malloc(ctypes.byref(p), nbytes))
        ptrs[name] = p
    ya.check(lib.yk_memcpy_h2d(ptrs["i"], idx.ctypes.data_as(vp), idx.nbytes))
    ya.check(lib.yk_pair_dist_f64(2, 50, 33, 6, ptrs["q"], ptrs["b"], ptrs["i"], ptrs["o"]))
    ya.check(lib.yk_memcpy_d2h(out.ctypes.data_as(vp), ptrs["o"], out.nbytes))
    for p in ptrs.values():
        ya.check(lib.yk_free(p))
    ref = oracle.pair_dist_f64(2, b, q, idx)
    assert np.array_equal(np.isnan(out), np.isnan(ref))
    assert np.array_equal(out[~np.isnan(out)], ref[~np.isnan(ref)])


def _both_forms(ya, q, b, k, typ=2):
    """knn_full with the precision form forced, then with the FP32 form only."""
    lib = ya.lib()
    old = lib.yael_knn_precision(2)
    try:
        i16, d16 = ya.knn(q, b, k, distance_type=typ)
        lib.yael_knn_precision(1)
        i32, d32 = ya.knn(q, b, k, distance_type=typ)
    finally:
        lib.yael_knn_precision(old)
    return i16, d16, i32, d32


@pytest.mark.parametrize("k,typ", [(1, 2), (10, 2), (100, 2), (7, 16)])
def test_knn_full_precision_form_same_bits(ya, k, typ):
    """knn_full through the precision form gives the FP32 form's bits."""
    rng = np.random.default_rng(100 + k)
    cen = rng.uniform(0, 255, (64, 48))
    b = np.clip(np.rint(cen[rng.integers(0, 64, 30000)] + 20 * rng.standard_normal((30000, 48))),
                0, 255).astype(np.float32)
    q = rng.standard_normal((300, 48)).astype(np.float32) * 30 + 120
    print("knn_full both forms k=%d type=%d" % (k, typ), flush=True)
    i16, d16, i32, d32 = _both_forms(ya, q, b, k, typ)
    assert np.array_equal(i16, i32)
    assert np.array_equal(d16.view(np.int32), d32.view(np.int32))


def test_knn_full_precision_form_nan_data(ya):
    """NaN values: the set statistics show them and knn_full keeps the FP32
    form (same bits, NaN rows never returned)."""
    rng = np.random.default_rng(7)
    b = rng.standard_normal((5000, 24)).astype(np.float32)
    b[17, 3] = np.nan
    q = rng.standard_normal((300, 24)).astype(np.float32)
    q[5, 0] = np.nan
    i16, d16, i32, d32 = _both_forms(ya, q, b, 12)
    assert np.array_equal(i16, i32)
    assert np.array_equal(d16.view(np.int32), d32.view(np.int32))
    assert not np.any(i16 == 17)


def test_knn_large_base_offsets(ya):
    """A 3 GB base (6e6 x 128 floats, a BF16 copy of 1.7 GB): rows far past
    2^31 bytes are found.  Each query is a copy of a base row (the last one
    included) on integer data with distinct rows, so its nearest neighbour is
    that row at distance exactly 0, in both forms; the two forms agree bit
    for bit on the whole lists."""
    rng = np.random.default_rng(606)
    nb, d, k = 6_000_000, 128, 10
    b = rng.integers(0, 256, (nb, d), dtype=np.uint8).astype(np.float32)
    rows = np.concatenate([np.linspace(0, nb - 1, 254).astype(np.int64), [nb - 2, nb - 1]])
    q = b[rows].copy()
    print("knn large base nb=%d, %d queries" % (nb, len(rows)), flush=True)
    i16, d16, i32, d32 = _both_forms(ya, q, b, k)
    del b
    assert np.array_equal(i32[:, 0], rows.astype(np.int32))
    assert np.all(d32[:, 0] == 0.0)
    assert np.array_equal(i16, i32)
    assert np.array_equal(d16.view(np.int32), d32.view(np.int32))


def test_empty_inputs(ya):
    """No queries, or an empty side of a cross-distance matrix: empty results
    (the reference's loops run zero times), in both kNN forms."""
    rng = np.random.default_rng(5)
    b = rng.integers(0, 9, (100, 16)).astype(np.float32)
    q0 = np.zeros((0, 16), np.float32)
    lib = ya.lib()
    for mode in (1, 2):
        old = lib.yael_knn_precision(mode)
        try:
            idx, dis = ya.knn(q0, b, 5)
        finally:
            lib.yael_knn_precision(old)
        assert idx.shape == (0, 5) and dis.shape == (0, 5)
    d0 = ya.cross_distances(q0, b)
    assert d0.shape == (100, 0)
    d1 = ya.cross_distances(b, q0)
    assert d1.shape == (0, 100)
    # and a call after them still works
    idx, dis = ya.knn(b[:3], b, 1)
    assert np.array_equal(idx[:, 0], np.arange(3)) and np.all(dis[:, 0] == 0)


@pytest.mark.parametrize("k", [300, 500])
def test_knn_full_precision_form_large_k(ya, k):
    """Large k through the precision form (k' = 2k: candidate lists and the
    sorts that trim them run past 1024 keys): the FP32 form's bits."""
    rng = np.random.default_rng(700 + k)
    cen = rng.uniform(0, 255, (128, 64))
    b = np.clip(np.rint(cen[rng.integers(0, 128, 60000)] + 25 * rng.standard_normal((60000, 64))),
                0, 255).astype(np.float32)
    q = np.clip(np.rint(cen[rng.integers(0, 128, 300)] + 25 * rng.standard_normal((300, 64))),
                0, 255).astype(np.float32)
    print("knn_full both forms k=%d" % k, flush=True)
    i16, d16, i32, d32 = _both_forms(ya, q, b, k)
    assert np.array_equal(d16.view(np.int32), d32.view(np.int32))
    # indices may differ only inside runs of equal distances (integer data)
    bad = np.nonzero((i16 != i32).any(1))[0]
    for i in bad:
        for dv in np.unique(d32[i][i16[i] != i32[i]]):
            assert sorted(i16[i][d16[i] == dv]) == sorted(i32[i][d32[i] == dv])


def test_reorder_shortlist_golden(ya):
    """knn_reorder_shortlist (nn.c:528-580) on the reference's golden case:
    float data, a shortlist cut short by -1 entries.  The reference forms
    each dot product with its BLAS's sgemv, whose summation order is that
    library's own; the GPU sums in component order.  So the distances agree
    within the FP32 rounding bound of the expanded form, and the order
    equals the reference's except between values within that bound."""
    g = load("shortlist")
    a, d = ya.knn_reorder_shortlist(g["b"], g["q"], g["assign_in"], np.zeros_like(g["dis"]))
    valid = g["assign"] >= 0
    assert np.array_equal(a < 0, ~valid)
    tol = parity.tolerance(2, g["q"], g["b"])
    err = np.abs(d.astype(np.float64) - g["dis"].astype(np.float64))
    assert np.all(err[valid] <= np.broadcast_to(tol[:, None], err.shape)[valid])
    for i in np.nonzero((a != g["assign"]).any(1))[0]:
        # a different order only among values within the bound of each other
        j = np.nonzero(a[i] != g["assign"][i])[0]
        assert np.ptp(d[i, j].astype(np.float64)) <= 2 * tol[i], i
    print("shortlist: max |gpu - reference| %.3g, order differs in %d of %d lists"
          % (err[valid].max(), int((a != g["assign"]).any(1).sum()), len(a)), flush=True)


def _recompute(ya, b, q, label0, kp, idx, dis, checked=True):
    L = ya.lib()
    b = np.ascontiguousarray(b, np.float32)
    q = np.ascontiguousarray(q, np.float32)
    idx = np.ascontiguousarray(idx, np.int32)
    kp = np.ascontiguousarray(kp, np.int32).copy()
    dis = np.ascontiguousarray(dis, np.float32).copy()
    nq, k = idx.shape
    rc = L.yael_knn_recompute_exact_dists_checked(nq, b.shape[0], b.shape[1], k,
                                                   b.ctypes.data, q.ctypes.data, label0,
                                                   kp.ctypes.data, idx.ctypes.data,
                                                   dis.ctypes.data)
    return rc, kp, dis


def test_recompute_exact_dists_golden(ya):
    """knn_recompute_exact_dists (nn.c:583-600) over three windows of the
    base in turn, against the reference's golden values: the advanced kp and
    every distance, bit for bit."""
    g = load("recompute")
    kp = np.zeros(len(g["q"]), np.int32)
    dis = np.full(g["idx"].shape, -1.0, np.float32)
    for w, (label0, nbw) in enumerate(g["windows"]):
        rc, kp, dis = _recompute(ya, g["b"][label0:label0 + nbw], g["q"], int(label0), kp,
                                 g["idx"], dis)
        assert rc == 0, ya.lib().yk_last_error()
        assert np.array_equal(kp, g["after"][w][:, 0].astype(np.int32)), w
        assert np.array_equal(dis.view(np.int32), g["after"][w][:, 1:].view(np.int32)), w


def test_recompute_exact_dists_vs_oracle(ya, oracle):
    """Long shortlists (k=300, beyond one wave's 64 lanes), d=200, random
    windows and start positions: the oracle's kp and distance bits; an
    entry below the window is an error, as the reference asserts."""
    rng = np.random.default_rng(17)
    nb_all, d, nq, k = 5000, 200, 70, 300
    b = rng.standard_normal((nb_all, d)).astype(np.float32)
    q = rng.standard_normal((nq, d)).astype(np.float32)
    idx = np.sort(np.stack([rng.choice(nb_all, k, replace=False) for _ in range(nq)]), 1)
    idx = idx.astype(np.int32)
    dis = rng.standard_normal((nq, k)).astype(np.float32)
    label0, nbw = 1200, 2100
    kp = np.array([np.searchsorted(r, label0) for r in idx], np.int32)
    kp[::7] = np.minimum(kp[::7] + 5, k)  # some queries start further in
    rc, gkp, gdis = _recompute(ya, b[label0:label0 + nbw], q, label0, kp, idx, dis)
    assert rc == 0
    okp, odis = oracle.knn_recompute_exact_dists(b[label0:label0 + nbw], q, label0, kp, idx, dis)
    assert np.array_equal(gkp, okp)
    assert np.array_equal(gdis.view(np.int32), odis.view(np.int32))
    bad = kp.copy()
    bad[3] = 0  # query 3 starts below the window
    assert idx[3, 0] < label0
    rc, _, _ = _recompute(ya, b[label0:label0 + nbw], q, label0, bad, idx, dis)
    assert rc != 0 and b"below" in ya.lib().yk_last_error()


@pytest.mark.parametrize("typ,k", [(2, 1025), (2, 3000), (16, 1500)])
def test_knn_large_k_vs_oracle(ya, oracle, typ, k):
    """k beyond the on-chip selection (k > 1024; the reference allows any
    k <= nb): distance rows, radix select, segmented sort.  Integer data:
    every distance exact, so the oracle's lists bit for bit (exact ties in
    its heap order aside)."""
    rng = np.random.default_rng(k + typ)
    b = mixture(rng, 30011, 64, 128, 20)
    q = mixture(rng, 257, 64, 128, 20)
    print("knn large k", typ, k, flush=True)
    gi, gd = ya.knn(q, b, k, distance_type=typ)
    oi, od = oracle.knn_full(typ, b, q, k, nt=16)
    parity.check_knn(typ, q, b, k, gi, gd, oi, od, exact=True, label="large k%d" % k)


def test_knn_large_k_weights_nan_and_k_equals_nb(ya, oracle):
    """Weighted search, NaN base rows (never selected) with fewer finite
    values than k (the tail padded with -1 and NaN bits 0xffffffff,
    nn.c:515-518), and k = nb."""
    rng = np.random.default_rng(9)
    b = mixture(rng, 1500, 16, 30, 20)
    b[::7, 3] = np.nan
    q = mixture(rng, 40, 16, 30, 20)
    w = rng.integers(1, 4, len(b)).astype(np.float32)
    for k in (1400, 1500):
        print("knn large k with NaN rows, k", k, flush=True)
        gi, gd = ya.knn(q, b, k, weights=w)
        oi, od = oracle.knn_full(2, b, q, k, w=w)
        nfin = int(np.isfinite(b).all(1).sum())
        assert np.all(gi[:, nfin:] == -1)
        assert np.all(gd[:, nfin:].view(np.uint32) == 0xffffffff)
        assert np.array_equal(np.sort(gd[:, :nfin], 1).view(np.int32),
                              np.sort(od[:, :nfin], 1).view(np.int32))
        parity.check_knn(2, q, b, k, gi, gd, oi, od, w=w, exact=True, label="nan k%d" % k)

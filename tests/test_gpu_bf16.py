"""GPU tests of the precision form (BF16 MFMA first pass + FP32 re-rank +
certificate + FP32 fallback).  The bar: the same indices and the same
distance bits as the FP32 form on every query (the re-rank recomputes
distances with the FP32 form's instruction sequence, and the certificate
guarantees the FP32 form's k nearest are among the candidates)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    import yael_amd
    from yael_amd import device
    yael_amd.init()
    return device


def mixture(rng, n, d, ncomp, sigma, integer=True):
    cen = rng.uniform(0, 255, (ncomp, d))
    x = np.clip(cen[rng.integers(0, ncomp, n)] + sigma * rng.standard_normal((n, d)), 0, 255)
    return (np.rint(x) if integer else x).astype(np.float32)


def normal_unit(rng, n, d):
    x = np.abs(rng.standard_normal((n, d))).astype(np.float32)
    return (x / np.linalg.norm(x, axis=1, keepdims=True)).astype(np.float32)


def same_as_fp32(dev, q, b, k, kp, typ=2):
    Q, B = dev.DeviceSet(q), dev.DeviceSet(b)
    i32, d32 = dev.knn_fp32(Q, B, k, typ)
    i16, d16, nfb = dev.knn_bf16(Q, B, k, kp, typ)
    Q.free()
    B.free()
    assert np.array_equal(i16, i32), "indices differ from the FP32 form"
    assert np.array_equal(d16.view(np.int32), d32.view(np.int32)), "distances differ"
    return nfb


@pytest.mark.parametrize("k,kp", [(1, 16), (10, 64), (100, 200)])
def test_integer_mixture(dev, k, kp):
    rng = np.random.default_rng(k)
    b = mixture(rng, 60000, 128, 256, 20)
    q = mixture(rng, 900, 128, 256, 20)
    nfb = same_as_fp32(dev, q, b, k, kp)
    print("k", k, "fallback queries", nfb, flush=True)
    assert nfb <= 0.05 * len(q)


def test_unit_vectors_d960(dev):
    rng = np.random.default_rng(5)
    b = normal_unit(rng, 40000, 960)
    q = normal_unit(rng, 300, 960)
    nfb = same_as_fp32(dev, q, b, 10, 64)
    print("fallback queries", nfb, flush=True)


@pytest.mark.parametrize("d,nq,nb", [(300, 77, 5003), (520, 300, 20011)])
def test_long_vectors_ragged(dev, d, nq, nb):
    """More than 16 chunks (the long-vector scan with fragment-major queries):
    query counts off the 256-query workgroup span, a partial last base block,
    integer data (ties) and both distance types."""
    rng = np.random.default_rng(d)
    b = mixture(rng, nb, d, 40, 20)
    q = mixture(rng, nq, d, 40, 20)
    same_as_fp32(dev, q, b, 20, 40, typ=2)
    same_as_fp32(dev, q, b, 20, 40, typ=16)


@pytest.mark.parametrize("d", [20, 200])
def test_float_dims_and_dot(dev, d):
    rng = np.random.default_rng(d)
    b = rng.standard_normal((20000, d)).astype(np.float32)
    q = rng.standard_normal((257, d)).astype(np.float32)
    same_as_fp32(dev, q, b, 8, 32, typ=2)
    same_as_fp32(dev, q, b, 8, 32, typ=16)


def test_escalation_certifies(dev):
    """k' = k: the first level can rarely certify; the uncertified queries run
    again with 8k' candidates, which certify them."""
    rng = np.random.default_rng(9)
    b = rng.standard_normal((5000, 64)).astype(np.float32)
    q = rng.standard_normal((100, 64)).astype(np.float32)
    nfb = same_as_fp32(dev, q, b, 20, 20)
    assert nfb == 0


def test_forced_fallback(dev):
    """k' = k = 1024 (no room to escalate): the certificate can rarely hold,
    the FP32 form answers the uncertified queries."""
    rng = np.random.default_rng(9)
    b = rng.standard_normal((5000, 64)).astype(np.float32)
    q = rng.standard_normal((40, 64)).astype(np.float32)
    nfb = same_as_fp32(dev, q, b, 1024, 1024)
    assert nfb > 0


@pytest.mark.parametrize("kp", [64, 300])
def test_massive_ties_overflow_the_candidate_lists(dev, kp):
    """16 distinct integer vectors repeated over 80,000 base rows (~5,000
    copies each): the filtered pass's bound value is shared by every copy of
    a query's nearest vector, so each query captures more than its
    4,096-key list; the overflow path (the k'-th captured key as a tighter
    bound, another pass) must still give the FP32 form's lists, exact ties
    ordered by index."""
    rng = np.random.default_rng(12)
    distinct = mixture(rng, 16, 64, 16, 30)
    b = distinct[rng.integers(0, 16, 80000)]
    q = mixture(rng, 300, 64, 16, 30)
    print("massive ties, k'", kp, flush=True)
    same_as_fp32(dev, q, b, 10, kp)


def test_starting_bound_underflow(dev):
    """The filtered pass starts from a bound estimated from the sketch's
    sample (an m-th smallest sampled minimum, m < k'); a query left with
    fewer than k' keys within it is scanned again without a bound, which
    overflows its list, and then with the bound the overflow gives (the
    k'-th smallest of the keys captured).  Here the only rows near the
    queries are the first rows of every 32nd block: all of them in blocks
    the stride-8 sketch samples, one per base range, so the sample
    over-represents them and the estimated bound lets through fewer than k'
    rows.  Both further passes must run (three filtered-pass launches, one
    sketch, so no escalation) and the lists must be the FP32 form's."""
    import ctypes
    import yael_amd
    rng = np.random.default_rng(21)
    nb, d, k, kp = 65536, 32, 10, 64
    # far rows not so far that the base's norms widen the certificate's
    # margin past the gap between the 10th and the 64th near row
    b = (12.0 + rng.standard_normal((nb, d))).astype(np.float32)
    near = np.arange(0, nb, 32 * 32)
    b[near] = rng.standard_normal((len(near), d)).astype(np.float32)
    q = rng.standard_normal((200, d)).astype(np.float32)
    lib = yael_amd.lib()
    lib.yk_prof_enable(1)
    lib.yk_prof_reset()
    try:
        same_as_fp32(dev, q, b, k, kp)
        counts = []
        for slot in (6, 7):  # filtered-pass scans, sketch scans
            tot, cnt = ctypes.c_double(), ctypes.c_long()
            lib.yk_prof_get(slot, ctypes.byref(tot), ctypes.byref(cnt))
            counts.append(cnt.value)
    finally:
        lib.yk_prof_enable(0)
    print("filtered-pass and sketch launches", counts, flush=True)
    assert counts == [3, 1]


def test_small_base(dev):
    rng = np.random.default_rng(10)
    b = mixture(rng, 50, 16, 4, 10)
    q = mixture(rng, 70, 16, 4, 10)
    same_as_fp32(dev, q, b, 5, 64)


def test_k1_top8(dev):
    """The k-means assign form: k = 1 from 8 re-ranked candidates."""
    rng = np.random.default_rng(11)
    b = rng.standard_normal((3000, 128)).astype(np.float32)
    q = rng.standard_normal((5000, 128)).astype(np.float32)
    nfb = same_as_fp32(dev, q, b, 1, 8)
    print("fallback queries", nfb, flush=True)


@pytest.mark.parametrize("k,kp", [(1, 8), (5, 16)])
def test_many_queries_direct_merge(dev, k, kp):
    """Enough query tiles for one base range: the lane pairs merge their
    candidate lists in registers and write them directly."""
    rng = np.random.default_rng(12 + k)
    b = mixture(rng, 2048, 32, 64, 25, integer=False)
    q = mixture(rng, 600000, 32, 64, 25, integer=False)
    nfb = same_as_fp32(dev, q, b, k, kp)
    print("k", k, "fallback queries", nfb, flush=True)


def test_k1_no_value_below_1e30(dev):
    """k = 1 takes only values below 1e30 (nn.c:404-440): queries whose every
    distance is larger get (-1, 1e30), as in the FP32 form."""
    rng = np.random.default_rng(13)
    b = rng.standard_normal((200, 16)).astype(np.float32)
    b[:100] *= np.float32(1e16)
    q = rng.standard_normal((300, 16)).astype(np.float32)
    q[:150] *= np.float32(1e16)
    same_as_fp32(dev, q, b, 1, 8)
    Q, B = dev.DeviceSet(q), dev.DeviceSet(b[:100])
    i32, d32 = dev.knn_fp32(Q, B, 1)
    i16, d16, nfb = dev.knn_bf16(Q, B, 1, 8)
    Q.free()
    B.free()
    assert np.array_equal(i16, i32) and np.array_equal(d16.view(np.int32), d32.view(np.int32))
    assert (i32[:, 0] == -1).any()


def _bf16_rne(x):
    """float32 -> bf16 bit patterns, round to nearest even (finite values)."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def _bf16_f(h):
    return (h.astype(np.uint32) << 16).view(np.float32)


def _norm_pieces(s2):
    """three bf16 pieces of float(s2), each the rounding of what is left"""
    nf = np.float32(s2)
    h1 = _bf16_rne(np.array([nf]))[0]
    r1 = np.float32(nf - _bf16_f(np.array([h1]))[0])
    h2 = _bf16_rne(np.array([r1]))[0]
    r2 = np.float32(r1 - _bf16_f(np.array([h2]))[0])
    h3 = _bf16_rne(np.array([r2]))[0]
    return [h1, h2, h3]


@pytest.mark.parametrize("d,typ", [(20, 2), (128, 2), (200, 2), (960, 2), (128, 16)])
def test_base_conversion_layout_stats_norms(dev, d, typ):
    """The fragment-major BF16 base copy decoded in numpy: block b, chunk c is
    the 1 KB fragment whose lane h*32 + r holds row 32b + r, columns 16c + 8h
    .. +7; columns past d hold the norm pieces and ones (L2), then zeros; rows
    past n are zero.  Statistics exact on integer data; the fused float norms
    carry the same bits as yk_norms'."""
    from yael_amd import lib
    rng = np.random.default_rng(d + typ)
    n = 70  # two full 32-row blocks and a partial one
    x = rng.integers(0, 256, (n, d)).astype(np.float32)
    S = dev.DeviceSet(x)
    S.compute_norms()
    ref_fn = S.fnorm.to_host(np.float32, (n,)).copy()
    h, st, mx = S.compute_bf16(typ, False, with_norms=True)
    fn = S.fnorm.to_host(np.float32, (n,))
    assert np.array_equal(fn.view(np.int32), ref_fn.view(np.int32)), "fused float norms differ"
    ldh = lib().yk_bf16_ld(d, typ)
    nch, nblk = ldh // 16, (n + 31) // 32
    raw = h.to_host(np.uint16, (nblk, nch, 2, 32, 8))
    got = raw.transpose(0, 3, 1, 2, 4).reshape(nblk * 32, ldh)
    want = np.zeros((nblk * 32, ldh), np.uint16)
    want[:n, :d] = _bf16_rne(x)
    s2 = (x.astype(np.float64) ** 2).sum(1)
    if typ == 2:
        for i in range(n):
            want[i, d:d + 6] = _norm_pieces(s2[i]) + [0x3F80] * 3
    assert np.array_equal(got, want), "BF16 base copy differs"
    stats = st.to_host(np.float32, (n, 4))
    exp = np.stack([np.sqrt(s2), np.sqrt(s2), np.zeros(n), s2], 1).astype(np.float32)
    assert np.array_equal(stats, exp)
    assert np.array_equal(mx.to_host(np.float32, (4,)), exp.max(0))
    S.free()


@pytest.mark.parametrize("d,typ", [(20, 2), (128, 2), (128, 16)])
def test_query_conversion_layout(dev, d, typ):
    """Queries are row-major BF16 rows of -2q (L2; q for dot product) with
    three ones and the norm pieces after column d."""
    from yael_amd import lib
    rng = np.random.default_rng(7 * d + typ)
    n = 45
    x = rng.integers(0, 256, (n, d)).astype(np.float32)
    S = dev.DeviceSet(x)
    h, st, _ = S.compute_bf16(typ, True)
    ldh = lib().yk_bf16_ld(d, typ)
    got = h.to_host(np.uint16, (n, ldh))
    want = np.zeros((n, ldh), np.uint16)
    want[:, :d] = _bf16_rne(x * (np.float32(-2) if typ == 2 else np.float32(1)))
    if typ == 2:
        s2 = (x.astype(np.float64) ** 2).sum(1)
        for i in range(n):
            want[i, d:d + 6] = [0x3F80] * 3 + _norm_pieces(s2[i])
    assert np.array_equal(got, want), "BF16 query copy differs"
    S.free()

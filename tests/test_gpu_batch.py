"""GPU tests of the batched form (many k-means instances in one grouped,
graph-captured run).  Every instance must give the bits of the
single-instance kmeans() with the same seed: the assign step is the same
kernel, the split runs on the device from host-precomputed streams with the
host code's arithmetic, and on integer-valued data every centroid sum is
exact in any order."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ya():
    import yael_amd
    yael_amd.init()
    return yael_amd


def mixture(rng, n, d, ncomp, sigma):
    cen = rng.uniform(0, 255, (ncomp, d))
    x = np.clip(cen[rng.integers(0, ncomp, n)] + sigma * rng.standard_normal((n, d)), 0, 255)
    return np.rint(x).astype(np.float32)


def zipf_sizes(ninst, total, lo, hi, s=1.1):
    w = 1.0 / np.arange(1, ninst + 1) ** s
    n = np.clip(np.floor(w / w.sum() * total), lo, hi).astype(np.int64)
    return n.astype(np.int32)


def single(ya, v, sizes, k, niter, seeds, init=None, flags=0x80000):
    """kmeans() per instance through the C-ABI (not ynumpy, which raises on
    the slightly negative objective the expanded distance can give on
    duplicated points); init: per-instance initial centroids
    (KMEANS_INIT_USER)."""
    import ctypes
    out = []
    off = 0
    d = v.shape[1]
    for i, (n, s) in enumerate(zip(sizes, seeds)):
        vi = np.ascontiguousarray(v[off:off + n])
        c = np.empty((k, d), np.float32) if init is None else init[i].copy()
        dis = np.empty(n, np.float32)
        asg = np.empty(n, np.int32)
        nas = np.empty(k, np.int32)
        q = ctypes.c_float()
        rc = ya.lib().yael_kmeans_checked(d, int(n), k, niter, vi.ctypes.data, 0x10000 | flags,
                                         int(s), 1, c.ctypes.data, dis.ctypes.data, asg.ctypes.data,
                                         nas.ctypes.data, ctypes.byref(q))
        ya.check(rc, "kmeans")
        out.append((c, q.value, dis, asg, nas))
        off += n
    return out


def compare(batched, singles, sizes):
    cent, ret, dis, assign, nassign = batched
    off = 0
    for i, (c, q, d1, a1, n1) in enumerate(singles):
        n = sizes[i]
        assert np.array_equal(assign[off:off + n], a1), "instance %d assign" % i
        assert np.array_equal(nassign[i], n1), "instance %d nassign" % i
        assert np.array_equal(cent[i].view(np.int32), c.view(np.int32)), "instance %d centroids" % i
        assert np.array_equal(dis[off:off + n].view(np.int32), d1.view(np.int32)), "instance %d dis" % i
        assert np.float32(ret[i]) == np.float32(q), "instance %d objective" % i
        off += n


@pytest.mark.parametrize("d", [32, 100, 128])
def test_uneven_instances_match_single_runs(ya, d):
    """d = 32: the float4-column update path (4 waves); d = 100: one thread
    per dimension; d = 128: groups of 32 dimensions per wave."""
    rng = np.random.default_rng(1 + d)
    sizes = zipf_sizes(24, 60000, 256, 20000)
    v = mixture(rng, int(sizes.sum()), d, 200, 18)
    seeds = np.arange(1, 25)
    print("d", d, "sizes", sizes.tolist(), flush=True)
    b = ya.kmeans_batched(v, sizes, 16, niter=12, seeds=seeds)
    compare(b, single(ya, v, sizes, 16, 12, seeds), sizes)


def test_splits_and_overflow_fallback(ya):
    """Duplicated points empty many clusters: splits on the device, and an
    instance with more empties in one iteration than precomputed draws is
    recomputed by the single-instance path."""
    rng = np.random.default_rng(2)
    blocks = []
    sizes = []
    for i in range(6):
        m = 40 + 10 * i
        pts = rng.integers(0, 4 + i, (m // 10 + 3, 8)).astype(np.float32)
        blocks.append(np.repeat(pts, 10, axis=0)[:m] if i % 2 else np.repeat(pts[:2], m // 2 + 1, axis=0)[:m])
        sizes.append(m)
    v = np.ascontiguousarray(np.concatenate(blocks))
    sizes = np.array(sizes, np.int32)
    seeds = np.arange(11, 17)
    b = ya.kmeans_batched(v, sizes, 12, niter=5, seeds=seeds)
    compare(b, single(ya, v, sizes, 12, 5, seeds), sizes)


def test_graph_and_direct_launches_agree(ya):
    import ctypes
    from yael_amd import device
    rng = np.random.default_rng(3)
    sizes = zipf_sizes(40, 50000, 300, 9000)
    v = mixture(rng, int(sizes.sum()), 64, 100, 20)
    V = device.DeviceSet(v)
    lib = ya.lib()
    outs = []
    for graph in (0, 1):
        cent = np.empty((40, 8, 64), np.float32)
        ret = np.empty(40, np.float32)
        it = np.empty(40, np.int32)
        chunk = ctypes.c_int(0)
        seeds = np.arange(100, 140, dtype=np.int64)
        rc = lib.yael_kmeans_batched_device(64, 40, sizes.ctypes.data, 8, 9, V.vec.ptr, 0x10000,
                                            seeds.ctypes.data, cent.ctypes.data, None, None, None,
                                            ret.ctypes.data, it.ctypes.data, graph, ctypes.byref(chunk))
        ya.check(rc, "batched")
        print("graph", graph, "chunk", chunk.value, "iters", it[:8].tolist(), flush=True)
        outs.append((cent, ret, it))
    V.free()
    assert np.array_equal(outs[0][0].view(np.int32), outs[1][0].view(np.int32))
    assert np.array_equal(outs[0][1], outs[1][1]) and np.array_equal(outs[0][2], outs[1][2])


def test_pq_subspaces_match_single_runs(ya):
    """yael_kmeans_subspaces_device: sub-space j (columns 8j..8j+7, padded to
    16 on the device) trained in one batched run equals kmeans() on that
    column block."""
    import ctypes
    from yael_amd import device
    rng = np.random.default_rng(4)
    n, nsub, dsub, k, niter = 5000, 4, 8, 24, 7
    v = mixture(rng, n, nsub * dsub, 50, 22)
    V = device.DeviceSet(v)
    cent = np.empty((nsub, k, dsub), np.float32)
    dis = np.empty(nsub * n, np.float32)
    asg = np.empty(nsub * n, np.int32)
    nas = np.empty((nsub, k), np.int32)
    ret = np.empty(nsub, np.float32)
    it = np.empty(nsub, np.int32)
    chunk = ctypes.c_int(0)
    seeds = np.arange(31, 31 + nsub, dtype=np.int64)
    rc = ya.lib().yael_kmeans_subspaces_device(nsub * dsub, n, nsub, k, niter, V.vec.ptr, 0x10000,
                                               seeds.ctypes.data, cent.ctypes.data,
                                               dis.ctypes.data, asg.ctypes.data, nas.ctypes.data,
                                               ret.ctypes.data, it.ctypes.data,
                                               ctypes.byref(chunk))
    ya.check(rc, "subspaces")
    V.free()
    sub = np.concatenate([v[:, j * dsub:(j + 1) * dsub] for j in range(nsub)])
    sizes = np.full(nsub, n, np.int32)
    compare((cent, ret, dis, asg, nas), single(ya, sub, sizes, k, niter, seeds), sizes)


@pytest.mark.parametrize("dsub", [6, 16, 32])
def test_pq_subspaces_large_instances(ya, dsub):
    """Instances of 40000 points: the update's partial sums take one wave per
    float4 column of the sub-vectors (dsub 16, 32) or the one-wave path
    (dsub 6, not a multiple of 4)."""
    import ctypes
    from yael_amd import device
    rng = np.random.default_rng(40 + dsub)
    n, nsub, k, niter = 40000, 2, 64, 6
    v = mixture(rng, n, nsub * dsub, 120, 20)
    V = device.DeviceSet(v)
    cent = np.empty((nsub, k, dsub), np.float32)
    dis = np.empty(nsub * n, np.float32)
    asg = np.empty(nsub * n, np.int32)
    nas = np.empty((nsub, k), np.int32)
    ret = np.empty(nsub, np.float32)
    it = np.empty(nsub, np.int32)
    chunk = ctypes.c_int(0)
    seeds = np.arange(7, 7 + nsub, dtype=np.int64)
    print("n", n, "dsub", dsub, flush=True)
    rc = ya.lib().yael_kmeans_subspaces_device(nsub * dsub, n, nsub, k, niter, V.vec.ptr, 0x10000,
                                               seeds.ctypes.data, cent.ctypes.data,
                                               dis.ctypes.data, asg.ctypes.data, nas.ctypes.data,
                                               ret.ctypes.data, it.ctypes.data,
                                               ctypes.byref(chunk))
    ya.check(rc, "subspaces")
    V.free()
    sub = np.concatenate([v[:, j * dsub:(j + 1) * dsub] for j in range(nsub)])
    sizes = np.full(nsub, n, np.int32)
    compare((cent, ret, dis, asg, nas), single(ya, sub, sizes, k, niter, seeds), sizes)


def test_profiled_batched_run_leaves_no_error(ya):
    """Library profiling switched on around a batched (graph-captured) run:
    no profiling event may be recorded inside the capture, and the next,
    unrelated calls must not see a stale HIP error."""
    import ctypes
    lib = ya.lib()
    rng = np.random.default_rng(12)
    sizes = np.array([3000, 2500, 4000], np.int32)
    v = mixture(rng, int(sizes.sum()), 32, 40, 15)
    lib.yk_prof_reset()
    lib.yk_prof_enable(1)
    try:
        b = ya.kmeans_batched(v, sizes, 8, niter=4, seeds=np.array([3, 4, 5]))
    finally:
        lib.yk_prof_enable(0)
    tot, cnt = ctypes.c_double(), ctypes.c_long()
    lib.yk_prof_get(5, ctypes.byref(tot), ctypes.byref(cnt))  # the batch slot: reads the events
    compare(b, single(ya, v, sizes, 8, 4, [3, 4, 5]), sizes)
    idx, dis = ya.knn(v[:5], v, 1)
    assert np.array_equal(idx[:, 0], np.arange(5))


@pytest.mark.parametrize("k,d", [(256, 128), (170, 100), (600, 36)])
def test_tables_larger_than_lds_match_single_runs(ya, k, d):
    """k x d beyond one LDS table (bf = 256 at d = 128 is 264 KB of FP64
    sums): the partial sums run in cluster slices, each instance still
    equal to its single-instance kmeans() bit for bit (d = 128: 32-column
    groups; d = 100: a thread per dimension; d = 36: float4 columns)."""
    rng = np.random.default_rng(k + d)
    sizes = zipf_sizes(6, 6 * 4 * k, 2 * k, 10 * k)
    v = mixture(rng, int(sizes.sum()), d, 400, 18)
    seeds = np.arange(11, 11 + len(sizes))
    print("k", k, "d", d, "sizes", sizes.tolist(), flush=True)
    b = ya.kmeans_batched(v, sizes, k, niter=6, seeds=seeds)
    compare(b, single(ya, v, sizes, k, 6, seeds), sizes)


def test_niter_zero_runs_to_convergence(ya):
    """niter = 0 means until convergence (at most 1,000,000 iterations,
    kmeans.c:345): the batched run goes window by window until every
    instance has stopped, and each instance equals its single-instance
    kmeans(niter = 0) bit for bit."""
    import ctypes
    rng = np.random.default_rng(77)
    sizes = zipf_sizes(10, 30000, 500, 9000)
    v = mixture(rng, int(sizes.sum()), 24, 60, 25)
    seeds = np.arange(21, 21 + len(sizes))
    b = ya.kmeans_batched(v, sizes, 20, niter=0, seeds=seeds)
    compare(b, single(ya, v, sizes, 20, 0, seeds), sizes)
    # the device-level entry reports the iterations each instance ran:
    # past one window of 16, and fewer than the cap
    from yael_amd.device import DeviceSet
    V = DeviceSet(v)
    n = int(sizes.sum())
    iters = np.empty(len(sizes), np.int32)
    chunk = ctypes.c_int(0)
    ya.check(ya.lib().yael_kmeans_batched_device(24, len(sizes), sizes.ctypes.data, 20, 0,
                                                 V.vec.ptr, 0x10000, seeds.ctypes.data, None,
                                                 None, None, None, None, iters.ctypes.data, 1,
                                                 ctypes.byref(chunk)))
    V.free()
    print("iterations per instance", iters.tolist(), flush=True)
    assert iters.max() > 16 and iters.max() < 1000


@pytest.mark.parametrize("d", [48, 1100])
def test_user_init_and_normalized_centroids_match_single_runs(ya, d):
    """KMEANS_INIT_USER (the caller's initial centroids; no initialisation
    draw from the seed) with KMEANS_NORMALIZE_CENTS (unit-norm centroids
    after every update, before the split): each instance equals its
    single-instance kmeans() with the same flags bit for bit (d = 1100: a
    row longer than the single path's 1024-column LDS row)."""
    rng = np.random.default_rng(d)
    sizes = zipf_sizes(5, 4000, 300, 2000)
    v = np.clip(np.rint(rng.uniform(0, 63, (int(sizes.sum()), d))), 0, 63).astype(np.float32)
    k = 12
    off = np.concatenate([[0], np.cumsum(sizes)])
    init = np.stack([v[off[i]:off[i] + k] for i in range(len(sizes))])
    seeds = np.arange(3, 3 + len(sizes))
    b = ya.kmeans_batched(v, sizes, k, niter=5, seeds=seeds, init_centroids=init,
                          normalize=True)
    compare(b, single(ya, v, sizes, k, 5, seeds, init=init, flags=0x100000 | 0x40000), sizes)
    np.testing.assert_allclose(np.linalg.norm(b[0].astype(np.float64), axis=2), 1.0, rtol=1e-6)

"""GPU test of yk_knn_f32_noalloc, the FP32 search that takes its scratch
from a static device array instead of allocating: through the C-ABI, its
results equal yk_knn_f32's bit for bit on the same device-resident inputs,
for k = 1, the register top-k (k <= 16) and the sketch + filtered path
(k > 16), L2, dot product and weighted searches, each on two calls in a
row, the second smaller than the first (the array keeps the first call's
contents).  Query counts are chosen so that the k <= 16 and k > 16 shapes
run in several passes."""
import ctypes
import zlib

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

vp = ctypes.c_void_p


def _mixture(rng, n, d, ncomp=50):
    cen = rng.uniform(0, 255, (ncomp, d))
    return np.clip(np.rint(cen[rng.integers(0, ncomp, n)] + 20 * rng.standard_normal((n, d))),
                   0, 255).astype(np.float32)


def _search(entry, typ, qs, bs, k, w=None):
    from yael_amd import lib, check
    from yael_amd.device import DeviceBuffer
    L = lib()
    oi, od = DeviceBuffer(qs.n * k * 4), DeviceBuffer(qs.n * k * 4)
    rc = getattr(L, entry)(typ, qs.n, bs.n, bs.d, k, qs.vec.ptr,
                           None if typ == 16 else qs.dnorm.ptr, bs.vec.ptr, bs.fnorm.ptr,
                           w.ptr if w is not None else None, oi.ptr, od.ptr)
    check(rc, entry)
    out = oi.to_host(np.int32, (qs.n, k)), od.to_host(np.float32, (qs.n, k))
    oi.free()
    od.free()
    return out


# (name, type, k, weighted, (nq, nb, d) first call, (nq, nb, d) second call)
CASES = [
    ("k1_l2", 2, 1, False, (6000, 40000, 128), (700, 9000, 40)),
    ("k10_l2", 2, 10, False, (3000, 30000, 128), (400, 5000, 24)),
    ("k100_l2", 2, 100, False, (2500, 50000, 64), (300, 8000, 128)),
    ("k5_dot", 16, 5, False, (2000, 20000, 96), (150, 3000, 16)),
    ("k7_weighted", 2, 7, True, (1500, 12000, 32), (200, 2500, 8)),
    ("k300_dot", 16, 300, False, (1200, 30000, 48), (90, 4000, 32)),
]


@pytest.mark.parametrize("name,typ,k,weighted,first,second", CASES, ids=[c[0] for c in CASES])
def test_noalloc_matches_allocating_search(name, typ, k, weighted, first, second):
    from yael_amd import lib
    from yael_amd.device import DeviceSet, DeviceBuffer
    rng = np.random.default_rng(zlib.crc32(name.encode()))
    for call, (nq, nb, d) in enumerate((first, second)):
        print("case", name, "call", call, "nq", nq, "nb", nb, "d", d, "k", k, flush=True)
        if typ == 16:
            b = rng.standard_normal((nb, d)).astype(np.float32)
            q = rng.standard_normal((nq, d)).astype(np.float32)
        else:
            b, q = _mixture(rng, nb, d), _mixture(rng, nq, d)
        bs, qs = DeviceSet(b), DeviceSet(q)
        bs.compute_norms()
        qs.compute_norms()
        w = None
        if weighted:
            wh = rng.uniform(0.5, 2.0, nb).astype(np.float32)
            w = DeviceBuffer(nb * 4)
            lib().yk_memcpy_h2d(w.ptr, wh.ctypes.data_as(vp), nb * 4)
        want_i, want_d = _search("yk_knn_f32", typ, qs, bs, k, w)
        got_i, got_d = _search("yk_knn_f32_noalloc", typ, qs, bs, k, w)
        assert np.array_equal(got_i, want_i), (name, call)
        assert np.array_equal(got_d.view(np.int32), want_d.view(np.int32)), (name, call)
        bs.free()
        qs.free()
        if w is not None:
            w.free()


def test_noalloc_rejects_what_the_search_rejects():
    from yael_amd import lib
    from yael_amd.device import DeviceSet, DeviceBuffer
    L = lib()
    bs = DeviceSet(np.zeros((50, 8), np.float32))
    qs = DeviceSet(np.zeros((4, 8), np.float32))
    oi, od = DeviceBuffer(4 * 60 * 4), DeviceBuffer(4 * 60 * 4)
    for typ, k in ((2, 60), (3, 1), (2, 0)):  # k > nb, a type with no GPU path, k < 1
        rc = L.yk_knn_f32_noalloc(typ, 4, 50, 8, k, qs.vec.ptr, qs.dnorm.ptr, bs.vec.ptr,
                                  bs.fnorm.ptr, None, oi.ptr, od.ptr)
        assert rc == -1, (typ, k)
        assert L.yk_last_error()
    for x in (bs, qs, oi, od):
        x.free()

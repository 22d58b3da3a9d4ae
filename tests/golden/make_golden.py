"""Write the golden fixtures of tests/golden/*.npz from the compiled reference.

Run in a tree where oracle/_ref/libyael_ref.so was built from yael's own
sources (oracle/Makefile, target 'ref').  Every fixture stores its inputs
and the reference's outputs; nothing of the reference's code is kept.

    python tests/golden/make_golden.py [fixture names]
"""
import ctypes
import os
import sys

os.environ.setdefault("OPENBLAS_NUM_THREADS", "1")
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import numpy as np  # noqa: E402

from oracle_bind import Reference, have_ref  # noqa: E402

QUIET = 0x10000
BERKELEY = 0x20000
NORMALIZE = 0x40000
INIT_USER = 0x100000


def knn_cases(ref, rng):
    out = {}
    # float data, tiles crossing the 256 boundary, k = 1 and k > 1
    b = rng.standard_normal((700, 20)).astype(np.float32)
    q = rng.standard_normal((300, 20)).astype(np.float32)
    for k in (1, 10):
        a, d = ref.knn_full(2, b, q, k)
        out["l2_k%d" % k] = dict(b=b, q=q, k=np.int32(k), typ=np.int32(2), assign=a, dis=d)
    # dot product, min-k
    a, d = ref.knn_full(16, b, q, 5)
    out["dot_k5"] = dict(b=b, q=q, k=np.int32(5), typ=np.int32(16), assign=a, dis=d)
    # many exact ties: small-integer data
    bi = rng.integers(0, 3, (600, 8)).astype(np.float32)
    qi = rng.integers(0, 3, (90, 8)).astype(np.float32)
    for k in (1, 12):
        a, d = ref.knn_full(2, bi, qi, k)
        out["ties_k%d" % k] = dict(b=bi, q=qi, k=np.int32(k), typ=np.int32(2), assign=a, dis=d)
    # weights
    w = rng.uniform(0.5, 2.0, 700).astype(np.float32)
    for k in (1, 7):
        a, d = ref.knn_full(2, b, q, k, w=w)
        out["w_k%d" % k] = dict(b=b, q=q, w=w, k=np.int32(k), typ=np.int32(2), assign=a, dis=d)
    # NaN rows and fewer finite values than k (padding)
    bn = rng.standard_normal((40, 6)).astype(np.float32)
    bn[5, 2] = np.nan
    bn[17, 0] = np.nan
    qn = rng.standard_normal((30, 6)).astype(np.float32)
    qn[3, 1] = np.nan
    for k in (1, 39):
        a, d = ref.knn_full(2, bn, qn, k)
        out["nan_k%d" % k] = dict(b=bn, q=qn, k=np.int32(k), typ=np.int32(2), assign=a, dis=d)
    # integer-valued mixture data (exact in FP32)
    cen = rng.uniform(0, 255, (16, 128))
    bm = np.clip(np.rint(cen[rng.integers(0, 16, 1000)] + 20 * rng.standard_normal((1000, 128))), 0, 255).astype(np.float32)
    qm = np.clip(np.rint(cen[rng.integers(0, 16, 100)] + 20 * rng.standard_normal((100, 128))), 0, 255).astype(np.float32)
    for k in (1, 100):
        a, d = ref.knn_full(2, bm, qm, k)
        out["mix_k%d" % k] = dict(b=bm, q=qm, k=np.int32(k), typ=np.int32(2), assign=a, dis=d)
    return out


def cross_cases(ref, rng):
    out = {}
    a = rng.standard_normal((77, 33)).astype(np.float32)
    b = rng.standard_normal((51, 33)).astype(np.float32)
    out["cross_l2"] = dict(a=a, b=b, typ=np.int32(2), dist=ref.cross_distances(a, b, 2))
    out["cross_dot"] = dict(a=a, b=b, typ=np.int32(16), dist=ref.cross_distances(a, b, 16))
    return out


def kmeans_cases(ref, rng):
    out = {}
    v = rng.standard_normal((2000, 16)).astype(np.float32)
    r, c, dis, asg, nas = ref.kmeans(v, 20, 10, flags=QUIET, seed=5)
    out["km_plain"] = dict(v=v, k=np.int32(20), niter=np.int32(10), flags=np.int32(QUIET), seed=np.int64(5),
                           redo=np.int32(1), ret=np.float32(r), cent=c, dis=dis, assign=asg, nassign=nas)
    r, c, dis, asg, nas = ref.kmeans(v, 20, 10, flags=QUIET, seed=9, redo=3)
    out["km_redo"] = dict(v=v, k=np.int32(20), niter=np.int32(10), flags=np.int32(QUIET), seed=np.int64(9),
                          redo=np.int32(3), ret=np.float32(r), cent=c, dis=dis, assign=asg, nassign=nas)
    r, c, dis, asg, nas = ref.kmeans(v, 20, 8, flags=QUIET | NORMALIZE, seed=3)
    out["km_norm"] = dict(v=v, k=np.int32(20), niter=np.int32(8), flags=np.int32(QUIET | NORMALIZE),
                          seed=np.int64(3), redo=np.int32(1), ret=np.float32(r), cent=c, dis=dis, assign=asg,
                          nassign=nas)
    # duplicates: many empty clusters, splits every iteration
    vd = np.repeat(rng.integers(0, 5, (40, 8)).astype(np.float32), 25, axis=0)
    r, c, dis, asg, nas = ref.kmeans(vd, 60, 6, flags=QUIET, seed=11)
    out["km_split"] = dict(v=vd, k=np.int32(60), niter=np.int32(6), flags=np.int32(QUIET), seed=np.int64(11),
                           redo=np.int32(1), ret=np.float32(r), cent=c, dis=dis, assign=asg, nassign=nas)
    # one step on integer data with 10 unreachable initial centroids: ten
    # empty clusters split in that step (all arithmetic before the split exact)
    cen = rng.uniform(0, 255, (40, 16))
    vi = np.clip(np.rint(cen[rng.integers(0, 40, 3000)] + 20 * rng.standard_normal((3000, 16))), 0, 255)
    vi = vi.astype(np.float32)
    init = np.concatenate([vi[rng.choice(3000, 30, replace=False)],
                           np.full((10, 16), 10000, np.float32) + np.arange(10, dtype=np.float32)[:, None]])
    r, c, dis, asg, nas = ref.kmeans(vi, 40, 1, flags=QUIET | INIT_USER, seed=13, init=init)
    out["km_split_step"] = dict(v=vi, k=np.int32(40), niter=np.int32(1), flags=np.int32(QUIET | INIT_USER),
                                seed=np.int64(13), redo=np.int32(1), init=init, ret=np.float32(r), cent=c,
                                dis=dis, assign=asg, nassign=nas)
    # one user-initialised step (the per-iteration parity lever)
    init = v[rng.choice(2000, 20, replace=False)].copy()
    r, c, dis, asg, nas = ref.kmeans(v, 20, 1, flags=QUIET | INIT_USER, seed=4, init=init)
    out["km_step"] = dict(v=v, k=np.int32(20), niter=np.int32(1), flags=np.int32(QUIET | INIT_USER),
                          seed=np.int64(4), redo=np.int32(1), init=init, ret=np.float32(r), cent=c, dis=dis,
                          assign=asg, nassign=nas)
    # a cluster above 46340 points: nassign^2 - 1 wraps in 32-bit int
    vb = np.zeros((100000, 4), np.float32)
    # quarter-integer values: every sum below is exact in FP32 and FP64
    vb[99000:] = 100 + np.rint(4 * rng.standard_normal((1000, 4))).astype(np.float32) / 4
    vb[:99000] += rng.integers(0, 2, (99000, 4)).astype(np.float32) * 0.25
    init = np.array([[0, 0, 0, 0], [100, 100, 100, 100], [1000, 1000, 1000, 1000],
                     [2000, 0, 0, 0]], np.float32)
    r, c, dis, asg, nas = ref.kmeans(vb, 4, 1, flags=QUIET | INIT_USER, seed=21, init=init)
    out["km_overflow"] = dict(v=vb, k=np.int32(4), niter=np.int32(1), flags=np.int32(QUIET | INIT_USER),
                              seed=np.int64(21), redo=np.int32(1), init=init, ret=np.float32(r), cent=c,
                              dis=dis, assign=asg, nassign=nas)
    # abandon: far more clusters than distinct points
    va = np.repeat(rng.standard_normal((30, 4)).astype(np.float32), 400, axis=0)
    r, c, dis, asg, nas = ref.kmeans(va, 2000, 5, flags=QUIET, seed=2)
    out["km_abandon"] = dict(v=va, k=np.int32(2000), niter=np.int32(5), flags=np.int32(QUIET),
                             seed=np.int64(2), redo=np.int32(1), ret=np.float32(r))
    return out


def kpp_cases(ref):
    """k-means++ seeding (own generator, so the other fixtures keep their
    inputs): a float-data run with redo, and one step on integer data from
    n <= 8192 points and from a subset of 8k points (n > 8k and n > 8192)."""
    rng = np.random.default_rng(20261203)
    out = {}
    v = rng.standard_normal((2000, 16)).astype(np.float32)
    r, c, dis, asg, nas = ref.kmeans(v, 20, 10, flags=QUIET | BERKELEY, seed=5, redo=2)
    out["km_kpp"] = dict(v=v, k=np.int32(20), niter=np.int32(10), flags=np.int32(QUIET | BERKELEY),
                         seed=np.int64(5), redo=np.int32(2), ret=np.float32(r), cent=c, dis=dis,
                         assign=asg, nassign=nas)
    cen = rng.uniform(0, 255, (50, 32))
    for name, n, k, seed in (("km_kpp_step", 3000, 30, 17), ("km_kpp_sub_step", 20000, 40, 19)):
        vi = np.clip(np.rint(cen[rng.integers(0, 50, n)] + 20 * rng.standard_normal((n, 32))), 0, 255)
        vi = vi.astype(np.float32)
        r, c, dis, asg, nas = ref.kmeans(vi, k, 1, flags=QUIET | BERKELEY, seed=seed)
        out[name] = dict(v=vi, k=np.int32(k), niter=np.int32(1), flags=np.int32(QUIET | BERKELEY),
                         seed=np.int64(seed), redo=np.int32(1), ret=np.float32(r), cent=c, dis=dis,
                         assign=asg, nassign=nas)
    return out


def hkm_cases(ref):
    """hierarchical k-means (own generator): a tree learned with one k-means
    step per node on integer data (exact everywhere) with its quantisation
    of new points and the bytes hkm_write gives; a float-data tree; and the
    descent through a tree of integer centroids (exact distances).  The
    node seeds come from lrand48(), seeded here with srand48."""
    import tempfile
    from oracle_bind import libc_srand48, make_hkm
    rng = np.random.default_rng(20261204)
    out = {}
    cen = rng.uniform(0, 255, (30, 16))
    v = np.clip(np.rint(cen[rng.integers(0, 30, 6000)] + 20 * rng.standard_normal((6000, 16))), 0, 255)
    v = v.astype(np.float32)
    q = np.clip(np.rint(cen[rng.integers(0, 30, 500)] + 20 * rng.standard_normal((500, 16))), 0, 255)
    q = q.astype(np.float32)
    libc_srand48(12345)
    levels, labels, h = ref.hkm_learn(v, 3, 4, 1)
    qidx = ref.hkm_quantize(h, q)
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "tree.hkm")
        ref.lib.hkm_write(path.encode(), h)
        raw = np.frombuffer(open(path, "rb").read(), np.uint8).copy()
    out["hkm_step"] = dict(v=v, nlevel=np.int32(3), bf=np.int32(4), niter=np.int32(1),
                           srand48=np.int64(12345), c0=levels[0], c1=levels[1], c2=levels[2],
                           labels=labels, q=q, qidx=qidx, hkm_file=raw)
    vf = rng.standard_normal((3000, 8)).astype(np.float32)
    libc_srand48(777)
    levels, labels, h = ref.hkm_learn(vf, 2, 5, 6)
    out["hkm_float"] = dict(v=vf, nlevel=np.int32(2), bf=np.int32(5), niter=np.int32(6),
                            srand48=np.int64(777), c0=levels[0], c1=levels[1], labels=labels)
    tl = [rng.integers(0, 64, (6 ** (l + 1), 12)).astype(np.float32) for l in range(3)]
    hq, _arr, _keep = make_hkm(tl, 6)
    qi = rng.integers(0, 64, (2000, 12)).astype(np.float32)
    qidx = ref.hkm_quantize(ctypes.pointer(hq), qi)
    out["hkm_quant"] = dict(c0=tl[0], c1=tl[1], c2=tl[2], bf=np.int32(6), q=qi, qidx=qidx)
    return out


def rng_cases(ref):
    lib = ref.lib
    out = {}
    f = np.empty(64, np.float32)
    lib.fvec_randn_r(f.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), ctypes.c_long(64), ctypes.c_uint(12345))
    lib.ivec_new_random_perm_r.restype = ctypes.POINTER(ctypes.c_int)
    p = lib.ivec_new_random_perm_r(1000, ctypes.c_uint(777))
    perm = np.ctypeslib.as_array(p, shape=(1000,)).copy()
    out["rng"] = dict(randn_seed=np.uint32(12345), randn=f, perm_seed=np.uint32(777), perm=perm)
    return out


def shortlist_cases(ref, rng):
    b = rng.standard_normal((500, 24)).astype(np.float32)
    q = rng.standard_normal((40, 24)).astype(np.float32)
    a = rng.integers(0, 500, (40, 9)).astype(np.int32)
    a[3, 5:] = -1
    d = np.zeros((40, 9), np.float32)
    a2, d2 = ref.knn_reorder_shortlist(b, q, a, d)
    out = {"shortlist": dict(b=b, q=q, assign_in=a, assign=a2, dis=d2)}
    # knn_recompute_exact_dists over three windows of the base: sorted
    # shortlists, each query starting where the previous window stopped
    bw = rng.standard_normal((600, 33)).astype(np.float32)
    qw = rng.standard_normal((50, 33)).astype(np.float32)
    idx = np.sort(np.stack([rng.choice(600, 12, replace=False) for _ in range(50)]), 1)
    idx = idx.astype(np.int32)
    kp = np.zeros(50, np.int32)
    dis = np.full((50, 12), -1.0, np.float32)
    steps = []
    for label0, nbw in ((0, 200), (200, 250), (450, 150)):
        kp, dis = ref.knn_recompute_exact_dists(bw[label0:label0 + nbw], qw, label0, kp, idx, dis)
        steps.append(np.concatenate([kp[:, None].astype(np.float32), dis], 1))
    out["recompute"] = dict(b=bw, q=qw, idx=idx, windows=np.array([[0, 200], [200, 250],
                                                                    [450, 150]], np.int32),
                            after=np.stack(steps))
    return out


def main():
    if not have_ref():
        sys.exit("oracle/_ref/libyael_ref.so missing: run make -C oracle ref")
    ref = Reference()
    rng = np.random.default_rng(20261202)
    cases = {}
    cases.update(knn_cases(ref, rng))
    cases.update(cross_cases(ref, rng))
    cases.update(kmeans_cases(ref, rng))
    cases.update(rng_cases(ref))
    cases.update(shortlist_cases(ref, rng))
    cases.update(kpp_cases(ref))
    cases.update(hkm_cases(ref))
    only = set(sys.argv[1:])  # fixture names to (re)write; none: all
    wrote = 0
    for name, arrs in cases.items():
        if only and name not in only:
            continue
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **arrs)
        wrote += 1
    print("wrote %d fixtures" % wrote)


if __name__ == "__main__":
    main()

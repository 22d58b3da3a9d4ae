"""Python 3 facade with the conventions of yael's ynumpy (yael/ynumpy.py:46-130).

Arrays are row-major (n, d) float32 -- the same bytes as yael's
column-major d x n fvec.  Every call goes through the C-ABI library
(libyael_mi355x.so) and therefore runs on the GPU; errors raise
``YaelGPUError`` instead of aborting.
"""
import ctypes

import numpy as np

from . import lib, check

KMEANS_QUIET = 0x10000
KMEANS_INIT_BERKELEY = 0x20000
KMEANS_NORMALIZE_CENTS = 0x40000
KMEANS_INIT_RANDOM = 0x80000
KMEANS_INIT_USER = 0x100000
KMEANS_L1 = 0x200000
KMEANS_CHI2 = 0x400000


def _rowf32(a, name):
    a = np.asarray(a)
    if a.dtype != np.float32 or a.ndim != 2 or not a.flags.c_contiguous:
        raise TypeError("%s must be a C-contiguous 2-D float32 array" % name)
    return a


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def knn(queries, base, nnn=1, distance_type=2, nt=1, weights=None):
    """k nearest neighbours of each query in base (ynumpy.py:46-67).

    Returns (idx, dis): int32 and float32 arrays of shape (nq, nnn), sorted
    by increasing distance.  distance_type 2/12: squared L2; 16: dot
    product (the nnn smallest).  nt is accepted for compatibility.
    """
    base = _rowf32(base, "base")
    queries = _rowf32(queries, "queries")
    n, d = base.shape
    nq, d2 = queries.shape
    if d != d2:
        raise ValueError("base and queries must have the same dimension (%d != %d)" % (d, d2))
    idx = np.empty((nq, nnn), np.int32)
    dis = np.empty((nq, nnn), np.float32)
    w = None
    if weights is not None:
        w = np.ascontiguousarray(weights, np.float32)
        if w.shape != (n,):
            raise ValueError("weights must have shape (nb,)")
    rc = lib().yael_knn_full_checked(distance_type, nq, n, d, nnn, _ptr(base), _ptr(queries),
                                     _ptr(w), _ptr(idx), _ptr(dis))
    check(rc, "knn")
    return idx, dis


def cross_distances(a, b, distance_type=12):
    """dis[j, i] = distance(a[i], b[j]), shape (nb, na) (ynumpy.py:69-84)."""
    a = _rowf32(a, "a")
    b = _rowf32(b, "b")
    na, d = a.shape
    nb, d2 = b.shape
    if d != d2:
        raise ValueError("dimension mismatch")
    dis = np.empty((nb, na), np.float32)
    rc = lib().yael_cross_distances_checked(distance_type, d, na, nb, _ptr(a), d, _ptr(b), d,
                                            _ptr(dis), na)
    check(rc, "cross_distances")
    return dis


def kmeans(v, k, distance_type=2, nt=1, niter=30, seed=0, redo=1, verbose=True,
           normalize=False, init="random", output="centroids", init_centroids=None):
    """k-means (ynumpy.py:87-130). output='all' returns
    (centroids, qerr, dis, assign, nassign).  init_centroids selects
    KMEANS_INIT_USER."""
    v = _rowf32(v, "v")
    n, d = v.shape
    cent = np.zeros((k, d), np.float32)
    dis = np.empty(n, np.float32)
    assign = np.empty(n, np.int32)
    nassign = np.empty(k, np.int32)
    flags = nt
    if not verbose:
        flags |= KMEANS_QUIET
    if distance_type == 1:
        flags |= KMEANS_L1
    elif distance_type == 3:
        flags |= KMEANS_CHI2
    if init == "random":
        flags |= KMEANS_INIT_RANDOM
    elif init == "kmeans++":
        flags |= KMEANS_INIT_BERKELEY
    if normalize:
        flags |= KMEANS_NORMALIZE_CENTS
    if init_centroids is not None:
        cent[:] = np.asarray(init_centroids, np.float32)
        flags |= KMEANS_INIT_USER
    ret = ctypes.c_float(0)
    rc = lib().yael_kmeans_checked(d, n, k, niter, _ptr(v), flags, seed, redo, _ptr(cent),
                                   _ptr(dis), _ptr(assign), _ptr(nassign), ctypes.byref(ret))
    check(rc, "kmeans")
    qerr = ret.value
    if qerr < 0:
        raise RuntimeError("kmeans: clustering failed. Is dataset diverse enough?")
    if output == "centroids":
        return cent
    return cent, qerr, dis, assign, nassign


def kmeans_batched(v, sizes, k, niter=30, seeds=None, verbose=False, init_centroids=None,
                   normalize=False):
    """ninst independent k-means instances on consecutive row ranges of v
    (instance i: the next sizes[i] rows), each with kmeans() semantics and its
    own seed (seed 0: lrand48(), in instance order).  One batched GPU run.
    init_centroids (ninst, k, d) selects KMEANS_INIT_USER; normalize
    KMEANS_NORMALIZE_CENTS.  Returns (centroids (ninst, k, d), qerr (ninst,),
    dis, assign, nassign (ninst, k))."""
    v = _rowf32(v, "v")
    sizes = np.ascontiguousarray(sizes, np.int32)
    ninst = len(sizes)
    n, d = v.shape
    if int(sizes.sum()) != n:
        raise ValueError("sizes must sum to the number of rows")
    seeds = np.zeros(ninst, np.int64) if seeds is None else np.ascontiguousarray(seeds, np.int64)
    cent = np.empty((ninst, k, d), np.float32)
    dis = np.empty(n, np.float32)
    assign = np.empty(n, np.int32)
    nassign = np.empty((ninst, k), np.int32)
    ret = np.empty(ninst, np.float32)
    flags = 0 if verbose else KMEANS_QUIET
    if init_centroids is not None:
        cent[:] = np.asarray(init_centroids, np.float32).reshape(ninst, k, d)
        flags |= KMEANS_INIT_USER
    else:
        flags |= KMEANS_INIT_RANDOM
    if normalize:
        flags |= KMEANS_NORMALIZE_CENTS
    rc = lib().yael_kmeans_batched(d, ninst, _ptr(sizes), k, niter, _ptr(v), flags, _ptr(seeds),
                                   _ptr(cent), _ptr(dis), _ptr(assign), _ptr(nassign), _ptr(ret))
    check(rc, "kmeans_batched")
    return cent, ret, dis, assign, nassign


def knn_reorder_shortlist(base, queries, idx, dis=None):
    """Re-rank shortlists idx (nq, k) by exact distance (nn.c:528-580)."""
    base = _rowf32(base, "base")
    queries = _rowf32(queries, "queries")
    idx = np.ascontiguousarray(idx, np.int32).copy()
    nq, k = idx.shape
    out = np.empty((nq, k), np.float32) if dis is None else np.ascontiguousarray(dis, np.float32).copy()
    rc = lib().yael_knn_reorder_shortlist_checked(nq, base.shape[0], base.shape[1], k,
                                                  _ptr(base), _ptr(queries), _ptr(idx),
                                                  _ptr(out))
    check(rc, "knn_reorder_shortlist")
    return idx, out

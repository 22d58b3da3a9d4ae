"""ctypes bindings to the parity checker under oracle/ (test infrastructure).

Two libraries are bound with one interface:
  * ``oracle/liboracle.so``  -- the C restatement (``oy_*`` symbols);
  * ``oracle/_ref/libyael_ref.so`` -- the reference's own sources compiled in
    this container (present only where the reference tree was available at
    build time; tests that need it skip otherwise).

All arrays follow yael's layout: a set of n vectors of dimension d is a
C-contiguous float32 array of shape (n, d) (the same bytes as yael's
column-major d x n ``fvec``).
"""
import ctypes
import glob
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libyael_ref.so")

_f32p = ctypes.POINTER(ctypes.c_float)
_i32p = ctypes.POINTER(ctypes.c_int)
_f64p = ctypes.POINTER(ctypes.c_double)


def _p(a, typ):
    if a is None:
        return None
    return a.ctypes.data_as(typ)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class _Api:
    """Common wrapper over the restatement (prefix 'oy_') and the compiled
    reference (no prefix)."""

    def __init__(self, path, prefix):
        self.lib = ctypes.CDLL(path)
        self.prefix = prefix
        self.path = path

    def fn(self, name):
        return getattr(self.lib, self.prefix + name)

    def knn_full(self, typ, b, q, k, w=None, nt=1):
        b, q = _f32(b), _f32(q)
        nb, d = b.shape
        nq = q.shape[0]
        assign = np.empty((nq, k), np.int32)
        dis = np.empty((nq, k), np.float32)
        wv = None if w is None else _f32(w)
        f = self.fn("knn_full_thread")
        f(ctypes.c_int(typ), nq, nb, d, k, _p(b, _f32p), _p(q, _f32p), _p(wv, _f32p),
          _p(assign, _i32p), _p(dis, _f32p), ctypes.c_int(nt))
        return assign, dis

    def cross_distances(self, a, b, typ=2):
        a, b = _f32(a), _f32(b)
        na, d = a.shape
        nb = b.shape[0]
        out = np.empty((nb, na), np.float32)  # dist2[i + na*j]
        if typ == 2:
            self.fn("cross_distances" if self.prefix else "compute_cross_distances")(
                d, na, nb, _p(a, _f32p), _p(b, _f32p), _p(out, _f32p))
        else:
            self.fn("cross_distances_alt" if self.prefix else "compute_cross_distances_alt")(
                typ, d, na, nb, _p(a, _f32p), _p(b, _f32p), _p(out, _f32p))
        return out

    def kmeans(self, v, k, niter, flags=0, seed=1, redo=1, init=None):
        v = _f32(v)
        n, d = v.shape
        cent = np.zeros((k, d), np.float32) if init is None else _f32(init).copy()
        dis = np.empty(n, np.float32)
        assign = np.empty(n, np.int32)
        nassign = np.empty(k, np.int32)
        f = self.fn("kmeans")
        f.restype = ctypes.c_float
        r = f(d, n, k, niter, _p(v, _f32p), ctypes.c_int(flags), ctypes.c_long(seed),
              redo, _p(cent, _f32p), _p(dis, _f32p), _p(assign, _i32p), _p(nassign, _i32p))
        return float(r), cent, dis, assign, nassign

    def knn_recompute_exact_dists(self, b, v, label0, kp, idx, dis):
        """The window re-computation (nn.c:583-600); returns (kp, dis) as
        updated (copies).  The oracle's twin returns -1 instead of asserting
        on an entry below the window; that is raised as ValueError."""
        b, v = _f32(b), _f32(v)
        idx = np.ascontiguousarray(idx, np.int32)
        kp = np.ascontiguousarray(kp, np.int32).copy()
        dis = np.ascontiguousarray(dis, np.float32).copy()
        nq, k = idx.shape
        f = self.fn("knn_recompute_exact_dists")
        r = f(nq, b.shape[0], b.shape[1], k, _p(b, _f32p), _p(v, _f32p), ctypes.c_int(label0),
              _p(kp, _i32p), _p(idx, _i32p), _p(dis, _f32p))
        if self.prefix and r != 0:
            raise ValueError("an entry below the window")
        return kp, dis

    def knn_reorder_shortlist(self, b, v, assign, dis):
        b, v = _f32(b), _f32(v)
        assign = np.ascontiguousarray(assign, np.int32).copy()
        dis = np.ascontiguousarray(dis, np.float32).copy()
        n, k = assign.shape
        self.fn("knn_reorder_shortlist")(n, b.shape[0], b.shape[1], k, _p(b, _f32p),
                                         _p(v, _f32p), _p(assign, _i32p), _p(dis, _f32p))
        return assign, dis


class Oracle(_Api):
    def __init__(self):
        super().__init__(ORACLE_SO, "oy_")
        self.lib.oy_fvec_sum.restype = ctypes.c_double
        self.lib.oy_kmeans_traced.restype = ctypes.c_float

    def set_blas(self, sgemm_addr, sgemv_addr):
        self.lib.oy_set_blas(ctypes.c_void_p(sgemm_addr), ctypes.c_void_p(sgemv_addr))

    def knn_f64(self, typ, b, q, k, w=None):
        b, q = _f32(b), _f32(q)
        nb, d = b.shape
        nq = q.shape[0]
        assign = np.empty((nq, k), np.int32)
        dis = np.empty((nq, k), np.float64)
        wv = None if w is None else _f32(w)
        self.lib.oy_knn_f64(typ, nq, nb, d, k, _p(b, _f32p), _p(q, _f32p), _p(wv, _f32p),
                            _p(assign, _i32p), _p(dis, _f64p))
        return assign, dis

    def pair_dist_f64(self, typ, b, q, assign):
        b, q = _f32(b), _f32(q)
        assign = np.ascontiguousarray(assign, np.int32)
        nq, k = assign.shape
        out = np.empty((nq, k), np.float64)
        self.lib.oy_pair_dist_f64(typ, nq, b.shape[1], k, _p(b, _f32p), _p(q, _f32p),
                                  _p(assign, _i32p), _p(out, _f64p))
        return out

    def kmeans_traced(self, v, k, niter, flags=0, seed=1, redo=1, init=None):
        v = _f32(v)
        n, d = v.shape
        cent = np.zeros((k, d), np.float32) if init is None else _f32(init).copy()
        dis = np.empty(n, np.float32)
        assign = np.empty(n, np.int32)
        nassign = np.empty(k, np.int32)
        size = self.lib.oy_trace_sizeof()
        buf = ctypes.create_string_buffer(size)
        r = self.lib.oy_kmeans_traced(d, n, k, niter, _p(v, _f32p), ctypes.c_int(flags),
                                      ctypes.c_long(seed), redo, _p(cent, _f32p),
                                      _p(dis, _f32p), _p(assign, _i32p), _p(nassign, _i32p),
                                      buf)
        raw = bytes(buf)
        niter_done, tot = np.frombuffer(raw[:8], np.int32)
        qerr = np.frombuffer(raw[8:8 + 8 * 1024], np.float64)[:niter_done].copy()
        nre = np.frombuffer(raw[8 + 8 * 1024:8 + 8 * 1024 + 4 * 1024], np.int32)[:niter_done].copy()
        return float(r), cent, dis, assign, nassign, dict(niter=int(niter_done), qerr=qerr,
                                                         nreassign=nre, nreassign_tot=int(tot))

    def hkm_learn(self, v, nlevel, bf, niter):
        """(levels, leaf labels) of the restatement (None where it fails)"""
        v = _f32(v)
        n, d = v.shape
        rows = sum(bf ** (l + 1) for l in range(nlevel))
        cent = np.zeros((rows, d), np.float32)
        labels = np.empty(n, np.int32)
        if self.lib.oy_hkm_learn(n, d, nlevel, bf, _p(v, _f32p), niter, _p(cent, _f32p),
                                 _p(labels, _i32p)) != 0:
            return None
        levels, off = [], 0
        for l in range(nlevel):
            levels.append(cent[off:off + bf ** (l + 1)].copy())
            off += bf ** (l + 1)
        return levels, labels

    def hkm_quantize(self, levels, bf, v):
        v = _f32(v)
        cent = np.ascontiguousarray(np.concatenate(levels), np.float32)
        idx = np.empty(v.shape[0], np.int32)
        self.lib.oy_hkm_quantize(len(levels), bf, cent.shape[1], _p(cent, _f32p), v.shape[0],
                                 _p(v, _f32p), _p(idx, _i32p))
        return idx

    def reassign_empty(self, cent, nassign, seed):
        cent = _f32(cent).copy()
        k, d = cent.shape
        nassign = np.ascontiguousarray(nassign, np.int32)
        n = self.lib.oy_reassign_empty(d, k, _p(cent, _f32p), _p(nassign, _i32p),
                                       ctypes.c_uint(seed))
        return n, cent

    def random_init(self, n, k, seed):
        sel = np.empty(k, np.int32)
        self.lib.oy_random_init(n, k, _p(sel, _i32p), ctypes.c_uint(seed))
        return sel

    def randn_r(self, n, seed):
        out = np.empty(n, np.float32)
        self.lib.oy_fvec_randn_r(_p(out, _f32p), ctypes.c_long(n), ctypes.c_uint(seed))
        return out


def _openblas_single_thread():
    """Run OpenBLAS single-threaded inside yael's OpenMP workers (yael's README
    recommends a sequential BLAS); the environment of a GPU host may ask for
    more threads, which deadlocks a pthreads OpenBLAS called from OpenMP."""
    os.environ["OPENBLAS_NUM_THREADS"] = "1"
    path = openblas_path()
    if path:
        lib = ctypes.CDLL(path)
        lib.scipy_openblas_set_num_threads(1)


class HkmT(ctypes.Structure):
    """hkm_t (hkm.h:10-16)."""
    _fields_ = [("nlevel", ctypes.c_int), ("bf", ctypes.c_int), ("k", ctypes.c_int),
                ("d", ctypes.c_int), ("centroids", ctypes.POINTER(ctypes.POINTER(ctypes.c_float)))]


def make_hkm(levels, bf):
    """An hkm_t over numpy level arrays (kept alive by the returned tuple)."""
    levels = [_f32(a) for a in levels]
    arr = (ctypes.POINTER(ctypes.c_float) * len(levels))(*[_p(a, _f32p) for a in levels])
    h = HkmT(len(levels), bf, bf ** len(levels), levels[0].shape[1], arr)
    return h, arr, levels


def libc_srand48(seed):
    """Seed the process's lrand48 stream (hkm's kmeans calls draw from it)."""
    ctypes.CDLL(None).srand48(ctypes.c_long(seed))


class Reference(_Api):
    def __init__(self):
        _openblas_single_thread()
        super().__init__(REF_SO, "")

    def hkm_learn(self, v, nlevel, bf, niter):
        """(levels: list of (bf^(l+1), d) arrays, leaf labels, hkm_t pointer)"""
        v = _f32(v)
        n, d = v.shape
        f = self.lib.hkm_learn
        f.restype = ctypes.POINTER(HkmT)
        asg = ctypes.POINTER(ctypes.c_int)()
        h = f(n, d, nlevel, bf, _p(v, _f32p), niter, 1, 0, ctypes.byref(asg))
        hc = h.contents
        levels = [np.ctypeslib.as_array(hc.centroids[l], shape=(bf ** (l + 1), d)).copy()
                  for l in range(nlevel)]
        labels = np.ctypeslib.as_array(asg, shape=(n,)).copy()
        ctypes.CDLL(None).free(asg)
        return levels, labels, h

    def hkm_quantize(self, hptr, v):
        v = _f32(v)
        idx = np.empty(v.shape[0], np.int32)
        self.lib.hkm_quantize(hptr, v.shape[0], _p(v, _f32p), _p(idx, _i32p))
        return idx


def openblas_path():
    import scipy
    libs = os.path.join(os.path.dirname(scipy.__file__), os.pardir, "scipy.libs")
    cands = sorted(glob.glob(os.path.join(libs, "libscipy_openblas-*.so")))
    return cands[0] if cands else None


def openblas_entry_points():
    """Addresses of the OpenBLAS sgemm_/sgemv_ the compiled reference links."""
    path = openblas_path()
    if path is None:
        return None
    _openblas_single_thread()
    lib = ctypes.CDLL(path)
    g = ctypes.cast(lib.scipy_sgemm_, ctypes.c_void_p).value
    v = ctypes.cast(lib.scipy_sgemv_, ctypes.c_void_p).value
    return g, v


def have_ref():
    return os.path.exists(REF_SO)

"""Hierarchical k-means (yael's hkm.h) on libyael_mi355x.

``hkm_learn(v, nlevel, bf, niter)`` builds the tree (hkm.c:36-116; every
level one batched k-means on the GPU) and returns an :class:`Hkm` and the
leaf label of every point; :meth:`Hkm.quantize` descends the tree on the
GPU (hkm.c:131-150); :meth:`Hkm.write` / :func:`hkm_read` use the
reference's file format (hkm.c:181-232).  Arrays are row-major (n, d)
float32, as in ynumpy.
"""
import ctypes

import numpy as np

from . import lib, check


class HkmT(ctypes.Structure):
    """hkm_t, same layout as hkm.h:10-16."""
    _fields_ = [("nlevel", ctypes.c_int), ("bf", ctypes.c_int), ("k", ctypes.c_int),
                ("d", ctypes.c_int), ("centroids", ctypes.POINTER(ctypes.POINTER(ctypes.c_float)))]


class Hkm:
    """A tree owned by the library (freed with hkm_delete)."""

    def __init__(self, ptr):
        self._p = ptr
        h = ctypes.cast(ptr, ctypes.POINTER(HkmT)).contents
        self.nlevel, self.bf, self.k, self.d = h.nlevel, h.bf, h.k, h.d

    @property
    def ptr(self):
        return self._p

    def centroids(self, level):
        """The bf^(level+1) centroids of a level, (rows, d) float32 copy; the
        children of node p are rows p*bf .. p*bf + bf - 1."""
        h = ctypes.cast(self._p, ctypes.POINTER(HkmT)).contents
        rows = self.bf ** (level + 1)
        return np.ctypeslib.as_array(h.centroids[level], shape=(rows, self.d)).copy()

    def quantize(self, v):
        v = np.ascontiguousarray(v, np.float32)
        n, d = v.shape
        if d != self.d:
            raise ValueError("dimension %d, tree has %d" % (d, self.d))
        idx = np.empty(n, np.int32)
        check(lib().yael_hkm_quantize_checked(self._p, n, v.ctypes.data, idx.ctypes.data),
              "hkm_quantize")
        return idx

    def write(self, path):
        lib().hkm_write(str(path).encode(), self._p)

    def __del__(self):
        if getattr(self, "_p", None):
            try:
                lib().hkm_delete(self._p)
            except Exception:
                pass
            self._p = None


def hkm_learn(v, nlevel, bf, niter=30, verbose=False):
    """(Hkm, leaf label of every point)."""
    v = np.ascontiguousarray(v, np.float32)
    n, d = v.shape
    out = ctypes.c_void_p()
    asg = ctypes.POINTER(ctypes.c_int)()
    check(lib().yael_hkm_learn_checked(n, d, nlevel, bf, v.ctypes.data, niter, 1, int(verbose),
                                       ctypes.byref(asg), ctypes.byref(out)), "hkm_learn")
    labels = np.ctypeslib.as_array(asg, shape=(n,)).copy()
    ctypes.CDLL(None).free(asg)
    return Hkm(out.value), labels


def hkm_read(path):
    p = lib().hkm_read(str(path).encode())
    if not p:
        raise OSError("hkm_read: cannot read %s" % path)
    return Hkm(p)

"""yael_amd -- yael's k-means / exhaustive-kNN path on one MI355X (gfx950).

The product is the shared library ``yael_amd/_lib/libyael_mi355x.so``:
yael's C API (``kmeans``, ``knn_full``, ``knn_full_thread``,
``compute_cross_distances`` ...; declared in ``include/yael_mi355x.h``) over
hand-written HIP kernels.  This package only loads that library and offers a
Python 3 facade with the conventions of yael's ``ynumpy`` module
(yael/ynumpy.py:46-130): row-major (n, d) float32 arrays, ``knn(queries,
base, nnn)`` argument order.

There is no CPU fallback: when the library is missing or no GPU is visible,
every entry point raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libyael_mi355x.so")

_lib = None


class YaelGPUError(RuntimeError):
    pass


def lib():
    """The loaded C-ABI library (ctypes.CDLL); raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise YaelGPUError(
                "libyael_mi355x.so is not built (run __graft_entry__.build() or "
                "make -C yael_amd/csrc)")
        _lib = ctypes.CDLL(LIB_PATH)
        from . import _abi
        _abi.declare(_lib)
    return _lib


def check(rc, what="yael_amd"):
    """Raise YaelGPUError for a negative yk_* return code."""
    if rc != 0:
        msg = lib().yk_last_error().decode(errors="replace")
        raise YaelGPUError("%s failed (%d): %s" % (what, rc, msg))


def init():
    check(lib().yk_init(), "yk_init")


def device_info():
    init()
    buf = ctypes.create_string_buffer(512)
    check(lib().yk_device_info(buf, 512), "yk_device_info")
    return buf.value.decode()


from .ynumpy import knn, cross_distances, kmeans, kmeans_batched, knn_reorder_shortlist  # noqa: E402,F401,E501

"""Device-resident vector sets and searches over the C-ABI (yk_* functions).

A ``DeviceSet`` holds n vectors in HBM in the library's padded layout (row
stride yk_ld(d) floats), with the norms the FP32 form needs and, on demand,
the BF16 copy and error statistics of the precision form.  Searches run on
the library's HIP stream and return numpy arrays.
"""
import ctypes

import numpy as np

from . import lib, check

vp = ctypes.c_void_p


class DeviceBuffer:
    def __init__(self, nbytes):
        self.ptr = vp()
        self.nbytes = int(nbytes)
        check(lib().yk_malloc(ctypes.byref(self.ptr), max(256, self.nbytes)), "yk_malloc")

    def free(self):
        if self.ptr:
            lib().yk_free(self.ptr)
            self.ptr = vp()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

    def to_host(self, dtype, shape):
        out = np.empty(shape, dtype)
        check(lib().yk_memcpy_d2h(out.ctypes.data_as(vp), self.ptr, out.nbytes), "d2h")
        return out


class DeviceSet:
    """n vectors of dimension d resident in HBM (float32, padded rows)."""

    def __init__(self, x=None, shape=None):
        """Upload host array x (n, d), or allocate an uninitialised set of
        shape (n, d) (filled on the device, e.g. by ``synth``)."""
        if x is not None:
            x = np.ascontiguousarray(x, np.float32)
            shape = x.shape
        self.n, self.d = int(shape[0]), int(shape[1])
        L = lib()
        self.ld = L.yk_ld(self.d)
        self.vec = DeviceBuffer(self.n * self.ld * 4)
        if x is not None:
            check(L.yk_upload_vectors(self.vec.ptr, x.ctypes.data_as(vp), self.n, self.d),
                  "upload")
        self.fnorm = DeviceBuffer(((self.n + 31) // 32 * 32) * 4)  # float-accumulated (base role)
        self.dnorm = DeviceBuffer(self.n * 8)                      # double-accumulated (query role)
        self.bf = {}

    @classmethod
    def synth(cls, n, d, kind, seed, lo=0.0, hi=1.0, centres=None, sigma=0.0, round_values=False,
              unit_rows=False):
        """A synthetic set generated in HBM (yk_synth): kind 0 uniform
        [lo, hi), 1 mixture around the rows of ``centres`` (a DeviceSet with
        ld == d) clipped to [lo, hi], 2 |N(0,1)|."""
        s = cls(shape=(n, d))
        cptr, ncomp = None, 0
        if centres is not None:
            if centres.ld != centres.d or centres.d != d:
                raise ValueError("centres must be a DeviceSet of dimension d with ld == d")
            cptr, ncomp = centres.vec.ptr, centres.n
        check(lib().yk_synth(s.vec.ptr, n, d, s.ld, kind, lo, hi, cptr, ncomp, sigma,
                             int(round_values), seed, int(unit_rows)), "synth")
        return s

    def compute_norms(self):
        check(lib().yk_norms(self.vec.ptr, self.n, self.d, self.fnorm.ptr, self.dnorm.ptr), "norms")

    def compute_bf16(self, typ, is_query, with_norms=False):
        """BF16 copy (+ error statistics) for the precision form; with_norms
        (base sets) also writes the float norms, as compute_norms does."""
        key = (typ, bool(is_query))
        if key not in self.bf:
            nbytes = lib().yk_bf16_bytes(self.n, self.d, typ, int(is_query))
            self.bf[key] = (DeviceBuffer(nbytes), DeviceBuffer(self.n * 16), DeviceBuffer(16))
        h, st, mx = self.bf[key]
        # the set maxima are needed for the base only (the certificate bounds
        # use each query's own statistics)
        check(lib().yk_to_bf16(self.vec.ptr, self.n, self.d, typ, int(is_query), h.ptr, st.ptr,
                               None if is_query else mx.ptr,
                               self.fnorm.ptr if with_norms else None), "to_bf16")
        return h, st, mx

    def free(self):
        for b in (self.vec, self.fnorm, self.dnorm):
            b.free()
        for t in self.bf.values():
            for b in t:
                b.free()


def knn_fp32(queries, base, k, typ=2, norms=True):
    """FP32 form: (idx, dis) of shape (nq, k).  queries/base are DeviceSets."""
    L = lib()
    if norms:
        queries.compute_norms()
        base.compute_norms()
    oi, od = DeviceBuffer(queries.n * k * 4), DeviceBuffer(queries.n * k * 4)
    check(L.yk_knn_f32(typ, queries.n, base.n, base.d, k, queries.vec.ptr, queries.dnorm.ptr,
                       base.vec.ptr, base.fnorm.ptr, None, oi.ptr, od.ptr), "knn_f32")
    return oi.to_host(np.int32, (queries.n, k)), od.to_host(np.float32, (queries.n, k))


def knn_bf16(queries, base, k, kprime, typ=2, prepare=True):
    """Precision form: BF16 first pass, FP32 re-rank of kprime candidates,
    certificate, FP32 fallback.  Returns (idx, dis, n_fallback)."""
    L = lib()
    if prepare:
        queries.compute_norms()
    qh, qst, _ = queries.compute_bf16(typ, True) if prepare else queries.bf[(typ, True)]
    # the base's float norms come out of its conversion pass
    bh, _, bmx = (base.compute_bf16(typ, False, with_norms=True) if prepare
                  else base.bf[(typ, False)])
    oi, od = DeviceBuffer(queries.n * k * 4), DeviceBuffer(queries.n * k * 4)
    nfb = ctypes.c_int(0)
    check(L.yk_knn_bf16(typ, queries.n, base.n, base.d, k, kprime, queries.vec.ptr,
                        queries.dnorm.ptr, qh.ptr, qst.ptr, base.vec.ptr, base.fnorm.ptr, bh.ptr,
                        bmx.ptr, oi.ptr, od.ptr, ctypes.byref(nfb)), "knn_bf16")
    return oi.to_host(np.int32, (queries.n, k)), od.to_host(np.float32, (queries.n, k)), nfb.value

"""Synthetic inputs of the measured workloads (host side, numpy).

* ``mixture``: vectors from a Gaussian mixture -- component centres uniform
  in [0, 255]^d, isotropic sigma, clipped to [0, 255], optionally rounded to
  integer values (float32).  The k-means and kNN workloads at d = 128 use it.
* ``unit_abs_normal``: i.i.d. |N(0, 1)| entries, each row scaled to unit L2
  norm (the d = 960 precision-form workload).
* ``zipf_sizes``: instance sizes proportional to rank^-s, water-filled into
  [lo, hi] so that they sum to a given total (the batched form's uneven
  instances: hkm tree nodes).

Large sets can also be generated directly in HBM (``device.DeviceSet.synth``).
"""
import numpy as np


def mixture(seed, n, d, ncomp, sigma, integer=True, centres_seed=None, lo=0.0, hi=255.0):
    """n vectors of dimension d from an ncomp-component mixture (float32).
    The centres come from centres_seed (default: seed), the draws from seed,
    generated in slices so that large n stays within a bounded working set."""
    crng = np.random.default_rng(centres_seed if centres_seed is not None else seed)
    cen = crng.uniform(lo, hi, (ncomp, d)).astype(np.float32)
    rng = np.random.default_rng(seed)
    out = np.empty((n, d), np.float32)
    step = 1 << 18
    for i in range(0, n, step):
        m = min(step, n - i)
        x = cen[rng.integers(0, ncomp, m)] + np.float32(sigma) * rng.standard_normal((m, d),
                                                                                     np.float32)
        np.clip(x, lo, hi, out=x)
        if integer:
            np.rint(x, out=x)
        out[i:i + m] = x
    return out


def unit_abs_normal(seed, n, d):
    """n rows of |N(0,1)| entries scaled to unit L2 norm (float32)."""
    x = np.abs(np.random.default_rng(seed).standard_normal((n, d), np.float32))
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    return x


def zipf_sizes(ninst, total, lo, hi, s=1.1):
    """ninst sizes proportional to rank^-s, clamped to [lo, hi], summing to
    total: the scale is found by bisection, the floor's remainder goes to the
    largest instance still below hi."""
    w = 1.0 / np.arange(1, ninst + 1) ** s
    a, b = 0.0, float(total) / w.min()
    for _ in range(200):
        lam = 0.5 * (a + b)
        if np.clip(lam * w, lo, hi).sum() > total:
            b = lam
        else:
            a = lam
    n = np.floor(np.clip(a * w, lo, hi)).astype(np.int64)
    free = np.nonzero(n < hi)[0]
    n[free[0]] += total - n.sum()
    return n.astype(np.int32)

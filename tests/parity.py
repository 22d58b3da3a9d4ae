"""Parity rules shared by the tests.

Bit-exact where the arithmetic is exact (integer-valued data with d*255^2 <
2^24: every norm, dot product and expanded distance is an exact FP32 integer,
whatever the summation order); otherwise the near-tie rule:

  a disagreement between two index results is a near-tie when the FP64
  distances of the two candidates differ by at most
      tol = 2 * gamma_d * (||x||^2 + max_c ||c||^2),  gamma_d = (d + 2) * 2^-24,
  the bound on the rounding error of the expanded FP32 distance.
Exact ties (equal values) are ordered by lower index on the GPU; the
reference's heap may order them differently, so k > 1 lists are compared as
(distance multiset, index set up to the boundary near-ties).
"""
import numpy as np


def gamma(d):
    return (d + 2) * 2.0 ** -24


def f64_dist(typ, q, b, idx, w=None):
    """FP64 distance of every (query, idx) pair; NaN where idx < 0."""
    q = q.astype(np.float64)
    b = b.astype(np.float64)
    idx = np.asarray(idx)
    safe = np.where(idx < 0, 0, idx)
    bb = b[safe]  # (nq, k, d)
    if typ == 16:
        v = np.einsum("qd,qkd->qk", q, bb)
    else:
        v = ((q[:, None, :] - bb) ** 2).sum(-1)
    if w is not None:
        v = v * np.asarray(w, np.float64)[safe]
    return np.where(idx < 0, np.nan, v)


def tolerance(typ, q, b, w=None):
    d = q.shape[1]
    q64 = q.astype(np.float64)
    b64 = b.astype(np.float64)
    qn = np.nan_to_num((q64 ** 2).sum(1), nan=0.0)
    bmax = np.nanmax((b64 ** 2).sum(1)) if len(b) else 0.0
    tol = 2 * gamma(d) * (qn + bmax)
    if typ == 16:
        tol = 2 * gamma(d) * np.sqrt(qn * bmax)
    if w is not None:
        tol = tol * float(np.nanmax(np.abs(w)))
    return tol  # per query


def check_knn(typ, q, b, k, idx_a, dis_a, idx_b, dis_b, w=None, exact=False, label=""):
    """Compare two kNN results (a = GPU, b = reference/oracle)."""
    idx_a = np.asarray(idx_a).reshape(len(q), k)
    idx_b = np.asarray(idx_b).reshape(len(q), k)
    dis_a = np.asarray(dis_a).reshape(len(q), k)
    dis_b = np.asarray(dis_b).reshape(len(q), k)
    if exact:
        same_i = np.array_equal(idx_a, idx_b)
        same_d = np.array_equal(dis_a.view(np.int32), dis_b.view(np.int32))
        if same_i and same_d:
            return dict(exact=True, near_ties=0)
        # exact ties may be ordered differently (k > 1): compare sorted values
        assert np.array_equal(np.sort(dis_a, 1).view(np.int32), np.sort(dis_b, 1).view(np.int32)), \
            "%s: distance multisets differ" % label
        bad = np.nonzero((idx_a != idx_b).any(1))[0]
        for i in bad:
            # differing indices must carry equal distances
            va = dict(zip(idx_a[i].tolist(), dis_a[i].tolist()))
            vb = dict(zip(idx_b[i].tolist(), dis_b[i].tolist()))
            diff = set(va) ^ set(vb)
            vals_a = sorted(va[j] for j in set(va) - set(vb))
            vals_b = sorted(vb[j] for j in set(vb) - set(va))
            assert vals_a == vals_b, "%s: query %d differs beyond exact ties: %s" % (label, i, diff)
        return dict(exact=False, exact_ties=len(bad), near_ties=0)
    tol = tolerance(typ, q, b, w)
    da = f64_dist(typ, q, b, idx_a, w)
    db = f64_dist(typ, q, b, idx_b, w)
    near = 0
    for i in np.nonzero((idx_a != idx_b).any(1))[0]:
        # both lists must hold the same FP64 distance multiset up to tol
        sa, sb = np.sort(da[i]), np.sort(db[i])
        fa, fb = np.isfinite(sa), np.isfinite(sb)
        assert np.array_equal(fa, fb), "%s: query %d: finite counts differ" % (label, i)
        assert np.all(np.abs(sa[fa] - sb[fb]) <= tol[i] + 1e-300), \
            "%s: query %d: distance multisets differ beyond the near-tie bound (%g > %g)" % (
                label, i, np.max(np.abs(sa[fa] - sb[fb])), tol[i])
        # every index present in only one list lies at the k-th boundary
        only = set(idx_a[i].tolist()) ^ set(idx_b[i].tolist())
        if only:
            kth = max(np.nanmax(da[i]), np.nanmax(db[i]))
            for j in only:
                dj = f64_dist(typ, q[i:i + 1], b, np.array([[j]]), w)[0, 0]
                assert abs(dj - kth) <= 2 * tol[i] + 1e-300, \
                    "%s: query %d: index %d is not a boundary near-tie" % (label, i, j)
        near += 1
    # float distances within the expanded-form rounding bound of each other
    both = np.isfinite(dis_a) & np.isfinite(dis_b) & (idx_a == idx_b)
    err = np.abs(dis_a.astype(np.float64) - dis_b.astype(np.float64))
    lim = np.broadcast_to(tol[:, None], err.shape)
    assert np.all(err[both] <= lim[both] + 1e-30), "%s: distances differ beyond tolerance" % label
    return dict(exact=False, near_ties=int(near))

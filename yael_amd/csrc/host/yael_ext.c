/*
 * yael_ext.c -- the shortlist re-rank functions on the GPU:
 * knn_reorder_shortlist (nn.h:102-104, nn.c:528-580) and
 * knn_recompute_exact_dists (nn.h:127-130, nn.c:583-600).
 */
#include <limits.h>
#include <stdlib.h>
#include <string.h>

#include "../yk.h"
#include "yael_host.h"

static int yh_shortlist(int n, int nb, int d, int k, const float *b, const float *v, int *idx,
                        float *dis) {
  if (n <= 0 || k <= 0) return YK_OK;
  int rc = yk_init();
  if (rc != YK_OK) return rc;
  int ld = yk_ld(d);
  float *db = (float *)yh_buf(YH_B, sizeof(float) * (size_t)nb * ld);
  float *dv = (float *)yh_buf(YH_Q, sizeof(float) * (size_t)n * ld);
  int *di = (int *)yh_buf(YH_OUTI, sizeof(int) * (size_t)n * k);
  float *dd = (float *)yh_buf(YH_OUTD, sizeof(float) * (size_t)n * k);
  if (!db || !dv || !di || !dd) return YK_ENOMEM;
  if ((rc = yk_upload_vectors(db, b, nb, d)) != YK_OK) return rc;
  if ((rc = yk_upload_vectors(dv, v, n, d)) != YK_OK) return rc;
  if ((rc = yk_memcpy_h2d(di, idx, sizeof(int) * (size_t)n * k)) != YK_OK) return rc;
  if ((rc = yk_memcpy_h2d(dd, dis, sizeof(float) * (size_t)n * k)) != YK_OK) return rc;
  for (long i = 0; i < (long)n * k; i++)
    if (idx[i] >= nb) return yk_set_error_c(YK_EINVAL, "shortlist label %d >= nb=%d", idx[i], nb);
  if ((rc = yk_rerank_shortlist(n, d, k, db, dv, di, dd)) != YK_OK) return rc;
  if ((rc = yk_memcpy_d2h(idx, di, sizeof(int) * (size_t)n * k)) != YK_OK) return rc;
  return yk_memcpy_d2h(dis, dd, sizeof(float) * (size_t)n * k);
}

int yael_knn_reorder_shortlist_checked(int n, int nb, int d, int k, const float *b,
                                       const float *v, int *idx, float *dis) {
  yh_lock();
  int rc = yh_shortlist(n, nb, d, k, b, v, idx, dis);
  yh_unlock();
  return rc;
}

void knn_reorder_shortlist(int n, int nb, int d, int k, const float *b, const float *v,
                           int *idx, float *dis) {
  if (yael_knn_reorder_shortlist_checked(n, nb, d, k, b, v, idx, dis) != YK_OK)
    yh_die("knn_reorder_shortlist");
}

static int yh_recompute(int nq, int nb, int d, int k, const float *b, const float *v, int label0,
                        int *kp, const int *idx, float *dis) {
  if (nq <= 0 || k <= 0) return YK_OK;
  for (int q = 0; q < nq; q++)
    if (kp[q] < 0) return yk_set_error_c(YK_EINVAL, "kp[%d] = %d < 0", q, kp[q]);
  int rc = yk_init();
  if (rc != YK_OK) return rc;
  const int ld = yk_ld(d);
  float *db = (float *)yh_buf(YH_B, sizeof(float) * (size_t)(nb > 0 ? nb : 1) * ld);
  float *dv = (float *)yh_buf(YH_Q, sizeof(float) * (size_t)nq * ld);
  int *di = (int *)yh_buf(YH_OUTI, sizeof(int) * ((size_t)nq * k + (size_t)nq + 64));
  float *dd = (float *)yh_buf(YH_OUTD, sizeof(float) * (size_t)nq * k);
  if (!db || !dv || !di || !dd) return YK_ENOMEM;
  int *dkp = di + (size_t)nq * k, *dbad = dkp + nq;
  const int none = INT_MAX;
  if (nb > 0 && (rc = yk_upload_vectors(db, b, nb, d)) != YK_OK) return rc;
  if ((rc = yk_upload_vectors(dv, v, nq, d)) != YK_OK) return rc;
  if ((rc = yk_memcpy_h2d(di, idx, sizeof(int) * (size_t)nq * k)) != YK_OK) return rc;
  if ((rc = yk_memcpy_h2d(dkp, kp, sizeof(int) * (size_t)nq)) != YK_OK) return rc;
  if ((rc = yk_memcpy_h2d(dbad, &none, sizeof(int))) != YK_OK) return rc;
  if ((rc = yk_memcpy_h2d(dd, dis, sizeof(float) * (size_t)nq * k)) != YK_OK) return rc;
  if ((rc = yk_recompute_exact_dists(nq, nb, d, k, db, dv, label0, dkp, di, dd, dbad)) != YK_OK)
    return rc;
  int bad = none;
  if ((rc = yk_memcpy_d2h(&bad, dbad, sizeof(int))) != YK_OK) return rc;
  if (bad != none)
    return yk_set_error_c(YK_EINVAL, "query %d: a shortlist entry lies below label0=%d", bad,
                          label0);
  if ((rc = yk_memcpy_d2h(kp, dkp, sizeof(int) * (size_t)nq)) != YK_OK) return rc;
  return yk_memcpy_d2h(dis, dd, sizeof(float) * (size_t)nq * k);
}

int yael_knn_recompute_exact_dists_checked(int nq, int nb, int d, int k, const float *b,
                                           const float *v, int label0, int *kp, const int *idx,
                                           float *dis) {
  yh_lock();
  int rc = yh_recompute(nq, nb, d, k, b, v, label0, kp, idx, dis);
  yh_unlock();
  return rc;
}

void knn_recompute_exact_dists(int nq, int nb, int d, int k, const float *b, const float *v,
                               int label0, int *kp, const int *idx, float *dis) {
  if (yael_knn_recompute_exact_dists_checked(nq, nb, d, k, b, v, label0, kp, idx, dis) != YK_OK)
    yh_die("knn_recompute_exact_dists");
}

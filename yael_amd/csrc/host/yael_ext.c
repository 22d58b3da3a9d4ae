/*
 * yael_ext.c -- knn_reorder_shortlist (nn.h:102-104, nn.c:528-580) on the GPU.
 */
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

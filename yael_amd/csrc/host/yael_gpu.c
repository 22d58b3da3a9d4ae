/*
 * yael_gpu.c -- yael's C API for the k-means / exhaustive-kNN path, served by
 * the HIP layer (yk.h).  Same symbols, signatures, argument order and layout
 * as yael's nn.h / kmeans.h; host pointers in and out.
 *
 * Conventions kept from the reference:
 *   - vectors are column-major d x n fvecs (row-major n x d in C terms);
 *   - misuse aborts with a message, like the reference's assert()/abort();
 *   - kmeans returns -1 when it abandons after too many empty-cluster splits.
 * The n_thread / flags&0xffff thread counts are accepted and ignored: the
 * GPU grid replaces yael's OpenMP query slices (nn.c:665-699).
 * Calls are serialised by one mutex (yael's functions are re-entrant, and
 * its SWIG layer releases the GIL, so concurrent callers must be safe).
 */
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../yk.h"
#include "yael_host.h"

static pthread_mutex_t g_mu = PTHREAD_MUTEX_INITIALIZER;

void yh_lock(void) { pthread_mutex_lock(&g_mu); }
void yh_unlock(void) { pthread_mutex_unlock(&g_mu); }

void yh_die(const char *where) {
  fprintf(stderr, "yael_mi355x: %s: %s\n", where, yk_last_error());
  abort();
}

/* grow-only device buffers, one per role */
static void *g_slot[YH_NSLOTS];
static size_t g_slot_bytes[YH_NSLOTS];

void *yh_buf(int slot, size_t bytes) {
  if (bytes == 0) bytes = 256;
  if (bytes > g_slot_bytes[slot]) {
    if (g_slot[slot]) yk_free(g_slot[slot]);
    g_slot[slot] = NULL;
    g_slot_bytes[slot] = 0;
    void *p = NULL;
    if (yk_malloc(&p, bytes) != YK_OK) return NULL;
    g_slot[slot] = p;
    g_slot_bytes[slot] = bytes;
  }
  return g_slot[slot];
}

void yh_release_all(void) {
  for (int i = 0; i < YH_NSLOTS; i++) {
    if (g_slot[i]) yk_free(g_slot[i]);
    g_slot[i] = NULL;
    g_slot_bytes[i] = 0;
  }
}

static long yh_pad32(long n) { return (n + 31) & ~31L; }

/* Upload n vectors of dimension d with host row stride lds into the padded
 * device layout of slot; returns the device pointer. */
static float *yh_upload(int slot, const float *src, long n, int d, long lds) {
  int ld = yk_ld(d);
  float *dst = (float *)yh_buf(slot, sizeof(float) * (size_t)(n > 0 ? n : 1) * ld);
  if (!dst) return NULL;
  if (n <= 0) return dst;
  if (lds == d) {
    if (yk_upload_vectors(dst, src, n, d) != YK_OK) return NULL;
  } else {
    /* strided source: pack on the host first */
    float *tmp = (float *)malloc(sizeof(float) * (size_t)n * d);
    if (!tmp) return NULL;
    for (long i = 0; i < n; i++) memcpy(tmp + i * d, src + i * lds, sizeof(float) * d);
    int rc = yk_upload_vectors(dst, tmp, n, d);
    free(tmp);
    if (rc != YK_OK) return NULL;
  }
  return dst;
}

/* ------------------------------------------------------------------------ */
/* kNN                                                                       */

/* kNN form: 0 = automatic (the precision form for large unweighted searches
 * on finite data), 1 = FP32 form only, 2 = the precision form whenever it
 * applies (unweighted, finite data, k <= YK_KMAX / 2). */
static int g_knn_precision = 0;

int yael_knn_precision(int mode) {
  const int old = g_knn_precision;
  if (mode >= 0 && mode <= 2) g_knn_precision = mode;
  return old;
}

/* The precision form (BF16 first pass, FP32 re-rank, certificate,
 * escalation, FP32 fallback) gives the FP32 form's results bit for bit; it
 * is chosen where its fixed costs (BF16 copies, sketch) are small against
 * the scan: at least 256 queries and 2^26 pairs. */
#define YH_KNN_BF16_MIN_NQ 256
#define YH_KNN_BF16_MIN_PAIRS (1L << 26)

static int yh_knn_bf16(int type, int nq, int nb, int d, int k, float *db, float *dq, float *bn,
                       double *qn, int *da, float *dd, int *used) {
  *used = 0;
  uint16_t *bh = (uint16_t *)yh_buf(YH_BH, (size_t)yk_bf16_bytes(nb, d, type, 0));
  uint16_t *qh = (uint16_t *)yh_buf(YH_QH, (size_t)yk_bf16_bytes(nq, d, type, 1));
  float *bst = (float *)yh_buf(YH_BST, 16 * (size_t)nb);
  float *qst = (float *)yh_buf(YH_QST, 16 * (size_t)nq);
  float *mx = (float *)yh_buf(YH_MX, 64);
  if (!bh || !qh || !bst || !qst || !mx) return YK_ENOMEM;
  /* the base's conversion also writes its float norms */
  int rc = yk_to_bf16(db, nb, d, type, 0, bh, bst, mx, type == 2 ? bn : NULL);
  if (rc == YK_OK) rc = yk_to_bf16(dq, nq, d, type, 1, qh, qst, mx + 4, NULL);
  float hm[8];
  if (rc == YK_OK) rc = yk_memcpy_d2h(hm, mx, sizeof hm);
  if (rc != YK_OK) return rc;
  for (int i = 0; i < 8; i++)
    if (!isfinite(hm[i])) return YK_OK; /* non-finite values: the FP32 form */
  if (type == 2 && (rc = yk_norms(dq, nq, d, NULL, qn)) != YK_OK) return rc;
  int kp = 2 * k < 32 ? 32 : 2 * k;
  if (kp > YK_KMAX) kp = YK_KMAX;
  if (kp > nb) kp = nb;
  int nfb = 0;
  rc = yk_knn_bf16(type, nq, nb, d, k, kp, dq, qn, qh, qst, db, bn, bh, mx, da, dd, &nfb);
  if (rc == YK_OK) *used = 1;
  return rc;
}

/* nn.c:451-525 semantics on the GPU; assign/dis are k x nq. */
int yh_knn(int type, int nq, int nb, int d, int k, const float *b, const float *q,
           const float *w, int *assign, float *dis) {
  if (type == 12) type = 2;
  if (type != 2 && type != 16) {
    return yk_set_error_c(YK_EINVAL, "distance type %d has no GPU path (2, 12, 16 only)", type);
  }
  if (k > nb) return yk_set_error_c(YK_EINVAL, "k=%d > nb=%d (nn.c:456)", k, nb);
  if (nq <= 0) return YK_OK;
  if (yk_init() != YK_OK) return YK_ENODEV;
  float *db = yh_upload(YH_B, b, nb, d, d);
  if (!db) return YK_ENOMEM;
  float *dq = yh_upload(YH_Q, q, nq, d, d);
  if (!dq) return YK_ENOMEM;
  float *bn = NULL;
  double *qn = NULL;
  if (type == 2) {
    bn = (float *)yh_buf(YH_BN, sizeof(float) * yh_pad32(nb));
    qn = (double *)yh_buf(YH_QN, sizeof(double) * nq);
    if (!bn || !qn) return YK_ENOMEM;
  }
  const int want_bf16 =
      !w && k <= YK_KMAX / 2 &&
      (g_knn_precision == 2 ||
       (g_knn_precision == 0 && nq >= YH_KNN_BF16_MIN_NQ && (long)nq * nb >= YH_KNN_BF16_MIN_PAIRS));
  if (want_bf16) {
    int *da = (int *)yh_buf(YH_OUTI, sizeof(int) * (size_t)nq * k);
    float *dd = (float *)yh_buf(YH_OUTD, sizeof(float) * (size_t)nq * k);
    if (!da || !dd) return YK_ENOMEM;
    int used = 0;
    int rc = yh_knn_bf16(type, nq, nb, d, k, db, dq, bn, qn, da, dd, &used);
    if (rc != YK_OK) return rc;
    if (used) {
      rc = yk_memcpy_d2h(assign, da, sizeof(int) * (size_t)nq * k);
      if (rc == YK_OK) rc = yk_memcpy_d2h(dis, dd, sizeof(float) * (size_t)nq * k);
      return rc;
    }
  }
  if (type == 2) {
    int rc = yk_norms(db, nb, d, bn, NULL);
    if (rc == YK_OK) rc = yk_norms(dq, nq, d, NULL, qn);
    if (rc != YK_OK) return rc;
  }
  float *dw = NULL;
  if (w) {
    dw = (float *)yh_buf(YH_W, sizeof(float) * yh_pad32(nb));
    if (!dw) return YK_ENOMEM;
    int rc = yk_memcpy_h2d(dw, w, sizeof(float) * nb);
    if (rc != YK_OK) return rc;
  }
  int *da = (int *)yh_buf(YH_OUTI, sizeof(int) * (size_t)nq * k);
  float *dd = (float *)yh_buf(YH_OUTD, sizeof(float) * (size_t)nq * k);
  if (!da || !dd) return YK_ENOMEM;
  int rc = yk_knn_f32(type, nq, nb, d, k, dq, qn, db, bn, dw, da, dd);
  if (rc != YK_OK) return rc;
  rc = yk_memcpy_d2h(assign, da, sizeof(int) * (size_t)nq * k);
  if (rc == YK_OK) rc = yk_memcpy_d2h(dis, dd, sizeof(float) * (size_t)nq * k);
  return rc;
}

/* nn.h:41-45 */
void knn_full(int distance_type, int nq, int nb, int d, int k, const float *b,
              const float *q, const float *b_weights, int *assign, float *dis) {
  yh_lock();
  int rc = yh_knn(distance_type, nq, nb, d, k, b, q, b_weights, assign, dis);
  yh_unlock();
  if (rc != YK_OK) yh_die("knn_full");
}

/* nn.h:49-54 */
void knn_full_thread(int distance_type, int nq, int nb, int d, int k, const float *b,
                     const float *q, const float *b_weights, int *assign, float *dis,
                     int n_thread) {
  (void)n_thread;
  knn_full(distance_type, nq, nb, d, k, b, q, b_weights, assign, dis);
}

/* nn.c:608-621 */
double nn(int npt, int nclust, int d, const float *codebook, const float *coords, int *vw) {
  float *dis = (float *)malloc(sizeof(float) * (npt > 0 ? npt : 1));
  knn_full(2, npt, nclust, d, 1, codebook, coords, NULL, vw, dis);
  double s = 0;
  for (long i = 0; i < npt; i++) s += dis[i];
  free(dis);
  return s;
}

/* nn.c:715-726 */
double nn_thread(int npt, int nclust, int d, const float *codebook, const float *coords,
                 int *vw, int n_thread) {
  (void)n_thread;
  return nn(npt, nclust, d, codebook, coords, vw);
}

/* nn.c:624-632 */
float *knn(int npt, int nclust, int d, int k, const float *codebook, const float *coords,
           int *vw) {
  float *dis = (float *)malloc(sizeof(float) * ((size_t)npt * k > 0 ? (size_t)npt * k : 1));
  knn_full(2, npt, nclust, d, k, codebook, coords, NULL, vw, dis);
  return dis;
}

/* nn.c:704-711 */
float *knn_thread(int npt, int nclust, int d, int k, const float *codebook,
                  const float *coords, int *vw, int n_thread) {
  (void)n_thread;
  return knn(npt, nclust, d, k, codebook, coords, vw);
}

/* ------------------------------------------------------------------------ */
/* cross distances                                                           */

int yh_cross(int type, int d, int na, int nb, const float *a, int lda, const float *b,
             int ldb, float *dist2, int ldd) {
  if (type == 12) type = 2;
  if (type != 2 && type != 16)
    return yk_set_error_c(YK_EINVAL, "distance type %d has no GPU path (2, 12, 16 only)", type);
  if (na <= 0 || nb <= 0) return YK_OK;
  if (yk_init() != YK_OK) return YK_ENODEV;
  float *da = yh_upload(YH_B, a, na, d, lda);
  float *db = yh_upload(YH_Q, b, nb, d, ldb);
  if (!da || !db) return YK_ENOMEM;
  float *an = NULL;
  double *bnd = NULL;
  if (type == 2) {
    an = (float *)yh_buf(YH_BN, sizeof(float) * yh_pad32(na));
    bnd = (double *)yh_buf(YH_QN, sizeof(double) * nb);
    if (!an || !bnd) return YK_ENOMEM;
    int rc = yk_norms(da, na, d, an, NULL);
    if (rc == YK_OK) rc = yk_norms(db, nb, d, NULL, bnd);
    if (rc != YK_OK) return rc;
  }
  float *dd = (float *)yh_buf(YH_OUTD, sizeof(float) * (size_t)na * nb);
  if (!dd) return YK_ENOMEM;
  int rc = yk_cross_distances(type, na, nb, d, da, an, db, bnd, dd, na);
  if (rc != YK_OK) return rc;
  if (ldd == na) return yk_memcpy_d2h(dist2, dd, sizeof(float) * (size_t)na * nb);
  float *tmp = (float *)malloc(sizeof(float) * (size_t)na * nb);
  if (!tmp) return yk_set_error_c(YK_ENOMEM, "host malloc");
  rc = yk_memcpy_d2h(tmp, dd, sizeof(float) * (size_t)na * nb);
  if (rc == YK_OK)
    for (long j = 0; j < nb; j++)
      memcpy(dist2 + j * (long)ldd, tmp + j * (long)na, sizeof(float) * na);
  free(tmp);
  return rc;
}

/* nn.h:141-143 */
void compute_cross_distances(int d, int na, int nb, const float *a, const float *b,
                             float *dist2) {
  yh_lock();
  int rc = yh_cross(2, d, na, nb, a, d, b, d, dist2, na);
  yh_unlock();
  if (rc != YK_OK) yh_die("compute_cross_distances");
}

/* nn.h:151-154 */
void compute_cross_distances_nonpacked(int d, int na, int nb, const float *a, int lda,
                                       const float *b, int ldb, float *dist2, int ldd) {
  yh_lock();
  int rc = yh_cross(2, d, na, nb, a, lda, b, ldb, dist2, ldd);
  yh_unlock();
  if (rc != YK_OK) yh_die("compute_cross_distances_nonpacked");
}

/* nn.h:156-159 */
void compute_cross_distances_thread(int d, int na, int nb, const float *a, const float *b,
                                    float *dist2, int nt) {
  (void)nt;
  compute_cross_distances(d, na, nb, a, b, dist2);
}

/* nn.h:174-176 */
void compute_cross_distances_alt(int distance_type, int d, int na, int nb, const float *a,
                                 const float *b, float *dist2) {
  yh_lock();
  int rc = yh_cross(distance_type, d, na, nb, a, d, b, d, dist2, na);
  yh_unlock();
  if (rc != YK_OK) yh_die("compute_cross_distances_alt");
}

/* nn.h:180-183 */
void compute_cross_distances_alt_nonpacked(int distance_type, int d, int na, int nb,
                                           const float *a, int lda, const float *b, int ldb,
                                           float *dist2, int ldd) {
  yh_lock();
  int rc = yh_cross(distance_type, d, na, nb, a, lda, b, ldb, dist2, ldd);
  yh_unlock();
  if (rc != YK_OK) yh_die("compute_cross_distances_alt_nonpacked");
}

/* nn.h:188-191 */
void compute_cross_distances_alt_thread(int distance_type, int d, int na, int nb,
                                        const float *a, const float *b, float *dist2,
                                        int nt) {
  (void)nt;
  compute_cross_distances_alt(distance_type, d, na, nb, a, b, dist2);
}

/* ------------------------------------------------------------------------ */
/* error-returning variants of the entry points above (same arguments), for
 * callers that want a status code instead of an abort (the Python facade). */

int yael_knn_full_checked(int distance_type, int nq, int nb, int d, int k, const float *b,
                          const float *q, const float *b_weights, int *assign, float *dis) {
  yh_lock();
  int rc = yh_knn(distance_type, nq, nb, d, k, b, q, b_weights, assign, dis);
  yh_unlock();
  return rc;
}

int yael_cross_distances_checked(int distance_type, int d, int na, int nb, const float *a,
                                 int lda, const float *b, int ldb, float *dist2, int ldd) {
  yh_lock();
  int rc = yh_cross(distance_type, d, na, nb, a, lda, b, ldb, dist2, ldd);
  yh_unlock();
  return rc;
}

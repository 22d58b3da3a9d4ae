/*
 * yael_kmeans.c -- kmeans() (kmeans.h:41-44) on the GPU.
 *
 * The Lloyd loop runs on the device: norms of the centroids, the fused
 * FP32 distance + argmin kernel (k = 1) as the assign step, the FP64
 * segmented update and the FP64 objective.  The host keeps what is O(k) and
 * sequential in the reference: the seed chain (kmeans.c:379-409), the random
 * initialisation (kmeans.c:15-20), the empty-cluster split with its rand_r /
 * gaussrand_r streams (kmeans.c:166-209), the abandon and stop rules
 * (kmeans.c:300-313) and the best-of-redo bookkeeping (kmeans.c:416-428).
 * Per iteration the host reads back k counts and one double; the centroid
 * matrix crosses the bus only in an iteration that splits a cluster.
 *
 * Assign precision: for large k the assign step takes the precision form
 * (yk_knn_bf16 with k = 1, k' = 8): a BF16 MFMA pass keeps each point's 8
 * best centroids, the FP32 form's own arithmetic re-ranks them and a bound
 * on the BF16 error certifies that the FP32 form's nearest centroid is among
 * them; uncertified points are searched again with the FP32 form.  The
 * result is the FP32 form's, bit for bit, either way.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../yk.h"
#include "yael_host.h"

/* kmeans.c:22-24 */
double yh_drand_r(unsigned *seed) { return rand_r(seed) / ((double)RAND_MAX + 1.0); }

/* vector.c:139-154: ratio-of-uniforms normal deviate */
double yh_gaussrand_r(unsigned *seed) {
  for (;;) {
    float u1 = (float)yh_drand_r(seed);
    float u2 = (float)yh_drand_r(seed);
    double z = 1.71552776992141 * (u1 - .5) / u2;
    float zz = (float)(z * z / 4.0);
    if (zz < -log(u2)) return z;
  }
}

/* L1 norm (double sum of |float|) then scale by the double reciprocal,
 * rounding every element to float (vector.c:2016-2025, 2180-2189). */
static void yh_normalize_l1(float *p, int k) {
  double s = 0;
  for (int i = 0; i < k; i++) s += fabs(p[i]);
  double inv = 1. / s;
  for (int i = 0; i < k; i++) p[i] = (float)(p[i] * inv);
}

int yh_reassign_empty(int d, int k, float *cent, long ldc, const int *nassign, unsigned seed) {
  int nre = 0;
  float *p = (float *)malloc(sizeof(float) * k);
  float *eps = (float *)malloc(sizeof(float) * (d > 0 ? d : 1));
  for (int c = 0; c < k; c++) {
    /* nassign^2 - 1 in 32-bit two's complement, as the reference's int
       arithmetic computes it (kmeans.c:175) */
    unsigned u = (unsigned)nassign[c];
    p[c] = nassign[c] < 2 ? 0.f : (float)(int)(u * u - 1u);
  }
  yh_normalize_l1(p, k);
  for (int c = 0; c < k; c++) {
    if (nassign[c] != 0) continue;
    nre++;
    double rd = yh_drand_r(&seed);
    int j = 0;
    while (j < k - 1) {
      rd -= p[j];
      if (rd < 0) break;
      j++;
    }
    float *cc = cent + (long)c * ldc, *cj = cent + (long)j * ldc;
    memmove(cc, cj, sizeof(float) * d);
    double nr = 0;
    for (int l = 0; l < d; l++) nr += cj[l] * cj[l]; /* float squares, double sum */
    double s = sqrt(nr) * 0.0000001;
    unsigned gs = (unsigned)rand_r(&seed);
    for (int l = 0; l < d; l++) eps[l] = (float)yh_gaussrand_r(&gs);
    for (int l = 0; l < d; l++) eps[l] = (float)(eps[l] * s);
    for (int l = 0; l < d; l++) cj[l] += eps[l];
    for (int l = 0; l < d; l++) cc[l] -= eps[l];
    p[j] = 0;
    yh_normalize_l1(p, k);
  }
  free(eps);
  free(p);
  return nre;
}

/* The first k entries of the Fisher-Yates shuffle of vector.c:226-253 are
 * final after step k-1 (later steps only touch slots >= k), so k draws give
 * the same selection as the reference's n-1 draws. */
void yh_random_init(int n, int k, int *sel, unsigned seed) {
  int *idx = (int *)malloc(sizeof(int) * (size_t)n);
  for (int i = 0; i < n; i++) idx[i] = i;
  int steps = k < n - 1 ? k : n - 1;
  for (int i = 0; i < steps; i++) {
    int j = i + rand_r(&seed) % (n - i);
    int t = idx[i];
    idx[i] = idx[j];
    idx[j] = t;
  }
  memcpy(sel, idx, sizeof(int) * k);
  free(idx);
}

/* k-means++ seeding (kmeans.c:27-82) among the first nsub points of the
 * device set v (double norms vn).  Per new centre the GPU updates every
 * point's distance to its nearest chosen centre (yk_kpp_update); the two
 * passes that turn those distances into the next draw stay on the host, in
 * the reference's order: the L1 normalisation (a sequential double sum, then
 * each value scaled by the double reciprocal, vector.c:2016-2026, 1792-1797)
 * and the walk of rd down the normalised values -- their rounding depends on
 * that order, which a parallel reduction or scan would not keep. */
static int yh_kmeanspp_init(const float *v, const double *vn, int d, int nsub, int k, int *sel,
                            unsigned seed) {
  float *best = NULL, *p = NULL;
  int rc = yk_host_alloc((void **)&p, sizeof(float) * (size_t)nsub);
  if (rc != YK_OK) return rc;
  if ((rc = yk_malloc((void **)&best, sizeof(float) * (size_t)nsub)) != YK_OK) {
    yk_host_free(p);
    return rc;
  }
  sel[0] = rand_r(&seed) % k;
  for (int i = 1; i < k && rc == YK_OK; i++) {
    /* the step kernel also writes the distances into the page-locked p */
    rc = yk_kpp_update(v, nsub, d, vn, sel[i - 1], best, i == 1, p);
    if (rc == YK_OK) rc = yk_sync();
    if (rc != YK_OK) break;
    double nr = 0;
    for (long j = 0; j < nsub; j++) nr += fabs(p[j]);
    const double scal = 1. / nr;
    double rd = rand_r(&seed) / ((double)RAND_MAX + 1.0);
    long j;
    /* each value scaled as the normalisation stores it, only as far as the
     * walk goes */
    for (j = 0; j < nsub - 1; j++) {
      rd -= (float)(p[j] * scal);
      if (rd < 0) break;
    }
    sel[i] = (int)j;
  }
  yk_free(best);
  yk_host_free(p);
  return rc;
}

typedef struct {
  float *v, *c, *cn, *dis;
  double *qn, *obj;
  int *assign, *counts, *sel;
  void *ws;
  /* precision form (bf16 != 0): BF16 copies and their statistics */
  int bf16;
  uint16_t *vh, *ch;
  float *vs, *cs, *csmax;
} yh_km_dev;

/* counters of the last kmeans call (yael_kmeans_last_stats) */
static long g_km_iters, g_km_assigned, g_km_fallback;

static int yh_km_alloc_bf16(yh_km_dev *m, long n, int k, int d) {
  const int ldh = yk_bf16_ld(d, 2);
  long kp = (k + 31) & ~31;
  m->vh = (uint16_t *)yh_buf(YH_KM_VH, sizeof(uint16_t) * (size_t)n * ldh);
  m->vs = (float *)yh_buf(YH_KM_QS, sizeof(float) * 4 * (size_t)n);
  m->ch = (uint16_t *)yh_buf(YH_KM_CH, sizeof(uint16_t) * (size_t)kp * ldh);
  m->cs = (float *)yh_buf(YH_KM_CS, sizeof(float) * 4 * (size_t)(kp + 1));
  if (!m->vh || !m->vs || !m->ch || !m->cs) return YK_ENOMEM;
  m->csmax = m->cs + 4 * kp;
  return YK_OK;
}

static int yh_km_alloc(yh_km_dev *m, long n, int k, int d, int v_on_device) {
  int ld = yk_ld(d);
  long kp = (k + 31) & ~31;
  m->v = v_on_device ? (float *)1 : (float *)yh_buf(YH_KM_V, sizeof(float) * (size_t)n * ld);
  m->sel = (int *)yh_buf(YH_KM_SEL, sizeof(int) * kp);
  m->qn = (double *)yh_buf(YH_KM_QN, sizeof(double) * n);
  m->c = (float *)yh_buf(YH_KM_C, sizeof(float) * (size_t)kp * ld);
  m->cn = (float *)yh_buf(YH_KM_CN, sizeof(float) * kp);
  m->assign = (int *)yh_buf(YH_KM_A, sizeof(int) * n);
  m->dis = (float *)yh_buf(YH_KM_D, sizeof(float) * n);
  m->counts = (int *)yh_buf(YH_KM_CNT, sizeof(int) * kp + sizeof(double) * 1040);
  m->ws = yh_buf(YH_KM_WS, yk_kmeans_workspace_size(n, k, d));
  if (!m->v || !m->sel || !m->qn || !m->c || !m->cn || !m->assign || !m->dis || !m->counts ||
      !m->ws)
    return YK_ENOMEM;
  m->obj = (double *)((char *)m->counts + sizeof(int) * kp);
  return YK_OK;
}

/* kmeans.c:213-329 on the device. Returns 0, or YK_EABANDON. */
static int yh_kmeans_core(int d, int n, int k, int niter, int flags, int verbose,
                          yh_km_dev *m, unsigned seed, int *hcounts, float *hcent,
                          double *qerr_out, long *iter_tot) {
  const int ld = yk_ld(d);
  double qerr = HUGE_VAL, qerr_old;
  int tot = 0;
  int rc;
  for (int iter = 1; iter <= niter; iter++) {
    (*iter_tot)++;
    if (!m->bf16 && (rc = yk_norms(m->c, k, d, m->cn, NULL)) != YK_OK) return rc;
    if (m->bf16) {
      int nfb = 0;
      /* the conversion also writes the centroids' float norms */
      if ((rc = yk_to_bf16(m->c, k, d, 2, 0, m->ch, m->cs, m->csmax, m->cn)) != YK_OK) return rc;
      if ((rc = yk_knn_bf16(2, n, k, d, 1, YH_KM_KPRIME, m->v, m->qn, m->vh, m->vs, m->c, m->cn,
                            m->ch, m->csmax, m->assign, m->dis, &nfb)) != YK_OK)
        return rc;
      g_km_fallback += nfb;
    } else if ((rc = yk_knn_f32(2, n, k, d, 1, m->v, m->qn, m->c, m->cn, NULL, m->assign,
                                m->dis)) != YK_OK) {
      return rc;
    }
    g_km_assigned += n;
    g_km_iters++;
    if ((rc = yk_kmeans_update_async(n, k, d, m->v, m->assign, m->c, m->counts,
                                     (flags & KMEANS_NORMALIZE_CENTS) != 0, m->ws)) != YK_OK)
      return rc;
    /* the counts, this assign step's objective (kmeans.c:300-305: the
       distances to the centroids before the update) and the update's check,
       one wait */
    qerr_old = qerr;
    if ((rc = yk_kmeans_readback_checked(m->counts, k, hcounts, m->dis, n, &qerr, m->obj,
                                         m->ws)) != YK_OK)
      return rc;
    int nempty = 0;
    for (int c = 0; c < k; c++) nempty += hcounts[c] == 0;
    unsigned split_seed = (unsigned)rand_r(&seed);
    int nre = 0;
    if (nempty) {
      if ((rc = yk_memcpy_d2h(hcent, m->c, sizeof(float) * (size_t)k * ld)) != YK_OK) return rc;
      nre = yh_reassign_empty(d, k, hcent, ld, hcounts, split_seed);
      if ((rc = yk_memcpy_h2d(m->c, hcent, sizeof(float) * (size_t)k * ld)) != YK_OK) return rc;
      if (verbose) fprintf(stderr, "# kmeans warning: %d empty clusters -> split\n", nre);
    }
    tot += nre;
    if (tot > n / 100 && tot > 1000) {
      fprintf(stderr, "# kmeans: reassigned %d times, abandoning\n", tot);
      return YK_EABANDON;
    }
    if (qerr_old == qerr && nre == 0) break;
    if (verbose) {
      printf(" -> %.3f", qerr / n);
      fflush(stdout);
    }
  }
  if (verbose) printf("\n");
  *qerr_out = qerr;
  return YK_OK;
}

int yh_kmeans(int d, int n, int k, int niter, const float *v, int v_on_device, int flags,
              long seed_in, int redo, float *cent_out, float *dis_out, int *assign_out,
              int *nassign_out, float *ret) {
  const int verbose = !(flags & KMEANS_QUIET);
  niter = niter == 0 ? 1000000 : niter;
  const int user = (flags & KMEANS_INIT_USER) != 0;
  if (flags & (KMEANS_L1 | KMEANS_CHI2))
    return yk_set_error_c(YK_EINVAL, "KMEANS_L1/KMEANS_CHI2 have no GPU path");
  if (user && !cent_out) return yk_set_error_c(YK_EINVAL, "KMEANS_INIT_USER needs centroids");
  if (user) redo = 1;
  if (k > n || k < 1 || d < 1)
    return yk_set_error_c(YK_EINVAL, "kmeans: need 1 <= k <= n (k=%d n=%d)", k, n);
  int rc = yk_init();
  if (rc != YK_OK) return rc;
  const int ld = yk_ld(d);
  yh_km_dev m = {0};
  if ((rc = yh_km_alloc(&m, n, k, d, v_on_device)) != YK_OK) return rc;
  if (v_on_device)
    m.v = (float *)v; /* caller's padded device vectors */
  else if ((rc = yk_upload_vectors(m.v, v, n, d)) != YK_OK)
    return rc;
  if ((rc = yk_norms(m.v, n, d, NULL, m.qn)) != YK_OK) return rc;
  g_km_iters = g_km_assigned = g_km_fallback = 0;
  m.bf16 = (flags & YAEL_KMEANS_BF16) ||
           (!(flags & YAEL_KMEANS_FP32) && k >= YH_KM_BF16_MIN_K);
  if (m.bf16) {
    /* the points' BF16 copy is made once per call */
    if ((rc = yh_km_alloc_bf16(&m, n, k, d)) != YK_OK) return rc;
    if ((rc = yk_to_bf16(m.v, n, d, 2, 1, m.vh, m.vs, NULL, NULL)) != YK_OK) return rc;
  }

  int *hcounts = (int *)malloc(sizeof(int) * k);
  float *hcent = (float *)malloc(sizeof(float) * (size_t)k * ld);
  int *sel = (int *)malloc(sizeof(int) * k);
  if (!hcounts || !hcent || !sel) {
    free(hcounts); free(hcent); free(sel);
    return yk_set_error_c(YK_ENOMEM, "host malloc");
  }
  if (seed_in == 0) seed_in = lrand48(); /* kmeans.c:379-380 */
  unsigned seed = (unsigned)seed_in;
  double qerr = 0, best = HUGE_VAL;
  long iter_tot = 0;
  int core = YK_OK;
  for (int run = 0; run < redo; run++) {
    if (verbose) printf("<><><><> kmeans / run %d <><><><><>\n", run);
    if (user) {
      if ((rc = yk_upload_vectors(m.c, cent_out, k, d)) != YK_OK) break;
    } else {
      /* the selected points are copied row by row on the device */
      if (flags & KMEANS_INIT_BERKELEY) {
        /* kmeans.c:393-400: at most 8k points when n > 8k and n > 8192 */
        const int nsub = (n > k * 8 && n > 8192) ? k * 8 : n;
        if ((rc = yh_kmeanspp_init(m.v, m.qn, d, nsub, k, sel, (unsigned)rand_r(&seed))) != YK_OK)
          break;
      } else {
        yh_random_init(n, k, sel, (unsigned)rand_r(&seed));
      }
      if ((rc = yk_memcpy_h2d(m.sel, sel, sizeof(int) * k)) != YK_OK) break;
      if ((rc = yk_gather_rows(m.v, m.sel, k, d, m.c)) != YK_OK) break;
    }
    core = yh_kmeans_core(d, n, k, niter, flags, verbose, &m, (unsigned)rand_r(&seed), hcounts,
                          hcent, &qerr, &iter_tot);
    if (core != YK_OK) break;
    if (qerr < best) {
      best = qerr;
      if (cent_out && (rc = yk_download_vectors(cent_out, m.c, k, d)) != YK_OK) break;
      if (dis_out && (rc = yk_memcpy_d2h(dis_out, m.dis, sizeof(float) * n)) != YK_OK) break;
      if (assign_out && (rc = yk_memcpy_d2h(assign_out, m.assign, sizeof(int) * n)) != YK_OK)
        break;
      if (nassign_out) memcpy(nassign_out, hcounts, sizeof(int) * k);
    }
  }
  if (verbose && core == YK_OK && rc == YK_OK) {
    printf("Total number of iterations: %ld\n", iter_tot);
  }
  free(sel);
  free(hcent);
  free(hcounts);
  if (rc != YK_OK) return rc;
  if (core == YK_EABANDON) {
    *ret = -1;
    return YK_OK;
  }
  if (core != YK_OK) return core;
  *ret = (float)(best / n);
  return YK_OK;
}

/* kmeans.h:41-44 */
float kmeans(int d, int n, int k, int niter, const float *v, int flags, long seed, int redo,
             float *centroids, float *dis, int *assign, int *nassign) {
  float ret = -1;
  yh_lock();
  int rc = yh_kmeans(d, n, k, niter, v, 0, flags, seed, redo, centroids, dis, assign, nassign,
                     &ret);
  yh_unlock();
  if (rc != YK_OK) yh_die("kmeans");
  return ret;
}

/* kmeans.c:452-482 */
float *clustering_kmeans_assign_with_score(int n, int d, const float *points, int k,
                                           int nb_iter_max, double normalize, int n_thread,
                                           double *score, int **clust_assign_out) {
  (void)normalize;
  (void)score;
  float *cent = (float *)malloc(sizeof(float) * (size_t)k * d);
  int *ca = clust_assign_out ? (int *)malloc(sizeof(int) * (size_t)n) : NULL;
  int nredo = 1;
  if (nb_iter_max / 100000 != 0) {
    nredo = nb_iter_max / 100000;
    nb_iter_max = nb_iter_max % 100000;
  }
  float r = kmeans(d, n, k, nb_iter_max, points, n_thread | KMEANS_INIT_RANDOM, 0, nredo, cent,
                   NULL, ca, NULL);
  if (r >= 0) {
    if (clust_assign_out) *clust_assign_out = ca;
    return cent;
  }
  free(cent);
  free(ca);
  if (clust_assign_out) *clust_assign_out = NULL;
  return NULL;
}

/* kmeans.c:484-490 */
float *clustering_kmeans_assign(int n, int d, const float *points, int k, int nb_iter_max,
                                double normalize, int **clust_assign_out) {
  return clustering_kmeans_assign_with_score(n, d, points, k, nb_iter_max, normalize, 1, NULL,
                                             clust_assign_out);
}

/* kmeans.c:492-504 */
float *clustering_kmeans(int n, int d, const float *points, int k, int nb_iter_max,
                         double normalize) {
  int *ca = NULL;
  float *c = clustering_kmeans_assign_with_score(n, d, points, k, nb_iter_max, normalize, 1,
                                                 NULL, &ca);
  free(ca);
  return c;
}

int yael_kmeans_checked(int d, int n, int k, int niter, const float *v, int flags, long seed,
                        int redo, float *centroids, float *dis, int *assign, int *nassign,
                        float *ret) {
  yh_lock();
  int rc = yh_kmeans(d, n, k, niter, v, 0, flags, seed, redo, centroids, dis, assign, nassign, ret);
  yh_unlock();
  return rc;
}

/* kmeans() on vectors already resident in HBM: v_dev holds n padded rows
 * (row stride yk_ld(d) floats, pad columns zero).  Outputs are host arrays
 * (any may be NULL), as for kmeans(). */
int yael_kmeans_device(int d, int n, int k, int niter, const float *v_dev, int flags, long seed,
                       int redo, float *centroids, float *dis, int *assign, int *nassign,
                       float *ret) {
  yh_lock();
  int rc = yh_kmeans(d, n, k, niter, v_dev, 1, flags, seed, redo, centroids, dis, assign, nassign,
                     ret);
  yh_unlock();
  return rc;
}

/* Counters of the last kmeans call on this process: Lloyd iterations run,
 * point assignments made, and assignments the precision form could not
 * certify (searched again with the FP32 form). */
int yael_kmeans_last_stats(long *iters, long *assigned, long *fallback) {
  yh_lock();
  if (iters) *iters = g_km_iters;
  if (assigned) *assigned = g_km_assigned;
  if (fallback) *fallback = g_km_fallback;
  yh_unlock();
  return 0;
}

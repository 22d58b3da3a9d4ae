/*
 * yael_oracle.c -- CPU restatement of yael's k-means / exhaustive-kNN path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
 * library in yael_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  Nothing in the product links it.
 *
 * Every function restates one reference function (cited as file:line inside
 * the yael source tree).  The numerics follow the reference exactly:
 *   - base-vector norms accumulated in float, query norms accumulated in
 *     double of float products, summed in double and rounded once to float
 *     (nn.c:108-124);
 *   - the cross term added by a GEMM with alpha=-2, beta=1 (nn.c:54-67);
 *   - argmin with strict '<' scanning ascending base index (nn.c:425-440);
 *   - k>1 selection through a max-heap with the reference's push/pop rules
 *     and a final sort by (value, heap slot) (binheap.c:85-156, 201-211,
 *     sorting.c:282-316);
 *   - centroid update as float sums in point order scaled by a double
 *     reciprocal (kmeans.c:278-288);
 *   - empty-cluster splitting driven by glibc rand_r (kmeans.c:166-209,
 *     vector.c:135-189);
 *   - k-means++ seeding with the reference's double-precision L1
 *     normalisation and prefix walk (kmeans.c:27-82, 393-400).
 *
 * The GEMM is pluggable: by default a plain loop (float accumulation in
 * ascending k, no fused multiply-add; this file is built with
 * -ffp-contract=off), or any Fortran-ABI sgemm_ handed in through
 * oy_set_sgemm() -- the tests plug the same OpenBLAS the compiled reference
 * links, which makes the restatement bit-comparable to it on any data.
 *
 * Parity pinning: see tests/test_oracle.py (golden vectors from the compiled
 * reference, closed forms, brute force on tiny inputs, S11 exactness).
 */
#include <assert.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define OY_QUIET 0x10000
#define OY_INIT_BERKELEY 0x20000
#define OY_NORMALIZE_CENTS 0x40000
#define OY_INIT_RANDOM 0x80000
#define OY_INIT_USER 0x100000
#define OY_L1 0x200000
#define OY_CHI2 0x400000

typedef int (*oy_sgemm_fn)(const char *, const char *, const int *, const int *,
                           const int *, const float *, const float *, const int *,
                           const float *, const int *, const float *, float *,
                           const int *);
typedef int (*oy_sgemv_fn)(const char *, const int *, const int *, const float *,
                           const float *, const int *, const float *, const int *,
                           const float *, float *, const int *);

static oy_sgemm_fn g_sgemm = NULL;
static oy_sgemv_fn g_sgemv = NULL;

/* Plug an external Fortran-ABI sgemm_/sgemv_ (NULL restores the plain loop). */
void oy_set_blas(void *sgemm, void *sgemv) {
  g_sgemm = (oy_sgemm_fn)sgemm;
  g_sgemv = (oy_sgemv_fn)sgemv;
}

/* C(m x n, column stride ldc) = alpha * A^T B + beta * C, with A given as m
 * columns of length kd (stride lda) and B as n columns of length kd (stride
 * ldb).  Same contract as sgemm_("T","N",...) in nn.c:64 and nn.c:273.       */
static void oy_gemm_tn(int m, int n, int kd, float alpha, const float *a, int lda,
                       const float *b, int ldb, float beta, float *c, int ldc) {
  if (g_sgemm) {
    g_sgemm("T", "N", &m, &n, &kd, &alpha, a, &lda, b, &ldb, &beta, c, &ldc);
    return;
  }
  for (long j = 0; j < n; j++) {
    const float *bj = b + j * (long)ldb;
    for (long i = 0; i < m; i++) {
      const float *ai = a + i * (long)lda;
      float acc = 0;
      for (int l = 0; l < kd; l++) acc += ai[l] * bj[l];
      float *cij = c + i + j * (long)ldc;
      if (beta == 0)
        *cij = alpha * acc;
      else
        *cij = beta * *cij + alpha * acc;
    }
  }
}

/* y(nb) = alpha * B^T x + beta * y, B given as nb columns of length d (stride
 * ldb).  Same contract as sgemv_("T",...) in nn.c:81.                        */
static void oy_gemv_t(int d, int nb, float alpha, const float *b, int ldb,
                      const float *x, float beta, float *y) {
  if (g_sgemv) {
    int one = 1;
    g_sgemv("T", &d, &nb, &alpha, b, &ldb, x, &one, &beta, y, &one);
    return;
  }
  for (long j = 0; j < nb; j++) {
    const float *bj = b + j * (long)ldb;
    float acc = 0;
    for (int l = 0; l < d; l++) acc += bj[l] * x[l];
    if (beta == 0)
      y[j] = alpha * acc;
    else
      y[j] = beta * y[j] + alpha * acc;
  }
}

/* ------------------------------------------------------------------------ */
/* distances                                                                 */

/* nn.c:100-129 */
void oy_cross_distances_nonpacked(int d, int na, int nb, const float *a, int lda,
                                  const float *b, int ldb, float *dist2, int ldd) {
  float *na2 = (float *)malloc(sizeof(float) * (na > 0 ? na : 1));
  for (long i = 0; i < na; i++) {
    const float *ai = a + i * (long)lda;
    float s = 0;
    for (int j = 0; j < d; j++) s += ai[j] * ai[j];
    na2[i] = s;
  }
  for (long i = 0; i < nb; i++) {
    const float *bi = b + i * (long)ldb;
    double s = 0;
    for (int j = 0; j < d; j++) s += bi[j] * bi[j]; /* float product, double sum */
    float *row = dist2 + i * (long)ldd;
    for (long j = 0; j < na; j++) row[j] = (float)(s + (double)na2[j]);
  }
  oy_gemm_tn(na, nb, d, -2.0f, a, lda, b, ldb, 1.0f, dist2, ldd);
  free(na2);
}

/* nn.c:92-97 */
void oy_cross_distances(int d, int na, int nb, const float *a, const float *b,
                        float *dist2) {
  oy_cross_distances_nonpacked(d, na, nb, a, d, b, d, dist2, na);
}

/* nn.c:280-350 for the types the GPU path serves: 2/12 (L2) and 16 (dot). */
void oy_cross_distances_alt_nonpacked(int type, int d, int na, int nb,
                                      const float *a, int lda, const float *b,
                                      int ldb, float *dist2, int ldd) {
  if (type == 2 || type == 12) {
    oy_cross_distances_nonpacked(d, na, nb, a, lda, b, ldb, dist2, ldd);
  } else if (type == 16) {
    oy_gemm_tn(na, nb, d, 1.0f, a, lda, b, ldb, 0.0f, dist2, ldd); /* nn.c:266-275 */
  } else {
    fprintf(stderr, "oracle: distance type %d not restated\n", type);
    abort();
  }
}

void oy_cross_distances_alt(int type, int d, int na, int nb, const float *a,
                            const float *b, float *dist2) {
  oy_cross_distances_alt_nonpacked(type, d, na, nb, a, d, b, d, dist2, na);
}

/* nn.c:132-154: one vector a against nb vectors b (double norms, sgemv). */
void oy_distances_1_nonpacked(int d, int nb, const float *a, const float *b, int ldb,
                              float *dist2) {
  double a2 = 0;
  for (int j = 0; j < d; j++) a2 += a[j] * a[j];
  for (long i = 0; i < nb; i++) {
    const float *bi = b + i * (long)ldb;
    double s = 0;
    for (int j = 0; j < d; j++) s += bi[j] * bi[j];
    dist2[i] = (float)(s + a2);
  }
  oy_gemv_t(d, nb, -2.0f, b, ldb, a, 1.0f, dist2);
}

/* ------------------------------------------------------------------------ */
/* max-heap of the k smallest (binheap.c:11-156), 1-based slots              */

typedef struct {
  float *val; /* val[1..cnt] */
  int *lab;
  int cnt, cap;
} oy_heap;

static void oy_heap_pop(oy_heap *h) { /* binheap.c:48-82 */
  float last = h->val[h->cnt];
  int i = 1;
  for (;;) {
    int l = 2 * i, r = l + 1;
    if (l > h->cnt) break;
    int child;
    if (r == h->cnt + 1 || h->val[l] > h->val[r])
      child = l;
    else
      child = r;
    if (last > h->val[child]) break;
    h->val[i] = h->val[child];
    h->lab[i] = h->lab[child];
    i = child;
  }
  h->val[i] = h->val[h->cnt];
  h->lab[i] = h->lab[h->cnt];
  h->cnt--;
}

static void oy_heap_push(oy_heap *h, int label, float v) { /* binheap.c:85-103 */
  int i = ++h->cnt;
  while (i > 1) {
    int up = i >> 1;
    if (h->val[up] >= v) break;
    h->val[i] = h->val[up];
    h->lab[i] = h->lab[up];
    i = up;
  }
  h->val[i] = v;
  h->lab[i] = label;
}

/* binheap.c:139-156 */
static void oy_heap_add_range(oy_heap *h, int n, int label0, const float *v) {
  int i = 0;
  for (; i < n && h->cnt < h->cap; i++)
    if (!isnan(v[i])) oy_heap_push(h, label0 + i, v[i]);
  float lim = h->val[1];
  for (; i < n; i++) {
    if (v[i] < lim) {
      oy_heap_pop(h);
      oy_heap_push(h, label0 + i, v[i]);
      lim = h->val[1];
    }
  }
}

/* sorting.c:282-316: indices ordered by value, ties by index. */
static __thread const float *oy_sort_tab;
static int oy_cmp_idx(const void *x, const void *y) {
  int i = *(const int *)x, j = *(const int *)y;
  float dt = oy_sort_tab[i] - oy_sort_tab[j];
  if (dt) return dt > 0 ? 1 : -1;
  return i - j;
}

void oy_sort_index(const float *tab, int n, int *perm) {
  for (int i = 0; i < n; i++) perm[i] = i;
  oy_sort_tab = tab;
  qsort(perm, n, sizeof(int), oy_cmp_idx);
}

/* binheap.c:201-211 */
static void oy_heap_sort(oy_heap *h, int *labels, float *vals) {
  int *perm = (int *)malloc(sizeof(int) * (h->cnt > 0 ? h->cnt : 1));
  oy_sort_index(h->val + 1, h->cnt, perm);
  for (int i = 0; i < h->cnt; i++) {
    labels[i] = h->lab[perm[i] + 1];
    vals[i] = h->val[perm[i] + 1];
  }
  free(perm);
}

/* ------------------------------------------------------------------------ */
/* kNN                                                                       */

#define OY_BLK 256 /* nn.c:371-372 */

static void oy_tile_dists(int type, int d, int mb, int mq, const float *b,
                          const float *q, float *dists) {
  if (type == 2)
    oy_cross_distances(d, mb, mq, b, q, dists);
  else
    oy_cross_distances_alt(type, d, mb, mq, b, q, dists);
}

/* nn.c:383-446 */
static void oy_nn_single(int type, int nq, int nb, int d, const float *b,
                         const float *q, const float *w, int *assign, float *dis) {
  int s1 = nq < OY_BLK ? nq : OY_BLK, s2 = nb < OY_BLK ? nb : OY_BLK;
  float *dists = (float *)malloc(sizeof(float) * (size_t)s1 * s2 + 1);
  for (long i1 = 0; i1 < nq; i1 += s1) {
    int m1 = (int)(nq - i1 < s1 ? nq - i1 : s1);
    for (int j = 0; j < m1; j++) {
      assign[i1 + j] = -1;
      dis[i1 + j] = 1e30f;
    }
    for (long i2 = 0; i2 < nb; i2 += s2) {
      int m2 = (int)(nb - i2 < s2 ? nb - i2 : s2);
      oy_tile_dists(type, d, m2, m1, b + i2 * d, q + i1 * d, dists);
      if (w)
        for (int j1 = 0; j1 < m1; j1++)
          for (int j2 = 0; j2 < m2; j2++) dists[j1 * m2 + j2] *= w[j2 + i2];
      for (int j1 = 0; j1 < m1; j1++) {
        const float *row = dists + (long)j1 * m2;
        int best = assign[i1 + j1];
        float bv = dis[i1 + j1];
        for (int j2 = 0; j2 < m2; j2++)
          if (row[j2] < bv) {
            best = (int)(j2 + i2);
            bv = row[j2];
          }
        assign[i1 + j1] = best;
        dis[i1 + j1] = bv;
      }
    }
  }
  free(dists);
}

/* nn.c:451-525 */
void oy_knn_full(int type, int nq, int nb, int d, int k, const float *b,
                 const float *q, const float *w, int *assign, float *dis) {
  assert(k <= nb);
  if (k == 1) {
    oy_nn_single(type, nq, nb, d, b, q, w, assign, dis);
    return;
  }
  int s1 = nq < OY_BLK ? nq : OY_BLK, s2 = nb < OY_BLK ? nb : OY_BLK;
  float *dists = (float *)malloc(sizeof(float) * (size_t)s1 * s2 + 1);
  oy_heap *heaps = (oy_heap *)malloc(sizeof(oy_heap) * (s1 > 0 ? s1 : 1));
  float *hval = (float *)malloc(sizeof(float) * (size_t)s1 * (k + 1) + 1);
  int *hlab = (int *)malloc(sizeof(int) * (size_t)s1 * (k + 1) + 1);
  for (int j = 0; j < s1; j++) {
    heaps[j].val = hval + (long)j * (k + 1);
    heaps[j].lab = hlab + (long)j * (k + 1);
    heaps[j].cap = k;
  }
  for (long i1 = 0; i1 < nq; i1 += s1) {
    int m1 = (int)(nq - i1 < s1 ? nq - i1 : s1);
    for (int j = 0; j < m1; j++) heaps[j].cnt = 0;
    for (long i2 = 0; i2 < nb; i2 += s2) {
      int m2 = (int)(nb - i2 < s2 ? nb - i2 : s2);
      oy_tile_dists(type, d, m2, m1, b + i2 * d, q + i1 * d, dists);
      if (w)
        for (int j1 = 0; j1 < m1; j1++)
          for (int j2 = 0; j2 < m2; j2++) dists[j1 * m2 + j2] *= w[j2 + i2];
      for (int j1 = 0; j1 < m1; j1++)
        oy_heap_add_range(&heaps[j1], m2, (int)i2, dists + (long)j1 * m2);
    }
    for (int j1 = 0; j1 < m1; j1++) {
      oy_heap *h = &heaps[j1];
      int *ao = assign + (i1 + j1) * k;
      float *dout = dis + (i1 + j1) * k;
      oy_heap_sort(h, ao, dout);
      if (h->cnt < k) { /* nn.c:515-518 */
        memset(ao + h->cnt, 0xff, sizeof(int) * (k - h->cnt));
        memset(dout + h->cnt, 0xff, sizeof(float) * (k - h->cnt));
      }
    }
  }
  free(hlab);
  free(hval);
  free(heaps);
  free(dists);
}

/* nn.c:665-699: query slices, one per thread (OpenMP here). */
void oy_knn_full_thread(int type, int nq, int nb, int d, int k, const float *b,
                        const float *q, const float *w, int *assign, float *dis,
                        int nt) {
  if (nq < nt || nt <= 1) {
    oy_knn_full(type, nq, nb, d, k, b, q, w, assign, dis);
    return;
  }
#pragma omp parallel for schedule(dynamic) num_threads(nt)
  for (int t = 0; t < nt; t++) {
    long n0 = (long)nq * t / nt, n1 = (long)nq * (t + 1) / nt;
    oy_knn_full(type, (int)(n1 - n0), nb, d, k, b, q + n0 * d, w, assign + n0 * k,
                dis + n0 * k);
  }
}

/* nn.c:528-580: rerank a shortlist with exact (double-norm) distances. */
void oy_knn_reorder_shortlist(int n, int nb, int d, int k, const float *b,
                              const float *v, int *assign, float *dists) {
  (void)nb;
  float *sub = (float *)malloc(sizeof(float) * (size_t)k * d + 1);
  float *tmp = (float *)malloc(sizeof(float) * (size_t)k + 1);
  int *perm = (int *)malloc(sizeof(int) * (size_t)k + 1);
  int *atmp = (int *)malloc(sizeof(int) * (size_t)k + 1);
  for (long i = 0; i < n; i++) {
    int *ai = assign + i * k;
    float *di = dists + i * k;
    int ki = 0;
    for (; ki < k; ki++) {
      if (ai[ki] < 0) break;
      memcpy(sub + (long)ki * d, b + (long)ai[ki] * d, sizeof(float) * d);
    }
    oy_distances_1_nonpacked(d, ki, v + i * d, sub, d, tmp);
    oy_sort_index(tmp, ki, perm);
    memcpy(atmp, ai, sizeof(int) * ki);
    for (int j = 0; j < ki; j++) {
      di[j] = tmp[perm[j]];
      ai[j] = atmp[perm[j]];
    }
  }
  free(atmp);
  free(perm);
  free(tmp);
  free(sub);
}

/* nn.c:583-600: distances of the shortlist entries that fall in a window of
 * the base (labels label0 .. label0+nb-1, rows of b), per query from entry
 * kp[q] on until the first entry past the window; difference form in double
 * (vector.c:2348-2359), rounded to float.  kp[q] advances to that entry.
 * Returns -1 where the reference's assert(j >= 0) would fire (an entry below
 * the window before the first one past it), 0 otherwise. */
int oy_knn_recompute_exact_dists(int nq, int nb, int d, int k, const float *b,
                                 const float *v, int label0, int *kp, const int *idx,
                                 float *dis) {
  for (long q = 0; q < nq; q++) {
    const float *vq = v + (long)d * q;
    long i;
    for (i = kp[q]; i < k; i++) {
      long j = (long)idx[q * k + i] - label0;
      if (j < 0) return -1;
      if (j >= nb) break;
      const float *bj = b + j * d;
      double s = 0;
      for (long c = 0; c < d; c++) {
        double a = (double)vq[c] - bj[c];
        s += a * a;
      }
      dis[q * k + i] = (float)s;
    }
    kp[q] = (int)i;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* FP64 brute force (not a reference function: the exact-arithmetic check)  */

/* Difference-form squared L2 (type 2/12) or dot (type 16) in double, times
 * the weight; keeps the k smallest by (value, index).                      */
void oy_knn_f64(int type, int nq, int nb, int d, int k, const float *b,
                const float *q, const float *w, int *assign, double *dis) {
#pragma omp parallel
  {
    double *row = (double *)malloc(sizeof(double) * (size_t)nb + 8);
    int *idx = (int *)malloc(sizeof(int) * (size_t)nb + 8);
#pragma omp for schedule(dynamic, 4)
    for (long i = 0; i < nq; i++) {
      const float *qi = q + i * d;
      for (long j = 0; j < nb; j++) {
        const float *bj = b + j * d;
        double s = 0;
        if (type == 16) {
          for (int l = 0; l < d; l++) s += (double)qi[l] * (double)bj[l];
        } else {
          for (int l = 0; l < d; l++) {
            double t = (double)qi[l] - (double)bj[l];
            s += t * t;
          }
        }
        if (w) s *= (double)w[j];
        row[j] = s;
      }
      /* partial selection sort of the k smallest (k is small in tests) */
      for (long j = 0; j < nb; j++) idx[j] = (int)j;
      for (int r = 0; r < k; r++) {
        long best = r;
        for (long j = r + 1; j < nb; j++) {
          double vj = row[idx[j]], vb = row[idx[best]];
          if (vj < vb || (vj == vb && idx[j] < idx[best])) best = j;
        }
        int t = idx[r];
        idx[r] = idx[best];
        idx[best] = t;
        assign[i * k + r] = idx[r];
        dis[i * k + r] = row[idx[r]];
      }
    }
    free(idx);
    free(row);
  }
}

/* FP64 distance of given (query, base) pairs: the error report's yardstick. */
void oy_pair_dist_f64(int type, int nq, int d, int k, const float *b,
                      const float *q, const int *assign, double *out) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < nq; i++) {
    for (int r = 0; r < k; r++) {
      int j = assign[i * k + r];
      if (j < 0) {
        out[i * k + r] = NAN;
        continue;
      }
      const float *qi = q + i * d, *bj = b + (long)j * d;
      double s = 0;
      for (int l = 0; l < d; l++) {
        if (type == 16)
          s += (double)qi[l] * (double)bj[l];
        else {
          double t = (double)qi[l] - (double)bj[l];
          s += t * t;
        }
      }
      out[i * k + r] = s;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* vector helpers with the reference's precision (vector.c)                  */

/* vector.c:2180-2212 for norm 1 and 2 */
static double oy_fvec_norm(const float *v, long n, double norm) {
  double s = 0;
  if (norm == 1) {
    for (long i = 0; i < n; i++) s += fabs(v[i]);
    return s;
  }
  for (long i = 0; i < n; i++) s += v[i] * v[i];
  return sqrt(s);
}

/* vector.c:1792-1797 */
static void oy_fvec_mul_by(float *v, long n, double s) {
  for (long i = 0; i < n; i++) v[i] = (float)(v[i] * s);
}

/* vector.c:2016-2025 */
static double oy_fvec_normalize(float *v, long n, double norm) {
  double nr = oy_fvec_norm(v, n, norm);
  oy_fvec_mul_by(v, n, 1. / nr);
  return nr;
}

/* vector.c:2066-2073 */
double oy_fvec_sum(const float *v, long n) {
  double s = 0;
  for (long i = 0; i < n; i++) s += v[i];
  return s;
}

/* kmeans.c:22-24 / vector.c:135-137 */
static double oy_drand_r(unsigned *seed) { return rand_r(seed) / ((double)RAND_MAX + 1.0); }

/* vector.c:139-154 (ratio of uniforms) */
static double oy_gaussrand_r(unsigned *seed) {
  double z;
  for (;;) {
    float u1 = (float)oy_drand_r(seed);
    float u2 = (float)oy_drand_r(seed);
    z = 1.71552776992141 * (u1 - .5) / u2;
    float zz = (float)(z * z / 4.0);
    if (zz < -log(u2)) break;
  }
  return z;
}

/* vector.c:184-189 */
void oy_fvec_randn_r(float *v, long n, unsigned seed) {
  for (long i = 0; i < n; i++) v[i] = (float)oy_gaussrand_r(&seed);
}

/* vector.c:226-253 + kmeans.c:15-20: first k of a partial Fisher-Yates. */
void oy_random_init(int n, int k, int *sel, unsigned seed) {
  int *idx = (int *)malloc(sizeof(int) * (size_t)n);
  for (int i = 0; i < n; i++) idx[i] = i;
  for (int i = 0; i < n - 1; i++) {
    int j = i + rand_r(&seed) % (n - i);
    int t = idx[i];
    idx[i] = idx[j];
    idx[j] = t;
  }
  memcpy(sel, idx, sizeof(int) * k);
  free(idx);
}

/* ------------------------------------------------------------------------ */
/* k-means                                                                   */

/* kmeans.c:166-209 */
int oy_reassign_empty(int d, int k, float *cent, const int *nassign, unsigned seed) {
  int nre = 0;
  float *p = (float *)malloc(sizeof(float) * k);
  float *eps = (float *)malloc(sizeof(float) * d);
  for (int c = 0; c < k; c++) {
    /* int arithmetic of the reference, with its two's-complement wrap made
       explicit (kmeans.c:175) */
    int na = nassign[c];
    int pc = na < 2 ? 0 : (int)((unsigned)na * (unsigned)na - 1u);
    p[c] = (float)pc;
  }
  oy_fvec_normalize(p, k, 1);
  for (int c = 0; c < k; c++) {
    if (nassign[c] != 0) continue;
    nre++;
    double rd = oy_drand_r(&seed);
    int j;
    for (j = 0; j < k - 1; j++) {
      rd -= p[j];
      if (rd < 0) break;
    }
    memmove(cent + (long)c * d, cent + (long)j * d, sizeof(float) * d);
    double s = oy_fvec_norm(cent + (long)j * d, d, 2) * 0.0000001;
    oy_fvec_randn_r(eps, d, (unsigned)rand_r(&seed));
    oy_fvec_mul_by(eps, d, s);
    for (int l = 0; l < d; l++) cent[(long)j * d + l] += eps[l];
    for (int l = 0; l < d; l++) cent[(long)c * d + l] -= eps[l];
    p[j] = 0;
    oy_fvec_normalize(p, k, 1);
  }
  free(eps);
  free(p);
  return nre;
}

/* Per-iteration trace for the tests (not part of the reference API). */
typedef struct {
  int niter_done;
  int nreassign_tot;
  double qerr[1024];
  int nreassign[1024];
} oy_kmeans_trace;

/* kmeans.c:213-329 (L2 only; NORMALIZE_CENTS honoured) */
static int oy_kmeans_core(int d, int n, int k, int niter, int nt, int flags,
                          float *cent, const float *v, unsigned seed, int *assign,
                          int *nassign, float *dis, double *qerr_out,
                          long *iter_tot, oy_kmeans_trace *tr) {
  double qerr = HUGE_VAL, qerr_old;
  int tot = 0;
  for (int iter = 1; iter <= niter; iter++) {
    (*iter_tot)++;
    oy_knn_full_thread(2, n, k, d, 1, cent, v, NULL, assign, dis, nt);
    memset(nassign, 0, sizeof(int) * k);
    for (long i = 0; i < n; i++) nassign[assign[i]]++;
    memset(cent, 0, sizeof(float) * (size_t)k * d);
    for (long i = 0; i < n; i++) {
      assert(assign[i] >= 0);
      float *c = cent + (long)assign[i] * d;
      const float *x = v + i * (long)d;
      for (int l = 0; l < d; l++) c[l] += x[l];
    }
    for (long c = 0; c < k; c++) oy_fvec_mul_by(cent + c * d, d, 1.0 / nassign[c]);
    if (flags & OY_NORMALIZE_CENTS)
      for (long c = 0; c < k; c++) oy_fvec_normalize(cent + c * d, d, 2.0);
    int nre = oy_reassign_empty(d, k, cent, nassign, (unsigned)rand_r(&seed));
    tot += nre;
    if (tr && iter <= 1024) tr->nreassign[iter - 1] = nre;
    if (tr) tr->nreassign_tot = tot;
    if (tot > n / 100 && tot > 1000) return -1;
    qerr_old = qerr;
    qerr = oy_fvec_sum(dis, n);
    if (tr) {
      if (iter <= 1024) tr->qerr[iter - 1] = qerr;
      tr->niter_done = iter;
    }
    if (qerr_old == qerr && nre == 0) break;
  }
  *qerr_out = qerr;
  return 0;
}

/* kmeans.c:27-82: k-means++ seeding among the first n points.  The first
 * centre is rand_r % k (as the reference draws it); each next one keeps the
 * running minimum of the squared distances to the centres chosen so far
 * (nn.c:132-154 against the newest centre), scales a copy to unit L1 norm in
 * place (vector.c:2016-2026, 1792-1797) and walks a draw rd in [0, 1) down it
 * in double: the first j with rd < 0, else n - 1. */
void oy_kmeanspp_init(int d, int n, int k, const float *v, int *sel, unsigned seed) {
  float *best = (float *)malloc(sizeof(float) * (size_t)n);
  float *tmp = (float *)malloc(sizeof(float) * (size_t)n);
  for (long j = 0; j < n; j++) best[j] = HUGE_VAL;
  sel[0] = rand_r(&seed) % k;
  for (int i = 1; i < k; i++) {
    oy_distances_1_nonpacked(d, n, v + (long)d * sel[i - 1], v, d, tmp);
    for (long j = 0; j < n; j++)
      if (tmp[j] < best[j]) best[j] = tmp[j];
    memcpy(tmp, best, sizeof(float) * (size_t)n);
    oy_fvec_normalize(tmp, n, 1);
    double rd = oy_drand_r(&seed);
    long j;
    for (j = 0; j < n - 1; j++) {
      rd -= tmp[j];
      if (rd < 0) break;
    }
    sel[i] = (int)j;
  }
  free(best);
  free(tmp);
}

/* kmeans.c:332-447 (random, k-means++ or user init) */
float oy_kmeans_traced(int d, int n, int k, int niter, const float *v, int flags,
                       long seed_in, int redo, float *cent_out, float *dis_out,
                       int *assign_out, int *nassign_out, oy_kmeans_trace *tr) {
  int nt = flags & 0xffff;
  if (nt == 0) nt = 1;
  niter = niter == 0 ? 1000000 : niter;
  int user = (flags & OY_INIT_USER) != 0;
  if (user) {
    assert(cent_out);
    redo = 1;
  }
  if (flags & (OY_L1 | OY_CHI2)) {
    fprintf(stderr, "oracle: kmeans flag not restated\n");
    abort();
  }
  float *cent = (float *)malloc(sizeof(float) * (size_t)k * d);
  float *dis = (float *)malloc(sizeof(float) * (size_t)n);
  int *assign = (int *)malloc(sizeof(int) * (size_t)n);
  int *nassign = (int *)malloc(sizeof(int) * (size_t)k);
  int *sel = (int *)malloc(sizeof(int) * (size_t)k);
  double qerr = 0, best = HUGE_VAL;
  long iter_tot = 0;
  assert(k <= n);
  if (seed_in == 0) seed_in = lrand48();
  unsigned seed = (unsigned)seed_in;
  int ret = 0;
  for (int run = 0; run < redo; run++) {
    if (user) {
      memcpy(cent, cent_out, sizeof(float) * (size_t)k * d);
    } else {
      if (flags & OY_INIT_BERKELEY) {
        /* kmeans.c:393-400: at most 8k points when n > 8k and n > 8192 */
        const int nsub = (n > k * 8 && n > 8192) ? k * 8 : n;
        oy_kmeanspp_init(d, nsub, k, v, sel, (unsigned)rand_r(&seed));
      } else {
        oy_random_init(n, k, sel, (unsigned)rand_r(&seed));
      }
      for (long i = 0; i < k; i++)
        memcpy(cent + i * d, v + (long)sel[i] * d, sizeof(float) * d);
    }
    ret = oy_kmeans_core(d, n, k, niter, nt, flags, cent, v, (unsigned)rand_r(&seed),
                         assign, nassign, dis, &qerr, &iter_tot, tr);
    if (ret < 0) break;
    if (qerr < best) {
      best = qerr;
      if (cent_out) memcpy(cent_out, cent, sizeof(float) * (size_t)k * d);
      if (dis_out) memcpy(dis_out, dis, sizeof(float) * (size_t)n);
      if (assign_out) memcpy(assign_out, assign, sizeof(int) * (size_t)n);
      if (nassign_out) memcpy(nassign_out, nassign, sizeof(int) * (size_t)k);
    }
  }
  free(sel);
  free(nassign);
  free(assign);
  free(dis);
  free(cent);
  if (ret < 0) return -1;
  return (float)(best / n);
}

float oy_kmeans(int d, int n, int k, int niter, const float *v, int flags,
                long seed, int redo, float *cent, float *dis, int *assign,
                int *nassign) {
  return oy_kmeans_traced(d, n, k, niter, v, flags, seed, redo, cent, dis, assign,
                          nassign, NULL);
}

/* ------------------------------------------------------------------------ */
/* hierarchical k-means (hkm.c)                                              */

static __thread const int *oy_lab_sort;
static int oy_cmp_lab(const void *x, const void *y) { /* sorting.c:324-338 */
  const int a = *(const int *)x, b = *(const int *)y;
  const int dt = oy_lab_sort[a] - oy_lab_sort[b];
  return dt ? dt : a - b;
}

/* hkm.c:36-116.  Level l clusters each of its bf^l nodes into bf children
 * with kmeans(d, nassign, bf, niter, v_node, KMEANS_INIT_RANDOM |
 * KMEANS_QUIET, 0, 1, ...) (seed 0: lrand48() per call), the node's points
 * taken in (label, index) order.  cent: the levels stacked (level l:
 * bf^(l+1) rows of d); assign: final labels.  Returns -1 where the
 * reference exits (an empty node). */
int oy_hkm_learn(int n, int d, int nlevel, int bf, const float *points, int niter, float *cent,
                 int *assign) {
  int *label = (int *)calloc((size_t)n, sizeof(int));
  int *perm = (int *)malloc(sizeof(int) * (size_t)n);
  int *ca = (int *)malloc(sizeof(int) * (size_t)n);
  float *v = (float *)malloc(sizeof(float) * (size_t)n * d);
  long k = 1, off = 0;
  int ret = 0;
  for (int l = 0; l < nlevel && ret == 0; l++) {
    for (int i = 0; i < n; i++) perm[i] = i;
    oy_lab_sort = label;
    qsort(perm, n, sizeof(int), oy_cmp_lab);
    for (int i = 0; i < n; i++) memcpy(v + (long)d * i, points + (long)d * perm[i], sizeof(float) * d);
    int pos = 0;
    for (long parent = 0; parent < k; parent++) {
      int nassign = 0;
      while (pos + nassign < n && label[perm[pos + nassign]] == parent) nassign++;
      if (nassign == 0) {
        ret = -1;
        break;
      }
      oy_kmeans(d, nassign, bf, niter, v + (long)d * pos, OY_INIT_RANDOM | OY_QUIET, 0, 1,
                cent + (off + parent * bf) * d, NULL, ca, NULL);
      for (int i = 0; i < nassign; i++) {
        const int t = perm[pos + i];
        label[t] = label[t] * bf + ca[i];
      }
      pos += nassign;
    }
    off += k * bf;
    k *= bf;
  }
  if (ret == 0) memcpy(assign, label, sizeof(int) * (size_t)n);
  free(label);
  free(perm);
  free(ca);
  free(v);
  return ret;
}

/* hkm.c:131-150: per point, nn() among the children of its node, level by
 * level (cent stacked as in oy_hkm_learn) */
void oy_hkm_quantize(int nlevel, int bf, int d, const float *cent, int n, const float *v,
                     int *idx) {
  for (int i = 0; i < n; i++) {
    long vw = 0, off = 0, k = 1;
    for (int l = 0; l < nlevel; l++) {
      int t;
      float dd;
      oy_knn_full(2, 1, bf, d, 1, cent + (off + vw * bf) * d, v + (long)d * i, NULL, &t, &dd);
      vw = vw * bf + t;
      off += k * bf;
      k *= bf;
    }
    idx[i] = (int)vw;
  }
}

/* Size of the trace struct, so Python can allocate it. */
int oy_trace_sizeof(void) { return (int)sizeof(oy_kmeans_trace); }

/* Thin rand_r wrapper so tests can check the stream arithmetic. */
int oy_rand_r(unsigned *seed) { return rand_r(seed); }

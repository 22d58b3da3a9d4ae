/*
 * yael_hkm.c -- hierarchical k-means (hkm.c): a tree of nlevel levels with
 * branching factor bf, every node clustered into bf children.
 *
 * hkm_learn (hkm.c:36-116) runs each level as ONE batched k-means over all
 * bf^l nodes of the level (yh_kmeans_batched, the GPU batched form) instead
 * of bf^l sequential kmeans() calls.  Each node's instance has the
 * semantics of the reference's call kmeans(d, nassign, bf, nb_iter_max,
 * v_node, KMEANS_INIT_RANDOM | KMEANS_QUIET, 0, 1, ...): seed 0, so each
 * node draws lrand48() in node order, and the node's points are in the
 * order of the reference's index sort of the node labels (hkm.c:57-62,
 * sorting.c:324-354: by label, then by point index -- a stable counting
 * sort here).  The points stay in HBM for the whole tree; per level the
 * host sorts the n labels and the device gathers the rows in that order.
 *
 * hkm_quantize (hkm.c:131-150) descends the tree for every point on the GPU
 * (yk_hkm_quantize: one thread per point, the children of its node compared
 * with the arithmetic of the reference's nn(): float centroid norms, double
 * point norms, -2 * dot added in float, strict '<' from (-1, 1e30)).
 *
 * hkm_write / hkm_read keep the reference's file format (hkm.c:181-232):
 * int nlevel, int bf, int d, then per level l an fvec record (int count,
 * count floats) of the bf^(l+1) * d centroid values.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../yk.h"
#include "yael_host.h"
#include "../../../include/yael_mi355x.h"

/* hkm.c:18-33 */
static hkm_t *yh_hkm_new(int d, int nlevel, int bf) {
  hkm_t *h = (hkm_t *)calloc(1, sizeof(*h));
  if (!h) return NULL;
  h->nlevel = nlevel;
  h->bf = bf;
  h->d = d;
  h->centroids = (float **)calloc((size_t)nlevel, sizeof(*h->centroids));
  if (!h->centroids) {
    free(h);
    return NULL;
  }
  long k = 1;
  for (int l = 0; l < nlevel; l++) {
    h->centroids[l] = (float *)malloc(sizeof(float) * (size_t)(k * bf) * d);
    if (!h->centroids[l]) {
      hkm_delete(h);
      return NULL;
    }
    k *= bf;
  }
  h->k = (int)k;
  return h;
}

/* hkm.c:119-126 */
void hkm_delete(hkm_t *h) {
  if (!h) return;
  if (h->centroids)
    for (int l = 0; l < h->nlevel; l++) free(h->centroids[l]);
  free(h->centroids);
  free(h);
}

/* hkm.c:154-157 */
float *hkm_get_centroids(const hkm_t *h, int l, int no) {
  return h->centroids[l] + (long)h->d * h->bf * no;
}

static int yh_hkm_learn(int n, int d, int nlevel, int bf, const float *points, int niter,
                        int verbose, int **assign_out, hkm_t **out) {
  *out = NULL;
  if (n < 1 || d < 1 || nlevel < 1 || bf < 1)
    return yk_set_error_c(YK_EINVAL, "hkm_learn: bad sizes (n=%d d=%d nlevel=%d bf=%d)", n, d,
                          nlevel, bf);
  long leaves = 1;
  for (int l = 0; l < nlevel; l++) {
    leaves *= bf;
    if (leaves > 0x7fffffffL)
      return yk_set_error_c(YK_EINVAL, "hkm_learn: bf^nlevel overflows int");
  }
  int rc = yk_init();
  if (rc != YK_OK) return rc;
  hkm_t *h = yh_hkm_new(d, nlevel, bf);
  const int ld = yk_ld(d);
  const long nodes_max = leaves / bf;
  int *label = (int *)calloc((size_t)n, sizeof(int)); /* node of each point */
  int *perm = (int *)malloc(sizeof(int) * (size_t)n);
  int *lassign = (int *)malloc(sizeof(int) * (size_t)n);
  int *count = (int *)malloc(sizeof(int) * (size_t)nodes_max);
  long *start = (long *)malloc(sizeof(long) * (size_t)(nodes_max + 1));
  float *ret = (float *)malloc(sizeof(float) * (size_t)nodes_max);
  float *v = NULL, *vg = NULL;
  int *perm_dev = NULL;
  if (!h || !label || !perm || !lassign || !count || !start || !ret)
    rc = yk_set_error_c(YK_ENOMEM, "hkm_learn: host malloc");
  if (rc == YK_OK) rc = yk_malloc((void **)&v, sizeof(float) * (size_t)n * ld);
  if (rc == YK_OK) rc = yk_malloc((void **)&vg, sizeof(float) * (size_t)n * ld);
  if (rc == YK_OK) rc = yk_malloc((void **)&perm_dev, sizeof(int) * (size_t)n);
  if (rc == YK_OK) rc = yk_upload_vectors(v, points, n, d);
  long k = 1; /* nodes at this level */
  for (int l = 0; l < nlevel && rc == YK_OK; l++) {
    /* points in (label, index) order: a stable counting sort */
    memset(count, 0, sizeof(int) * (size_t)k);
    for (long i = 0; i < n; i++) count[label[i]]++;
    start[0] = 0;
    for (long p = 0; p < k; p++) start[p + 1] = start[p] + count[p];
    for (long p = 0; p < k && rc == YK_OK; p++) {
      if (count[p] == 0) /* hkm.c:77-80 */
        rc = yk_set_error_c(YK_EINVAL, "hkm_learn: no enough vectors in a node (level %d, "
                                       "node %ld)", l, p);
      else if (count[p] < bf) /* kmeans.c:377 */
        rc = yk_set_error_c(YK_EINVAL, "hkm_learn: node %ld of level %d has %d < bf=%d points",
                            p, l, count[p], bf);
    }
    if (rc != YK_OK) break;
    {
      long *fill = (long *)malloc(sizeof(long) * (size_t)k);
      if (!fill) {
        rc = yk_set_error_c(YK_ENOMEM, "hkm_learn: host malloc");
        break;
      }
      memcpy(fill, start, sizeof(long) * (size_t)k);
      for (long i = 0; i < n; i++) perm[fill[label[i]]++] = (int)i;
      free(fill);
    }
    if ((rc = yk_memcpy_h2d(perm_dev, perm, sizeof(int) * (size_t)n)) != YK_OK) break;
    if ((rc = yk_gather_rows(v, perm_dev, n, d, vg)) != YK_OK) break;
    /* every node of the level in one batched run; seeds NULL: each node
     * draws lrand48() in node order, as kmeans(..., seed 0, ...) does */
    rc = yh_kmeans_batched(d, (int)k, count, bf, niter, vg, 1, KMEANS_INIT_RANDOM | KMEANS_QUIET,
                           NULL, h->centroids[l], NULL, lassign, NULL, ret, NULL, 1, NULL);
    if (rc != YK_OK) break;
    for (long p = 0; p < k; p++) {
      if (verbose)
        fprintf(stderr, "[Level %d | Parent %ld] nassign=%d | pos=%ld-> err = %.3f\n", l, p,
                count[p], start[p], ret[p]);
      if (ret[p] < 0) /* the reference would go on with uninitialised centroids */
        rc = yk_set_error_c(YK_EABANDON, "hkm_learn: k-means of node %ld (level %d) abandoned",
                            p, l);
    }
    if (rc != YK_OK) break;
    for (long i = 0; i < n; i++) label[perm[i]] = label[perm[i]] * bf + lassign[i];
    k *= bf;
  }
  if (v) yk_free(v);
  if (vg) yk_free(vg);
  if (perm_dev) yk_free(perm_dev);
  free(perm);
  free(lassign);
  free(count);
  free(start);
  free(ret);
  if (rc != YK_OK) {
    free(label);
    hkm_delete(h);
    return rc;
  }
  if (assign_out)
    *assign_out = label;
  else
    free(label);
  *out = h;
  return YK_OK;
}

static int yh_hkm_quantize(const hkm_t *h, int n, const float *v, int *idx) {
  if (n <= 0) return YK_OK;
  if (!h || h->nlevel < 1 || h->bf < 1 || h->d < 1)
    return yk_set_error_c(YK_EINVAL, "hkm_quantize: bad tree");
  int rc = yk_init();
  if (rc != YK_OK) return rc;
  const int d = h->d, ld = yk_ld(d), bf = h->bf;
  long rows = 0, k = 1;
  for (int l = 0; l < h->nlevel; l++) {
    k *= bf;
    rows += k;
  }
  float *vd = NULL, *cd = NULL, *cn = NULL;
  double *vn = NULL;
  int *id = NULL;
  rc = yk_malloc((void **)&vd, sizeof(float) * (size_t)n * ld);
  if (rc == YK_OK) rc = yk_malloc((void **)&vn, sizeof(double) * (size_t)n);
  if (rc == YK_OK) rc = yk_malloc((void **)&cd, sizeof(float) * (size_t)rows * ld);
  if (rc == YK_OK) rc = yk_malloc((void **)&cn, sizeof(float) * (size_t)rows);
  if (rc == YK_OK) rc = yk_malloc((void **)&id, sizeof(int) * (size_t)n);
  if (rc == YK_OK) rc = yk_upload_vectors(vd, v, n, d);
  k = 1;
  long off = 0;
  for (int l = 0; l < h->nlevel && rc == YK_OK; l++) {
    k *= bf;
    rc = yk_upload_vectors(cd + off * ld, h->centroids[l], k, d);
    off += k;
  }
  if (rc == YK_OK) rc = yk_norms(vd, n, d, NULL, vn);
  if (rc == YK_OK) rc = yk_norms(cd, rows, d, cn, NULL);
  if (rc == YK_OK) rc = yk_hkm_quantize(vd, vn, n, d, h->nlevel, bf, cd, cn, id);
  if (rc == YK_OK) rc = yk_memcpy_d2h(idx, id, sizeof(int) * (size_t)n);
  if (vd) yk_free(vd);
  if (vn) yk_free(vn);
  if (cd) yk_free(cd);
  if (cn) yk_free(cn);
  if (id) yk_free(id);
  return rc;
}

/* hkm.c:181-199 */
static int yh_hkm_write(const char *filename, const hkm_t *h) {
  FILE *f = fopen(filename, "wb");
  if (!f) return yk_set_error_c(YK_EINVAL, "hkm_write: cannot open %s", filename);
  int ok = fwrite(&h->nlevel, sizeof(int), 1, f) == 1 && fwrite(&h->bf, sizeof(int), 1, f) == 1 &&
           fwrite(&h->d, sizeof(int), 1, f) == 1;
  long k = h->bf;
  for (int l = 0; l < h->nlevel && ok; l++) {
    const int cnt = (int)(k * h->d);
    ok = fwrite(&cnt, sizeof(int), 1, f) == 1 &&
         fwrite(h->centroids[l], sizeof(float), (size_t)cnt, f) == (size_t)cnt;
    k *= h->bf;
  }
  if (fclose(f) != 0) ok = 0;
  return ok ? YK_OK : yk_set_error_c(YK_EINVAL, "hkm_write: write error on %s", filename);
}

/* hkm.c:202-232 */
static int yh_hkm_read(const char *filename, hkm_t **out) {
  *out = NULL;
  FILE *f = fopen(filename, "rb");
  if (!f) return yk_set_error_c(YK_EINVAL, "hkm_read: cannot open %s", filename);
  int hdr[3];
  if (fread(hdr, sizeof(int), 3, f) != 3 || hdr[0] < 1 || hdr[1] < 1 || hdr[2] < 1) {
    fclose(f);
    return yk_set_error_c(YK_EINVAL, "hkm_read: unable to read the hkm file %s", filename);
  }
  hkm_t *h = yh_hkm_new(hdr[2], hdr[0], hdr[1]);
  if (!h) {
    fclose(f);
    return yk_set_error_c(YK_ENOMEM, "hkm_read: host malloc");
  }
  long k = h->bf;
  int ok = 1;
  for (int l = 0; l < h->nlevel && ok; l++) {
    int cnt = 0;
    ok = fread(&cnt, sizeof(int), 1, f) == 1 && cnt == k * h->d &&
         fread(h->centroids[l], sizeof(float), (size_t)cnt, f) == (size_t)cnt;
    k *= h->bf;
  }
  fclose(f);
  if (!ok) {
    hkm_delete(h);
    return yk_set_error_c(YK_EINVAL, "hkm_read: unable to read the hkm file %s", filename);
  }
  *out = h;
  return YK_OK;
}

/* ---------------- public entry points ---------------- */

int yael_hkm_learn_checked(int n, int d, int nlevel, int bf, const float *v, int nb_iter_max,
                           int nt, int verbose, int **clust_assign_out, hkm_t **out) {
  (void)nt;
  yh_lock();
  int rc = yh_hkm_learn(n, d, nlevel, bf, v, nb_iter_max, verbose, clust_assign_out, out);
  yh_unlock();
  return rc;
}

int yael_hkm_quantize_checked(const hkm_t *hkm, int n, const float *v, int *idx) {
  yh_lock();
  int rc = yh_hkm_quantize(hkm, n, v, idx);
  yh_unlock();
  return rc;
}

/* hkm.h:21-23 */
hkm_t *hkm_learn(int n, int d, int nlevel, int bf, const float *v, int nb_iter_max, int nt,
                 int verbose, int **clust_assign_out) {
  hkm_t *h = NULL;
  if (yael_hkm_learn_checked(n, d, nlevel, bf, v, nb_iter_max, nt, verbose, clust_assign_out,
                             &h) != YK_OK)
    yh_die("hkm_learn");
  return h;
}

/* hkm.h:29 */
void hkm_quantize(const hkm_t *hkm, int n, const float *v, int *idx) {
  if (yael_hkm_quantize_checked(hkm, n, v, idx) != YK_OK) yh_die("hkm_quantize");
}

/* hkm.h:32: the reference prints a message and returns on a write error */
void hkm_write(const char *filename, const hkm_t *hkm) {
  if (yh_hkm_write(filename, hkm) != YK_OK) fprintf(stderr, "# %s\n", yk_last_error());
}

/* hkm.h:33: NULL (with a message) when the file cannot be read */
hkm_t *hkm_read(const char *filename) {
  hkm_t *h = NULL;
  if (yh_hkm_read(filename, &h) != YK_OK) fprintf(stderr, "# %s\n", yk_last_error());
  return h;
}

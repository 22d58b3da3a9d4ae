/*
 * yael_batch.c -- many independent k-means instances in one batched run
 * (the hkm_learn pattern, hkm.c:66-103, where every tree node is clustered
 * by its own kmeans() call with KMEANS_INIT_RANDOM, redo 1).
 *
 * Each instance i has the semantics of
 *     kmeans(d, n_i, k, niter, v_i, flags, seeds[i], 1, ...)
 * (seeds[i] == 0 draws lrand48() in instance order, as successive kmeans()
 * calls would).  The host does what depends only on the seeds: the seed
 * chains (kmeans.c:379-409), the random initialisation (kmeans.c:15-20) and
 * the split streams (kmeans.c:183-195) -- for every instance and iteration
 * the first YH_ECAP (draw, gaussian vector) pairs are precomputed, a window
 * of iterations at a time -- plus the work-item tables.  The device then
 * runs the window's iterations without the host (yk_batch.hip), captured
 * once in a hipGraph and replayed; between windows the host reads the
 * instances' done flags and stops once every instance has stopped, so
 * niter = 0 ("until convergence", at most 1,000,000 iterations, kmeans.c:345)
 * costs the iterations actually run.  An instance that empties more than
 * YH_ECAP clusters in one iteration is flagged and recomputed by the
 * single-instance path, so the results never depend on the cap.
 *
 * The chunk size of the update's partial sums is chosen by a small
 * autotuner: each candidate is timed on the actual data, the faster of two
 * runs of three iterations from the same initial state; the fastest
 * candidate is kept and remembered in the tuning cache keyed by the problem
 * shape.
 */
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../yk.h"
#include "yael_host.h"

#define YH_ECAP 4
#define YH_TILE 128 /* points per assign work item (4 waves x 32) */
/* split-table window: at most this many bytes of draws on the device, and
   when niter exceeds YH_WIN_LONG, windows of YH_WIN_CHECK iterations (how
   often a long run checks whether every instance has stopped) */
#define YH_WIN_BYTES (256L << 20)
#define YH_WIN_LONG 64
#define YH_WIN_CHECK 16

typedef struct {
  void *p[32];
  int n;
} yh_allocs;

static void *yh_dmalloc(yh_allocs *a, size_t bytes) {
  void *p = NULL;
  if (a->n == 32 || yk_malloc(&p, bytes ? bytes : 256) != YK_OK) return NULL;
  a->p[a->n++] = p;
  return p;
}

static void yh_dfree_all(yh_allocs *a) {
  for (int i = 0; i < a->n; i++) yk_free(a->p[i]);
  a->n = 0;
}

/* ---- the batched run ---- */

typedef struct {
  int ninst, d, ld, k, kpad, niter, win, normalize;
  long ntot;
  long *off; /* host offsets */
  float *cent_user; /* KMEANS_INIT_USER: the initial centroids (device) */
  /* device */
  long *d_off;
  const float *v;
  double *vn;
  float *cent, *cn, *dis;
  int *assign, *sel, *nassign, *flags;
  double *rd, *qerr2;
  float *eps, *pbuf;
  int *item_inst, *item_tile;
  long nitems;
} yh_bstate;

/* update-chunk tables for chunk size P, and the batch object */
static int yh_batch_bind(yh_bstate *s, int P, yh_allocs *al, yk_batch **out) {
  const int ninst = s->ninst;
  long nu = 0;
  int *uf = (int *)malloc(sizeof(int) * (ninst + 1));
  for (int i = 0; i < ninst; i++) {
    uf[i] = (int)nu;
    nu += (s->off[i + 1] - s->off[i] + P - 1) / P;
  }
  uf[ninst] = (int)nu;
  int *ui = (int *)malloc(sizeof(int) * (nu > 0 ? nu : 1));
  for (int i = 0; i < ninst; i++)
    for (long u = uf[i]; u < uf[i + 1]; u++) ui[u] = i;
  int *d_uf = (int *)yh_dmalloc(al, sizeof(int) * (ninst + 1));
  int *d_ui = (int *)yh_dmalloc(al, sizeof(int) * nu);
  /* partial sums, then one objective partial per chunk */
  double *psum = (double *)yh_dmalloc(al, sizeof(double) * ((size_t)nu * s->k * s->d + nu));
  int *pcnt = (int *)yh_dmalloc(al, sizeof(int) * (size_t)nu * s->k);
  int rc = (!d_uf || !d_ui || !psum || !pcnt) ? YK_ENOMEM : YK_OK;
  if (rc == YK_OK) rc = yk_memcpy_h2d(d_uf, uf, sizeof(int) * (ninst + 1));
  if (rc == YK_OK) rc = yk_memcpy_h2d(d_ui, ui, sizeof(int) * nu);
  free(uf);
  free(ui);
  if (rc != YK_OK) return rc;
  return yk_batch_create(out, ninst, s->d, s->k, s->kpad, s->niter, s->win, YH_ECAP, P, s->ntot,
                         s->v,
                         s->vn, s->cent, s->cn, s->assign, s->dis, s->d_off, s->nitems,
                         s->item_inst, s->item_tile, nu, d_ui, d_uf, psum, pcnt, s->nassign,
                         s->pbuf, s->rd, s->eps, s->flags, s->qerr2, s->normalize);
}

/* the initial centroids: the caller's (KMEANS_INIT_USER) or the rows the
   seeded random initialisation selected */
static int yh_batch_init_state(yh_bstate *s) {
  if (s->cent_user)
    return yk_memcpy_d2d(s->cent, s->cent_user,
                         sizeof(float) * (size_t)s->ninst * s->kpad * s->ld);
  return yk_gather_rows(s->v, s->sel, (long)s->ninst * s->kpad, s->d, s->cent);
}

/* host precomputation for instances [i0, ninst) in steps of step.  First
   call (win == 0): the initial selection (kmeans.c:15-20, 402) and the
   state of the instance's per-iteration seed chain (kmeans.c:408-409).
   Window calls: for the next win iterations, the split stream's first
   YH_ECAP (draw, gaussian vector) pairs (kmeans.c:183-195, 296), advancing
   the chain */
typedef struct {
  int ninst, win, k, kpad, d, user_init;
  const int *n_inst;
  const long *off, *seed_used;
  unsigned *chain;
  int *sel;
  double *rd;
  float *eps;
  int i0, step, failed;
} yh_prep;

static void *yh_prep_worker(void *arg) {
  yh_prep *p = (yh_prep *)arg;
  const int win = p->win, d = p->d;
  if (win == 0) {
    int *tmp = (int *)malloc(sizeof(int) * p->k);
    if (!tmp) {
      p->failed = 1;
      return NULL;
    }
    for (int i = p->i0; i < p->ninst; i += p->step) {
      unsigned seed = (unsigned)p->seed_used[i];
      if (!p->user_init) { /* user centroids draw nothing (kmeans.c:388-389) */
        yh_random_init(p->n_inst[i], p->k, tmp, (unsigned)rand_r(&seed));
        for (int c = 0; c < p->kpad; c++)
          p->sel[(long)i * p->kpad + c] = (int)(p->off[i] + tmp[c < p->k ? c : 0]);
      }
      p->chain[i] = (unsigned)rand_r(&seed);
    }
    free(tmp);
    return NULL;
  }
  for (int i = p->i0; i < p->ninst; i += p->step) {
    unsigned cs = p->chain[i];
    for (int it = 0; it < win; it++) {
      unsigned S = (unsigned)rand_r(&cs); /* kmeans.c:296 */
      for (int e = 0; e < YH_ECAP; e++) {
        long t = ((long)i * win + it) * YH_ECAP + e;
        p->rd[t] = yh_drand_r(&S);
        unsigned gs = (unsigned)rand_r(&S);
        for (int l = 0; l < d; l++) p->eps[t * d + l] = (float)yh_gaussrand_r(&gs);
      }
    }
    p->chain[i] = cs;
  }
  return NULL;
}

/* runs yh_prep_worker over the instances; the results do not depend on the
   number of threads (a thread that cannot start leaves its share to the
   calling thread) */
static void yh_prep_run(yh_prep *p) {
  long work = 0;
  for (int i = 0; i < p->ninst; i++)
    work += p->win ? (long)p->win * YH_ECAP * p->d : p->n_inst[i];
  int nt = work < 200000 ? 1 : 16;
  if (nt > p->ninst) nt = p->ninst;
  yh_prep job[16];
  pthread_t th[16];
  int started[16] = {0};
  for (int t = 0; t < nt; t++) {
    job[t] = *p;
    job[t].i0 = t;
    job[t].step = nt;
    job[t].failed = 0;
    if (t > 0) started[t] = pthread_create(&th[t], NULL, yh_prep_worker, &job[t]) == 0;
  }
  yh_prep_worker(&job[0]);
  for (int t = 1; t < nt; t++) {
    if (started[t])
      pthread_join(th[t], NULL);
    else
      yh_prep_worker(&job[t]);
  }
  p->failed = 0;
  for (int t = 0; t < nt; t++) p->failed |= job[t].failed;
}

int yh_kmeans_batched(int d, int ninst, const int *n_inst, int k, int niter, const float *v,
                      int v_on_device, int flags, const long *seeds, float *cent_out,
                      float *dis_out, int *assign_out, int *nassign_out, float *ret_out,
                      int *iters_out, int use_graph, int *chunk_used) {
  if (flags & (KMEANS_L1 | KMEANS_CHI2 | KMEANS_INIT_BERKELEY))
    return yk_set_error_c(YK_EINVAL,
                          "batched k-means: L2 with random or user initialisation only");
  const int user_init = (flags & KMEANS_INIT_USER) != 0;
  if (user_init && !cent_out)
    return yk_set_error_c(YK_EINVAL, "batched k-means: KMEANS_INIT_USER needs the centroids");
  if (ninst < 1 || k < 1 || d < 1 || niter < 0)
    return yk_set_error_c(YK_EINVAL, "batched k-means: bad sizes");
  const int niter_in = niter;
  if (niter == 0) niter = 1000000; /* kmeans.c:345 */
  int rc = yk_init();
  if (rc != YK_OK) return rc;
  yh_bstate s;
  memset(&s, 0, sizeof s);
  s.ninst = ninst;
  s.d = d;
  s.ld = yk_ld(d);
  s.k = k;
  s.kpad = (k + 31) & ~31;
  s.niter = niter;
  s.normalize = (flags & KMEANS_NORMALIZE_CENTS) != 0;
  s.off = (long *)malloc(sizeof(long) * (ninst + 1));
  s.off[0] = 0;
  for (int i = 0; i < ninst; i++) {
    if (n_inst[i] < k) {
      free(s.off);
      return yk_set_error_c(YK_EINVAL, "batched k-means: instance %d has %d < k points", i,
                            n_inst[i]);
    }
    s.off[i + 1] = s.off[i] + n_inst[i];
  }
  s.ntot = s.off[ninst];
  const int ld = s.ld, kpad = s.kpad;
  /* the split-table window */
  {
    const long per_iter = (long)ninst * YH_ECAP * (long)(sizeof(double) + sizeof(float) * d);
    long w = YH_WIN_BYTES / per_iter;
    if (niter > YH_WIN_LONG && w > YH_WIN_CHECK) w = YH_WIN_CHECK;
    if (w > niter) w = niter;
    s.win = w < 1 ? 1 : (int)w;
  }
  const int win = s.win;

  /* seed chains (lrand48 in instance order, as successive kmeans() calls
     draw them), then per instance, on up to 16 host threads, the initial
     selection and the split streams (each a function of its seed only) */
  int *sel = (int *)malloc(sizeof(int) * (size_t)ninst * kpad);
  double *rd = (double *)malloc(sizeof(double) * (size_t)ninst * win * YH_ECAP);
  float *eps = (float *)malloc(sizeof(float) * (size_t)ninst * win * YH_ECAP * d);
  long *seed_used = (long *)malloc(sizeof(long) * ninst);
  unsigned *chain = (unsigned *)malloc(sizeof(unsigned) * ninst);
  int host_ok = sel && rd && eps && seed_used && chain;
  yh_prep prep = {ninst,    0,         k,     kpad, d,   user_init, n_inst, s.off,
                  seed_used, chain, sel, rd, eps, 0, 1, 0};
  if (host_ok) {
    for (int i = 0; i < ninst; i++) {
      long si = seeds ? seeds[i] : 0;
      if (si == 0) si = lrand48(); /* kmeans.c:379-380 */
      seed_used[i] = si;
    }
    yh_prep_run(&prep); /* selections and seed chains */
    host_ok = !prep.failed;
    prep.win = win;
    if (host_ok) yh_prep_run(&prep); /* the first window's split draws */
    host_ok = host_ok && !prep.failed;
  }

  /* assign work items: (instance, tile of 128 points) */
  long nitems = 0;
  for (int i = 0; i < ninst; i++) nitems += (n_inst[i] + YH_TILE - 1) / YH_TILE;
  int *ii = (int *)malloc(sizeof(int) * nitems), *it_ = (int *)malloc(sizeof(int) * nitems);
  host_ok = host_ok && ii && it_;
  for (int i = 0, t = 0; host_ok && i < ninst; i++)
    for (int j = 0; j < (n_inst[i] + YH_TILE - 1) / YH_TILE; j++, t++) {
      ii[t] = i;
      it_[t] = j;
    }
  s.nitems = nitems;

  yh_allocs al;
  al.n = 0;
  s.d_off = (long *)yh_dmalloc(&al, sizeof(long) * (ninst + 1));
  float *vdev = v_on_device ? (float *)v : (float *)yh_dmalloc(&al, sizeof(float) * s.ntot * ld);
  s.v = vdev;
  s.vn = (double *)yh_dmalloc(&al, sizeof(double) * s.ntot);
  s.cent = (float *)yh_dmalloc(&al, sizeof(float) * (size_t)ninst * kpad * ld);
  s.cn = (float *)yh_dmalloc(&al, sizeof(float) * (size_t)ninst * kpad);
  s.dis = (float *)yh_dmalloc(&al, sizeof(float) * s.ntot);
  s.assign = (int *)yh_dmalloc(&al, sizeof(int) * s.ntot);
  s.sel = (int *)yh_dmalloc(&al, sizeof(int) * (size_t)ninst * kpad);
  s.nassign = (int *)yh_dmalloc(&al, sizeof(int) * (size_t)ninst * k);
  s.flags = (int *)yh_dmalloc(&al, sizeof(int) * (7L * ninst + 8));
  s.qerr2 = (double *)yh_dmalloc(&al, sizeof(double) * 2L * ninst);
  s.rd = (double *)yh_dmalloc(&al, sizeof(double) * (size_t)ninst * win * YH_ECAP);
  s.eps = (float *)yh_dmalloc(&al, sizeof(float) * (size_t)ninst * win * YH_ECAP * d);
  s.pbuf = (float *)yh_dmalloc(&al, sizeof(float) * (size_t)ninst * k);
  s.item_inst = (int *)yh_dmalloc(&al, sizeof(int) * nitems);
  s.item_tile = (int *)yh_dmalloc(&al, sizeof(int) * nitems);
  rc = host_ok ? YK_OK : yk_set_error_c(YK_ENOMEM, "batched k-means: host malloc");
  if (rc == YK_OK && (!s.d_off || !vdev || !s.vn || !s.cent || !s.cn || !s.dis || !s.assign || !s.sel ||
      !s.nassign || !s.flags || !s.qerr2 || !s.rd || !s.eps || !s.pbuf || !s.item_inst ||
      !s.item_tile))
    rc = YK_ENOMEM;
  if (rc == YK_OK) rc = yk_memcpy_h2d(s.d_off, s.off, sizeof(long) * (ninst + 1));
  if (rc == YK_OK && !v_on_device) rc = yk_upload_vectors(vdev, v, s.ntot, d);
  if (rc == YK_OK) rc = yk_memcpy_h2d(s.sel, sel, sizeof(int) * (size_t)ninst * kpad);
  float *cent_init = NULL; /* KMEANS_INIT_USER: the caller's centroids, kept for the
                              single-instance path (cent_out receives the results) */
  if (rc == YK_OK && user_init) {
    cent_init = (float *)malloc(sizeof(float) * (size_t)ninst * k * d);
    if (cent_init) memcpy(cent_init, cent_out, sizeof(float) * (size_t)ninst * k * d);
    /* padded rows; the pad rows past k repeat centroid 0 of their
       instance, as the random initialisation's selection does */
    float *hc = (float *)malloc(sizeof(float) * (size_t)ninst * kpad * d);
    s.cent_user = (float *)yh_dmalloc(&al, sizeof(float) * (size_t)ninst * kpad * ld);
    if (!hc || !s.cent_user || !cent_init)
      rc = yk_set_error_c(YK_ENOMEM, "batched k-means: user centroids");
    for (int i = 0; rc == YK_OK && i < ninst; i++)
      for (int c = 0; c < kpad; c++)
        memcpy(hc + ((size_t)i * kpad + c) * d, cent_out + ((size_t)i * k + (c < k ? c : 0)) * d,
               sizeof(float) * d);
    if (rc == YK_OK) rc = yk_upload_vectors(s.cent_user, hc, (long)ninst * kpad, d);
    free(hc);
  }
  if (rc == YK_OK) rc = yk_memcpy_h2d(s.rd, rd, sizeof(double) * (size_t)ninst * win * YH_ECAP);
  if (rc == YK_OK)
    rc = yk_memcpy_h2d(s.eps, eps, sizeof(float) * (size_t)ninst * win * YH_ECAP * d);
  if (rc == YK_OK) rc = yk_memcpy_h2d(s.item_inst, ii, sizeof(int) * nitems);
  if (rc == YK_OK) rc = yk_memcpy_h2d(s.item_tile, it_, sizeof(int) * nitems);
  if (rc == YK_OK) rc = yk_norms(vdev, s.ntot, d, NULL, s.vn);
  free(ii);
  free(it_);
  free(sel);

  /* autotune the update chunk (timed from the initial state, then reset) */
  int P = 0;
  char key[256];
  snprintf(key, sizeof key, "batch_chunk:d=%d:k=%d:ninst=%d:ntot=%ld", d, k, ninst, s.ntot);
  if (rc == YK_OK && !yk_tune_get(key, &P)) {
    /* candidates: enough chunks to give every CU work (>= 512), partial
       sums within 8 GB; the smallest chunk is always a candidate */
    const int cand[7] = {256, 512, 1024, 2048, 4096, 8192, 16384};
    float best = 1e30f;
    for (int c = 0; c < 7 && rc == YK_OK; c++) {
      long nch = 0;
      for (int i = 0; i < ninst; i++) nch += (n_inst[i] + cand[c] - 1) / cand[c];
      if (c > 0 && (nch < 512 || (double)nch * k * d * 8 > 8e9)) continue;
      yh_allocs tal;
      tal.n = 0;
      yk_batch *bt = NULL;
      float ms = 0;
      rc = yh_batch_bind(&s, cand[c], &tal, &bt);
      if (rc == YK_OK) rc = yh_batch_init_state(&s);
      if (rc == YK_OK) rc = yk_batch_reset(bt);
      if (rc == YK_OK) rc = yk_batch_run(bt, 1, 0); /* warm-up */
      /* the faster of two timings of three iterations each */
      ms = 1e30f;
      for (int rep = 0; rep < 2 && rc == YK_OK; rep++) {
        float t = 0;
        rc = yh_batch_init_state(&s);
        if (rc == YK_OK) rc = yk_batch_reset(bt);
        if (rc == YK_OK) rc = yk_timer_start();
        if (rc == YK_OK) rc = yk_batch_run(bt, 3, 0);
        if (rc == YK_OK) rc = yk_timer_stop(&t);
        if (rc == YK_OK && t < ms) ms = t;
      }
      yk_batch_destroy(bt);
      yk_sync();
      yh_dfree_all(&tal);
      if (rc == YK_OK && ms < best) {
        best = ms;
        P = cand[c];
      }
    }
    if (rc == YK_OK) yk_tune_put(key, P);
  }
  if (chunk_used) *chunk_used = P;

  /* the run */
  yk_batch *bt = NULL;
  yh_allocs bal;
  bal.n = 0;
  if (rc == YK_OK) rc = yh_batch_bind(&s, P, &bal, &bt);
  if (rc == YK_OK) rc = yh_batch_init_state(&s);
  if (rc == YK_OK) rc = yk_batch_reset(bt);
  /* window by window: while the device runs one, the host draws the next
     one's splits; then, past the last window or once every instance has
     stopped, the run ends */
  int *hflags = (int *)malloc(sizeof(int) * (7L * ninst + 8));
  if (rc == YK_OK && !hflags) rc = yk_set_error_c(YK_ENOMEM, "batched k-means: host malloc");
  for (long w0 = 0; rc == YK_OK && w0 < niter; w0 += win) {
    const int nw = (int)(niter - w0 < win ? niter - w0 : win);
    if (w0 > 0) {
      int w0i = (int)w0;
      rc = yk_memcpy_h2d(s.rd, rd, sizeof(double) * (size_t)ninst * win * YH_ECAP);
      if (rc == YK_OK)
        rc = yk_memcpy_h2d(s.eps, eps, sizeof(float) * (size_t)ninst * win * YH_ECAP * d);
      if (rc == YK_OK) rc = yk_memcpy_h2d(s.flags + 6L * ninst + 1, &w0i, sizeof(int));
    }
    if (rc == YK_OK) rc = yk_batch_run(bt, nw, use_graph);
    if (rc != YK_OK || w0 + nw >= niter) break;
    yh_prep_run(&prep); /* the next window's draws, while the device runs */
    if (prep.failed) rc = yk_set_error_c(YK_ENOMEM, "batched k-means: host malloc");
    if (rc == YK_OK) rc = yk_memcpy_d2h(hflags, s.flags, sizeof(int) * (size_t)ninst);
    int all_done = 1;
    for (int i = 0; rc == YK_OK && i < ninst && all_done; i++) all_done = hflags[i] != 0;
    if (all_done) break;
  }
  if (rc == YK_OK) rc = yk_sync();
  free(rd);
  free(eps);
  free(chain);

  /* results */
  double *hq = (double *)malloc(sizeof(double) * 2L * ninst);
  if (rc == YK_OK) rc = yk_memcpy_d2h(hflags, s.flags, sizeof(int) * (7L * ninst + 8));
  if (rc == YK_OK) rc = yk_memcpy_d2h(hq, s.qerr2, sizeof(double) * 2L * ninst);
  if (rc == YK_OK && cent_out) {
    float *hc = (float *)malloc(sizeof(float) * (size_t)ninst * kpad * ld);
    rc = yk_memcpy_d2h(hc, s.cent, sizeof(float) * (size_t)ninst * kpad * ld);
    for (int i = 0; rc == YK_OK && i < ninst; i++)
      for (int c = 0; c < k; c++)
        memcpy(cent_out + ((long)i * k + c) * d, hc + ((long)i * kpad + c) * ld,
               sizeof(float) * d);
    free(hc);
  }
  if (rc == YK_OK && dis_out) rc = yk_memcpy_d2h(dis_out, s.dis, sizeof(float) * s.ntot);
  if (rc == YK_OK && assign_out) rc = yk_memcpy_d2h(assign_out, s.assign, sizeof(int) * s.ntot);
  if (rc == YK_OK && nassign_out)
    rc = yk_memcpy_d2h(nassign_out, s.nassign, sizeof(int) * (size_t)ninst * k);
  for (int i = 0; rc == YK_OK && i < ninst; i++) {
    const int *overflow = hflags + ninst, *iters = hflags + 4L * ninst,
              *aband = hflags + 5L * ninst;
    const long n = s.off[i + 1] - s.off[i];
    if (iters_out) iters_out[i] = iters[i];
    if (ret_out) ret_out[i] = aband[i] ? -1.f : (float)(hq[i] / n);
    if (overflow[i]) {
      /* more empties in one iteration than precomputed draws: the exact
         single-instance path recomputes this instance */
      float *vi = (float *)malloc(sizeof(float) * n * d);
      if (v_on_device)
        rc = yk_download_vectors(vi, vdev + s.off[i] * ld, n, d);
      else
        memcpy(vi, v + s.off[i] * d, sizeof(float) * n * d);
      float *ci = (float *)malloc(sizeof(float) * (size_t)k * d);
      if (ci && cent_init) memcpy(ci, cent_init + (size_t)i * k * d, sizeof(float) * (size_t)k * d);
      float r = 0;
      if (rc == YK_OK)
        rc = yh_kmeans(d, (int)n, k, niter_in, vi, 0, flags | KMEANS_QUIET, seed_used[i], 1, ci,
                       dis_out ? dis_out + s.off[i] : NULL,
                       assign_out ? assign_out + s.off[i] : NULL,
                       nassign_out ? nassign_out + (long)i * k : NULL, &r);
      if (rc == YK_OK) {
        if (cent_out) memcpy(cent_out + (long)i * k * d, ci, sizeof(float) * (size_t)k * d);
        if (ret_out) ret_out[i] = r;
        if (iters_out) iters_out[i] = -1;
      }
      free(ci);
      free(vi);
    }
  }
  free(hflags);
  free(hq);
  free(seed_used);
  free(cent_init);
  yk_batch_destroy(bt);
  yk_sync();
  yh_dfree_all(&bal);
  yh_dfree_all(&al);
  free(s.off);
  return rc;
}

/* Public entry points (declared in include/yael_mi355x.h). */
int yael_kmeans_batched(int d, int ninst, const int *n_inst, int k, int niter, const float *v,
                        int flags, const long *seeds, float *centroids, float *dis, int *assign,
                        int *nassign, float *ret) {
  yh_lock();
  int rc = yh_kmeans_batched(d, ninst, n_inst, k, niter, v, 0, flags, seeds, centroids, dis,
                             assign, nassign, ret, NULL, 1, NULL);
  yh_unlock();
  return rc;
}

int yael_kmeans_batched_device(int d, int ninst, const int *n_inst, int k, int niter,
                               const float *v_dev, int flags, const long *seeds,
                               float *centroids, float *dis, int *assign, int *nassign,
                               float *ret, int *iters, int use_graph, int *chunk_used) {
  yh_lock();
  int rc = yh_kmeans_batched(d, ninst, n_inst, k, niter, v_dev, 1, flags, seeds, centroids, dis,
                             assign, nassign, ret, iters, use_graph, chunk_used);
  yh_unlock();
  return rc;
}

/* Product-quantiser training: nsub k-means, sub-space j being columns
 * [j*dsub, (j+1)*dsub) of the n device-resident points (padded rows of
 * yk_ld(d) floats, d = nsub*dsub), all advanced in one batched run (instance
 * j = sub-space j, seed seeds[j]).  Outputs as yael_kmeans_batched_device
 * with instances sub-space-major: centroids nsub*k*dsub, dis/assign nsub*n,
 * nassign nsub*k, ret/iters nsub. */
int yael_kmeans_subspaces_device(int d, int n, int nsub, int k, int niter, const float *v_dev,
                                 int flags, const long *seeds, float *centroids, float *dis,
                                 int *assign, int *nassign, float *ret, int *iters,
                                 int *chunk_used) {
  if (nsub < 1 || d % nsub != 0 || n < 1)
    return yk_set_error_c(YK_EINVAL, "subspaces: d=%d is not a multiple of nsub=%d", d, nsub);
  const int dsub = d / nsub, ld = yk_ld(d), lds = yk_ld(dsub);
  yh_lock();
  int rc = yk_init();
  float *vs = NULL;
  int *sizes = (int *)malloc(sizeof(int) * nsub);
  if (rc == YK_OK && !sizes) rc = yk_set_error_c(YK_ENOMEM, "host malloc");
  if (rc == YK_OK) rc = yk_malloc((void **)&vs, sizeof(float) * (size_t)nsub * n * lds);
  if (rc == YK_OK && lds != dsub) rc = yk_memset(vs, 0, sizeof(float) * (size_t)nsub * n * lds);
  for (int j = 0; j < nsub && rc == YK_OK; j++) {
    sizes[j] = n;
    rc = yk_copy_cols(vs + (size_t)j * n * lds, lds, v_dev + (size_t)j * dsub, ld, n, dsub);
  }
  if (rc == YK_OK)
    rc = yh_kmeans_batched(dsub, nsub, sizes, k, niter, vs, 1, flags, seeds, centroids, dis,
                           assign, nassign, ret, iters, 1, chunk_used);
  if (vs) yk_free(vs);
  free(sizes);
  yh_unlock();
  return rc;
}

/*
 * yk.h -- C-ABI between the C host layer (the host/ sources) and the HIP layer
 * (*.hip).  Plain pointers and sizes only.  Every function returns 0 on
 * success or a negative YK_E* code; yk_last_error() gives the message.
 *
 * Device layout ("padded fvec"): a set of n vectors of dimension d is stored
 * row-major with a row stride ld = yk_ld(d) floats (d rounded up to 16), the
 * pad columns zero.  Row i, component j sits at x[i*ld + j].  This is yael's
 * column-major d x n fvec with a padded leading dimension.
 */
#ifndef YK_H_INCLUDED
#define YK_H_INCLUDED

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  YK_OK = 0,
  YK_EINVAL = -1,   /* bad argument (size, type, k > limit) */
  YK_ENOMEM = -2,   /* device allocation failed */
  YK_EHIP = -3,     /* HIP runtime error */
  YK_ENODEV = -4,   /* no usable GPU */
  YK_EABANDON = -5  /* k-means gave up after too many empty-cluster splits */
};

/* Distance types served on the GPU (nn.h:165-171): 2 and 12 squared L2,
 * 16 dot product (min-k). */
#define YK_L2 2
#define YK_DOT 16

/* Largest k the fused selection keeps on chip. */
#define YK_KMAX 1024

const char *yk_last_error(void);
/* set the message returned by yk_last_error; returns code */
int yk_set_error_c(int code, const char *fmt, ...);
int yk_init(void);
int yk_ld(int d);

/* device memory and copies (the library's own stream) */
int yk_malloc(void **p, size_t bytes);
int yk_free(void *p);
int yk_upload_vectors(float *dst, const float *src, long n, int d);  /* host n x d -> padded */
int yk_download_vectors(float *dst, const float *src, long n, int d); /* padded -> host n x d */
/* page-locked, coherent, mapped host memory (fast device-to-host copies;
 * kernels may write it directly) */
int yk_host_alloc(void **p, size_t bytes);
int yk_host_free(void *p);
int yk_memcpy_h2d(void *dst, const void *src, size_t bytes);
int yk_memcpy_d2h(void *dst, const void *src, size_t bytes);
int yk_memcpy_d2d(void *dst, const void *src, size_t bytes);
int yk_memset(void *dst, int value, size_t bytes);
/* device copy by a plain kernel (16 bytes per lane; bytes % 16 == 0) */
int yk_copy_kernel(void *dst, const void *src, size_t bytes);
/* device 2-D copy: columns [0, ncols) of n rows, row strides ldd / lds floats */
int yk_copy_cols(float *dst, int ldd, const float *src, int lds, long n, int ncols);
int yk_sync(void);

/* Norm kernel: out_f[i] = float-accumulated ||x_i||^2 (nn.c:108-114) when
 * out_f != NULL; out_d[i] = double-accumulated sum of float squares
 * (nn.c:116-120) when out_d != NULL. */
int yk_norms(const float *x, long n, int d, float *out_f, double *out_d);

/* Fused distance + k-smallest selection, FP32 on MFMA, device pointers.
 * q: nq padded queries with double norms qn; b: nb padded base vectors with
 * float norms bn (ignored for type 16); w: optional base weights.
 * Output row-major nq x k: assign (int32) and dis (float32), sorted.      */
int yk_knn_f32(int type, long nq, long nb, int d, int k, const float *q,
               const double *qn, const float *b, const float *bn, const float *w,
               int *assign, float *dis);

/* Precision form.  yk_to_bf16 writes the BF16 copy of n padded vectors
 * (row stride yk_bf16_ld(d, type) uint16; for L2 the squared norm is appended
 * in three BF16 pieces, queries are stored as -2q) and per-vector statistics
 * stats[4*i..] = {|x|, |bf(x)|, |x - bf(x)|, |x|^2}; stats_max (4 floats,
 * optional) receives their maxima over the set (NaN if any value is NaN);
 * fnorm (base sets only,
 * optional) the float norms exactly as yk_norms writes them.
 * yk_knn_bf16: BF16 MFMA first pass keeps the kp best candidates per query,
 * re-ranks them with the FP32 form's arithmetic (bit-identical values) and
 * certifies the result; uncertified queries are searched with the FP32 form
 * (*n_fallback of them).  Output as yk_knn_f32. */
int yk_bf16_ld(int d, int type);
/* bytes of a BF16 copy of n vectors: queries row-major (n rows of
 * yk_bf16_ld), base fragment-major (rows padded to a multiple of 32; block
 * b, chunk c is the 1 KB MFMA fragment at ((b * ldh/16 + c) * 64) * 8
 * elements, lane h*32 + r holding row 32b + r, columns 16c + 8h .. +7) */
long yk_bf16_bytes(long n, int d, int type, int is_query);
int yk_to_bf16(const float *x, long n, int d, int type, int is_query, uint16_t *out,
               float *stats, float *stats_max, float *fnorm);
int yk_knn_bf16(int type, long nq, long nb, int d, int k, int kp, const float *q,
                const double *qn, const uint16_t *qh, const float *qstats, const float *b,
                const float *bn, const uint16_t *bh, const float *bstats_max, int *assign,
                float *dis, int *n_fallback);

/* Synthetic sets generated in HBM (counter-based, reproducible from seed):
 * n rows of stride ld, pad columns zero.  kind 0: uniform [lo, hi); 1:
 * mixture -- centre of a hashed component (centres: ncomp x d on the device)
 * + sigma * N(0,1), clipped to [lo, hi], rounded when round_values; 2:
 * |N(0,1)|.  unit_rows scales every row to unit L2 norm. */
int yk_synth(float *out, long n, int d, int ld, int kind, float lo, float hi,
             const float *centres, int ncomp, float sigma, int round_values,
             unsigned long long seed, int unit_rows);

/* Cross distances, dist2[i + ldd*j] for a (na, float norms) x b (nb, double
 * norms), types 2/12/16 (nn.c:92-129, 266-306). */
int yk_cross_distances(int type, long na, long nb, int d, const float *a,
                       const float *an, const float *b, const double *bnd,
                       float *dist2, long ldd);

/* FP64 pair distances for the error report: out[i*k+r] = exact distance of
 * query i to base assign[i*k+r] (NaN where assign < 0). */
int yk_pair_dist_f64(int type, long nq, int d, int k, const float *q,
                     const float *b, const int *assign, double *out);

/* Shortlist re-rank (nn.c:528-580) in place: idx/dis are n x k device arrays,
 * b and v padded device vectors. */
int yk_rerank_shortlist(long n, int d, int k, const float *b, const float *v, int *idx,
                        float *dis);

/* Window re-computation (nn.c:583-600): for each query q, entries kp[q] ..
 * of its sorted shortlist idx (nq x k labels) that lie in [label0, label0 +
 * nb) -- b holds those nb rows, padded -- get their difference-form distance
 * (double accumulation, rounded to float) in dis; kp[q] advances to the
 * first entry past the window.  *bad (device int, set to INT_MAX by the
 * caller) receives the smallest query whose next entry lies below the
 * window (the reference asserts there). */
int yk_recompute_exact_dists(long nq, long nb, int d, int k, const float *b, const float *v,
                             int label0, int *kp, const int *idx, float *dis, int *bad);

/* k-means device steps (kmeans.c:242-310).
 * update: counts[k], centroids[k x ld] = (float)(FP64 sum / count) in point
 * order of each cluster; FP64 objective of dis.  ws: workspace from
 * yk_kmeans_workspace_size. */
size_t yk_kmeans_workspace_size(long n, int k, int d);
int yk_kmeans_update(long n, int k, int d, const float *v, const int *assign,
                     float *cent, int *counts, int normalize, void *ws);
int yk_sum_f64(const float *x, long n, double *out_host, void *ws);
/* The update and the readback of an iteration with one synchronisation:
 * yk_kmeans_update_async enqueues the update without waiting (its check
 * that every point has a nearest centroid stays on the device), and
 * yk_kmeans_readback_checked reads that check back with counts_host[k] =
 * counts and *obj_host = the FP64 objective of dis (as yk_sum_f64, ws its
 * workspace; upd_ws: the update's workspace, for n points and k clusters),
 * returning YK_EINVAL as yk_kmeans_update does when it failed.  One kernel
 * writes the values into a mapped pinned buffer of the library's (one per
 * process, grown to the largest k, calls serialised by a lock). */
int yk_kmeans_update_async(long n, int k, int d, const float *v, const int *assign,
                           float *cent, int *counts, int normalize, void *ws);
int yk_kmeans_readback_checked(const int *counts, int k, int *counts_host, const float *dis,
                               long n, double *obj_host, void *ws, void *upd_ws);
/* k-means++ step: best[j] = min(best[j], squared distance of point j to
 * point s) for j < n (first: best starts at HUGE_VAL); vn: the points'
 * double norms (yk_norms out_d); host_copy (optional, from yk_host_alloc)
 * receives best as well, valid after yk_sync */
int yk_kpp_update(const float *v, long n, int d, const double *vn, long s, float *best, int first,
                  float *host_copy);
/* hkm tree descent: idx[i] = leaf of point i (v padded, vn double norms);
 * cent/cn: the levels' centroids stacked (level l: bf^(l+1) padded rows)
 * and their float norms */
int yk_hkm_quantize(const float *v, const double *vn, long n, int d, int nlevel, int bf,
                    const float *cent, const float *cn, int *idx);
/* dst row i = src row idx[i] (padded rows; idx is a device array) */
int yk_gather_rows(const float *src, const int *idx, long n, int d, float *dst);

/* Batched form (yk_batch.hip): many k-means instances advanced together.
 * yk_batch_assign: the FP32 assign step over work items (instance, 128-point
 * tile), centroid sets kpad rows apart.  yk_batch_create binds device
 * arrays laid out by the caller (see host/yael_batch.c; psum holds
 * nuitems * (k * d + 1) doubles: the chunks' partial sums, then their
 * objective partials; flags holds 7 x ninst ints, then the iteration
 * counter, then w0).  The split draws rd / eps cover a window of win
 * iterations per instance (row it - w0 of instance i, w0 the window's first
 * iteration); the caller refills them and w0 between windows.  k * d tables
 * larger than the LDS are summed in cluster slices; normalize scales every
 * centroid to unit L2 norm after its update.  yk_batch_run queues
 * niter iterations, replaying one captured hipGraph when use_graph. */
typedef struct yk_batch yk_batch;
int yk_batch_assign(long nitems, const int *item_inst, const int *item_tile, const long *inst_off,
                    int d, int kc, int kpad, const float *v, const double *vn, const float *cent,
                    const float *cn, const int *done, int *assign, float *dis);
size_t yk_batch_lds_bytes(int k, int d);
int yk_batch_create(yk_batch **out, int ninst, int d, int k, int kpad, int niter, int win,
                    int ecap, int chunk, long ntot, const float *v, const double *vn, float *cent,
                    float *cn, int *assign, float *dis, const long *inst_off, long nitems,
                    const int *item_inst, const int *item_tile, long nuitems,
                    const int *uitem_inst, const int *uitem_first, double *psum, int *pcnt,
                    int *nassign, float *pbuf, const double *rd, const float *eps, int *flags,
                    double *qerr2, int normalize);
int yk_batch_reset(yk_batch *bt);
int yk_batch_run(yk_batch *bt, int niter, int use_graph);
int yk_batch_destroy(yk_batch *bt);

/* Autotuning cache of launch plans: "key value" lines in a text file
 * ($YAEL_MI355X_TUNE_CACHE, default ~/.cache/yael_mi355x_tune.txt), held in
 * memory after the first read.  The plans it records are schedules (work
 * splits, sampling strides, update chunks): every plan gives the same bits.
 * yk_tune_enabled: 0 when YAEL_MI355X_TUNE=0 (default plans, no timing). */
int yk_tune_get(const char *key, int *val);
void yk_tune_put(const char *key, int val);
int yk_tune_enabled(void);

/* Timing of the library's stream (HIP events recorded on it). */
int yk_timer_start(void);
int yk_timer_stop(float *ms);
/* Per-kernel timing on the library stream: when enabled, launches are
 * bracketed by events and summed per slot: 0 FP32 scan (filtered pass, k=1
 * and register top-k), 1 norms, 2 k-means update, 3 merges/selects,
 * 4 re-rank, 5 batched iteration, 6 BF16 scan (filtered pass, top-k),
 * 7 BF16 sketch pass, 8 FP32 sketch pass. */
int yk_prof_enable(int on);
int yk_prof_get(int which, double *total_ms, long *count);
int yk_prof_reset(void);

/* Descriptive strings (device name, build) */
int yk_device_info(char *buf, int len);

/* Test hook: the chosen launch configuration for a kNN shape. */
int yk_knn_plan(long nq, long nb, int d, int k, int *nsplit, int *waves,
                int *nblk);

#ifdef __cplusplus
}
#endif
#endif

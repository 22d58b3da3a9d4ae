/*
 * yael_mi355x.h -- public C-ABI of libyael_mi355x.so (yael_amd/_lib/).
 *
 * Part 1 is yael's own API for the k-means / exhaustive-kNN path: same
 * names, signatures, argument order (base before queries), layout
 * (column-major d x n fvecs = row-major n x d floats) and error behaviour
 * (abort on misuse, kmeans returns -1 on abandon).  Each declaration cites
 * the reference declaration it replaces.  Thread-count arguments are
 * accepted and ignored: the GPU grid replaces yael's OpenMP slices.
 *
 * Part 2 are status-returning twins of part 1 (same arguments) for callers
 * that want an error code instead of an abort.
 *
 * Part 3 is the device-level layer (device pointers, the library's HIP
 * stream), used by the benchmark and by callers that keep data resident in
 * HBM.  It is declared in yael_amd/csrc/yk.h, which this header includes.
 */
#ifndef YAEL_MI355X_H_INCLUDED
#define YAEL_MI355X_H_INCLUDED

#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------- part 1: yael's API ---------------- */

/* kmeans flags, kmeans.h:11-18 */
#define KMEANS_QUIET 0x10000
#define KMEANS_INIT_BERKELEY 0x20000 /* k-means++ (kmeans.c:27-82) */
#define KMEANS_NORMALIZE_CENTS 0x40000
#define KMEANS_INIT_RANDOM 0x80000
#define KMEANS_INIT_USER 0x100000
#define KMEANS_L1 0x200000   /* not served on the GPU: error */
#define KMEANS_CHI2 0x400000 /* not served on the GPU: error */
/* Assign-step precision; no counterpart in kmeans.h.  YAEL_KMEANS_BF16: BF16
 * MFMA first pass keeping each point's 8 best centroids, exact FP32 re-rank
 * and a certificate, FP32 search for uncertified points -- the result is the
 * FP32 form's bit for bit.  YAEL_KMEANS_FP32: FP32 form only.  Neither: BF16
 * when k >= 4096. */
#define YAEL_KMEANS_BF16 0x1000000
#define YAEL_KMEANS_FP32 0x2000000

/* replaces kmeans.h:41-44 (kmeans.c:332-447) */
float kmeans(int d, int n, int k, int niter, const float *v, int flags, long seed, int redo,
             float *centroids, float *dis, int *assign, int *nassign);

/* replaces kmeans.h:49-51, 56-59, 63-66 (kmeans.c:452-504) */
float *clustering_kmeans(int n, int d, const float *points, int k, int nb_iter_max,
                         double normalize);
float *clustering_kmeans_assign(int n, int d, const float *points, int k, int nb_iter_max,
                                double normalize, int **clust_assign_out);
float *clustering_kmeans_assign_with_score(int n, int d, const float *points, int k,
                                           int nb_iter_max, double normalize, int n_thread,
                                           double *score_out, int **clust_assign_out);

/* replaces nn.h:41-45 (nn.c:451-525); distance_type 2/12 (L2) or 16 (dot) */
void knn_full(int distance_type, int nq, int nb, int d, int k, const float *b, const float *q,
              const float *b_weights, int *assign, float *dis);

/* replaces nn.h:49-54 (nn.c:679-699) */
void knn_full_thread(int distance_type, int nq, int nb, int d, int k, const float *b,
                     const float *q, const float *b_weights, int *assign, float *dis,
                     int n_thread);

/* replace nn.h:60-83 (nn.c:608-632, 704-726) */
double nn(int n, int nb, int d, const float *b, const float *v, int *assign);
double nn_thread(int n, int nb, int d, const float *b, const float *v, int *assign,
                 int n_thread);
float *knn(int n, int nb, int d, int k, const float *b, const float *v, int *assign);
float *knn_thread(int nq, int nb, int d, int k, const float *b, const float *v, int *assign,
                  int n_thread);

/* replaces nn.h:102-104 (nn.c:528-580) */
void knn_reorder_shortlist(int n, int nb, int d, int k, const float *b, const float *v,
                           int *idx, float *dis);

/* replaces nn.h:127-130 (nn.c:583-600): the shortlist distances of a
 * partial base b (labels label0 .. label0 + nb - 1), per query from entry
 * kp[q] until the first entry past that window; kp advances there */
void knn_recompute_exact_dists(int nq, int nb, int d, int k, const float *b, const float *v,
                               int label0, int *kp, const int *idx, float *dis);

/* replaces nn.h:141-143 (nn.c:92-97) */
void compute_cross_distances(int d, int na, int nb, const float *a, const float *b,
                             float *dist2);
/* replaces nn.h:151-154 (nn.c:100-129) */
void compute_cross_distances_nonpacked(int d, int na, int nb, const float *a, int lda,
                                       const float *b, int ldb, float *dist2, int ldd);
/* replaces nn.h:156-159 (nn.c:777-792) */
void compute_cross_distances_thread(int d, int na, int nb, const float *a, const float *b,
                                    float *dist2, int nt);
/* replaces nn.h:174-176, 180-183, 188-191 (nn.c:280-356, 795-810); types 2/12/16 */
void compute_cross_distances_alt(int distance_type, int d, int na, int nb, const float *a,
                                 const float *b, float *dist2);
void compute_cross_distances_alt_nonpacked(int distance_type, int d, int na, int nb,
                                           const float *a, int lda, const float *b, int ldb,
                                           float *dist2, int ldd);
void compute_cross_distances_alt_thread(int distance_type, int d, int na, int nb,
                                        const float *a, const float *b, float *dist2, int nt);

/* hierarchical k-means: the tree, same layout as hkm.h:10-16.  Level l
 * holds bf^(l+1) centroids (d floats each); the children of node p of
 * level l are its rows p*bf .. p*bf + bf - 1. */
typedef struct hkm_s {
  int nlevel;        /* number of levels */
  int bf;            /* branching factor */
  int k;             /* number of leaves, bf^nlevel */
  int d;             /* dimension of the vectors */
  float **centroids; /* nlevel arrays */
} hkm_t;

/* replaces hkm.h:21-23 (hkm.c:36-116); every level is one batched k-means
 * on the GPU (node seeds drawn from lrand48() in node order, as the
 * reference's kmeans(seed 0) calls draw them) */
hkm_t *hkm_learn(int n, int d, int nlevel, int bf, const float *v, int nb_iter_max, int nt,
                 int verbose, int **clust_assign_out);
/* replaces hkm.h:25 (hkm.c:119-126) */
void hkm_delete(hkm_t *hkm);
/* replaces hkm.h:28 (hkm.c:131-150): tree descent on the GPU */
void hkm_quantize(const hkm_t *hkm, int n, const float *v, int *idx);
/* replace hkm.h:31-32 (hkm.c:181-232): the reference's file format */
void hkm_write(const char *filename, const hkm_t *hkm);
hkm_t *hkm_read(const char *filename);
/* replaces hkm.h:35 (hkm.c:154-157) */
float *hkm_get_centroids(const hkm_t *hkm, int l, int no);

/* vector files (doc/file_format.rst; vector.h I/O): fvecs = per vector int
 * d then d floats, ivecs = int d then d ints, bvecs = int d then d bytes,
 * text = whitespace-separated numbers.  Host-side; same return values as
 * vector.c (vectors read or written, -1 on error). */
/* replace vector.c:628-641 */
long fvecs_fsize(const char *fname, int *d_out, int *n_out);
long ivecs_fsize(const char *fname, int *d_out, int *n_out);
long bvecs_fsize(const char *fname, int *d_out, int *n_out);
/* replaces vector.c:882-920 */
int fvecs_read(const char *fname, int d, int n, float *v);
/* replaces vector.c:923-936 (bvecs read as floats) */
int b2fvecs_read(const char *fname, int d, int n, float *v);
/* replaces vector.c:939-962 */
int fvecs_read_txt(const char *fname, int d, int n, float *v);
/* replace vector.c:651-666, 1307-1322 (malloc'd arrays) */
int fvecs_new_read(const char *fname, int *d_out, float **vf_out);
int ivecs_new_read(const char *fname, int *d_out, int **vi_out);
/* replace vector.c:1369-1383, 1446-1456, 1459-1472, 1475-1492 */
int fvec_fwrite(FILE *fo, const float *v, int d);
int fvecs_fwrite(FILE *fo, int d, int n, const float *vf);
int fvecs_write(const char *fname, int d, int n, const float *vf);
int fvecs_write_txt(const char *fname, int d, int n, const float *vf);
/* replace vector.c:1494-1508, 1513-1520, 1524-1537, 1279-1295 */
int ivec_fwrite(FILE *f, const int *v, int d);
int ivecs_fwrite(FILE *f, int d, int n, const int *v);
int ivecs_write(const char *fname, int d, int n, const int *v);
int ivecs_write_txt(const char *fname, int d, int n, const int *v);
/* replaces vector.c:1955-1964 */
long fvec_purge_nans(float *v, long n, float replace_value);

/* ---------------- part 2: status-returning twins (0 = ok, < 0 = YK_E*) ---------------- */

/* hkm_learn / hkm_quantize returning an error code (the tree in *out) */
int yael_hkm_learn_checked(int n, int d, int nlevel, int bf, const float *v, int nb_iter_max,
                           int nt, int verbose, int **clust_assign_out, hkm_t **out);
int yael_hkm_quantize_checked(const hkm_t *hkm, int n, const float *v, int *idx);

/* Form of knn_full / knn_full_thread / yael_knn_full_checked (process-wide;
 * returns the previous mode): 0 automatic -- the precision form (BF16 first
 * pass, exact FP32 re-rank, certificate, fallback; the FP32 form's results
 * bit for bit) for unweighted searches of at least 256 queries and 2^26
 * pairs on finite data, k <= 512 --, 1 the FP32 form only, 2 the precision
 * form whenever it applies.  No counterpart in the reference. */
int yael_knn_precision(int mode);
int yael_knn_full_checked(int distance_type, int nq, int nb, int d, int k, const float *b,
                          const float *q, const float *b_weights, int *assign, float *dis);
int yael_cross_distances_checked(int distance_type, int d, int na, int nb, const float *a,
                                 int lda, const float *b, int ldb, float *dist2, int ldd);
int yael_kmeans_checked(int d, int n, int k, int niter, const float *v, int flags, long seed,
                        int redo, float *centroids, float *dis, int *assign, int *nassign,
                        float *ret);
/* kmeans() on vectors resident in HBM (v_dev: n rows of stride yk_ld(d) floats,
 * pad columns zero); outputs are host arrays as for kmeans().  No
 * counterpart in the reference: for callers that keep the training set on
 * the device across calls. */
int yael_kmeans_device(int d, int n, int k, int niter, const float *v_dev, int flags, long seed,
                       int redo, float *centroids, float *dis, int *assign, int *nassign,
                       float *ret);
/* Counters of the last kmeans call in this process: Lloyd iterations,
 * point assignments, assignments the BF16 first pass could not certify
 * (searched again with the FP32 form).  No counterpart in the reference. */
int yael_kmeans_last_stats(long *iters, long *assigned, long *fallback);
int yael_knn_reorder_shortlist_checked(int n, int nb, int d, int k, const float *b,
                                       const float *v, int *idx, float *dis);
int yael_knn_recompute_exact_dists_checked(int nq, int nb, int d, int k, const float *b,
                                           const float *v, int label0, int *kp, const int *idx,
                                           float *dis);

/* Batched form: ninst independent k-means instances, instance i being
 * kmeans(d, n_inst[i], k, niter, v_i, flags, seeds[i], 1, ...) on the rows
 * v_i of the concatenated point set v (sum n_inst rows), as hkm_learn
 * clusters every tree node (hkm.c:66-103).  Outputs concatenated in instance
 * order: centroids ninst*k rows, dis/assign per point (assign local to the
 * instance), nassign ninst*k, ret ninst values (-1 = abandoned).  No
 * counterpart in the reference (which loops over kmeans()); flags: QUIET,
 * INIT_RANDOM and the thread bits only. */
int yael_kmeans_batched(int d, int ninst, const int *n_inst, int k, int niter, const float *v,
                        int flags, const long *seeds, float *centroids, float *dis, int *assign,
                        int *nassign, float *ret);
/* the same on device-resident points (padded rows of yk_ld(d) floats);
 * iters (per instance, may be NULL), use_graph (1: replay one captured
 * hipGraph per iteration), chunk_used (the autotuned update chunk) */
int yael_kmeans_batched_device(int d, int ninst, const int *n_inst, int k, int niter,
                               const float *v_dev, int flags, const long *seeds,
                               float *centroids, float *dis, int *assign, int *nassign,
                               float *ret, int *iters, int use_graph, int *chunk_used);
/* Product-quantiser training: nsub k-means, sub-space j being the columns
 * [j*d/nsub, (j+1)*d/nsub) of the n device-resident points, run as one
 * batched k-means (instance j = sub-space j, seed seeds[j]).  Outputs as
 * yael_kmeans_batched_device, sub-space-major.  No counterpart in the
 * reference. */
int yael_kmeans_subspaces_device(int d, int n, int nsub, int k, int niter, const float *v_dev,
                                 int flags, const long *seeds, float *centroids, float *dis,
                                 int *assign, int *nassign, float *ret, int *iters,
                                 int *chunk_used);

#ifdef __cplusplus
}
#endif

/* ---------------- part 3: device-level layer ---------------- */
#include "../yael_amd/csrc/yk.h"

/* ---------------- part 4: search without allocation ---------------- */

#ifdef __cplusplus
extern "C" {
#endif

/* Size of the static device array the no-allocation search takes its
 * scratch from (every GPU the library runs on holds one copy until the
 * process ends). */
#define YK_NOALLOC_SCRATCH_BYTES (64u << 20)

/* yk_knn_f32 without allocating: the same arguments (device pointers, the
 * library's stream) and results bit for bit, with the search's scratch
 * (partial keys, thresholds, counters, candidate lists) taken from the
 * static 64 MiB array above instead of library-allocated buffers; no
 * hipMalloc, hipFree or allocating library call is made (the first call on
 * a device looks up that device's copy of the array).  Queries go in passes
 * sized to fit the array; a shape whose scratch cannot fit even in a pass
 * of 32 queries is rejected with YK_EINVAL.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The array keeps the last call's contents and is shared by
 * every call on that GPU: calls are ordered on the library's stream and
 * serialised by the library lock; kernels queued from another stream
 * while a call's kernels run would race on it. */
int yk_knn_f32_noalloc(int type, long nq, long nb, int d, int k, const float *q,
                       const double *qn, const float *b, const float *bn, const float *w,
                       int *assign, float *dis);

#ifdef __cplusplus
}
#endif

#endif

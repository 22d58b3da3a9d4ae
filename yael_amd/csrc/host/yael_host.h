/* yael_host.h -- internal declarations of the C host layer. */
#ifndef YAEL_HOST_H_INCLUDED
#define YAEL_HOST_H_INCLUDED

#include <stddef.h>

/* device buffer roles */
enum {
  YH_B = 0, YH_Q, YH_BN, YH_QN, YH_W, YH_OUTI, YH_OUTD,
  YH_KM_V, YH_KM_QN, YH_KM_C, YH_KM_CN, YH_KM_A, YH_KM_D, YH_KM_CNT, YH_KM_WS,
  YH_KM_BEST_C, YH_KM_SEL, YH_KM_VH, YH_KM_QS, YH_KM_CH, YH_KM_CS,
  YH_BH, YH_QH, YH_BST, YH_QST, YH_MX,
  YH_NSLOTS
};

void yh_lock(void);
void yh_unlock(void);
void yh_die(const char *where);
void *yh_buf(int slot, size_t bytes);
void yh_release_all(void);

int yh_knn(int type, int nq, int nb, int d, int k, const float *b, const float *q,
           const float *w, int *assign, float *dis);
int yh_cross(int type, int d, int na, int nb, const float *a, int lda, const float *b,
             int ldb, float *dist2, int ldd);

int yh_kmeans(int d, int n, int k, int niter, const float *v, int v_on_device, int flags,
              long seed_in, int redo, float *cent_out, float *dis_out, int *assign_out,
              int *nassign_out, float *ret);

/* kmeans flags (kmeans.h:11-18) */
#define KMEANS_QUIET 0x10000
#define KMEANS_INIT_BERKELEY 0x20000
#define KMEANS_NORMALIZE_CENTS 0x40000
#define KMEANS_INIT_RANDOM 0x80000
#define KMEANS_INIT_USER 0x100000
#define KMEANS_L1 0x200000
#define KMEANS_CHI2 0x400000
/* assign-step precision (not in kmeans.h): BF16 first pass + exact FP32
 * re-rank of the 8 best (results identical to the FP32 form), or the FP32
 * form only; neither set: BF16 when k >= YH_KM_BF16_MIN_K */
#define YAEL_KMEANS_BF16 0x1000000
#define YAEL_KMEANS_FP32 0x2000000
#define YH_KM_BF16_MIN_K 4096
#define YH_KM_KPRIME 8

/* the batched form (host/yael_batch.c) without the library lock */
int yh_kmeans_batched(int d, int ninst, const int *n_inst, int k, int niter, const float *v,
                      int v_on_device, int flags, const long *seeds, float *cent_out,
                      float *dis_out, int *assign_out, int *nassign_out, float *ret_out,
                      int *iters_out, int use_graph, int *chunk_used);

/* empty-cluster split (kmeans.c:166-209) on host copies of the centroids */
int yh_reassign_empty(int d, int k, float *cent, long ldc, const int *nassign, unsigned seed);
/* first k entries of the partial Fisher-Yates of kmeans.c:15-20 */
void yh_random_init(int n, int k, int *sel, unsigned seed);
/* kmeans.c:22-24 and vector.c:139-154 */
double yh_drand_r(unsigned *seed);
double yh_gaussrand_r(unsigned *seed);

#endif

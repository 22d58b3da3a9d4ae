// yk_dev.h -- internal helpers shared by the HIP translation units.
#ifndef YK_DEV_H_INCLUDED
#define YK_DEV_H_INCLUDED

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "yk.h"

int yk_set_error(int code, const char *fmt, ...);
int yk_hip_fail(hipError_t e, const char *what);
hipStream_t yk_stream(void);
// dst = up to 256 bytes of device memory at src, after the stream's queued
// work, with one wait
int yk_read_small(void *dst, const void *src, size_t bytes);
int yk_prof_setup(void);
// Workgroups of `block` threads that fit on the whole device at once for
// kernel `fn` (occupancy x CU count), at least 8 and a multiple of 8.
int yk_resident_wgs(const void *fn, int block);
int yk_prof_begin(int slot);
void yk_prof_end(int tok);

// The library lock (recursive): the device-level searches take it for the
// whole call, because they share process-wide scratch buffers; the host
// layer's calls nest inside it.
void yk_lock(void);
void yk_unlock(void);
struct YkGuard {
  YkGuard() { yk_lock(); }
  ~YkGuard() { yk_unlock(); }
  YkGuard(const YkGuard &) = delete;
  YkGuard &operator=(const YkGuard &) = delete;
};

// The batched assign (yk_batch_assign) with an optional index table: point
// j of the instances' concatenation is row qidx[j] of v / vn (outputs stay
// in concatenation order).
int yk_batch_assign_idx(long nitems, const int *item_inst, const int *item_tile,
                        const long *inst_off, int d, int kc, int kpad, const float *v,
                        const double *vn, const int *qidx, const float *cent, const float *cn,
                        const int *done, int *assign, float *dis);

// The batched assign over the update chunks (chunk points of one instance
// per item, uitem_* as in yk_batch_create) with the centroids held in
// registers; 1 when the shape (d, kc) has no such kernel.
int yk_batch_assign_res(int ninst, int d, int kc, int kpad, int chunk, const float *v,
                        const double *vn, const float *cent, const float *cn,
                        const long *inst_off, long nuitems, const int *uitem_inst,
                        const int *uitem_first, const int *done, int *assign, float *dis);

// Exact kNN for k > YK_KMAX (yk_bigk.hip): distance rows per query batch,
// per-query radix select and segmented sort; arguments as yk_knn_f32.
int yk_knn_bigk(int type, long nq, long nb, int d, int k, const float *q, const double *qn,
                const float *b, const float *bn, const float *w, int *assign, float *dis);

// Milliseconds of run(arg) on the library stream (HIP events around it).
int yk_time_ms(int (*run)(void *), void *arg, float *ms);

// The persistent scans' per-XCD work-item ticket counters (8 ints used),
// zeroed on the library stream before the caller's next launch.
int yk_item_tickets(int **out);

// The filtered pass's starting bound (both forms, k' > 16): the m-th
// smallest sketch minimum, m the least count with stride * (m - 5 sqrt(m)) >=
// kp.  Over a sample of one block in `stride`, a bound with m sampled rows
// below it has about stride * m rows of the whole base below it, give or
// take stride * sqrt(m): five standard deviations above the kp needed (about
// 2 kp candidates per query where the kp-th minimum, the proven bound, lets
// through about 4.5 kp at stride 4).  A query left with fewer than kp keys
// is scanned again without a bound (then with the bound its overflow
// gives), so no result depends on m.
static inline int yk_bound_rank(int kp, long stride) {
  const double t = (double)kp / (double)stride;
  const double r = 0.5 * (5.0 + __builtin_sqrt(25.0 + 4.0 * t));
  const long m = (long)__builtin_ceil(r * r);
  return (int)(m < kp ? m : kp);
}

// profiling slots
enum { YK_PROF_FUSED = 0, YK_PROF_NORM = 1, YK_PROF_UPDATE = 2, YK_PROF_MERGE = 3,
       YK_PROF_RERANK = 4, YK_PROF_BATCH = 5, YK_PROF_BF16 = 6, YK_PROF_BF16_SKETCH = 7,
       YK_PROF_F32_SKETCH = 8 };

#define YK_TRY(x)                 \
  do {                            \
    int _rc = (x);                \
    if (_rc != YK_OK) return _rc; \
  } while (0)

#define YK_HIP(x)                                    \
  do {                                               \
    hipError_t _e = (x);                             \
    if (_e != hipSuccess) return yk_hip_fail(_e, #x); \
  } while (0)

// Order-preserving 32-bit image of a float (after -0 -> +0); NaN is never
// passed in.  Ascending key <=> ascending float.
__device__ __forceinline__ uint32_t yk_fkey(float v) {
  uint32_t u = __float_as_uint(v + 0.0f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float yk_fkey_inv(uint32_t k) {
  uint32_t u = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  return __uint_as_float(u);
}
// (value, index) packed so that unsigned order is lexicographic order.
__device__ __forceinline__ uint64_t yk_key(float v, int idx) {
  return ((uint64_t)yk_fkey(v) << 32) | (uint32_t)idx;
}

#define YK_KEY_NONE 0xffffffffffffffffull

#endif

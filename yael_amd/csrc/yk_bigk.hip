// yk_bigk.hip -- exact k-nearest neighbours for k beyond what the fused
// scans keep on chip (k > YK_KMAX; knn_full allows any k <= nb, nn.c:456).
//
// Per batch of queries the whole distance rows are formed by the cross-
// distance kernel (yk_cross_distances: the FP32 form's MFMA sequence and
// epilogue, so the values are the ones the fused scans compute), weights
// applied as the scans apply them (dist * w, in float).  Then per query, one
// workgroup:
//   * a radix select over the order-preserving 32-bit images of the values
//     (NaN excluded: it never enters, nn.c / binheap.c:139-156), four 8-bit
//     digits from the top, finds the kk-th smallest image T (kk = k, or the
//     number of non-NaN values when fewer);
//   * a pass in index order compacts every value below T and, of the values
//     equal to T, the ones with the smallest indices, into the query's kk
//     slots as 64-bit (value, index) keys;
// then one segmented radix sort (rocPRIM) orders every query's keys, and the
// output pass writes them, padded with (-1, NaN bits 0xffffffff) past kk
// (nn.c:515-518).  The ordering is the library's (value, index) order, the
// same as the fused scans' for k <= YK_KMAX.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_segmented_radix_sort.hpp>

#include "yk.h"
#include "yk_dev.h"

#define YK_BK_THREADS 1024

__device__ __forceinline__ uint32_t yk_bk_image(const float *row, const float *w, long i,
                                                bool &valid) {
  float v = row[i];
  if (w) v *= w[i];
  valid = v == v;
  return valid ? yk_fkey(v) : 0xffffffffu;
}

// exclusive prefix of x over the workgroup (in thread order); total in *tot
__device__ __forceinline__ int yk_bk_scan(int x, int *wsum, int *tot) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  int incl = x;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(incl, o);
    if (lane >= o) incl += y;
  }
  if (lane == 63) wsum[wave] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    int run = 0;
    for (int u = 0; u < nw; u++) {
      const int t = wsum[u];
      wsum[u] = run;
      run += t;
    }
    wsum[nw] = run;
  }
  __syncthreads();
  const int ex = wsum[wave] + incl - x;
  *tot = wsum[nw];
  __syncthreads();  // wsum is reused by the next call
  return ex;
}

__global__ __launch_bounds__(YK_BK_THREADS) void yk_bigk_select_kernel(
    const float *__restrict__ dist, long ldd, long nb, int k, const float *__restrict__ w,
    uint64_t *__restrict__ keys, int *__restrict__ kk_out) {
  __shared__ int hist[256];
  __shared__ int wsum[YK_BK_THREADS / 64 + 1];
  __shared__ uint32_t s_prefix;
  __shared__ int s_rank;
  const long q = blockIdx.x;
  const float *row = dist + q * ldd;
  const int tid = threadIdx.x;
  // non-NaN count
  int nv = 0;
  for (long i = tid; i < nb; i += blockDim.x) {
    bool ok;
    (void)yk_bk_image(row, w, i, ok);
    nv += ok;
  }
  int nvalid;
  (void)yk_bk_scan(nv, wsum, &nvalid);
  const int kk = nvalid < k ? nvalid : k;
  if (tid == 0) {
    s_prefix = 0;
    s_rank = kk;  // 1-based rank of T among the valid images
    kk_out[q] = kk;
  }
  __syncthreads();
  if (kk == 0) return;
  // radix select, most significant digit first
  uint32_t mask = 0;
  for (int shift = 24; shift >= 0; shift -= 8) {
    for (int i = tid; i < 256; i += blockDim.x) hist[i] = 0;
    __syncthreads();
    const uint32_t prefix = s_prefix;
    for (long i = tid; i < nb; i += blockDim.x) {
      bool ok;
      const uint32_t img = yk_bk_image(row, w, i, ok);
      if (ok && (img & mask) == prefix) atomicAdd(&hist[(img >> shift) & 255], 1);
    }
    __syncthreads();
    if (tid == 0) {
      int rank = s_rank, dg = 0;
      while (dg < 255 && rank > hist[dg]) {
        rank -= hist[dg];
        dg++;
      }
      s_rank = rank;
      s_prefix = prefix | ((uint32_t)dg << shift);
    }
    mask |= 255u << shift;
    __syncthreads();
  }
  const uint32_t T = s_prefix;
  const int need_eq = s_rank, nless = kk - need_eq;
  // compaction in index order
  uint64_t *out = keys + q * (long)k;
  int run_lt = 0, run_eq = 0;
  for (long c0 = 0; c0 < nb; c0 += blockDim.x) {
    const long i = c0 + tid;
    bool ok = false;
    const uint32_t img = i < nb ? yk_bk_image(row, w, i, ok) : 0u;
    const int lt = ok && img < T, eq = ok && img == T;
    int tlt, teq;
    const int plt = yk_bk_scan(lt, wsum, &tlt);
    const int peq = yk_bk_scan(eq, wsum, &teq);
    if (lt) out[run_lt + plt] = ((uint64_t)img << 32) | (uint32_t)i;
    if (eq && run_eq + peq < need_eq) out[nless + run_eq + peq] = ((uint64_t)img << 32) | (uint32_t)i;
    run_lt += tlt;
    run_eq += teq;
  }
}

__global__ void yk_bigk_bounds_kernel(const int *__restrict__ kk, long nq, int k,
                                      int *__restrict__ beg, int *__restrict__ end) {
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= nq) return;
  beg[q] = (int)(q * k);
  end[q] = (int)(q * k + kk[q]);
}

__global__ void yk_bigk_out_kernel(const uint64_t *__restrict__ keys, const int *__restrict__ kk,
                                   long nq, int k, int *__restrict__ assign,
                                   float *__restrict__ dis) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= nq * k) return;
  const long q = t / k;
  const int r = (int)(t - q * k);
  if (r < kk[q]) {
    const uint64_t key = keys[t];
    assign[t] = (int)(uint32_t)key;
    dis[t] = yk_fkey_inv((uint32_t)(key >> 32));
  } else {
    assign[t] = -1;
    dis[t] = __uint_as_float(0xffffffffu);
  }
}

// q: nq padded queries with double norms qn; b: nb padded base vectors with
// float norms bn; w optional weights; outputs nq x k.
int yk_knn_bigk(int type, long nq, long nb, int d, int k, const float *q, const double *qn,
                const float *b, const float *bn, const float *w, int *assign, float *dis) {
  YK_TRY(yk_init());
  if (nq <= 0) return YK_OK;
  if ((long)nq * k > 0x7fffffffL || nb > 0x7fffffffL)
    return yk_set_error(YK_EINVAL, "k=%d: nq*k exceeds int32", k);
  hipStream_t st = yk_stream();
  const int ld = yk_ld(d);
  // queries per batch: distance rows within 2 GB
  long qb = (2L << 30) / (nb * 4);
  if (qb < 1) qb = 1;
  if (qb > nq) qb = nq;
  float *D = nullptr;
  uint64_t *kin = nullptr, *kout = nullptr;
  int *kk = nullptr, *beg = nullptr, *end = nullptr;
  void *tmp = nullptr;
  size_t tmp_bytes = 0;
  int rc = yk_malloc((void **)&D, sizeof(float) * (size_t)qb * nb);
  if (rc == YK_OK) rc = yk_malloc((void **)&kin, sizeof(uint64_t) * (size_t)qb * k);
  if (rc == YK_OK) rc = yk_malloc((void **)&kout, sizeof(uint64_t) * (size_t)qb * k);
  if (rc == YK_OK) rc = yk_malloc((void **)&kk, sizeof(int) * (size_t)qb * 3);
  if (rc == YK_OK) {
    beg = kk + qb;
    end = kk + 2 * qb;
    if (rocprim::segmented_radix_sort_keys(nullptr, tmp_bytes, kin, kout, (unsigned)(qb * k),
                                           (unsigned)qb, beg, end, 0, 64, st) != hipSuccess) {
      (void)hipGetLastError();
      rc = yk_set_error(YK_EHIP, "segmented radix sort: temporary size query failed");
    }
  }
  if (rc == YK_OK) rc = yk_malloc(&tmp, tmp_bytes);
  for (long q0 = 0; q0 < nq && rc == YK_OK; q0 += qb) {
    const long n = nq - q0 < qb ? nq - q0 : qb;
    rc = yk_cross_distances(type, nb, n, d, b, bn, q + q0 * ld, qn ? qn + q0 : nullptr, D, nb);
    if (rc != YK_OK) break;
    hipLaunchKernelGGL(yk_bigk_select_kernel, dim3((unsigned)n), dim3(YK_BK_THREADS), 0, st, D,
                       nb, nb, k, w, kin, kk);
    hipLaunchKernelGGL(yk_bigk_bounds_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       st, kk, n, k, beg, end);
    if (hipGetLastError() != hipSuccess) {
      rc = yk_set_error(YK_EHIP, "large-k select launch failed");
      break;
    }
    const hipError_t e = rocprim::segmented_radix_sort_keys(
        tmp, tmp_bytes, kin, kout, (unsigned)(n * k), (unsigned)n, beg, end, 0, 64, st);
    if (e != hipSuccess) {
      rc = yk_hip_fail(e, "segmented radix sort");
      break;
    }
    const long tot = n * k;
    hipLaunchKernelGGL(yk_bigk_out_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st,
                       kout, kk, n, k, assign + q0 * k, dis + q0 * k);
    if (hipGetLastError() != hipSuccess) rc = yk_set_error(YK_EHIP, "large-k output launch failed");
  }
  const int rs = yk_sync();
  yk_free(tmp);
  yk_free(kk);
  yk_free(kout);
  yk_free(kin);
  yk_free(D);
  return rc != YK_OK ? rc : rs;
}

// yk_kmeans.hip -- device side of the k-means iteration (kmeans.c:242-310):
// cluster histogram, stable cluster-major permutation of the points, FP64
// segmented centroid sums in point order, scaling by the double reciprocal
// of the count, optional L2 normalisation, and the FP64 objective.
//
// Determinism: the histograms use integer atomics (order-free); the
// permutation is a stable sort of the points by cluster (a counting sort
// for small k, else an LSD radix sort of (cluster, point) pairs), so every
// cluster sums its points in ascending point order, like the reference's
// loop (kmeans.c:280-283), only in FP64 instead of FP32; the objective is a
// fixed-shape tree reduction.  Two runs on the same inputs give the same bits.
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <rocprim/device/device_radix_sort.hpp>

#include "yk.h"
#include "yk_dev.h"

#define YK_OBJ_BLOCKS 1024

struct KmWs {
  int *offsets;    // k + 1
  int *keys_out;   // n
  int *perm;       // n
  int *iota;       // n
  double *obj;     // YK_OBJ_BLOCKS + 1
  int *err;        // 1
  int *mat;        // counting sort: k x nseg segment counts (yk_cs_ok shapes only)
  void *sort_tmp;  // rocprim temp
  size_t sort_bytes;
};

// The update's cluster-major permutation by a stable counting sort where
// the per-segment count matrix is small (k <= 2048 clusters, k x segments
// <= 2^17 entries, segments of YK_CS_SEG points, one wave each), else by
// rocprim's LSD radix sort of (cluster, point) pairs.  Both list each
// cluster's points in ascending order.
#define YK_CS_SEG 1024
static long yk_cs_nseg(long n) { return (n + YK_CS_SEG - 1) / YK_CS_SEG; }
static bool yk_cs_ok(long n, int k) { return k <= 2048 && (long)k * yk_cs_nseg(n) <= (1L << 17); }

static size_t yk_align(size_t x) { return (x + 255) & ~(size_t)255; }

static int yk_bits_for(int k) {
  int b = 1;
  while (b < 31 && (1L << b) < k) b++;
  return b;
}

static size_t yk_sort_tmp_bytes(long n, int k) {
  size_t bytes = 0;
  (void)rocprim::radix_sort_pairs((void *)nullptr, bytes, (int *)nullptr, (int *)nullptr,
                            (int *)nullptr, (int *)nullptr, (size_t)n, 0, yk_bits_for(k));
  return bytes;
}

static KmWs yk_ws_layout(void *base, long n, int k) {
  KmWs w;
  char *p = (char *)base;
  w.offsets = (int *)p;
  p += yk_align(sizeof(int) * (k + 1));
  w.keys_out = (int *)p;
  p += yk_align(sizeof(int) * n);
  w.perm = (int *)p;
  p += yk_align(sizeof(int) * n);
  w.iota = (int *)p;
  p += yk_align(sizeof(int) * n);
  w.obj = (double *)p;
  p += yk_align(sizeof(double) * (YK_OBJ_BLOCKS + 1));
  w.err = (int *)p;
  p += 256;
  w.mat = (int *)p;
  if (yk_cs_ok(n, k)) p += yk_align(sizeof(int) * (size_t)k * yk_cs_nseg(n));
  w.sort_bytes = yk_sort_tmp_bytes(n, k);
  w.sort_tmp = p;
  return w;
}

extern "C" size_t yk_kmeans_workspace_size(long n, int k, int d) {
  (void)d;
  return yk_align(sizeof(int) * (k + 1)) + 3 * yk_align(sizeof(int) * n) +
         yk_align(sizeof(double) * (YK_OBJ_BLOCKS + 1)) + 256 +
         (yk_cs_ok(n, k) ? yk_align(sizeof(int) * (size_t)k * yk_cs_nseg(n)) : 0) +
         yk_align(yk_sort_tmp_bytes(n, k));
}

// Counting sort, pass 1: wave w of the grid counts the clusters of segment
// w (points [w YK_CS_SEG, +YK_CS_SEG), all its loads issued at once) in its
// own LDS histogram and writes them segment-major, mat[w * k + c]; an
// out-of-range cluster sets err.
__global__ __launch_bounds__(256) void yk_cs_hist_kernel(const int *__restrict__ assign, long n,
                                                         int k, long nseg, int *__restrict__ mat,
                                                         int *err) {
  extern __shared__ int lh[];  // 4 waves x k
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long seg = blockIdx.x * 4L + wave;
  if (seg >= nseg) return;
  int *h = lh + wave * k;
  for (int c = lane; c < k; c += 64) h[c] = 0;
  const long p0 = seg * YK_CS_SEG;
  int a[YK_CS_SEG / 64];
#pragma unroll
  for (int j = 0; j < YK_CS_SEG / 64; j++) {
    const long p = p0 + 64 * j + lane;
    a[j] = p < n ? assign[p] : 0;
  }
  bool bad = false;
#pragma unroll
  for (int j = 0; j < YK_CS_SEG / 64; j++) {
    if (p0 + 64 * j + lane >= n) continue;
    if ((unsigned)a[j] < (unsigned)k)
      atomicAdd(&h[a[j]], 1);
    else
      bad = true;
  }
  if (__any(bad) && lane == 0) atomicOr(err, 1);
  for (int c = lane; c < k; c += 64) mat[seg * k + c] = h[c];
}

// Pass 2, one workgroup: per cluster (thread c, and c + 1024) the prefix of
// its counts over the segments, in place (loads eight segments at a time),
// then the clusters' starts offsets[0..k] (exclusive scan of their totals,
// in LDS) and sizes counts[0..k).
__global__ __launch_bounds__(1024) void yk_cs_scan_kernel(int *__restrict__ mat, int k, long nseg,
                                                          int *__restrict__ offsets,
                                                          int *__restrict__ counts) {
  __shared__ int part[2048];
  const int t = threadIdx.x;
  for (int c = t; c < 2048; c += 1024) {
    int run = 0;
    if (c < k) {
      long s = 0;
      for (; s + 8 <= nseg; s += 8) {
        int v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = mat[(s + u) * k + c];
#pragma unroll
        for (int u = 0; u < 8; u++) {
          mat[(s + u) * k + c] = run;
          run += v[u];
        }
      }
      for (; s < nseg; s++) {
        const int v = mat[s * k + c];
        mat[s * k + c] = run;
        run += v;
      }
      counts[c] = run;
    }
    part[c] = run;
  }
  __syncthreads();
  // exclusive scan of part[0..2048): each thread two entries
  const int x0 = part[2 * t], x1 = part[2 * t + 1];
  __syncthreads();
  int v = x0 + x1;
  part[t] = v;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const int u = t >= o ? part[t - o] : 0;
    __syncthreads();
    part[t] += u;
    __syncthreads();
  }
  const int before = part[t] - v;  // exclusive prefix of the pair
  if (2 * t < k) offsets[2 * t] = before;
  if (2 * t + 1 < k) offsets[2 * t + 1] = before + x0;
  if (t == 1023) offsets[k] = part[1023];
}

// Pass 3: wave w places segment w's points.  Per 64-point chunk (all the
// segment's clusters loaded at once), a bitonic sort of the keys (cluster
// << 6 | lane) puts the chunk's points with equal clusters next to each
// other in lane order; a point's place in its cluster is the cluster's
// start, plus the earlier segments' counts of it, plus this segment's
// running count of it (LDS, per wave, updated chunk by chunk in order),
// plus its rank in its run: the sort is stable.  The chunks' ranks are
// computed first and their gathers and stores issued together.
__global__ __launch_bounds__(256) void yk_cs_scatter_kernel(const int *__restrict__ assign, long n,
                                                            int k, long nseg,
                                                            const int *__restrict__ mat,
                                                            const int *__restrict__ offsets,
                                                            int *__restrict__ perm) {
  extern __shared__ int lr[];  // 4 waves x k
  constexpr int NJ = YK_CS_SEG / 64;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long seg = blockIdx.x * 4L + wave;
  if (seg >= nseg) return;
  int *run = lr + wave * k;
  for (int c = lane; c < k; c += 64) run[c] = 0;
  const long p0 = seg * YK_CS_SEG;
  unsigned key[NJ];
#pragma unroll
  for (int j = 0; j < NJ; j++) {
    const long p = p0 + 64 * j + lane;
    const int a = p < n ? assign[p] : -1;
    key[j] = (unsigned)a < (unsigned)k ? ((unsigned)a << 6) | (unsigned)lane : 0xffffffffu;
  }
  int pos[NJ];
#pragma unroll
  for (int j = 0; j < NJ; j++) {
    unsigned kk = key[j];
#pragma unroll
    for (int size = 2; size <= 64; size <<= 1) {
#pragma unroll
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
        const unsigned other = __shfl_xor(kk, stride);
        const bool up = (lane & size) == 0, lower = (lane & stride) == 0;
        kk = (lower == up) ? min(kk, other) : max(kk, other);
      }
    }
    key[j] = kk;
    const bool valid = kk != 0xffffffffu;
    const int c = (int)(kk >> 6);
    const int prev = __shfl_up(c, 1), next = __shfl_down(c, 1);
    const bool start = lane == 0 || prev != c, end = lane == 63 || next != c;
    int st = start ? lane : 0;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int v = __shfl_up(st, o);
      if (lane >= o) st = max(st, v);
    }
    const int rank = lane - st;
    int base = 0;
    if (valid) base = run[c];
    if (valid && end) run[c] = base + rank + 1;
    pos[j] = base + rank;
  }
#pragma unroll
  for (int j = 0; j < NJ; j++) {
    if (key[j] == 0xffffffffu) continue;
    const int c = (int)(key[j] >> 6);
    perm[offsets[c] + mat[seg * k + c] + pos[j]] = (int)(p0 + 64 * j + (key[j] & 63));
  }
}

// the update's set-up in one launch: counts[0..k) = 0, *err = 0,
// iota[i] = i for i < n (the sort's values)
__global__ void yk_upd_prep_kernel(int *counts, int k, int *err, int *iota, long n) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < k) counts[i] = 0;
  if (i < n) iota[i] = (int)i;
  if (i == 0) *err = 0;
}

__global__ void yk_hist_kernel(const int *__restrict__ assign, long n, int k, int *counts,
                               int *err) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  int c = assign[i];
  if (c < 0 || c >= k) {
    atomicOr(err, 1);
    return;
  }
  atomicAdd(&counts[c], 1);
}

// exclusive scan of counts into offsets[0..k], one workgroup
__global__ __launch_bounds__(1024) void yk_scan_kernel(const int *__restrict__ counts, int k,
                                                       int *__restrict__ offsets) {
  __shared__ int part[1024];
  const int t = threadIdx.x;
  const int per = (k + 1023) / 1024;
  const int lo = t * per, hi = min(k, lo + per);
  int s = 0;
  for (int i = lo; i < hi; i++) s += counts[i];
  part[t] = s;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    int v = t >= off ? part[t - off] : 0;
    __syncthreads();
    part[t] += v;
    __syncthreads();
  }
  int run = t == 0 ? 0 : part[t - 1];
  for (int i = lo; i < hi; i++) {
    offsets[i] = run;
    run += counts[i];
  }
  if (t == 1023) offsets[k] = part[1023];
}

// One wave per cluster: FP64 sums over the cluster's points in ascending
// point order, then centroid = (float)(sum * (1.0 / count)) (kmeans.c:286-288;
// count 0 gives 0 * inf = NaN as in the reference).  normalize: L2 norm with
// float squares summed in double, then c *= 1/nr (vector.c:2016-2025).
__global__ __launch_bounds__(256) void yk_segsum_kernel(const float *__restrict__ v, int d,
                                                        int ld, int k,
                                                        const int *__restrict__ offsets,
                                                        const int *__restrict__ perm,
                                                        float *__restrict__ cent,
                                                        int normalize) {
  __shared__ float row[4][1024 + 16];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const long c = (long)blockIdx.x * 4 + wave;
  if (c >= k) return;
  const int beg = offsets[c], end = offsets[c + 1];
  const double inv = 1.0 / (double)(end - beg);
  for (int j0 = 0; j0 < d; j0 += 128) {
    const int ja = j0 + 2 * lane;
    const int jl = ja < d ? ja : 0;  // lanes past d load column 0 and discard it
    double s0 = 0.0, s1 = 0.0;
    // the cluster's point indices 64 at a time (one coalesced load, each
    // then read from its lane), rows 8 at a time: no row load waits for an
    // index load; the sums still take the rows in point order
    for (int pb = beg; pb < end; pb += 64) {
      const int nb = end - pb < 64 ? end - pb : 64;
      const int pidx = lane < nb ? perm[pb + lane] : 0;
      int t = 0;
      for (; t + 8 <= nb; t += 8) {
        float2 x[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const long i = __shfl(pidx, t + u);
          x[u] = *(const float2 *)(v + i * ld + jl);
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
          s0 += (double)x[u].x;
          s1 += (double)x[u].y;
        }
      }
      for (; t < nb; t++) {
        const long i = __shfl(pidx, t);
        const float2 x0 = *(const float2 *)(v + i * ld + jl);
        s0 += (double)x0.x;
        s1 += (double)x0.y;
      }
    }
    if (ja < d) {
      float c0 = (float)(s0 * inv), c1 = (float)(s1 * inv);
      if (normalize) {
        if (ja < 1024) row[wave][ja] = c0;
        if (ja + 1 < 1024) row[wave][ja + 1] = c1;
      }
      cent[c * ld + ja] = c0;
      if (ja + 1 < d) cent[c * ld + ja + 1] = c1;
    }
  }
  if (normalize && d <= 1024) {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    double nr = 0.0;
    if (lane == 0) {
      for (int j = 0; j < d; j++) {
        float x = row[wave][j];
        nr = __dadd_rn(nr, (double)__fmul_rn(x, x));
      }
      nr = sqrt(nr);
    }
    nr = __shfl(nr, 0);
    const double sc = 1.0 / nr;
    for (int j = lane; j < d; j += 64) cent[c * ld + j] = (float)((double)row[wave][j] * sc);
  }
}

// The same sums with a cluster's columns split over waves, 64 per wave
// (one per lane; nsl = ceil(d / 64) waves per cluster): at d = 128 twice the
// waves of yk_segsum_kernel for the same rows, each column still summed in
// point order, so the bits are that kernel's.  No normalisation (that needs
// the whole row in one wave).
__global__ __launch_bounds__(256) void yk_segsum_cols_kernel(const float *__restrict__ v, int d,
                                                             int ld, int k, int nsl,
                                                             const int *__restrict__ offsets,
                                                             const int *__restrict__ perm,
                                                             float *__restrict__ cent) {
  const int lane = threadIdx.x & 63;
  const long g = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (g >= (long)k * nsl) return;
  const long c = g / nsl;
  const int j = (int)(g % nsl) * 64 + lane;
  const int jl = j < d ? j : 0;  // lanes past d load column 0 and discard it
  const int beg = offsets[c], end = offsets[c + 1];
  const double inv = 1.0 / (double)(end - beg);
  double s = 0.0;
  for (int pb = beg; pb < end; pb += 64) {
    const int nb = end - pb < 64 ? end - pb : 64;
    const int pidx = lane < nb ? perm[pb + lane] : 0;
    int t = 0;
    for (; t + 8 <= nb; t += 8) {
      float x[8];
#pragma unroll
      for (int u = 0; u < 8; u++) x[u] = v[(long)__shfl(pidx, t + u) * ld + jl];
#pragma unroll
      for (int u = 0; u < 8; u++) s += (double)x[u];
    }
    for (; t < nb; t++) s += (double)v[(long)__shfl(pidx, t) * ld + jl];
  }
  if (j < d) cent[c * ld + j] = (float)(s * inv);
}

// L2 normalisation of k rows of d > 1024 floats (the segmented-sum kernel
// keeps at most 1024 columns of a row in LDS): one wave per row, 64 columns
// per coalesced load, the squares then added by lane 0 in column order
// through lane reads, so the double sum rounds as fvec_norm's sequential
// loop does (vector.c:2180-2197); c *= 1/nr as fvec_mul_by (vector.c:1792).
__global__ __launch_bounds__(256) void yk_normalize_rows_kernel(float *__restrict__ cent, int d,
                                                                int ld, int k) {
  const int lane = threadIdx.x & 63;
  const long c = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (c >= k) return;
  float *row = cent + c * ld;
  double nr = 0.0;
  for (int j0 = 0; j0 < d; j0 += 64) {
    const float x = j0 + lane < d ? row[j0 + lane] : 0.0f;
    const float sq = __fmul_rn(x, x);
    const int m = d - j0 < 64 ? d - j0 : 64;
    for (int t = 0; t < m; t++) nr = __dadd_rn(nr, (double)__shfl(sq, t));
  }
  const double sc = 1.0 / sqrt(nr);
  for (int j = lane; j < d; j += 64) row[j] = (float)((double)row[j] * sc);
}

// FP64 sum of n floats, fixed reduction shape.
__global__ __launch_bounds__(256) void yk_sum_stage1(const float *__restrict__ x, long n,
                                                     double *__restrict__ part) {
  __shared__ double s[256];
  const long per = (n + gridDim.x - 1) / gridDim.x;
  const long lo = blockIdx.x * per, hi = min(n, lo + per);
  double acc = 0.0;
  for (long i = lo + threadIdx.x; i < hi; i += 256) acc += (double)x[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ __launch_bounds__(256) void yk_sum_stage2(double *part, int nparts) {
  __shared__ double s[256];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nparts; i += 256) acc += part[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) s[threadIdx.x] += s[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[nparts] = s[0];
}

int yk_sum_f64_async(const float *x, long n, double *part) {
  hipLaunchKernelGGL(yk_sum_stage1, dim3(YK_OBJ_BLOCKS), dim3(256), 0, yk_stream(), x, n, part);
  hipLaunchKernelGGL(yk_sum_stage2, dim3(1), dim3(256), 0, yk_stream(), part, YK_OBJ_BLOCKS);
  YK_HIP(hipGetLastError());
  return YK_OK;
}

extern "C" int yk_sum_f64(const float *x, long n, double *out_host, void *ws) {
  YK_TRY(yk_init());
  double *part = (double *)ws;
  YK_TRY(yk_sum_f64_async(x, n, part));
  YK_TRY(yk_read_small(out_host, part + YK_OBJ_BLOCKS, sizeof(double)));
  return YK_OK;
}

// the update's kernels, enqueued; the no-nearest-centroid flag left in w.err
static int yk_kmeans_update_enqueue(long n, int k, int d, const float *v, const int *assign,
                                    float *cent, int *counts, int normalize, void *ws) {
  YK_TRY(yk_init());
  KmWs w = yk_ws_layout(ws, n, k);
  hipStream_t st = yk_stream();
  const int ld = yk_ld(d);
  int tok = yk_prof_begin(YK_PROF_UPDATE);
  const bool cs = yk_cs_ok(n, k);
  const long niota = cs ? 0 : n;  // the counting sort needs no sort values
  const long nprep = niota > k ? niota : k;
  hipLaunchKernelGGL(yk_upd_prep_kernel, dim3((unsigned)((nprep + 255) / 256)), dim3(256), 0, st,
                     counts, k, w.err, w.iota, niota);
  if (cs) {
    const long nseg = yk_cs_nseg(n);
    const unsigned g = (unsigned)((nseg + 3) / 4);
    const size_t lds = sizeof(int) * 4 * (size_t)k;
    if (nseg > 0)
      hipLaunchKernelGGL(yk_cs_hist_kernel, dim3(g), dim3(256), lds, st, assign, n, k, nseg, w.mat,
                         w.err);
    hipLaunchKernelGGL(yk_cs_scan_kernel, dim3(1), dim3(1024), 0, st, w.mat, k, nseg, w.offsets,
                       counts);
    if (nseg > 0)
      hipLaunchKernelGGL(yk_cs_scatter_kernel, dim3(g), dim3(256), lds, st, assign, n, k, nseg,
                         w.mat, w.offsets, w.perm);
  } else {
    hipLaunchKernelGGL(yk_hist_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, assign,
                       n, k, counts, w.err);
    hipLaunchKernelGGL(yk_scan_kernel, dim3(1), dim3(1024), 0, st, counts, k, w.offsets);
    size_t bytes = w.sort_bytes;
    YK_HIP(rocprim::radix_sort_pairs(w.sort_tmp, bytes, assign, w.keys_out, w.iota, w.perm,
                                     (size_t)n, 0, yk_bits_for(k), st));
  }
  if (normalize) {
    hipLaunchKernelGGL(yk_segsum_kernel, dim3((k + 3) / 4), dim3(256), 0, st, v, d, ld, k,
                       w.offsets, w.perm, cent, normalize);
  } else {
    const int nsl = (d + 63) / 64;
    hipLaunchKernelGGL(yk_segsum_cols_kernel, dim3((unsigned)(((long)k * nsl + 3) / 4)),
                       dim3(256), 0, st, v, d, ld, k, nsl, w.offsets, w.perm, cent);
  }
  if (normalize && d > 1024)
    hipLaunchKernelGGL(yk_normalize_rows_kernel, dim3((k + 3) / 4), dim3(256), 0, st, cent, d,
                       ld, k);
  yk_prof_end(tok);
  YK_HIP(hipGetLastError());
  return YK_OK;
}

// pinned landing buffer of the per-iteration readback (grown on demand,
// kept until the process ends)
static std::mutex g_rb_mu;
static void *g_rb = nullptr, *g_rb_dev = nullptr;  // host and device addresses
static size_t g_rb_bytes = 0;

// counts, the objective and the update's check into the readback buffer
// (layout as in yk_kmeans_readback_checked)
__global__ __launch_bounds__(256) void yk_rb_gather_kernel(const int *__restrict__ counts, int k,
                                                           const double *__restrict__ obj,
                                                           const int *__restrict__ err,
                                                           char *__restrict__ rb) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < k) ((int *)(rb + 16))[i] = counts[i];
  if (i == 0) {
    *(double *)rb = *obj;
    *(int *)(rb + 8) = *err;
  }
}

static int yk_kmeans_nearest_error(void) {
  return yk_set_error(YK_EINVAL,
                      "a point has no nearest centroid (NaN in the input?), kmeans.c:281");
}

extern "C" int yk_kmeans_update(long n, int k, int d, const float *v, const int *assign,
                                float *cent, int *counts, int normalize, void *ws) {
  YK_TRY(yk_kmeans_update_enqueue(n, k, d, v, assign, cent, counts, normalize, ws));
  int err = 0;
  YK_TRY(yk_read_small(&err, yk_ws_layout(ws, n, k).err, sizeof(int)));
  return err ? yk_kmeans_nearest_error() : YK_OK;
}

extern "C" int yk_kmeans_update_async(long n, int k, int d, const float *v, const int *assign,
                                      float *cent, int *counts, int normalize, void *ws) {
  return yk_kmeans_update_enqueue(n, k, d, v, assign, cent, counts, normalize, ws);
}

extern "C" int yk_kmeans_readback_checked(const int *counts, int k, int *counts_host,
                                          const float *dis, long n, double *obj_host, void *ws,
                                          void *upd_ws) {
  YK_TRY(yk_init());
  double *part = (double *)ws;
  YK_TRY(yk_sum_f64_async(dis, n, part));
  // The three values are written by one kernel straight into mapped pinned
  // host memory (three copies would be three more launches, and copies into
  // pageable memory go through a bounce buffer one at a time).  Layout:
  // objective (8 bytes), error flag (4), pad (4), counts (4 k).
  std::lock_guard<std::mutex> lock(g_rb_mu);
  const size_t need = 16 + sizeof(int) * (size_t)k;
  if (need > g_rb_bytes) {
    if (g_rb) YK_HIP(hipHostFree(g_rb));
    g_rb = g_rb_dev = nullptr;
    g_rb_bytes = 0;
    const size_t sz = need < 4096 ? 4096 : need;
    YK_HIP(hipHostMalloc(&g_rb, sz, hipHostMallocCoherent | hipHostMallocMapped));
    g_rb_bytes = sz;
    YK_HIP(hipHostGetDevicePointer(&g_rb_dev, g_rb, 0));
  }
  char *rb = (char *)g_rb;
  hipLaunchKernelGGL(yk_rb_gather_kernel, dim3((unsigned)((k + 255) / 256)), dim3(256), 0,
                     yk_stream(), counts, k, part + YK_OBJ_BLOCKS, yk_ws_layout(upd_ws, n, k).err,
                     (char *)g_rb_dev);
  YK_HIP(hipGetLastError());
  YK_HIP(hipStreamSynchronize(yk_stream()));
  int err;
  memcpy(obj_host, rb, sizeof(double));
  memcpy(&err, rb + 8, sizeof(int));
  memcpy(counts_host, rb + 16, sizeof(int) * (size_t)k);
  return err ? yk_kmeans_nearest_error() : YK_OK;
}

// workspace pointer of the objective partials (for callers that keep the
// objective on the device)
double *yk_ws_obj(void *ws, long n, int k) { return yk_ws_layout(ws, n, k).obj; }

// dst[i] = src[idx[i]] for padded rows (initial centroids from selected points)
// k-means++ step (kmeans.c:51-62 with nn.c:132-154): the squared distance
// of every point j < n to point s -- float(||v_j||^2 + ||v_s||^2) from the
// double norms (double sums of float squares), then the cross term added
// as -2 * (float dot product) -- and the running minimum kept in best
// (first: best starts at HUGE_VAL; a NaN distance never replaces it).
__global__ __launch_bounds__(256) void yk_kpp_kernel(const float *__restrict__ v, int ld, int d,
                                                     const double *__restrict__ vn, long n, long s,
                                                     float *__restrict__ best, int first,
                                                     float *__restrict__ host_copy) {
  const long j = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (j >= n) return;
  const float *a = v + s * ld, *b = v + j * ld;
  float dot = 0.f;
  for (int l = 0; l < d; l++) dot = dot + b[l] * a[l];
  float dist = (float)(vn[j] + vn[s]);
  dist = dist + -2.0f * dot;
  float cur = first ? HUGE_VALF : best[j];
  if (dist < cur) cur = dist;
  best[j] = cur;
  if (host_copy) host_copy[j] = cur;
}

extern "C" int yk_kpp_update(const float *v, long n, int d, const double *vn, long s, float *best,
                             int first, float *host_copy) {
  YK_TRY(yk_init());
  if (n <= 0) return YK_OK;
  if (s < 0 || s >= n) return yk_set_error(YK_EINVAL, "yk_kpp_update: centre %ld of %ld", s, n);
  float *hdev = nullptr;
  if (host_copy) YK_HIP(hipHostGetDevicePointer((void **)&hdev, host_copy, 0));
  hipLaunchKernelGGL(yk_kpp_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, yk_stream(),
                     v, yk_ld(d), d, vn, n, s, best, first, hdev);
  YK_HIP(hipGetLastError());
  return YK_OK;
}

// hkm tree descent (hkm_quantize, hkm.c:131-150), a level at a time: the
// points are ordered by their current node (a counting sort on the device:
// histogram, exclusive scan, scatter of point indices), every node is one
// instance of the batched FP32 assign (yk_batch_assign_idx: k = 1 over the
// node's bf children with the FP32 form's arithmetic, nn.c:383-446, points
// read through the index table), and each point's node becomes
// node * bf + child.  A point with no child below 1e30 gets -1 and leaves
// the descent (the reference would index outside the level).

__global__ void yk_hq_hist_kernel(const int *__restrict__ node, long n, int *__restrict__ cnt) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n && node[i] >= 0) atomicAdd(&cnt[node[i]], 1);
}

// one workgroup: off[p] (points before node p, p <= nn), cursor = off, and
// the work items: itoff[p] = items before node p (ceil(count / 128) each)
__global__ __launch_bounds__(1024) void yk_hq_scan_kernel(const int *__restrict__ cnt, long nn,
                                                          long *__restrict__ off,
                                                          long *__restrict__ cursor,
                                                          long *__restrict__ itoff) {
  __shared__ long sp[1024], si[1024];
  const int t = threadIdx.x;
  const long per = (nn + 1023) / 1024, lo = t * per, hi = min(nn, lo + per);
  long a = 0, b = 0;
  for (long p = lo; p < hi; p++) {
    a += cnt[p];
    b += (cnt[p] + 127) / 128;
  }
  sp[t] = a;
  si[t] = b;
  __syncthreads();
  if (t == 0) {  // exclusive prefix over the 1024 partials
    long ra = 0, rb = 0;
    for (int u = 0; u < 1024; u++) {
      const long xa = sp[u], xb = si[u];
      sp[u] = ra;
      si[u] = rb;
      ra += xa;
      rb += xb;
    }
    off[nn] = ra;
    itoff[nn] = rb;
  }
  __syncthreads();
  a = sp[t];
  b = si[t];
  for (long p = lo; p < hi; p++) {
    off[p] = a;
    cursor[p] = a;
    itoff[p] = b;
    a += cnt[p];
    b += (cnt[p] + 127) / 128;
  }
}

__global__ void yk_hq_scatter_kernel(const int *__restrict__ node, long n, long *cursor,
                                     int *__restrict__ perm) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n && node[i] >= 0) perm[atomicAdd((unsigned long long *)&cursor[node[i]], 1ull)] = (int)i;
}

// work item j -> (node, tile): the node by binary search over itoff
__global__ void yk_hq_items_kernel(const long *__restrict__ itoff, long nn, long nitems,
                                   int *__restrict__ item_inst, int *__restrict__ item_tile) {
  const long j = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (j >= nitems) return;
  long lo = 0, hi = nn;  // itoff[lo] <= j < itoff[hi]
  while (hi - lo > 1) {
    const long mid = (lo + hi) >> 1;
    if (itoff[mid] <= j)
      lo = mid;
    else
      hi = mid;
  }
  item_inst[j] = (int)lo;
  item_tile[j] = (int)(j - itoff[lo]);
}

// the level's centroids with each node's bf children padded to kpad rows
// (pad rows repeat child 0; the assign never selects them), and their norms
__global__ void yk_hq_pad_kernel(const float *__restrict__ cl, const float *__restrict__ cnl,
                                 long nn, int bf, int kpad, int ld, float *__restrict__ dst,
                                 float *__restrict__ dstn) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long rows = nn * kpad, row = t / (ld / 4);
  if (row >= rows) return;
  const int c4 = (int)(t % (ld / 4));
  const long p = row / kpad;
  const int c = (int)(row % kpad);
  const long src = p * bf + (c < bf ? c : 0);
  ((float4 *)(dst + row * ld))[c4] = ((const float4 *)(cl + src * ld))[c4];
  if (c4 == 0) dstn[row] = cnl[src];
}

__global__ void yk_hq_update_kernel(const int *__restrict__ perm, const int *__restrict__ child,
                                    long m, int bf, int *__restrict__ node) {
  const long j = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (j >= m) return;
  const int i = perm[j], c = child[j];
  node[i] = c < 0 ? -1 : node[i] * bf + c;
}

__global__ void yk_hq_fill_kernel(int *p, long n, int v) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

extern "C" int yk_hkm_quantize(const float *v, const double *vn, long n, int d, int nlevel, int bf,
                               const float *cent, const float *cn, int *idx) {
  YK_TRY(yk_init());
  if (n <= 0) return YK_OK;
  if (nlevel < 1 || bf < 1) return yk_set_error(YK_EINVAL, "yk_hkm_quantize: bad tree");
  if (n > 0x7fffffffL) return yk_set_error(YK_EINVAL, "yk_hkm_quantize: n exceeds int32");
  const int ld = yk_ld(d), kpad = (bf + 31) & ~31;
  hipStream_t st = yk_stream();
  // the widest level's parent nodes: bf^(nlevel-1)
  long nn_max = 1;
  for (int l = 1; l < nlevel; l++) nn_max *= bf;
  void *mem = nullptr;
  const size_t n_i = sizeof(int) * (size_t)n, nn_l = sizeof(long) * (size_t)(nn_max + 1);
  const size_t items_max = (size_t)(n / 128 + nn_max + 1);
  const size_t bytes = 4 * n_i + 3 * nn_l + sizeof(int) * (size_t)nn_max +
                       2 * sizeof(int) * items_max +
                       sizeof(float) * (size_t)nn_max * kpad * (ld + 1) + 16 * 256;
  YK_TRY(yk_malloc(&mem, bytes));
  char *q = (char *)mem;
  auto carve = [&](size_t b) {
    char *r = q;
    q += (b + 255) & ~(size_t)255;
    return (void *)r;
  };
  int *perm = (int *)carve(n_i), *child = (int *)carve(n_i);
  float *dis = (float *)carve(n_i);
  int *cnt = (int *)carve(sizeof(int) * (size_t)nn_max);
  long *off = (long *)carve(nn_l), *cursor = (long *)carve(nn_l), *itoff = (long *)carve(nn_l);
  int *item_inst = (int *)carve(sizeof(int) * items_max);
  int *item_tile = (int *)carve(sizeof(int) * items_max);
  float *pc = (float *)carve(sizeof(float) * (size_t)nn_max * kpad * ld);
  float *pcn = (float *)carve(sizeof(float) * (size_t)nn_max * kpad);
  int rc = YK_OK;
  const unsigned gb = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(yk_hq_fill_kernel, dim3(gb), dim3(256), 0, st, idx, n, 0);  // node 0
  long lvl = 0, nn = 1;  // first centroid row of the level, its parent nodes
  for (int l = 0; l < nlevel && rc == YK_OK; l++) {
    hipLaunchKernelGGL(yk_hq_fill_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, st,
                       cnt, nn, 0);
    hipLaunchKernelGGL(yk_hq_hist_kernel, dim3(gb), dim3(256), 0, st, idx, n, cnt);
    hipLaunchKernelGGL(yk_hq_scan_kernel, dim3(1), dim3(1024), 0, st, cnt, nn, off, cursor,
                       itoff);
    hipLaunchKernelGGL(yk_hq_scatter_kernel, dim3(gb), dim3(256), 0, st, idx, n, cursor, perm);
    long tot[2] = {0, 0};  // points still descending, work items
    rc = yk_memcpy_d2h(&tot[0], off + nn, sizeof(long));
    if (rc == YK_OK) rc = yk_memcpy_d2h(&tot[1], itoff + nn, sizeof(long));
    if (rc != YK_OK || tot[0] == 0) break;
    hipLaunchKernelGGL(yk_hq_items_kernel, dim3((unsigned)((tot[1] + 255) / 256)), dim3(256), 0,
                       st, itoff, nn, tot[1], item_inst, item_tile);
    const long prow = nn * kpad * (ld / 4);
    hipLaunchKernelGGL(yk_hq_pad_kernel, dim3((unsigned)((prow + 255) / 256)), dim3(256), 0, st,
                       cent + lvl * ld, cn + lvl, nn, bf, kpad, ld, pc, pcn);
    YK_HIP(hipGetLastError());
    rc = yk_batch_assign_idx(tot[1], item_inst, item_tile, off, d, bf, kpad, v, vn, perm, pc,
                             pcn, nullptr, child, dis);
    if (rc != YK_OK) break;
    hipLaunchKernelGGL(yk_hq_update_kernel, dim3((unsigned)((tot[0] + 255) / 256)), dim3(256), 0,
                       st, perm, child, tot[0], bf, idx);
    YK_HIP(hipGetLastError());
    lvl += nn * bf;
    nn *= bf;
  }
  const int rs = yk_sync();
  yk_free(mem);
  return rc != YK_OK ? rc : rs;
}

__global__ void yk_gather_kernel(const float *__restrict__ src, const int *__restrict__ idx,
                                 long n, int ld, float *__restrict__ dst) {
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  long row = t / (ld / 4), c4 = t % (ld / 4);
  if (row >= n) return;
  ((float4 *)(dst + row * ld))[c4] = ((const float4 *)(src + (long)idx[row] * ld))[c4];
}

extern "C" int yk_gather_rows(const float *src, const int *idx, long n, int d, float *dst) {
  YK_TRY(yk_init());
  if (n <= 0) return YK_OK;
  const int ld = yk_ld(d);
  long total = n * (ld / 4);
  hipLaunchKernelGGL(yk_gather_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                     yk_stream(), src, idx, n, ld, dst);
  YK_HIP(hipGetLastError());
  return YK_OK;
}

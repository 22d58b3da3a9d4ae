// yk_batch.hip -- the batched form: many independent k-means instances of
// uneven size (hierarchical k-means nodes, product-quantisation sub-spaces)
// advanced together, one grouped launch per step, the whole iteration free
// of host synchronisation so it can be captured once in a hipGraph and
// replayed niter times.
//
// Per iteration (all instances, instance-major work items):
//   norms of every centroid           yk_norms
//   assign                            yk_batch_assign (yk_select.hip): the
//                                     FP32 scan kernel with a work-item ->
//                                     (instance, 128-point tile) table
//   update, partial sums              yk_bupd_partial_kernel: one workgroup
//                                     per (instance, chunk of P points) adds
//                                     the chunk's points in point order into
//                                     an LDS FP64 table [k][d], and sums the
//                                     chunk's objective
//   update, reduce + scale            yk_bupd_reduce_kernel: chunks summed in
//                                     chunk order, (float)(sum * (1/count))
//   empty-cluster split               yk_bsplit_kernel: kmeans.c:166-209 on
//                                     the device, its random draws taken from
//                                     host-precomputed per-(instance,
//                                     iteration) tables (the streams depend
//                                     only on the seed, kmeans.c:296)
//   objective, abandon and stop rules yk_bobj_kernel (kmeans.c:300-313):
//                                     chunk objectives in chunk order
//   iteration counter                 yk_bstep_kernel
// Finished instances are skipped by every kernel (done flags), so their
// outputs are those of the iteration at which the reference would stop.
// Every reduction has a fixed order: results are deterministic.
#include <hip/hip_runtime.h>
#include <math.h>

#include "yk.h"
#include "yk_dev.h"

struct BatchDev {
  int ninst, d, ld, k, kpad, niter, ecap, chunk;
  int kslice, nslice;  // the update's clusters per LDS table, and the tables per chunk
  int win;             // iterations per window of the split tables
  int normalize;       // KMEANS_NORMALIZE_CENTS: centroids scaled to unit L2 norm
  long ntot;
  const float *v;   // ntot x ld
  const double *vn;  // ntot
  float *cent;       // ninst x kpad x ld
  float *cn;         // ninst x kpad
  int *assign;       // ntot
  float *dis;        // ntot
  const long *inst_off;  // ninst + 1
  long nitems;
  const int *item_inst, *item_tile;
  long nuitems;
  const int *uitem_inst, *uitem_first;  // update chunks; first chunk of each instance
  double *psum;  // nuitems x k x d
  double *pobj;  // nuitems (chunk objective partials, after psum)
  int *pcnt;     // nuitems x k
  int *nassign;  // ninst x k
  float *pbuf;   // ninst x k (split probabilities)
  const double *rd;   // ninst x win x ecap: the current window's draws
  const float *eps;   // ninst x win x ecap x d
  int *done, *overflow, *nre, *tot, *iters, *abandoned;
  double *qerr, *qerr_old;
  int *it;  // iteration counter
  int *w0;  // first iteration of the current window (set by the host)
};

// Row stride (doubles) of the LDS table: the column paths (d < 64 with d a
// multiple of 4, d >= 64 a multiple of 32) pad the rows to an odd length so
// that the lanes of one instruction, each on its own cluster's row, spread
// over the LDS banks.
__host__ __device__ static inline int yk_bupd_stride(int d) {
  return ((d < 64 && (d & 3) == 0) || (d & 31) == 0) ? d + 1 : d;
}

// One workgroup per update chunk adds the chunk's points into an LDS FP64
// table [k][d].  d >= 64 a multiple of 32: wave w owns the groups of 8 float4
// columns (32 dimensions) w, w + 4, ... (no two waves share a sum); inside
// a group, lane 8 p + c takes column c of the step's point p, so one load
// instruction reads 8 whole 128-byte row segments, eight steps (64 points)
// in flight.  Other d >= 64: one thread per dimension (up to 4 waves), each
// (cluster, dimension) sum owned by one thread and accumulated in point
// order.  d < 64, d a multiple of 4: wave w owns the float4 columns
// w, w + 4, ... of the rows, a lane takes one point per step, four steps in
// flight; d < 64 otherwise: one wave whose lanes split into 64/d groups of
// d.  The column paths use ds_add_f64: the LDS atomics of one wave apply in
// instruction order and, inside one instruction, in a fixed order, so every
// sum is accumulated in point order.  Deterministic, whatever the grid.  The chunk's objective (its points' distances, FP64,
// thread-strided then a fixed tree) goes to pobj[u].
__global__ __launch_bounds__(256) void yk_bupd_partial_kernel(BatchDev b) {
  extern __shared__ double lsum[];  // ks x sd
  __shared__ double wobj[4];
  const int d = b.d, ld = b.ld, sd = yk_bupd_stride(d);
  const int tid = threadIdx.x, nthr = blockDim.x, lane = tid & 63, wave = tid >> 6;
  const long u = blockIdx.x;
  // this workgroup's clusters [kc0, kc0 + ks): tables larger than LDS are
  // taken a slice at a time, each slice's workgroup reading every
  // assignment of the chunk and only its own clusters' rows
  const int kc0 = blockIdx.y * b.kslice, ks = min(b.kslice, b.k - kc0);
  int *lcnt = (int *)(lsum + (size_t)ks * sd);
  const int inst = b.uitem_inst[u];
  if (b.done[inst]) return;
  const long off = b.inst_off[inst], n = b.inst_off[inst + 1] - off;
  const long c0 = (u - b.uitem_first[inst]) * (long)b.chunk;
  const long p0 = off + c0, p1 = off + (c0 + b.chunk < n ? c0 + b.chunk : n);
  // a point's row in this slice's table, or -1 when its cluster is another
  // slice's
  auto slot = [&](int a) { return (unsigned)(a - kc0) < (unsigned)ks ? a - kc0 : -1; };
  for (int i = tid; i < ks * sd; i += nthr) lsum[i] = 0.0;
  for (int i = tid; i < ks; i += nthr) lcnt[i] = 0;
  // a point without a nearest centroid (NaN input) has assign -1: the
  // instance is handed to the single-instance path, which reports it
  bool bad = false;
  for (long p = p0 + tid; p < p1; p += nthr) bad |= (unsigned)b.assign[p] >= (unsigned)b.k;
  if (__syncthreads_or(bad)) {
    if (tid == 0) b.overflow[inst] = 1;
    return;
  }
  if (blockIdx.y == 0) {  // the chunk objective
    double obj = 0.0;
    for (long p = p0 + tid; p < p1; p += nthr) obj += (double)b.dis[p];
    for (int o = 32; o > 0; o >>= 1) obj += __shfl_xor(obj, o);
    if (lane == 0) wobj[wave] = obj;
  }
  if (d >= 64 && (d & 31) == 0) {
    // Within each 32-dimension group the table's columns are permuted,
    // dimension 4 c + e at column 8 e + c: the lanes of one ds_add_f64 (the
    // e-th component of their float4) then hit 8 consecutive doubles of a
    // row, not every fourth pair of banks.  Each step's loads are issued a
    // step ahead (at indices clamped into the chunk: always in bounds), so
    // the adds of one step overlap the next step's loads.
    const int nwave = nthr >> 6, pl = lane >> 3, cl = lane & 7;
    for (int g = wave; g < (d >> 5); g += nwave) {
      const int j0 = g * 32 + cl * 4;
      int an[8];
      float4 xn[8];
      auto load = [&](long p) {
#pragma unroll
        for (int s = 0; s < 8; s++) {
          const long q = p + 8 * s < p1 ? p + 8 * s : p1 - 1;
          an[s] = b.assign[q];
          xn[s] = *(const float4 *)(b.v + q * ld + j0);
        }
      };
      if (p0 + pl < p1) load(p0 + pl);
      for (long p = p0 + pl; p < p1; p += 64) {
        int a[8];
        float4 x[8];
#pragma unroll
        for (int s = 0; s < 8; s++) {
          a[s] = an[s];
          x[s] = xn[s];
        }
        load(p + 64 < p1 ? p + 64 : p);
#pragma unroll
        for (int s = 0; s < 8; s++) {
          a[s] = p + 8 * s < p1 ? slot(a[s]) : -1;
          if (a[s] < 0) continue;
          double *row = lsum + a[s] * sd + g * 32 + cl;
          __hip_atomic_fetch_add(row, (double)x[s].x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          __hip_atomic_fetch_add(row + 8, (double)x[s].y, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_WORKGROUP);
          __hip_atomic_fetch_add(row + 16, (double)x[s].z, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_WORKGROUP);
          __hip_atomic_fetch_add(row + 24, (double)x[s].w, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_WORKGROUP);
          if (g == 0 && cl == 0)
            __hip_atomic_fetch_add(&lcnt[a[s]], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
      }
    }
  } else if (d >= 64) {
    // a thread owns its dimensions: plain read-add-write, four points in
    // flight
    for (int t = tid; t < d; t += nthr) {
      long p = p0;
      for (; p + 4 <= p1; p += 4) {
        int a[4];
        float x[4];
#pragma unroll
        for (int s = 0; s < 4; s++) {
          a[s] = b.assign[p + s];
          x[s] = b.v[(p + s) * ld + t];
        }
#pragma unroll
        for (int s = 0; s < 4; s++) {
          a[s] = slot(a[s]);
          if (a[s] < 0) continue;
          lsum[a[s] * d + t] += (double)x[s];
          if (t == 0) lcnt[a[s]]++;
        }
      }
      for (; p < p1; p++) {
        const int a = slot(b.assign[p]);
        if (a < 0) continue;
        lsum[a * d + t] += (double)b.v[p * ld + t];
        if (t == 0) lcnt[a]++;
      }
    }
  } else if ((d & 3) == 0) {
    const int nwave = nthr >> 6;
    for (int col = wave; col < (d >> 2); col += nwave) {
      const int j0 = col * 4;
      for (long p = p0 + lane; p < p1; p += 256) {
        int a[4];
        float4 x[4];
#pragma unroll
        for (int s = 0; s < 4; s++) {
          const long q = p + 64 * s < p1 ? p + 64 * s : p1 - 1;  // clamped, as above
          a[s] = b.assign[q];
          x[s] = *(const float4 *)(b.v + q * ld + j0);
        }
#pragma unroll
        for (int s = 0; s < 4; s++) {
          a[s] = p + 64 * s < p1 ? slot(a[s]) : -1;
          if (a[s] < 0) continue;
          double *row = lsum + a[s] * sd + j0;
          __hip_atomic_fetch_add(row, (double)x[s].x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          __hip_atomic_fetch_add(row + 1, (double)x[s].y, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_WORKGROUP);
          __hip_atomic_fetch_add(row + 2, (double)x[s].z, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_WORKGROUP);
          __hip_atomic_fetch_add(row + 3, (double)x[s].w, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_WORKGROUP);
          if (col == 0)
            __hip_atomic_fetch_add(&lcnt[a[s]], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
      }
    }
  } else {
    const int ppp = 64 / d, g = tid / d, t = tid - g * d;
    if (g < ppp) {
      for (long p = p0 + g; p < p1; p += ppp) {
        const int a = slot(b.assign[p]);
        if (a < 0) continue;
        __hip_atomic_fetch_add(&lsum[a * d + t], (double)b.v[p * ld + t], __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);
        if (t == 0)
          __hip_atomic_fetch_add(&lcnt[a], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      }
    }
  }
  __syncthreads();
  if (tid == 0 && blockIdx.y == 0) {
    double obj = 0.0;
    for (int w = 0; w < (nthr >> 6); w++) obj += wobj[w];
    b.pobj[u] = obj;
  }
  double *ps = b.psum + (u * (size_t)b.k + kc0) * b.d;
  const bool perm = d >= 64 && (d & 31) == 0;  // the column order of that path
  for (int i = tid; i < ks * d; i += nthr) {
    const int j = i % d;
    ps[i] = lsum[(i / d) * sd + (perm ? (j & ~31) | ((j & 3) << 3) | ((j >> 2) & 7) : j)];
  }
  for (int i = tid; i < ks; i += nthr) b.pcnt[u * b.k + kc0 + i] = lcnt[i];
}

// One wave per (instance, cluster): chunk sums in chunk order, then
// centroid = (float)(sum * (1.0 / count)) (kmeans.c:278-288; an empty
// cluster gets 0 * inf = NaN until the split overwrites it).
__global__ __launch_bounds__(256) void yk_bupd_reduce_kernel(BatchDev b) {
  const int lane = threadIdx.x & 63;
  const long g = blockIdx.x * 4L + (threadIdx.x >> 6);
  if (g >= (long)b.ninst * b.k) return;
  const int inst = (int)(g / b.k), c = (int)(g % b.k);
  if (b.done[inst]) return;
  const long u0 = b.uitem_first[inst], u1 = b.uitem_first[inst + 1];
  int cnt = 0;
#pragma unroll 8
  for (long u = u0; u < u1; u++) cnt += b.pcnt[u * b.k + c];
  const double inv = 1.0 / (double)cnt;
  float *crow = b.cent + ((long)inst * b.kpad + c) * b.ld;
  for (int t = lane; t < b.d; t += 64) {
    double s = 0.0;
#pragma unroll 8
    for (long u = u0; u < u1; u++) s += b.psum[(u * b.k + c) * (size_t)b.d + t];
    crow[t] = (float)(s * inv);
  }
  if (b.normalize) {
    // L2 normalisation with the single-instance update's arithmetic: float
    // squares added in double in column order (each lane re-reads only its
    // own columns; lane reads hand the squares over in order), c *= 1/nr
    // (vector.c:1792-1797, 2016-2025, 2180-2197)
    double nr = 0.0;
    for (int j0 = 0; j0 < b.d; j0 += 64) {
      const float x = j0 + lane < b.d ? crow[j0 + lane] : 0.0f;
      const float sq = __fmul_rn(x, x);
      const int m = b.d - j0 < 64 ? b.d - j0 : 64;
      for (int t = 0; t < m; t++) nr = __dadd_rn(nr, (double)__shfl(sq, t));
    }
    const double sc = 1.0 / sqrt(nr);
    for (int t = lane; t < b.d; t += 64) crow[t] = (float)((double)crow[t] * sc);
  }
  if (lane == 0) b.nassign[(long)inst * b.k + c] = cnt;
}

// L1 normalisation of the split probabilities (vector.c:2016-2025)
__device__ void yk_l1_normalize(float *p, int k) {
  double s = 0.0;
  for (int i = 0; i < k; i++) s += fabs((double)p[i]);
  const double inv = 1. / s;
  for (int i = 0; i < k; i++) p[i] = (float)((double)p[i] * inv);
}

// kmeans.c:166-209 for every unfinished instance, one thread per instance.
__global__ void yk_bsplit_kernel(BatchDev b) {
  const int inst = blockIdx.x * blockDim.x + threadIdx.x;
  if (inst >= b.ninst || b.done[inst]) return;
  // the iteration's row of the current window's tables (tuning runs, which
  // may go past the window, reuse its last row)
  const int it = min(max(*b.it - *b.w0, 0), b.win - 1);
  const int k = b.k, d = b.d;
  const int *na = b.nassign + (long)inst * k;
  // the split probabilities are needed only when a cluster is empty (the
  // draws are taken per empty cluster)
  bool any_empty = false;
  for (int c = 0; c < k; c++) any_empty |= na[c] == 0;
  if (!any_empty) {
    b.nre[inst] = 0;
    return;
  }
  float *p = b.pbuf + (long)inst * k;
  for (int c = 0; c < k; c++) {
    const unsigned u = (unsigned)na[c];
    p[c] = na[c] < 2 ? 0.f : (float)(int)(u * u - 1u);  // 32-bit wrap as in the reference
  }
  yk_l1_normalize(p, k);
  const long tab = ((long)inst * b.win + it) * b.ecap;
  int e = 0;
  for (int c = 0; c < k; c++) {
    if (na[c] != 0) continue;
    if (e == b.ecap) {  // more empties than precomputed draws: redo on the host
      b.overflow[inst] = 1;
      break;
    }
    double rd = b.rd[tab + e];
    int j = 0;
    while (j < k - 1) {
      rd -= p[j];
      if (rd < 0) break;
      j++;
    }
    float *cc = b.cent + ((long)inst * b.kpad + c) * b.ld;
    float *cj = b.cent + ((long)inst * b.kpad + j) * b.ld;
    double nr = 0.0;
    for (int l = 0; l < d; l++) {
      cc[l] = cj[l];
      nr += (double)(cj[l] * cj[l]);  // float squares, double sum
    }
    const double s = __dsqrt_rn(nr) * 0.0000001;
    const float *ep = b.eps + (tab + e) * d;
    for (int l = 0; l < d; l++) {
      const float el = (float)((double)ep[l] * s);
      cj[l] += el;
      cc[l] -= el;
    }
    p[j] = 0;
    yk_l1_normalize(p, k);
    e++;
  }
  b.nre[inst] = e;
}

// Objective (FP64, fixed shape) and the abandon / stop rules; one
// workgroup per instance.
__global__ __launch_bounds__(256) void yk_bobj_kernel(BatchDev b) {
  __shared__ double s[256];
  const int inst = blockIdx.x;
  if (b.done[inst]) return;
  const long n = b.inst_off[inst + 1] - b.inst_off[inst];
  // the chunks' objective partials, in chunk order per thread, then a
  // fixed tree
  double acc = 0.0;
  for (long u = b.uitem_first[inst] + threadIdx.x; u < b.uitem_first[inst + 1]; u += 256)
    acc += b.pobj[u];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x) return;
  b.iters[inst]++;
  if (b.overflow[inst]) {
    b.done[inst] = 1;
    return;
  }
  const int tot = b.tot[inst] + b.nre[inst];
  b.tot[inst] = tot;
  if (tot > n / 100 && tot > 1000) {
    b.abandoned[inst] = 1;
    b.done[inst] = 1;
    return;
  }
  const double q = s[0], qold = b.qerr_old[inst];
  b.qerr[inst] = q;
  b.qerr_old[inst] = q;
  if (qold == q && b.nre[inst] == 0) b.done[inst] = 1;
}

__global__ void yk_bstep_kernel(int *it) {
  if (threadIdx.x == 0) *it += 1;
}

__global__ void yk_binit_kernel(BatchDev b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0) {
    *b.it = 0;
    *b.w0 = 0;
  }
  if (i < b.ninst) {
    b.done[i] = 0;
    b.overflow[i] = 0;
    b.nre[i] = 0;
    b.tot[i] = 0;
    b.iters[i] = 0;
    b.abandoned[i] = 0;
    b.qerr[i] = 0.0;
    b.qerr_old[i] = __longlong_as_double(0x7ff0000000000000ll);  // HUGE_VAL
  }
}

// ---------------------------------------------------------------------------
// host side of the HIP layer

struct yk_batch {
  BatchDev dev;
  void *mem;
  hipGraph_t graph;
  hipGraphExec_t exec;
  int captured;
};

// the update's LDS table: clusters per slice and bytes of a slice of ks
#define YK_BUPD_LDS_MAX (160 * 1024)
static size_t yk_bupd_lds(int ks, int d) {
  return sizeof(double) * (size_t)ks * yk_bupd_stride(d) + sizeof(int) * (size_t)ks;
}

// clusters per slice: all k when the table fits, else k split into equal
// slices that do
static int yk_bupd_kslice(int k, int d) {
  const int kmax = (int)(YK_BUPD_LDS_MAX / yk_bupd_lds(1, d));
  if (kmax < 1) return 0;
  const int nslice = (k + kmax - 1) / kmax;
  return (k + nslice - 1) / nslice;
}

extern "C" size_t yk_batch_lds_bytes(int k, int d) {
  const int ks = yk_bupd_kslice(k, d);
  return ks ? yk_bupd_lds(ks, d) : (size_t)-1;
}

// The device state of one batched run.  All pointers are device pointers
// laid out by the caller (host/yael_batch.c) in one allocation.
extern "C" int yk_batch_create(yk_batch **out, int ninst, int d, int k, int kpad, int niter,
                               int win, int ecap, int chunk, long ntot, const float *v,
                               const double *vn, float *cent, float *cn, int *assign, float *dis,
                               const long *inst_off, long nitems, const int *item_inst,
                               const int *item_tile, long nuitems, const int *uitem_inst,
                               const int *uitem_first, double *psum, int *pcnt, int *nassign,
                               float *pbuf, const double *rd, const float *eps, int *flags,
                               double *qerr2, int normalize) {
  YK_TRY(yk_init());
  const int kslice = yk_bupd_kslice(k, d);
  if (kslice < 1)
    return yk_set_error(YK_EINVAL, "batched k-means: d = %d too large for one LDS table row", d);
  if (win < 1) return yk_set_error(YK_EINVAL, "batched k-means: window of %d iterations", win);
  yk_batch *bt = new yk_batch();
  BatchDev &b = bt->dev;
  b.ninst = ninst; b.d = d; b.ld = yk_ld(d); b.k = k; b.kpad = kpad; b.niter = niter;
  b.ecap = ecap; b.chunk = chunk; b.ntot = ntot;
  b.kslice = kslice; b.nslice = (k + kslice - 1) / kslice; b.win = win;
  b.normalize = normalize;
  b.v = v; b.vn = vn; b.cent = cent; b.cn = cn; b.assign = assign; b.dis = dis;
  b.inst_off = inst_off; b.nitems = nitems; b.item_inst = item_inst; b.item_tile = item_tile;
  b.nuitems = nuitems; b.uitem_inst = uitem_inst; b.uitem_first = uitem_first;
  b.psum = psum; b.pobj = psum + (size_t)nuitems * k * d; b.pcnt = pcnt; b.nassign = nassign; b.pbuf = pbuf; b.rd = rd; b.eps = eps;
  // flags: 7 int arrays of ninst, the iteration counter, the window start
  b.done = flags; b.overflow = flags + ninst; b.nre = flags + 2L * ninst;
  b.tot = flags + 3L * ninst; b.iters = flags + 4L * ninst; b.abandoned = flags + 5L * ninst;
  b.it = flags + 6L * ninst;
  b.w0 = flags + 6L * ninst + 1;
  b.qerr = qerr2; b.qerr_old = qerr2 + ninst;
  bt->captured = 0;
  *out = bt;
  size_t lds = yk_bupd_lds(kslice, d);
  if (lds > 65536)
    YK_HIP(hipFuncSetAttribute((const void *)yk_bupd_partial_kernel,
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  return YK_OK;
}

extern "C" int yk_batch_reset(yk_batch *bt) {
  BatchDev &b = bt->dev;
  hipLaunchKernelGGL(yk_binit_kernel, dim3((b.ninst + 255) / 256), dim3(256), 0, yk_stream(), b);
  YK_HIP(hipGetLastError());
  return YK_OK;
}

// One iteration of every unfinished instance, queued on the library stream
// with no host synchronisation (so it can be captured).
static int yk_batch_enqueue(yk_batch *bt) {
  BatchDev &b = bt->dev;
  hipStream_t st = yk_stream();
  YK_TRY(yk_norms(b.cent, (long)b.ninst * b.kpad, b.d, b.cn, nullptr));
  // small centroid sets: centroids in registers, points streamed per chunk
  const int rc = yk_batch_assign_res(b.ninst, b.d, b.k, b.kpad, b.chunk, b.v, b.vn, b.cent, b.cn,
                                     b.inst_off, b.nuitems, b.uitem_inst, b.uitem_first, b.done,
                                     b.assign, b.dis);
  if (rc < 0) return rc;
  if (rc == 1)
    YK_TRY(yk_batch_assign(b.nitems, b.item_inst, b.item_tile, b.inst_off, b.d, b.k, b.kpad, b.v,
                           b.vn, b.cent, b.cn, b.done, b.assign, b.dis));
  // d >= 64 a multiple of 32: a wave per 32-dimension group (at most 4);
  // d < 64: a wave per float4 column (at most 4), or one wave when d is not
  // a multiple of 4; else one thread per dimension (at most 256)
  const int pthreads = (b.d >= 64 && (b.d & 31) == 0) ? 64 * (b.d / 32 < 4 ? b.d / 32 : 4)
                       : b.d < 64      ? ((b.d & 3) == 0 ? 64 * (b.d / 4 < 4 ? b.d / 4 : 4) : 64)
                                       : (b.d > 256 ? 256 : ((b.d + 63) / 64) * 64);
  hipLaunchKernelGGL(yk_bupd_partial_kernel, dim3((unsigned)b.nuitems, (unsigned)b.nslice),
                     dim3(pthreads), yk_bupd_lds(b.kslice, b.d), st, b);
  hipLaunchKernelGGL(yk_bupd_reduce_kernel, dim3((unsigned)(((long)b.ninst * b.k + 3) / 4)),
                     dim3(256), 0, st, b);
  hipLaunchKernelGGL(yk_bsplit_kernel, dim3((b.ninst + 63) / 64), dim3(64), 0, st, b);
  hipLaunchKernelGGL(yk_bobj_kernel, dim3(b.ninst), dim3(256), 0, st, b);
  hipLaunchKernelGGL(yk_bstep_kernel, dim3(1), dim3(64), 0, st, b.it);
  YK_HIP(hipGetLastError());
  return YK_OK;
}

// Run niter iterations: with use_graph, the iteration is captured once and
// the graph replayed; otherwise launched directly.
extern "C" int yk_batch_run(yk_batch *bt, int niter, int use_graph) {
  hipStream_t st = yk_stream();
  int tok = yk_prof_begin(YK_PROF_BATCH);
  if (use_graph) {
    if (!bt->captured) {
      YK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      int rc = yk_batch_enqueue(bt);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(st, &g);
      if (rc != YK_OK) return rc;
      if (e != hipSuccess) return yk_hip_fail(e, "hipStreamEndCapture");
      bt->graph = g;
      YK_HIP(hipGraphInstantiate(&bt->exec, g, nullptr, nullptr, 0));
      bt->captured = 1;
    }
    for (int i = 0; i < niter; i++) YK_HIP(hipGraphLaunch(bt->exec, st));
  } else {
    for (int i = 0; i < niter; i++) YK_TRY(yk_batch_enqueue(bt));
  }
  yk_prof_end(tok);
  return YK_OK;
}

extern "C" int yk_batch_destroy(yk_batch *bt) {
  if (!bt) return YK_OK;
  hipError_t e = hipSuccess, e2 = hipSuccess;
  if (bt->captured) {
    e = hipGraphExecDestroy(bt->exec);
    e2 = hipGraphDestroy(bt->graph);
  }
  delete bt;
  if (e != hipSuccess) return yk_hip_fail(e, "hipGraphExecDestroy");
  if (e2 != hipSuccess) return yk_hip_fail(e2, "hipGraphDestroy");
  return YK_OK;
}

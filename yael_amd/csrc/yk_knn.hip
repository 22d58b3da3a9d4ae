// yk_knn.hip -- vector norms, fused FP32 distance + k-smallest selection on
// MFMA, split merge, cross-distance tile writer, FP64 pair distances.
//
// Reference semantics (yael nn.c):
//   dist(q,b) = float( double(sum_j q_j*q_j) [float products, double sum]
//                    + double(float-accumulated sum_j b_j*b_j) )  - 2 * q.b
//   (nn.c:108-126), optionally times w[b] (nn.c:497-500); type 16 is the
//   plain dot product (nn.c:266-306).  k = 1: strict '<' from (-1, 1e30)
//   (nn.c:404-440).  k > 1: the k smallest non-NaN values, sorted, padded
//   with (-1, 0xffffffff) (nn.c:463-525).  Ties are broken by lower base
//   index here, which is what the reference's ascending scan gives for k = 1.
//
// Kernel structure (gfx950, wave64):
//   * a workgroup of W waves; each wave owns 32*NB queries (MFMA columns)
//     and streams the workgroup's base range in blocks of 32 rows (MFMA
//     rows) straight from global memory into registers;
//   * v_mfma_f32_32x32x2_f32: lane (r = lane%32, h = lane/32) supplies row r
//     / column r and the k-slice h; per 16-wide k-chunk a lane loads 8
//     contiguous floats (two dwordx4) of its row, so the A and B operands
//     use the same k permutation and every 64-byte row segment is read once;
//   * the 32x32 accumulator gives each lane one query and 16 base rows:
//     selection needs no cross-lane traffic until the end (k = 1), or an
//     insertion into a per-query sorted list in LDS when a value beats the
//     query's current k-th best (k > 1);
//   * the base range is split over blockIdx.y when there are too few query
//     tiles to fill 256 CUs; partial lists are merged by a second kernel.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>

#include "yk.h"
#include "yk_dev.h"

typedef float f32x16 __attribute__((ext_vector_type(16)));

// ---------------------------------------------------------------------------
// norms

template <bool WF, bool WD>
__global__ __launch_bounds__(256) void yk_norms_kernel(const float *__restrict__ x, long n,
                                                       int ld, int d, float *__restrict__ of,
                                                       double *__restrict__ od) {
  // 256 rows per workgroup, taken 32 columns (one 128-byte line per row) at
  // a time: coalesced float4 loads into an LDS tile, then one thread per row
  // sums its 32 values in column order.  The next chunk's loads are issued
  // before the current chunk is summed.
  constexpr int CW = 32, NL = CW / 4;  // columns per chunk, float4 per row segment
  __shared__ float tile[256][CW + 1];
  const long row0 = (long)blockIdx.x * 256;
  const int t = threadIdx.x;
  float sf = 0.f;
  double sd = 0.0;
  float4 buf[NL];
  auto load = [&](int c0) {
#pragma unroll
    for (int i = 0; i < NL; i++) {
      const int f = t + 256 * i, r = f / NL, cg = f % NL;
      const long row = row0 + r;
      buf[i] = make_float4(0.f, 0.f, 0.f, 0.f);
      if (row < n && c0 + cg * 4 < ld) buf[i] = *(const float4 *)(x + row * ld + c0 + cg * 4);
    }
  };
  load(0);
  for (int c0 = 0; c0 < d; c0 += CW) {
#pragma unroll
    for (int i = 0; i < NL; i++) {
      const int f = t + 256 * i, r = f / NL, cg = f % NL;
      tile[r][cg * 4 + 0] = buf[i].x;
      tile[r][cg * 4 + 1] = buf[i].y;
      tile[r][cg * 4 + 2] = buf[i].z;
      tile[r][cg * 4 + 3] = buf[i].w;
    }
    __syncthreads();
    if (c0 + CW < d) load(c0 + CW);
    const int cmax = d - c0 < CW ? d - c0 : CW;
    for (int j = 0; j < cmax; j++) {
      float v = tile[t][j];
      float p = __fmul_rn(v, v);  // float product, never fused
      if (WF) sf = __fadd_rn(sf, p);
      if (WD) sd = __dadd_rn(sd, (double)p);
    }
    __syncthreads();
  }
  long row = row0 + t;
  if (row < n) {
    if (WF) of[row] = sf;
    if (WD) od[row] = sd;
  }
}

extern "C" int yk_norms(const float *x, long n, int d, float *out_f, double *out_d) {
  YK_TRY(yk_init());
  if (n <= 0 || (!out_f && !out_d)) return YK_OK;
  int ld = yk_ld(d);
  dim3 grid((unsigned)((n + 255) / 256));
  int tok = yk_prof_begin(YK_PROF_NORM);
  if (out_f && out_d)
    hipLaunchKernelGGL((yk_norms_kernel<true, true>), grid, dim3(256), 0, yk_stream(), x, n, ld,
                       d, out_f, out_d);
  else if (out_f)
    hipLaunchKernelGGL((yk_norms_kernel<true, false>), grid, dim3(256), 0, yk_stream(), x, n, ld,
                       d, out_f, out_d);
  else
    hipLaunchKernelGGL((yk_norms_kernel<false, true>), grid, dim3(256), 0, yk_stream(), x, n, ld,
                       d, out_f, out_d);
  yk_prof_end(tok);
  YK_HIP(hipGetLastError());
  return YK_OK;
}

// ---------------------------------------------------------------------------
// cross distances: dist2[i + ldd*j], a = "base" role (float norms, MFMA
// columns), b = "query" role (double norms, MFMA rows); a wave computes 32
// rows of b by 64 columns of a, stores are 128-byte rows.

__global__ __launch_bounds__(256) void yk_cross_kernel(const float *__restrict__ A,
                                                       const float *__restrict__ an, long na,
                                                       const float *__restrict__ B,
                                                       const double *__restrict__ bnd, long nb,
                                                       int ld, int type, float *__restrict__ out,
                                                       long ldd) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = lane & 31, h = lane >> 5;
  const long j0 = ((long)blockIdx.y * 4 + wave) * 32;  // rows of b
  const long i0 = (long)blockIdx.x * 64;               // columns of a
  if (j0 >= nb) return;
  const int nch = ld >> 4;
  const long jrow = min(j0 + r, nb - 1);
  long icol[2];
  icol[0] = min(i0 + r, na - 1);
  icol[1] = min(i0 + 32 + r, na - 1);
  f32x16 acc[2];
#pragma unroll
  for (int t = 0; t < 2; t++)
#pragma unroll
    for (int e = 0; e < 16; e++) acc[t][e] = 0.f;
  for (int c = 0; c < nch; c++) {
    const float *bp = B + jrow * ld + c * 16 + h * 8;
    float4 a0 = *(const float4 *)bp, a1 = *(const float4 *)(bp + 4);
#pragma unroll
    for (int t = 0; t < 2; t++) {
      const float *ap = A + icol[t] * ld + c * 16 + h * 8;
      float4 b0 = *(const float4 *)ap, b1 = *(const float4 *)(ap + 4);
      acc[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.x, b0.x, acc[t], 0, 0, 0);
      acc[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.y, b0.y, acc[t], 0, 0, 0);
      acc[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.z, b0.z, acc[t], 0, 0, 0);
      acc[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.w, b0.w, acc[t], 0, 0, 0);
      acc[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.x, b1.x, acc[t], 0, 0, 0);
      acc[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.y, b1.y, acc[t], 0, 0, 0);
      acc[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.z, b1.z, acc[t], 0, 0, 0);
      acc[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.w, b1.w, acc[t], 0, 0, 0);
    }
  }
#pragma unroll
  for (int t = 0; t < 2; t++) {
    const long i = i0 + 32 * t + r;
    if (i >= na) continue;
    const float ani = type == YK_DOT ? 0.f : an[i];
#pragma unroll
    for (int e = 0; e < 16; e++) {
      const long j = j0 + 8 * (e >> 2) + 4 * h + (e & 3);
      if (j >= nb) continue;
      float v;
      if (type == YK_DOT)
        v = acc[t][e];
      else
        v = (float)(bnd[j] + (double)ani) - 2.0f * acc[t][e];
      out[j * ldd + i] = v;
    }
  }
}

extern "C" int yk_cross_distances(int type, long na, long nb, int d, const float *a,
                                  const float *an, const float *b, const double *bnd,
                                  float *dist2, long ldd) {
  YK_TRY(yk_init());
  if (type == 12) type = YK_L2;
  if (type != YK_L2 && type != YK_DOT)
    return yk_set_error(YK_EINVAL, "distance type %d has no GPU path (2, 12, 16 only)", type);
  if (na <= 0 || nb <= 0) return YK_OK;
  dim3 grid((unsigned)((na + 63) / 64), (unsigned)((nb + 127) / 128));
  int tok = yk_prof_begin(YK_PROF_FUSED);
  hipLaunchKernelGGL(yk_cross_kernel, grid, dim3(256), 0, yk_stream(), a, an, na, b, bnd, nb,
                     yk_ld(d), type, dist2, ldd);
  yk_prof_end(tok);
  YK_HIP(hipGetLastError());
  return YK_OK;
}

// ---------------------------------------------------------------------------
// FP64 pair distances (error report)

__global__ void yk_pair_f64_kernel(int type, long nq, int d, int ld, int k,
                                   const float *__restrict__ q, const float *__restrict__ b,
                                   const int *__restrict__ assign, double *__restrict__ out) {
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= nq * k) return;
  long qi = t / k;
  int j = assign[t];
  if (j < 0) {
    out[t] = __longlong_as_double(0x7ff8000000000000ll);
    return;
  }
  const float *x = q + qi * ld, *y = b + (long)j * ld;
  double s = 0.0;
  // product then sum, each rounded (a fused form would round once, which
  // differs when (x - y)^2 needs more than 53 bits)
  for (int l = 0; l < d; l++) {
    double u = type == YK_DOT ? (double)x[l] : (double)x[l] - (double)y[l];
    double p = type == YK_DOT ? __dmul_rn(u, (double)y[l]) : __dmul_rn(u, u);
    s = __dadd_rn(s, p);
  }
  out[t] = s;
}

extern "C" int yk_pair_dist_f64(int type, long nq, int d, int k, const float *q,
                                const float *b, const int *assign, double *out) {
  YK_TRY(yk_init());
  if (type == 12) type = YK_L2;
  long n = nq * k;
  if (n <= 0) return YK_OK;
  hipLaunchKernelGGL(yk_pair_f64_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     yk_stream(), type, nq, d, yk_ld(d), k, q, b, assign, out);
  YK_HIP(hipGetLastError());
  return YK_OK;
}

// ---------------------------------------------------------------------------
// shortlist re-rank (nn.c:528-580): for each query, the leading non-negative
// entries of its list get the distance of compute_distances_1 (double norms,
// nn.c:132-154: float(double|b|^2 + double|q|^2) - 2 q.b) and are re-sorted
// by (distance, position in the list) like fvec_sort_index (sorting.c:282-316).
// One wave per query; lists up to YK_KMAX entries.

__global__ __launch_bounds__(64) void yk_shortlist_kernel(const float *__restrict__ b,
                                                          const float *__restrict__ v, int d,
                                                          int ld, int k, int *__restrict__ idx,
                                                          float *__restrict__ dis) {
  __shared__ uint64_t key[YK_KMAX];
  __shared__ int lab[YK_KMAX];
  const long q = blockIdx.x;
  const int lane = threadIdx.x;
  int *row = idx + q * k;
  // ki = leading run of non-negative labels
  int ki = k;
  for (int j0 = 0; j0 < k; j0 += 64) {
    int j = j0 + lane;
    bool neg = j < k && row[j] < 0;
    unsigned long long m = __ballot(neg);
    if (m) {
      ki = j0 + __ffsll((long long)m) - 1;
      break;
    }
  }
  const float *x = v + q * ld;
  double qn = 0.0;
  for (int l = 0; l < d; l++) qn = __dadd_rn(qn, (double)__fmul_rn(x[l], x[l]));
  int p2 = 1;
  while (p2 < ki) p2 <<= 1;
  for (int j = lane; j < p2; j += 64) {
    if (j < ki) {
      int lb = row[j];
      lab[j] = lb;
      const float *y = b + (long)lb * ld;
      double bn = 0.0;
      float dot = 0.f;
      for (int l = 0; l < d; l++) {
        bn = __dadd_rn(bn, (double)__fmul_rn(y[l], y[l]));
        dot = __fadd_rn(dot, __fmul_rn(y[l], x[l]));
      }
      float pre = (float)(bn + qn);
      float val = __fadd_rn(pre, -2.0f * dot);
      uint32_t kv = val != val ? 0xffffffffu : yk_fkey(val);
      key[j] = ((uint64_t)kv << 32) | (uint32_t)j;
    } else {
      key[j] = YK_KEY_NONE;
    }
  }
  __syncthreads();
  for (int size = 2; size <= p2; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = lane; t < p2 / 2; t += 64) {
        int lo = 2 * t - (t & (stride - 1)), hi = lo + stride;
        bool up = (lo & size) == 0;
        uint64_t a = key[lo], c = key[hi];
        if ((a > c) == up) {
          key[lo] = c;
          key[hi] = a;
        }
      }
      __syncthreads();
    }
  for (int j = lane; j < ki; j += 64) {
    uint64_t kk = key[j];
    uint32_t kv = (uint32_t)(kk >> 32);
    row[j] = lab[(uint32_t)kk];
    dis[q * k + j] = kv == 0xffffffffu ? __uint_as_float(0x7fc00000u) : yk_fkey_inv(kv);
  }
}

extern "C" int yk_rerank_shortlist(long n, int d, int k, const float *b, const float *v,
                                   int *idx, float *dis) {
  YK_TRY(yk_init());
  if (k > YK_KMAX) return yk_set_error(YK_EINVAL, "shortlist of %d > %d entries", k, YK_KMAX);
  if (n <= 0 || k <= 0) return YK_OK;
  int tok = yk_prof_begin(YK_PROF_RERANK);
  hipLaunchKernelGGL(yk_shortlist_kernel, dim3((unsigned)n), dim3(64), 0, yk_stream(), b, v, d,
                     yk_ld(d), k, idx, dis);
  yk_prof_end(tok);
  YK_HIP(hipGetLastError());
  return YK_OK;
}

// ---------------------------------------------------------------------------
// Distances of the shortlist entries inside a window of the base
// (knn_recompute_exact_dists, nn.c:583-600): for query q, from entry kp[q]
// on, every entry whose label lies in [label0, label0 + nb) up to the first
// one past the window gets the difference-form distance, accumulated in
// double in component order and rounded to float (vector.c:2348-2359); kp[q]
// advances to that first entry.  An entry below the window met before it is
// the reference's assert: the smallest such query is recorded in *bad (the
// host reports it).  One wave per query; a lane per entry, its row read 16
// bytes at a time (the zero pad columns add exact zeros).

__global__ __launch_bounds__(256) void yk_recompute_kernel(const float *__restrict__ b,
                                                          const float *__restrict__ v, long nq,
                                                          long nb, int ld, int k, int label0,
                                                          int *__restrict__ kp,
                                                          const int *__restrict__ idx,
                                                          float *__restrict__ dis, int *bad) {
  const int lane = threadIdx.x & 63;
  const long q = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (q >= nq) return;
  const int k0 = kp[q];
  const int *row = idx + q * k;
  // first entry from k0 on that leaves the window (below or past it)
  int stop = k;
  bool below = false;
  for (int i0 = k0; i0 < k; i0 += 64) {
    const int i = i0 + lane;
    const long j = i < k ? (long)row[i] - label0 : 0;
    const unsigned long long m = __ballot(i < k && (j < 0 || j >= nb));
    if (m) {
      const int f = __ffsll((long long)m) - 1;
      stop = i0 + f;
      below = __shfl(j, f) < 0;
      break;
    }
  }
  if (below) {
    if (lane == 0) atomicMin(bad, (int)q);
    return;
  }
  const float *vq = v + q * ld;
  for (int i = k0 + lane; i < stop; i += 64) {
    const float *bj = b + ((long)row[i] - label0) * ld;
    double s = 0.0;
    for (int c = 0; c < ld; c += 4) {
      const float4 x = *(const float4 *)(vq + c), y = *(const float4 *)(bj + c);
      double a = (double)x.x - (double)y.x;
      s = __dadd_rn(s, __dmul_rn(a, a));
      a = (double)x.y - (double)y.y;
      s = __dadd_rn(s, __dmul_rn(a, a));
      a = (double)x.z - (double)y.z;
      s = __dadd_rn(s, __dmul_rn(a, a));
      a = (double)x.w - (double)y.w;
      s = __dadd_rn(s, __dmul_rn(a, a));
    }
    dis[q * k + i] = (float)s;
  }
  if (lane == 0 && stop > k0) kp[q] = stop;
}

extern "C" int yk_recompute_exact_dists(long nq, long nb, int d, int k, const float *b,
                                        const float *v, int label0, int *kp, const int *idx,
                                        float *dis, int *bad) {
  YK_TRY(yk_init());
  if (nq <= 0 || k <= 0) return YK_OK;
  int tok = yk_prof_begin(YK_PROF_RERANK);
  hipLaunchKernelGGL(yk_recompute_kernel, dim3((unsigned)((nq + 3) / 4)), dim3(256), 0,
                     yk_stream(), b, v, nq, nb, yk_ld(d), k, label0, kp, idx, dis, bad);
  yk_prof_end(tok);
  YK_HIP(hipGetLastError());
  return YK_OK;
}

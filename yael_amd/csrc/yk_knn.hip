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
  __shared__ float tile[256][17];
  const long row0 = (long)blockIdx.x * 256;
  const int t = threadIdx.x;
  float sf = 0.f;
  double sd = 0.0;
  for (int c0 = 0; c0 < d; c0 += 16) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
      int f = t + 256 * i, r = f >> 2, cg = f & 3;
      long row = row0 + r;
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      if (row < n) v = *(const float4 *)(x + row * ld + c0 + cg * 4);
      tile[r][cg * 4 + 0] = v.x;
      tile[r][cg * 4 + 1] = v.y;
      tile[r][cg * 4 + 2] = v.z;
      tile[r][cg * 4 + 3] = v.w;
    }
    __syncthreads();
    int cmax = d - c0 < 16 ? d - c0 : 16;
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
// fused distance + selection

struct FusedArgs {
  const float *q;
  const double *qn;
  long nq;
  const float *b;
  const float *bn;  // >= round_up(nb, 32) entries
  const float *w;   // optional, >= round_up(nb, 32) entries
  long nb;
  int ld;
  int type;
  long brange;  // base rows per split, multiple of 32
  int k;
  int nsplit;
  int tiles_x;  // query tiles
  int total;    // query tiles * splits
  int per_xcd;  // work items per XCD slot
  int *out_i;
  float *out_d;
  uint64_t *part;  // [nq][nsplit][k]
};

// Work-item index for this workgroup: the hardware hands consecutive block
// ids to the 8 XCDs in turn; give each XCD a contiguous run of work items
// (split-major), so workgroups sharing a base range share that XCD's L2.
__device__ __forceinline__ int yk_xcd_item(int per_xcd) {
  int id = blockIdx.x;
  return (id & 7) * per_xcd + (id >> 3);
}

__device__ __forceinline__ void yk_wave_fence() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

template <int NB, bool QRES, int MODE>
__global__ __launch_bounds__(256) void yk_fused_f32(FusedArgs a) {
  extern __shared__ uint64_t lds_keys[];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  const int r = lane & 31, h = lane >> 5;
  const int item = yk_xcd_item(a.per_xcd);
  if (item >= a.total) return;
  const int tile = item % a.tiles_x, split = item / a.tiles_x;
  const long qbase = ((long)tile * nwave + wave) * (32 * NB);
  const long bbeg = (long)split * a.brange;
  const long bend = min(a.nb, bbeg + a.brange);
  const int nch = a.ld >> 4;
  const int ld = a.ld;
  const bool dot = (a.type == YK_DOT);

  long qrow[NB];
  double qnd[NB];
#pragma unroll
  for (int j = 0; j < NB; j++) {
    long qi = qbase + j * 32 + r;
    qrow[j] = qi < a.nq ? qi : a.nq - 1;
    qnd[j] = dot ? 0.0 : a.qn[qrow[j]];
  }

  float4 qreg[NB][QRES ? 16 : 1];
  if (QRES) {
#pragma unroll
    for (int j = 0; j < NB; j++)
#pragma unroll
      for (int c = 0; c < 8; c++)
        if (c < nch) {
          const float *p = a.q + qrow[j] * ld + c * 16 + h * 8;
          qreg[j][2 * c] = *(const float4 *)p;
          qreg[j][2 * c + 1] = *(const float4 *)(p + 4);
        }
  }

  // selection state
  float bv[NB];
  int bi[NB];
  uint64_t thr[NB];
  uint64_t *keys = nullptr;
  int *cnts = nullptr;
  const int k = a.k;
#pragma unroll
  for (int j = 0; j < NB; j++) {
    bv[j] = 1e30f;
    bi[j] = -1;
    thr[j] = YK_KEY_NONE;
  }
  if (MODE == 1) {
    keys = lds_keys + (size_t)wave * (32 * NB) * k;
    cnts = (int *)(lds_keys + (size_t)nwave * (32 * NB) * k) + wave * (32 * NB);
    if (lane < 32 * NB) cnts[lane] = 0;
    yk_wave_fence();
  }

  for (long bb = bbeg; bb < bend; bb += 32) {
    const long brow = min(bb + r, a.nb - 1);
    const float *bp = a.b + brow * ld + h * 8;
    f32x16 acc[NB];
#pragma unroll
    for (int j = 0; j < NB; j++)
#pragma unroll
      for (int e = 0; e < 16; e++) acc[j][e] = 0.f;

    if (QRES) {
      float4 av[16];
#pragma unroll
      for (int c = 0; c < 8; c++)
        if (c < nch) {
          av[2 * c] = *(const float4 *)(bp + c * 16);
          av[2 * c + 1] = *(const float4 *)(bp + c * 16 + 4);
        }
#pragma unroll
      for (int c = 0; c < 8; c++)
        if (c < nch) {
#pragma unroll
          for (int j = 0; j < NB; j++) {
            float4 a0 = av[2 * c], a1 = av[2 * c + 1];
            float4 b0 = qreg[j][QRES ? 2 * c : 0], b1 = qreg[j][QRES ? 2 * c + 1 : 0];
            acc[j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.x, b0.x, acc[j], 0, 0, 0);
            acc[j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.y, b0.y, acc[j], 0, 0, 0);
            acc[j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.z, b0.z, acc[j], 0, 0, 0);
            acc[j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.w, b0.w, acc[j], 0, 0, 0);
            acc[j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.x, b1.x, acc[j], 0, 0, 0);
            acc[j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.y, b1.y, acc[j], 0, 0, 0);
            acc[j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.z, b1.z, acc[j], 0, 0, 0);
            acc[j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.w, b1.w, acc[j], 0, 0, 0);
          }
        }
    } else {
      for (int c = 0; c < nch; c++) {
        float4 a0 = *(const float4 *)(bp + c * 16);
        float4 a1 = *(const float4 *)(bp + c * 16 + 4);
#pragma unroll
        for (int j = 0; j < NB; j++) {
          const float *qp = a.q + qrow[j] * ld + c * 16 + h * 8;
          float4 b0 = *(const float4 *)qp, b1 = *(const float4 *)(qp + 4);
          acc[j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.x, b0.x, acc[j], 0, 0, 0);
          acc[j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.y, b0.y, acc[j], 0, 0, 0);
          acc[j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.z, b0.z, acc[j], 0, 0, 0);
          acc[j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.w, b0.w, acc[j], 0, 0, 0);
          acc[j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.x, b1.x, acc[j], 0, 0, 0);
          acc[j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.y, b1.y, acc[j], 0, 0, 0);
          acc[j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.z, b1.z, acc[j], 0, 0, 0);
          acc[j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.w, b1.w, acc[j], 0, 0, 0);
        }
      }
    }

    // epilogue: lane holds query column r and base rows 8g + 4h + e
    float vv[NB][16];
#pragma unroll
    for (int g = 0; g < 4; g++) {
      const long b4 = bb + 8 * g + 4 * h;
      float4 n4 = dot ? make_float4(0.f, 0.f, 0.f, 0.f) : *(const float4 *)(a.bn + b4);
      float4 w4 = a.w ? *(const float4 *)(a.w + b4) : make_float4(1.f, 1.f, 1.f, 1.f);
      float nn[4] = {n4.x, n4.y, n4.z, n4.w};
      float ww[4] = {w4.x, w4.y, w4.z, w4.w};
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int i = 4 * g + e;
#pragma unroll
        for (int j = 0; j < NB; j++) {
          float v;
          if (dot)
            v = acc[j][i];
          else
            v = (float)(qnd[j] + (double)nn[e]) - 2.0f * acc[j][i];
          if (a.w) v *= ww[e];
          vv[j][i] = v;
        }
      }
    }

    if (MODE == 0) {
#pragma unroll
      for (int i = 0; i < 16; i++) {
        const long bidx = bb + 8 * (i >> 2) + 4 * h + (i & 3);
#pragma unroll
        for (int j = 0; j < NB; j++)
          if (bidx < bend && vv[j][i] < bv[j]) {
            bv[j] = vv[j][i];
            bi[j] = (int)bidx;
          }
      }
    } else {
      uint32_t mask[NB];
      bool any = false;
#pragma unroll
      for (int j = 0; j < NB; j++) {
        mask[j] = 0;
#pragma unroll
        for (int i = 0; i < 16; i++) {
          const long bidx = bb + 8 * (i >> 2) + 4 * h + (i & 3);
          float v = vv[j][i];
          if (bidx < bend && v == v && yk_key(v, (int)bidx) < thr[j]) mask[j] |= 1u << i;
        }
        any |= mask[j] != 0;
      }
      if (__any(any)) {
        for (int turn = 0; turn < 2; turn++) {
          if (h == turn) {
#pragma unroll
            for (int j = 0; j < NB; j++) {
              if (!mask[j]) continue;
              const int ql = j * 32 + r;
              uint64_t *L = keys + (size_t)ql * k;
              int cnt = cnts[ql];
              uint64_t T = thr[j];
              for (int i = 0; i < 16; i++) {
                if (!(mask[j] >> i & 1)) continue;
                const long bidx = bb + 8 * (i >> 2) + 4 * h + (i & 3);
                uint64_t key = yk_key(vv[j][i], (int)bidx);
                if (key >= T) continue;
                int pos = cnt < k ? cnt++ : k - 1;
                while (pos > 0 && L[pos - 1] > key) {
                  L[pos] = L[pos - 1];
                  pos--;
                }
                L[pos] = key;
                T = cnt == k ? L[k - 1] : YK_KEY_NONE;
              }
              cnts[ql] = cnt;
            }
          }
          yk_wave_fence();
          // both lanes of each query pick up the new threshold
#pragma unroll
          for (int j = 0; j < NB; j++) {
            const int ql = j * 32 + r;
            int cnt = cnts[ql];
            thr[j] = cnt == k ? keys[(size_t)ql * k + k - 1] : YK_KEY_NONE;
          }
          yk_wave_fence();
        }
      }
    }
  }

  // write results
  if (MODE == 0) {
#pragma unroll
    for (int j = 0; j < NB; j++) {
      float ov = __shfl_xor(bv[j], 32);
      int oi = __shfl_xor(bi[j], 32);
      if (ov < bv[j] || (ov == bv[j] && oi < bi[j])) {
        bv[j] = ov;
        bi[j] = oi;
      }
      const long qi = qbase + j * 32 + r;
      if (h == 0 && qi < a.nq) {
        if (a.nsplit == 1) {
          a.out_d[qi] = bv[j];
          a.out_i[qi] = bi[j];
        } else {
          a.part[qi * a.nsplit + split] = bi[j] < 0 ? YK_KEY_NONE : yk_key(bv[j], bi[j]);
        }
      }
    }
  } else {
#pragma unroll
    for (int j = 0; j < NB; j++) {
      const long qi = qbase + j * 32 + r;
      if (qi >= a.nq) continue;
      const int ql = j * 32 + r;
      const int cnt = cnts[ql];
      const uint64_t *L = keys + (size_t)ql * k;
      for (int s = h; s < k; s += 2) {
        uint64_t key = s < cnt ? L[s] : YK_KEY_NONE;
        if (a.nsplit == 1) {
          a.out_i[qi * k + s] = key == YK_KEY_NONE ? -1 : (int)(uint32_t)key;
          a.out_d[qi * k + s] =
              key == YK_KEY_NONE ? __uint_as_float(0xffffffffu) : yk_fkey_inv((uint32_t)(key >> 32));
        } else {
          a.part[(qi * a.nsplit + split) * (long)k + s] = key;
        }
      }
    }
  }
}

// k = 1 merge of per-split minima
__global__ void yk_merge1_kernel(const uint64_t *__restrict__ part, long nq, int nsplit,
                                 int *out_i, float *out_d) {
  long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (q >= nq) return;
  uint64_t best = YK_KEY_NONE;
  for (int s = 0; s < nsplit; s++) {
    uint64_t v = part[q * nsplit + s];
    best = v < best ? v : best;
  }
  if (best == YK_KEY_NONE) {
    out_i[q] = -1;
    out_d[q] = 1e30f;
  } else {
    out_i[q] = (int)(uint32_t)best;
    out_d[q] = yk_fkey_inv((uint32_t)(best >> 32));
  }
}

// k > 1 merge: bitonic sort of the nsplit*k keys of one query in LDS.
__global__ __launch_bounds__(256) void yk_mergek_kernel(const uint64_t *__restrict__ part,
                                                        long nq, int nsplit, int k, int p2,
                                                        int *out_i, float *out_d) {
  extern __shared__ uint64_t s[];
  const long q = blockIdx.x;
  const int n = nsplit * k;
  const uint64_t *src = part + q * (long)n;
  for (int i = threadIdx.x; i < p2; i += blockDim.x) s[i] = i < n ? src[i] : YK_KEY_NONE;
  __syncthreads();
  for (int size = 2; size <= p2; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = threadIdx.x; t < p2 / 2; t += blockDim.x) {
        int lo = 2 * t - (t & (stride - 1));
        int hi = lo + stride;
        bool up = (lo & size) == 0;
        uint64_t x = s[lo], y = s[hi];
        if ((x > y) == up) {
          s[lo] = y;
          s[hi] = x;
        }
      }
      __syncthreads();
    }
  }
  for (int i = threadIdx.x; i < k; i += blockDim.x) {
    uint64_t key = s[i];
    out_i[q * k + i] = key == YK_KEY_NONE ? -1 : (int)(uint32_t)key;
    out_d[q * k + i] =
        key == YK_KEY_NONE ? __uint_as_float(0xffffffffu) : yk_fkey_inv((uint32_t)(key >> 32));
  }
}

// ---------------------------------------------------------------------------
// planning

struct YkPlan {
  int nb_blocks;  // NB template value
  int waves;      // waves per workgroup
  int nsplit;
  long brange;
  int tiles;
  size_t lds;
};

static size_t g_lds_max = 0;

static int yk_lds_limit(void) {
  if (!g_lds_max) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, 0) != hipSuccess) return 65536;
    g_lds_max = p.sharedMemPerBlock;
    if (g_lds_max < 65536) g_lds_max = 65536;
  }
  return (int)g_lds_max;
}

static YkPlan yk_make_plan(long nq, long nb, int d, int k) {
  YkPlan p;
  p.nb_blocks = 1;
  p.waves = 4;
  const int lds_max = yk_lds_limit();
  if (k > 1) {
    size_t per_wave = (size_t)32 * p.nb_blocks * k * 8 + 32 * p.nb_blocks * 4;
    while (p.waves > 1 && per_wave * p.waves > (size_t)lds_max) p.waves >>= 1;
    p.lds = per_wave * p.waves;
  } else {
    p.lds = 0;
  }
  const long qt = 32L * p.nb_blocks * p.waves;
  p.tiles = (int)((nq + qt - 1) / qt);
  // enough workgroups for 256 CUs (4 per CU), but at least 512 base rows per
  // split, and split lists small enough for the merge kernel
  long want = (1024 + p.tiles - 1) / p.tiles;
  long maxs = (nb + 511) / 512;
  if (want > maxs) want = maxs;
  if (k > 1) {
    long cap = 4096 / k;
    if (want > cap) want = cap;
  }
  if (want < 1) want = 1;
  long br = (nb + want - 1) / want;
  br = (br + 31) & ~31L;
  p.brange = br;
  p.nsplit = (int)((nb + br - 1) / br);
  (void)d;
  return p;
}

extern "C" int yk_knn_plan(long nq, long nb, int d, int k, int *nsplit, int *waves,
                           int *nblk) {
  YK_TRY(yk_init());
  YkPlan p = yk_make_plan(nq, nb, d, k);
  *nsplit = p.nsplit;
  *waves = p.waves;
  *nblk = p.nb_blocks;
  return YK_OK;
}

// grow-only scratch for split partials
static void *g_part = nullptr;
static size_t g_part_bytes = 0;

static int yk_part_buffer(size_t bytes, void **out) {
  if (bytes > g_part_bytes) {
    if (g_part) (void)hipFree(g_part);
    g_part = nullptr;
    g_part_bytes = 0;
    hipError_t e = hipMalloc(&g_part, bytes);
    if (e != hipSuccess) return yk_hip_fail(e, "hipMalloc(partials)");
    g_part_bytes = bytes;
  }
  *out = g_part;
  return YK_OK;
}

template <int MODE, bool QRES>
static int yk_launch_fused(const YkPlan &p, FusedArgs &a) {
  const int total = p.tiles * p.nsplit;
  const int per = (total + 7) / 8;
  a.tiles_x = p.tiles;
  a.total = total;
  a.per_xcd = per;
  auto kern = yk_fused_f32<1, QRES, MODE>;
  if (p.lds > 65536) {
    YK_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)p.lds));
  }
  int tok = yk_prof_begin(YK_PROF_FUSED);
  hipLaunchKernelGGL(kern, dim3(per * 8), dim3(64 * p.waves), p.lds, yk_stream(), a);
  yk_prof_end(tok);
  YK_HIP(hipGetLastError());
  return YK_OK;
}

extern "C" int yk_knn_f32(int type, long nq, long nb, int d, int k, const float *q,
                          const double *qn, const float *b, const float *bn, const float *w,
                          int *assign, float *dis) {
  YK_TRY(yk_init());
  if (type == 12) type = YK_L2;
  if (type != YK_L2 && type != YK_DOT)
    return yk_set_error(YK_EINVAL, "distance type %d has no GPU path (2, 12, 16 only)", type);
  if (k < 1 || k > nb || k > YK_KMAX)
    return yk_set_error(YK_EINVAL, "k=%d out of range (nb=%ld, max %d)", k, nb, YK_KMAX);
  if (nq <= 0) return YK_OK;
  if (nb > 0x7fffffffL || nq > 0x7fffffffL) return yk_set_error(YK_EINVAL, "sizes exceed int32");
  YkPlan p = yk_make_plan(nq, nb, d, k);
  if (k > 1 && p.lds > (size_t)yk_lds_limit())
    return yk_set_error(YK_EINVAL, "k=%d needs %zu bytes of LDS", k, p.lds);
  FusedArgs a;
  a.q = q;
  a.qn = qn;
  a.nq = nq;
  a.b = b;
  a.bn = bn;
  a.w = w;
  a.nb = nb;
  a.ld = yk_ld(d);
  a.type = type;
  a.brange = p.brange;
  a.k = k;
  a.nsplit = p.nsplit;
  a.out_i = assign;
  a.out_d = dis;
  a.part = nullptr;
  if (p.nsplit > 1) {
    void *buf;
    YK_TRY(yk_part_buffer(sizeof(uint64_t) * (size_t)nq * p.nsplit * k, &buf));
    a.part = (uint64_t *)buf;
  }
  const bool qres = a.ld <= 128;
  if (k == 1) {
    if (qres) YK_TRY((yk_launch_fused<0, true>(p, a)));
    else YK_TRY((yk_launch_fused<0, false>(p, a)));
  } else {
    if (qres) YK_TRY((yk_launch_fused<1, true>(p, a)));
    else YK_TRY((yk_launch_fused<1, false>(p, a)));
  }
  if (p.nsplit > 1) {
    int tok = yk_prof_begin(YK_PROF_MERGE);
    if (k == 1) {
      hipLaunchKernelGGL(yk_merge1_kernel, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0,
                         yk_stream(), a.part, nq, p.nsplit, assign, dis);
    } else {
      int n = p.nsplit * k, p2 = 1;
      while (p2 < n) p2 <<= 1;
      hipLaunchKernelGGL(yk_mergek_kernel, dim3((unsigned)nq), dim3(256), sizeof(uint64_t) * p2,
                         yk_stream(), a.part, nq, p.nsplit, k, p2, assign, dis);
    }
    yk_prof_end(tok);
    YK_HIP(hipGetLastError());
  }
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
  for (int l = 0; l < d; l++) {
    if (type == YK_DOT) {
      s = __fma_rn((double)x[l], (double)y[l], s);
    } else {
      double u = (double)x[l] - (double)y[l];
      s = __fma_rn(u, u, s);
    }
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

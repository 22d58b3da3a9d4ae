// yk_bf16.hip -- precision form of the exhaustive kNN: a BF16 MFMA first pass
// selects k' candidates per query, an FP32 re-rank recomputes their
// distances exactly as the FP32 form does, and a certificate proves that the
// FP32 form's k nearest are among the candidates (queries that fail it are
// searched again with the FP32 form).
//
// First pass.  Vectors are stored as BF16 with the squared norms appended
// as three BF16 pieces (hi + mid + lo carries 24 bits), so one MFMA
// contraction yields the whole expanded distance:
//     base  row  [ bf(b_0) .. bf(b_d-1) | nb_hi nb_mid nb_lo | 1 1 1 ]
//     query col  [ bf(-2q_0) .. bf(-2q_d-1) | 1 1 1 | nq_hi nq_mid nq_lo ]
//     acc = ||q||^2 + ||b||^2 - 2 bf(q).bf(b)        (FP32 accumulation)
// (type 16: plain bf(q).bf(b)).  The epilogue is a compare and a rare
// append: v_mfma_f32_32x32x16_bf16 does 16 k-steps per instruction.
//
// Layout.  Query copies are row-major (row stride d').  Base copies are
// fragment-major, the order in which the MFMA consumes them: the rows are
// taken in blocks of 32 (the set padded with zero rows to a multiple of 32);
// block b, 16-wide chunk c is one 1 KB fragment at ((b * nch + c) * 64) * 8
// elements, lane h * 32 + r of a wave holding row 32b + r, columns 16c + 8h
// .. 16c + 8h + 7.  A wave reads a fragment with one fully coalesced 1 KB
// load (a row-major copy would touch 32 rows' 32-byte pieces per load).
//
// Re-rank.  Each candidate's distance is recomputed with the FP32 form's own
// instruction sequence (v_mfma_f32_32x32x2_f32, same k order, same
// norm/epilogue arithmetic), so re-ranked values are bit-identical to what the
// FP32 form computes for the same pair.
//
// Certificate.  With T = the k'-th smallest first-pass value, every
// non-candidate has first-pass value >= T, so its exact distance is >=
// T - E_bf and its FP32-form value >= T - E_bf - E_32.  If the k-th re-ranked
// value F_k < T - E_bf - E_32 the FP32 form's top k are all candidates.
//   E_bf  = 2 (|e_q| max|b| + |bf q| max|e_b|)                 (Cauchy-Schwarz
//           on the BF16 residuals e = x - bf(x))
//         + gamma (2 |bf q| max|bf b| + |q|^2 + max|b|^2)     (FP32 accumulation)
//   E_32  = gamma (2 |q| max|b| + |q|^2 + max|b|^2)
// with gamma = 2 (d' + 8) 2^-23 (d' = padded length, truncating adds allowed).
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "yk.h"
#include "yk_dev.h"

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
// word 3 of a raw (untyped, stride 0) buffer descriptor on gfx9-family GPUs
#define YK_BUF_DWORD3 0x00020000

#define YK_CAND_CAP 4096
#define YK_REG_R 16
#define YK_CSTAGE 8  // MODE 3: keys staged in LDS per (lane, query tile)
#define YK_RING 128   // MODE 3 (register-resident queries): parked tiles per wave

// ---------------------------------------------------------------------------
// conversion

__device__ __forceinline__ uint16_t yk_bf16_rne(float x) {
  uint32_t u = __float_as_uint(x);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40);  // NaN stays NaN
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float yk_bf16_f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

extern "C" int yk_bf16_ld(int d, int type) {
  int extra = (type == YK_DOT) ? 0 : 6;
  return (d + extra + 15) & ~15;
}

extern "C" long yk_bf16_bytes(long n, int d, int type, int is_query) {
  const long rows = is_query ? n : (n + 31) & ~31L;
  return rows * yk_bf16_ld(d, type) * (long)sizeof(uint16_t);
}


// Query copies (row-major), one wave per vector (coalesced): BF16 row
// (augmented for L2) + error
// statistics stats[i] = {|x|, |bf(x)|, |x - bf(x)|, |x|^2} (float); the
// per-set maxima go to smax (uint bit patterns of non-negative floats, whose
// integer order is their float order) by atomicMax.
// max that keeps a NaN (fmaxf drops it): the set maxima of the statistics
// then show a NaN anywhere in the set
__device__ __forceinline__ float yk_fmaxn(float a, float b) {
  return a != a ? a : (b != b ? b : fmaxf(a, b));
}

__global__ __launch_bounds__(256) void yk_to_bf16_kernel(const float *__restrict__ x, long n,
                                                         int d, int ld, int ldh, int type,
                                                         int is_query,
                                                         uint16_t *__restrict__ out,
                                                         float4 *__restrict__ stats,
                                                         unsigned *smax) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __shared__ float4 wmax[4];
  float4 mx = make_float4(0.f, 0.f, 0.f, 0.f);
  for (long i = blockIdx.x * 4L + wave; i < n; i += gridDim.x * 4L) {
  const float *row = x + i * ld;
  uint16_t *o = out + i * ldh;
  double s2 = 0, sb = 0, se = 0;
  const float scale = (type != YK_DOT && is_query) ? -2.0f : 1.0f;
  for (int j = lane; j < d; j += 64) {
    float v = row[j];
    uint16_t hb = yk_bf16_rne(v * scale);
    float vb = yk_bf16_f(hb) / scale;  // exact: scale is a power of two
    o[j] = hb;
    s2 += (double)v * v;
    sb += (double)vb * vb;
    se += ((double)v - vb) * ((double)v - vb);
  }
  for (int off = 32; off > 0; off >>= 1) {
    s2 += __shfl_xor(s2, off);
    sb += __shfl_xor(sb, off);
    se += __shfl_xor(se, off);
  }
  const uint16_t one = 0x3f80;
  uint16_t tail[6] = {0, 0, 0, 0, 0, 0};
  int extra = 0;
  if (type != YK_DOT) {
    float nf = (float)s2;
    uint16_t h1 = yk_bf16_rne(nf);
    float r1 = nf - yk_bf16_f(h1);
    uint16_t h2 = yk_bf16_rne(r1);
    float r2 = r1 - yk_bf16_f(h2);
    uint16_t h3 = yk_bf16_rne(r2);
    if (is_query) {
      tail[0] = one; tail[1] = one; tail[2] = one; tail[3] = h1; tail[4] = h2; tail[5] = h3;
    } else {
      tail[0] = h1; tail[1] = h2; tail[2] = h3; tail[3] = one; tail[4] = one; tail[5] = one;
    }
    extra = 6;
  }
  for (int j = d + lane; j < ldh; j += 64) o[j] = (j - d) < extra ? tail[j - d] : 0;
  float4 s = make_float4((float)sqrt(s2), (float)sqrt(sb), (float)sqrt(se), (float)s2);
  if (lane == 0) stats[i] = s;
  mx = make_float4(yk_fmaxn(mx.x, s.x), yk_fmaxn(mx.y, s.y), yk_fmaxn(mx.z, s.z), yk_fmaxn(mx.w, s.w));
  }  // rows
  if (lane == 0) wmax[wave] = mx;
  __syncthreads();
  if (smax && threadIdx.x == 0) {
    for (int w = 1; w < 4; w++)
      mx = make_float4(yk_fmaxn(mx.x, wmax[w].x), yk_fmaxn(mx.y, wmax[w].y), yk_fmaxn(mx.z, wmax[w].z),
                       yk_fmaxn(mx.w, wmax[w].w));
    atomicMax(smax + 0, __float_as_uint(mx.x));
    atomicMax(smax + 1, __float_as_uint(mx.y));
    atomicMax(smax + 2, __float_as_uint(mx.z));
    atomicMax(smax + 3, __float_as_uint(mx.w));
  }
}

__global__ void yk_zero4_kernel(unsigned *p) {
  if (threadIdx.x < 4) p[threadIdx.x] = 0;
}

// The base copy (fragment-major): one wave per 32-row block; lane h*32 + r
// owns row 32b + r, columns 16c + 8h .. +7 of every chunk c, so each chunk is
// one coalesced 1 KB store.  One read of the rows: the chunks holding only
// columns below d are stored as they are converted; the chunk holding column
// d (its data part kept in registers) and the chunks after it follow once the
// row statistics are summed (the lane pair combines its halves by a
// shuffle), with the norm pieces and ones after column d.  Conversion, norm
// pieces and statistics are those of yk_to_bf16_kernel.  With fnorm, the
// float norm of each row is summed in column order exactly as yk_norms does
// (float squares, float sums): both lanes of a pair run the same sequence,
// taking the other lane's 8 values of each chunk by a shuffle.
__global__ __launch_bounds__(256) void yk_to_bf16_frag_kernel(const float *__restrict__ x, long n,
                                                              int d, int ld, int ldh, int type,
                                                              uint16_t *__restrict__ out,
                                                              float4 *__restrict__ stats,
                                                              unsigned *smax,
                                                              float *__restrict__ fnorm) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = lane & 31, h = lane >> 5;
  const int nch = ldh >> 4;
  const int cd = d >> 4;  // first chunk holding a column >= d
  const long nblk = (n + 31) >> 5;
  __shared__ float4 wmax[4];
  float4 mx = make_float4(0.f, 0.f, 0.f, 0.f);
  for (long b = blockIdx.x * 4L + wave; b < nblk; b += gridDim.x * 4L) {
    const long row = b * 32 + r;
    const bool valid = row < n;
    const float *xr = x + (valid ? row : 0) * ld;
    uint4 *ob = (uint4 *)(out + (b * nch * 64 + lane) * 8);
    // the lane's 8 values of chunk c (columns past d read as zero; the FP32
    // rows are padded to ld, a multiple of 16, so full 8-column runs below
    // d are two aligned float4 loads)
    // (a uniform test per chunk; rows past n load row 0 and are zeroed by
    // the caller, so that the loads of full chunks carry no per-lane
    // condition)
    auto vals = [&](int c, float (&v)[8]) {
      const int c0 = 16 * c + 8 * h;
      if (16 * c + 16 <= d) {
        const float4 p = *(const float4 *)(xr + c0), q = *(const float4 *)(xr + c0 + 4);
        v[0] = p.x; v[1] = p.y; v[2] = p.z; v[3] = p.w;
        v[4] = q.x; v[5] = q.y; v[6] = q.z; v[7] = q.w;
      } else {
#pragma unroll
        for (int t = 0; t < 8; t++) v[t] = c0 + t < d ? xr[c0 + t] : 0.f;
      }
    };
    auto pack = [](const uint16_t (&e)[8]) {
      return make_uint4((uint32_t)e[0] | ((uint32_t)e[1] << 16),
                        (uint32_t)e[2] | ((uint32_t)e[3] << 16),
                        (uint32_t)e[4] | ((uint32_t)e[5] << 16),
                        (uint32_t)e[6] | ((uint32_t)e[7] << 16));
    };
    double s2 = 0, sb = 0, se = 0;
    float sf = 0.f;
    uint16_t keep[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    // chunks holding data, four at a time: their loads are issued together
    const int cdl = min(nch, (d + 15) >> 4);
    for (int cg = 0; cg < cdl; cg += 4) {
    float vg[4][8];
#pragma unroll
    for (int u = 0; u < 4; u++)
      if (cg + u < cdl) vals(cg + u, vg[u]);
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int c = cg + u;
      if (c >= cdl) break;
      float v[8];
#pragma unroll
      for (int t = 0; t < 8; t++) v[t] = valid ? vg[u][t] : 0.f;
      uint16_t e[8];
#pragma unroll
      for (int t = 0; t < 8; t++) {
        e[t] = yk_bf16_rne(v[t]);  // zero past d
        const float vb = yk_bf16_f(e[t]);
        s2 += (double)v[t] * v[t];
        sb += (double)vb * vb;
        se += ((double)v[t] - vb) * ((double)v[t] - vb);
      }
      if (fnorm) {
        float o[8];
#pragma unroll
        for (int t = 0; t < 8; t++) o[t] = __shfl_xor(v[t], 32);
        const int cmax = min(16, d - 16 * c);
#pragma unroll
        for (int t = 0; t < 16; t++) {
          const float u = (t < 8) == (h == 0) ? v[t & 7] : o[t & 7];
          if (t < cmax) sf = __fadd_rn(sf, __fmul_rn(u, u));
        }
      }
      if (c < cd) {
        ob[c * 64] = pack(e);
      } else {
#pragma unroll
        for (int t = 0; t < 8; t++) keep[t] = e[t];
      }
    }
    }
    s2 += __shfl_xor(s2, 32);
    sb += __shfl_xor(sb, 32);
    se += __shfl_xor(se, 32);
    const uint16_t one = 0x3f80;
    uint16_t tail[6] = {0, 0, 0, 0, 0, 0};
    if (type != YK_DOT && valid) {
      const float nf = (float)s2;
      const uint16_t h1 = yk_bf16_rne(nf);
      const float r1 = nf - yk_bf16_f(h1);
      const uint16_t h2 = yk_bf16_rne(r1);
      const float r2 = r1 - yk_bf16_f(h2);
      const uint16_t h3 = yk_bf16_rne(r2);
      tail[0] = h1; tail[1] = h2; tail[2] = h3; tail[3] = one; tail[4] = one; tail[5] = one;
    }
    for (int c = cd; c < nch; c++) {
      const int c0 = 16 * c + 8 * h;
      uint16_t e[8];
#pragma unroll
      for (int t = 0; t < 8; t++) {
        const int col = c0 + t;
        e[t] = col < d ? keep[t] : ((col - d) < 6 ? tail[col - d] : 0);
      }
      ob[c * 64] = pack(e);
    }
    if (valid) {
      const float4 st = make_float4((float)sqrt(s2), (float)sqrt(sb), (float)sqrt(se), (float)s2);
      if (h == 0) {
        stats[row] = st;
        if (fnorm) fnorm[row] = sf;
      }
      mx = make_float4(yk_fmaxn(mx.x, st.x), yk_fmaxn(mx.y, st.y), yk_fmaxn(mx.z, st.z), yk_fmaxn(mx.w, st.w));
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    mx.x = yk_fmaxn(mx.x, __shfl_xor(mx.x, o));
    mx.y = yk_fmaxn(mx.y, __shfl_xor(mx.y, o));
    mx.z = yk_fmaxn(mx.z, __shfl_xor(mx.z, o));
    mx.w = yk_fmaxn(mx.w, __shfl_xor(mx.w, o));
  }
  if (lane == 0) wmax[wave] = mx;
  __syncthreads();
  if (smax && threadIdx.x == 0) {
    for (int w = 1; w < 4; w++)
      mx = make_float4(yk_fmaxn(mx.x, wmax[w].x), yk_fmaxn(mx.y, wmax[w].y), yk_fmaxn(mx.z, wmax[w].z),
                       yk_fmaxn(mx.w, wmax[w].w));
    atomicMax(smax + 0, __float_as_uint(mx.x));
    atomicMax(smax + 1, __float_as_uint(mx.y));
    atomicMax(smax + 2, __float_as_uint(mx.z));
    atomicMax(smax + 3, __float_as_uint(mx.w));
  }
}

extern "C" int yk_to_bf16(const float *x, long n, int d, int type, int is_query, uint16_t *out,
                          float *stats, float *stats_max, float *fnorm) {
  YK_TRY(yk_init());
  if (type == 12) type = YK_L2;
  if (n <= 0) return YK_OK;
  if (is_query && fnorm)
    return yk_set_error(YK_EINVAL, "yk_to_bf16: float norms are written for base sets only");
  const int ld = yk_ld(d), ldh = yk_bf16_ld(d, type);
  if (stats_max) hipLaunchKernelGGL(yk_zero4_kernel, dim3(1), dim3(64), 0, yk_stream(),
                                    (unsigned *)stats_max);
  if (is_query) {
    long blocks = (n + 3) / 4;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(yk_to_bf16_kernel, dim3((unsigned)blocks), dim3(256), 0, yk_stream(), x, n,
                       d, ld, ldh, type, is_query, out, (float4 *)stats, (unsigned *)stats_max);
  } else {
    // a grid-stride grid of at most 2048 workgroups (8 per CU): enough waves
    // in flight to stream the set, few enough that the per-workgroup maxima
    // do not queue on 4 atomics
    long blocks = ((n + 31) / 32 + 3) / 4;
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(yk_to_bf16_frag_kernel, dim3((unsigned)blocks), dim3(256), 0, yk_stream(),
                       x, n, d, ld, ldh, type, out, (float4 *)stats, (unsigned *)stats_max, fnorm);
  }
  YK_HIP(hipGetLastError());
  return YK_OK;
}

// ---------------------------------------------------------------------------
// BF16 scan

struct Bf16Args {
  const uint16_t *q;  // nq x ldh
  long nq;
  const uint16_t *b;  // nb x ldh
  long nb;
  int ldh;
  long brange, bstep;
  int nsplit, tiles_x, total, per_xcd;
  uint64_t *part;       // MODE 1: [nq][nsplit][2][2]; MODE 2: [nq][nsplit][2][R]
  const uint64_t *thr;  // MODE 3
  int *cnt;
  uint64_t *cand;       // MODE 3: candidate lists; MODE 2 with kp_direct: [nq][kp]
  int kp_direct;        // MODE 2, nsplit == 1: the lane pair's kp <= R smallest keys
  const uint16_t *qf;   // long vectors: the queries fragment-major (as the base)
  int *tick;            // per XCD, zero at launch: item tickets (nullptr: static order)
};

template <int R>
__device__ __forceinline__ void yk_reg_insert16(uint64_t (&t)[R], uint64_t x) {
#pragma unroll
  for (int i = R - 1; i > 0; i--) {
    uint64_t lo = t[i - 1], cur = t[i];
    t[i] = x < lo ? lo : (x < cur ? x : cur);
  }
  t[0] = x < t[0] ? x : t[0];
}

// Per-wave selection state over NB 32-query column tiles: lane (r, h) owns
// query r of each tile and, per 32-row block, rows 8g + 4h + e (g, e < 4).
template <int NB, int MODE, int R, int CST = YK_CSTAGE, int RING = 0>
struct BfSel {
  long qrow[NB];
  bool qok[NB];
  uint64_t top[NB][MODE == 2 ? R : 1];
  float bv[NB][MODE == 1 ? 2 : 1];  // MODE 1: minima of the even and odd blocks
  uint64_t T[NB];
  float Tv[NB];  // MODE 3: the bound's value; MODE 2: the value of top[j][R - 1] (or INF)
  // MODE 3: candidate keys are staged in LDS, CST per (query tile,
  // lane), and appended to the query's global list a region at a time (one
  // atomic per region instead of one per key)
  uint64_t *cs;
  int ncs[NB];
  int lane;
  float *vs;  // MODES 2, 3: this lane's 16-value LDS slot
  // MODE 3 with RING > 0: a tile whose lane minimum passes the bound is
  // parked in the wave's ring of RING entries (its 16 values, then the
  // query, the lane's first row and the query's bound key as its value and
  // row: five float4 each); the ring is drained 64 entries at a time, one
  // entry per lane, so the per-value work of the rare passing tiles runs on
  // full waves instead of once per tile for one or two lanes.  rhead, rtail:
  // wave-uniform entry counters; rend: the end of the item's row range.
  float4 *ring;
  int rhead, rtail;
  long rend;

  __device__ __forceinline__ void init(const Bf16Args &a, long qbase, int r, int lane_,
                                       uint64_t *cs_wave, float *vs_lane,
                                       float4 *ring_wave = nullptr) {
    lane = lane_;
    cs = cs_wave;
    vs = vs_lane;
    ring = ring_wave;
    rhead = rtail = 0;
#pragma unroll
    for (int j = 0; j < NB; j++) {
      ncs[j] = 0;
      const long qi = qbase + j * 32 + r;
      qok[j] = qi < a.nq;
      qrow[j] = qok[j] ? qi : a.nq - 1;
      bv[j][0] = INFINITY;
      if (MODE == 1) bv[j][MODE == 1 ? 1 : 0] = INFINITY;
#pragma unroll
      for (int s = 0; s < (MODE == 2 ? R : 1); s++) top[j][s] = YK_KEY_NONE;
      T[j] = MODE == 3 ? a.thr[qrow[j]] : 0;
      // MODE 3: a NaN bound for a query past the end or already done (thr
      // 0), which no value passes, whatever it is: the pass test is then the
      // bare compare
      Tv[j] = (T[j] == 0 || !qok[j])
                  ? __builtin_nanf("")
                  : (T[j] == YK_KEY_NONE ? INFINITY : yk_fkey_inv((uint32_t)(T[j] >> 32)));
      if (MODE == 2) Tv[j] = INFINITY;  // the empty list's bound
    }
  }

  // MODE 3: does any owned query still need the filtered scan?
  __device__ __forceinline__ bool active() const {
    bool act = false;
#pragma unroll
    for (int j = 0; j < NB; j++) act |= qok[j] && T[j] != 0;
    return act;
  }

  // one 32-row block starting at row bb: a min over the lane's 16 values
  // decides whether any can enter (fminf drops NaN, which never enters);
  // only then, per value, a float compare against the entry bound before
  // any key is built.  Lane row i is bb + 8(i/4) + (i%4) + 4h; only the last
  // block of the base can hold rows past bend (the zero pad rows of the
  // fragment-major copy), masked with one integer compare against
  // rl = bend - bb - 4h.
  __device__ __forceinline__ void block(const Bf16Args &a, const f32x16 (&acc)[NB], long bb,
                                        long bend, int h, int sub) {
    // a uniform branch: the per-row mask exists only in the (rare) partial
    // block, so the common path carries no per-value row compares (MODE 2
    // keeps one masked copy: its insertion network is large enough that a
    // second copy costs the register budget of two waves per SIMD)
    const int rl = (int)min(bend - bb, 64L) - 4 * h;
    if (MODE != 2 && bb + 32 <= bend)
      block_impl<true>(a, acc, bb, rl, h, sub);
    else
      block_impl<false>(a, acc, bb, rl, h, sub);
  }

  template <bool FULL>
  __device__ __forceinline__ void block_impl(const Bf16Args &a, const f32x16 (&acc)[NB], long bb,
                                             int rl, int h, int sub) {
#pragma unroll
    for (int j = 0; j < NB; j++) {
      if (MODE == 1) {
        // sketch: the lane's minimum value only, kept apart for the even and
        // odd blocks of the item (sub, a constant after inlining; see finish)
        // (selects, not an index: a run-time index would place bv in
        // scratch memory)
        float m = sub ? bv[j][MODE == 1 ? 1 : 0] : bv[j][0];
#pragma unroll
        for (int i = 0; i < 16; i++)
          if (FULL || 8 * (i >> 2) + (i & 3) < rl) m = fminf(m, acc[j][i]);
        if (sub)
          bv[j][MODE == 1 ? 1 : 0] = m;
        else
          bv[j][0] = m;
        continue;
      }
      // (a chain started from +inf compiles to eight v_min3 with no operand
      // canonicalisation; NaN is still dropped)
      float m = INFINITY;
#pragma unroll
      for (int i = 0; i < 16; i++) m = fminf(m, acc[j][i]);
      // (MODE 2: the list bound's value is kept in Tv, updated on insertion)
      const float lim = Tv[j];
      if (MODE == 3 && RING > 0) {
        // park the passing lanes' tiles (the row limit goes with them: the
        // minimum may come from a pad row of the last block)
        const bool pass = m <= lim;
        const unsigned long long bal = __builtin_amdgcn_ballot_w64(pass);
        if (bal) {
          if (pass) {
            const int rank = __builtin_amdgcn_mbcnt_hi(
                (unsigned)(bal >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)bal, 0u));
            float4 *e = ring + ((rhead + rank) & (RING - 1)) * 5;
#pragma unroll
            for (int q4 = 0; q4 < 4; q4++)
              e[q4] = make_float4(acc[j][4 * q4], acc[j][4 * q4 + 1], acc[j][4 * q4 + 2],
                                  acc[j][4 * q4 + 3]);
            e[4] = make_float4(__int_as_float((int)qrow[j]), __int_as_float((int)bb + 4 * h), lim,
                               __uint_as_float((uint32_t)T[j]));
          }
          rhead += __popcll(bal);
          if (rhead - rtail >= 64) drain(a, 64);
        }
        continue;
      }
      if (!(m <= lim)) continue;
      // A lane that may take a value: a 16-bit mask of its passing values
      // (branch-free), its 16 values parked in its LDS slot, then one
      // iteration per set bit -- a handful of iterations for the wave
      // instead of a branch per value.
      unsigned msk = 0;
#pragma unroll
      for (int i = 0; i < 16; i++) {
        const bool p = acc[j][i] <= lim && (FULL || 8 * (i >> 2) + (i & 3) < rl);
        msk |= p ? (1u << i) : 0u;
      }
      if (MODE == 3 && !qok[j]) msk = 0;
      if (msk == 0) continue;
      float4 *slot = (float4 *)vs;
#pragma unroll
      for (int q4 = 0; q4 < 4; q4++)
        slot[q4] = make_float4(acc[j][4 * q4], acc[j][4 * q4 + 1], acc[j][4 * q4 + 2],
                               acc[j][4 * q4 + 3]);
      do {
        const int i = __builtin_ctz(msk);
        msk &= msk - 1;
        const uint64_t key = yk_key(vs[i], (int)bb + 8 * (i >> 2) + (i & 3) + 4 * h);
        if (MODE == 2) {
          if (key < top[j][R - 1]) {
            yk_reg_insert16<R>(top[j], key);
            Tv[j] = top[j][R - 1] == YK_KEY_NONE ? INFINITY
                                                 : yk_fkey_inv((uint32_t)(top[j][R - 1] >> 32));
          }
        } else if (key <= T[j]) {
          const int n = ncs[j];
          cs[(j * 64 + lane) * CST + n] = key;
          if (n + 1 == CST) {
            flush(a, j, CST);
            ncs[j] = 0;
          } else {
            ncs[j] = n + 1;
          }
        }
      } while (msk);
    }
  }

  // the n oldest ring entries, lane i taking entry rtail + i: the values of
  // rows in the item's range whose keys (value, row) are not above the
  // query's bound key are appended to its list -- counted first, so that one
  // atomic per lane reserves their places (an atomic with return per key,
  // and a load of the bound, made the wave wait out a round trip to the L2
  // for each)
  __device__ __forceinline__ void drain(const Bf16Args &a, int n) {
    if (lane < n) {
      const float4 *e = ring + ((rtail + lane) & (RING - 1)) * 5;
      const float4 md = e[4];
      const int q = __float_as_int(md.x), row0 = __float_as_int(md.y);
      const float lim = md.z;  // the bound key's value (INFINITY for no bound)
      const uint32_t trow = __float_as_uint(md.w);  // and its row
      const int rl = (int)min(rend - row0, 64L);
      unsigned msk = 0, eq = 0;
#pragma unroll
      for (int q4 = 0; q4 < 4; q4++) {
        const float4 v = e[q4];
        const float w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int t = 0; t < 4; t++) {
          const bool in = 8 * q4 + t < rl;
          msk |= (w[t] <= lim && in) ? (1u << (4 * q4 + t)) : 0u;
          eq |= (w[t] == lim && in) ? (1u << (4 * q4 + t)) : 0u;
        }
      }
      // below the bound's value the key is below the bound key; at equal
      // values (the same order-preserving image: -0 is taken as +0) the row
      // decides
      for (unsigned m = eq; m;) {
        const int i = __builtin_ctz(m);
        m &= m - 1;
        if ((uint32_t)(row0 + 8 * (i >> 2) + (i & 3)) > trow) msk &= ~(1u << i);
      }
      if (msk) {
        const float *ev = (const float *)e;
        const long base = (long)q * YK_CAND_CAP;
        int b = atomicAdd(&a.cnt[q], __popc(msk));
        do {
          const int i = __builtin_ctz(msk);
          msk &= msk - 1;
          if (b < YK_CAND_CAP) a.cand[base + b] = yk_key(ev[i], row0 + 8 * (i >> 2) + (i & 3));
          b++;
        } while (msk);
      }
    }
    rtail += n;
  }

  // append n staged keys of tile j to the query's list (the count keeps
  // growing past YK_CAND_CAP: the select kernel detects the overflow)
  __device__ __forceinline__ void flush(const Bf16Args &a, int j, int n) {
    const int base = atomicAdd(&a.cnt[qrow[j]], n);
    const uint64_t *src = cs + (j * 64 + lane) * CST;
    for (int s = 0; s < n; s++)
      if (base + s < YK_CAND_CAP) a.cand[qrow[j] * YK_CAND_CAP + base + s] = src[s];
  }

  // end of a work item: MODE 1 the lane minima, MODE 2 the lane top-R lists
  // (or, with kp_direct, the pair-merged kp smallest keys), MODE 3 the keys
  // still staged
  __device__ __forceinline__ void finish(const Bf16Args &a, long qbase, int split, int r, int h) {
    if (MODE == 3 && RING > 0) {
      while (rtail < rhead) drain(a, min(64, rhead - rtail));
    } else if (MODE == 3) {
#pragma unroll
      for (int j = 0; j < NB; j++)
        if (ncs[j] > 0) flush(a, j, ncs[j]);
    } else if (MODE == 1) {
      // Each lane's minima over its rows of the item's even and odd blocks:
      // the (split, lane half, parity) subsets are disjoint row sets, so the
      // k-th smallest of these values is attained by k distinct rows; the
      // key (value, 0xffffffff) therefore bounds the k-th smallest key from
      // above without the row index being known.  (More, smaller subsets
      // bring the bound closer to the sample's own k-th value.)
#pragma unroll
      for (int j = 0; j < NB; j++) {
        const long qi = qbase + j * 32 + r;
        if (!qok[j]) continue;
#pragma unroll
        for (int u = 0; u < 2; u++) {
          const float v = bv[j][MODE == 1 ? u : 0];
          a.part[((qi * a.nsplit + split) * 2 + h) * 2 + u] =
              v < INFINITY ? (((uint64_t)yk_fkey(v) << 32) | 0xffffffffu) : YK_KEY_NONE;
        }
      }
    } else if (MODE == 2 && a.kp_direct) {
      // lanes r and r + 32 hold sorted top-R lists of disjoint rows of the
      // same query: the first half-cleaner of the bitonic sequence (mine,
      // partner's reversed) keeps the R smallest of the union, R/2..1
      // cleaners sort them
#pragma unroll
      for (int j = 0; j < NB; j++) {
        uint64_t o[R], m[R];
#pragma unroll
        for (int s = 0; s < R; s++) {
          const uint32_t lo = __shfl_xor((uint32_t)top[j][s], 32);
          const uint32_t hi = __shfl_xor((uint32_t)(top[j][s] >> 32), 32);
          o[s] = ((uint64_t)hi << 32) | lo;
        }
#pragma unroll
        for (int s = 0; s < R; s++) m[s] = top[j][s] < o[R - 1 - s] ? top[j][s] : o[R - 1 - s];
#pragma unroll
        for (int w = R / 2; w >= 1; w >>= 1)
#pragma unroll
          for (int s = 0; s < R; s++)
            if ((s & w) == 0) {
              const uint64_t x = m[s], y = m[s + w];
              m[s] = x < y ? x : y;
              m[s + w] = x < y ? y : x;
            }
        const long qi = qbase + j * 32 + r;
        if (h == 0 && qok[j]) {
#pragma unroll
          for (int s = 0; s < R; s++)
            if (s < a.kp_direct) a.cand[qi * a.kp_direct + s] = m[s];
        }
      }
    } else if (MODE == 2) {
#pragma unroll
      for (int j = 0; j < NB; j++) {
        if (!qok[j]) continue;
        const long qi = qbase + j * 32 + r;
        uint64_t *dst = a.part + ((qi * a.nsplit + split) * 2 + h) * R;
#pragma unroll
        for (int s = 0; s < R; s++) dst[s] = top[j][s];
      }
    }
  }
};

// Streaming scan, each wave on its own: NCH > 0 queries register-resident
// (exactly NCH = d'/16 chunks, so every loop is unrolled without guards),
// base fragments loaded straight to registers one block ahead; NCH = 0 both
// operands streamed from memory (long vectors).
template <int NB, int NCH, int MODE, int R>
__device__ __forceinline__ void yk_scan_bf16_body(const Bf16Args &a) {
  constexpr bool QRES = NCH > 0;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  const int r = lane & 31, h = lane >> 5;
  // persistent grid, XCD-contiguous runs of work items (as in yk_select.hip)
  const int xcd = blockIdx.x & 7, local = blockIdx.x >> 3, wpx = gridDim.x >> 3;
  const int ibeg = xcd * a.per_xcd, iend = min(a.total, ibeg + a.per_xcd);
  const int nch = a.ldh >> 4;
  const int ldh = a.ldh;
  // MODE 3 parks passing tiles in a per-wave ring (YK_RING entries)
  constexpr int RING = MODE == 3 ? YK_RING : 0;
  __shared__ uint64_t cstage[MODE == 3 && RING == 0 ? 4 * NB * 64 * YK_CSTAGE : 1];
  __shared__ float vstage[MODE == 2 || (MODE == 3 && RING == 0) ? 4 * 64 * 16 : 1];
  __shared__ float4 rstage[RING > 0 ? 4 * RING * 5 : 1];
  // The first item of each workgroup is fixed, the later ones are drawn in
  // order from the XCD's ticket counter: the workgroups of an XCD then keep
  // to a window of consecutive items (items are split-major, so one or two
  // base ranges), which its L2 holds while every query tile reads it --
  // with a fixed stride between a workgroup's items they drift apart and
  // the ranges are fetched from HBM several times over.
  __shared__ int s_item;
  auto next_item = [&](int it) {
    if (!a.tick) return it + wpx;
    __syncthreads();  // every wave has read the previous ticket
    if (threadIdx.x == 0) s_item = ibeg + wpx + atomicAdd(&a.tick[xcd], 1);
    __syncthreads();
    return s_item;
  };
  for (int item = ibeg + local; item < iend; item = next_item(item)) {
    const int tile = item % a.tiles_x, split = item / a.tiles_x;
    const long qbase = ((long)tile * nwave + wave) * (32 * NB);
    const long bbeg = (long)split * a.brange;
    const long bend = min(a.nb, bbeg + a.brange);
    BfSel<NB, MODE, R, YK_CSTAGE, RING> sel;
    sel.init(a, qbase, r, lane, cstage + wave * (MODE == 3 && RING == 0 ? NB * 64 * YK_CSTAGE : 0),
             vstage + (MODE == 2 || (MODE == 3 && RING == 0) ? (wave * 64 + lane) * 16 : 0),
             rstage + wave * RING * 5);
    sel.rend = bend;
    if (MODE == 3 && !__any(sel.active())) continue;

    bf16x8 qreg[NB][QRES ? NCH : 1];
    if (QRES) {
#pragma unroll
      for (int j = 0; j < NB; j++)
#pragma unroll
        for (int c = 0; c < NCH; c++)
          qreg[j][c] = *(const bf16x8 *)(a.q + sel.qrow[j] * ldh + c * 16 + h * 8);
    }
    // fragment c of the block starting at row bb: bfrag(bb) + c * 512
    auto bfrag = [&](long bb) { return a.b + ((bb >> 5) * nch * 64 + lane) * 8; };
    if (QRES) {
      // Buffer loads: the item's fragments through a descriptor based at its
      // first block, the block/chunk offset in a scalar register and the
      // lane's 16 bytes in one constant vector register -- no per-block
      // 64-bit vector address arithmetic, and no address register for the
      // allocator to recycle as a load destination (which forces a wait for
      // the outstanding loads before the next address can be formed).
      const long nbk = (a.nb + 31) >> 5;  // blocks of the padded base copy
      const long rbytes = (nbk - (bbeg >> 5)) * (long)NCH * 1024;
      const uint16_t *ibase = a.b + (bbeg >> 5) * (long)NCH * 512;
      const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
          (void *)ibase, 0, (int)min(rbytes, 0x7fffffffL), YK_BUF_DWORD3);
      const int voff = lane * 16;
      auto soff_of = [&](long bb) { return (int)(((bb - bbeg) >> 5) * NCH * 1024); };
      // One rolling register buffer: as soon as the MFMAs of chunk c of block
      // bb have been issued, chunk c of the next block loads into the same
      // registers, so every load has a whole block of MFMAs to arrive in and
      // the buffer costs NCH registers rather than two blocks' worth (which
      // is what lets three query tiles fit at two waves per SIMD).  The
      // scheduling barriers keep each load after the MFMAs that read its
      // registers.  The prefetch is unconditional (past the range it reloads
      // the current block): a conditional load would join two paths with
      // different outstanding-load counts, and the wait before the next MFMAs
      // would then also wait for the loads just issued.
      bf16x8 buf[QRES ? NCH : 1];
      if (bbeg < bend) {
        const int s0 = soff_of(bbeg);
#pragma unroll
        for (int c = 0; c < NCH; c++)
          buf[c] = __builtin_bit_cast(
              bf16x8, __builtin_amdgcn_raw_buffer_load_b128(rs, voff, s0 + c * 1024, 0));
      }
      int sub = 0;
      for (long bb = bbeg; bb < bend; bb += a.bstep, sub ^= 1) {
        const int sn = soff_of(bb + a.bstep < bend ? bb + a.bstep : bb);
        f32x16 acc[NB];
#pragma unroll
        for (int j = 0; j < NB; j++)
#pragma unroll
          for (int e = 0; e < 16; e++) acc[j][e] = 0.f;
#pragma unroll
        for (int c = 0; c < NCH; c++) {
#pragma unroll
          for (int j = 0; j < NB; j++)
            acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(buf[c], qreg[j][QRES ? c : 0], acc[j],
                                                             0, 0, 0);
          __builtin_amdgcn_sched_barrier(0);
          buf[c] = __builtin_bit_cast(
              bf16x8, __builtin_amdgcn_raw_buffer_load_b128(rs, voff, sn + c * 1024, 0));
          __builtin_amdgcn_sched_barrier(0);
        }
        sel.block(a, acc, bb, bend, h, sub);
      }
    } else {
      for (long bb = bbeg; bb < bend; bb += a.bstep) {
        const uint16_t *bp = bfrag(bb);
        f32x16 acc[NB];
#pragma unroll
        for (int j = 0; j < NB; j++)
#pragma unroll
          for (int e = 0; e < 16; e++) acc[j][e] = 0.f;
        for (int c = 0; c < nch; c++) {
          bf16x8 av = *(const bf16x8 *)(bp + c * 512);
#pragma unroll
          for (int j = 0; j < NB; j++) {
            bf16x8 bq = *(const bf16x8 *)(a.q + sel.qrow[j] * ldh + c * 16 + h * 8);
            acc[j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bq, acc[j], 0, 0, 0);
          }
        }
        sel.block(a, acc, bb, bend, h, 0);
      }
    }
    sel.finish(a, qbase, split, r, h);
  }  // work items
}

template <int NB, int NCH, int MODE, int R>
__global__ __launch_bounds__(256) void yk_scan_bf16(Bf16Args a) {
  yk_scan_bf16_body<NB, NCH, MODE, R>(a);
}

// The register top-R scan with three query tiles, held to two waves per SIMD
// (the insertion network's state for three tiles spills a few registers to
// scratch; the fast test keeps that path rare once the lists have filled).
template <int NB, int NCH, int MODE, int R>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void
yk_scan_bf16_w2(Bf16Args a) {
  yk_scan_bf16_body<NB, NCH, MODE, R>(a);
}


// Long vectors, LDS-tiled (d'/16 > 16 chunks; the sketch and the filtered
// pass).  Neither operand fits in registers for the whole K, so the
// workgroup's YK_TW = 8 waves share YK_TQ = 4 query tiles (128 queries)
// staged through LDS YK_TS = 3 chunks at a time, in two buffers used in
// turn (one barrier per stage: a stage's buffer is written only after every
// wave has left the stage that last read it), and each wave streams its own
// YK_TG = 2 base blocks per step (the workgroup 16 blocks, 512 rows) straight
// to registers with buffer loads issued a stage ahead, in a ring of YK_TS
// slots whose slot is the chunk's position in its stage (a compile-time
// index): a slot reloads as soon as the MFMAs reading it have issued.  Per
// chunk a wave reads 4 fragments from LDS and 2 from memory for 8 MFMAs
// (the per-wave kernel this replaces read 5 fragments from memory for 6
// MFMAs, and its workgroup's waves read the same base blocks through the
// vector L1, whose line rate bounded it at 15% of the MFMA peak).  What
// bounds this kernel is the data brought to the CUs: every 512-row step
// re-reads the 250 KB query tile (8 waves per workgroup: 19.4 ms per
// filtered pass at d=960 against 21.4 ms with 4, the query tile staged once
// per 256 rows).  The queries of a stage come from the fragment-major copy
// (qf), wave w < 4 loading tile w's fragments one stage ahead into
// registers and writing them to LDS at the stage's end.  Selection as in
// yk_scan_bf16: every wave of the workgroup holds the same queries (so
// MODE 3's "any active" test is uniform over the workgroup); the sketch's
// disjoint row subsets are (split, wave, lane half, block parity), so
// MODE 1 writes nsplit x 8 (the waves) partial minima per query and half.
#define YK_TQ 4
#define YK_TG 2
#define YK_TS 3
#define YK_TCST 4  // MODE 3 keys staged per (query tile, lane)
#define YK_TW 8    // waves per workgroup

template <int MODE, int W>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(2, 2))) void
yk_scan_bf16_tiled(Bf16Args a) {
  constexpr int NB = YK_TQ, G = YK_TG, S = YK_TS;
  // the wave index as a scalar: the base offsets derived from it then stay
  // in scalar registers (a per-lane offset turns every buffer load into a
  // readfirstlane loop)
  const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int r = lane & 31, h = lane >> 5;
  const int xcd = blockIdx.x & 7, local = blockIdx.x >> 3, wpx = gridDim.x >> 3;
  const int ibeg = xcd * a.per_xcd, iend = min(a.total, ibeg + a.per_xcd);
  const int nch = a.ldh >> 4;
  const int nstage = (nch + S - 1) / S;
  __shared__ bf16x8 qs[2][S][NB][64];
  __shared__ uint64_t cstage[MODE == 3 ? W * NB * 64 * YK_TCST : 1];
  __shared__ float vstage[MODE == 3 ? W * 64 * 16 : 1];
  // the sketch's partial minima: nsplit x W subsets per query and half
  Bf16Args aw = a;
  aw.nsplit = a.nsplit * W;
  // items split-major: the workgroups an XCD runs at once share one base
  // range (measured: grouping a few query tiles per XCD instead, to keep
  // their fragments in L2, fetched 23-49% more bytes and ran 8-21% slower)
  for (int item = ibeg + local; item < iend; item += wpx) {
    const int tile = item % a.tiles_x, split = item / a.tiles_x;
    const long qbase = (long)tile * (32 * NB);
    const long bbeg = (long)split * a.brange;
    const long bend = min(a.nb, bbeg + a.brange);
    BfSel<NB, MODE, 1, YK_TCST> sel;
    sel.init(a, qbase, r, lane, cstage + wave * (MODE == 3 ? NB * 64 * YK_TCST : 0),
             vstage + (MODE == 3 ? (wave * 64 + lane) * 16 : 0));
    if (MODE == 3 && !__any(sel.active())) continue;  // the same in every wave
    // waves 0 .. NB-1 stage the queries, wave w tile (qbase/32 + w)
    const bool qloader = wave < NB;
    const uint16_t *qfw = a.qf + (((qbase >> 5) + (qloader ? wave : 0)) * nch * 64 + lane) * 8;
    const long nbk = (a.nb + 31) >> 5;
    const long rbytes = (nbk - (bbeg >> 5)) * (long)nch * 1024;
    const uint16_t *ibase = a.b + (bbeg >> 5) * (long)nch * 512;
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        (void *)ibase, 0, (int)min(rbytes, 0x7fffffffL), YK_BUF_DWORD3);
    const int voff = lane * 16;
    const int bstride = (int)((a.bstep >> 5) * nch * 1024);  // bytes between sampled blocks
    for (long bb0 = bbeg; bb0 < bend; bb0 += W * G * a.bstep) {
      const long bw = bb0 + (long)wave * G * a.bstep;  // this wave's first block
      const int sw = (int)(((bw - bbeg) >> 5) * nch * 1024);
      f32x16 acc[G][NB];
#pragma unroll
      for (int g = 0; g < G; g++)
#pragma unroll
        for (int j = 0; j < NB; j++)
#pragma unroll
          for (int e = 0; e < 16; e++) acc[g][j][e] = 0.f;
      // query fragments of stage st for this wave's tile (chunk indices
      // clamped: a stage past the last chunk reloads it, never used)
      bf16x8 qn[S];
      auto qload = [&](int st) {
        if (qloader) {
#pragma unroll
          for (int t = 0; t < S; t++) {
            const int c = min(st * S + t, nch - 1);
            qn[t] = *(const bf16x8 *)(qfw + c * 512);
          }
        }
      };
      auto qstore = [&](int buf) {
        if (qloader) {
#pragma unroll
          for (int t = 0; t < S; t++) qs[buf][t][wave][lane] = qn[t];
        }
      };
      // base fragments, a ring of S chunk slots, chunk c in slot c % S (past
      // the base end the buffer bounds return zeros; past the range, rows
      // are not selected)
      bf16x8 ring[S][G];
      auto bload = [&](int slot, int c) {
        const int cc = min(c, nch - 1);
#pragma unroll
        for (int g = 0; g < G; g++)
          ring[slot][g] = __builtin_bit_cast(
              bf16x8, __builtin_amdgcn_raw_buffer_load_b128(rs, voff, sw + g * bstride + cc * 1024,
                                                            0));
      };
      qload(0);
#pragma unroll
      for (int t = 0; t < S; t++) bload(t, t);
      qstore(0);
      qload(1);
      __syncthreads();
      for (int st = 0; st < nstage; st++) {
        const int buf = st & 1;
        // the stage's query fragments read from LDS one chunk ahead of the
        // MFMAs that use them (two register sets used in turn)
        bf16x8 qv[2][NB];
#pragma unroll
        for (int j = 0; j < NB; j++) qv[0][j] = qs[buf][0][j][lane];
#pragma unroll
        for (int t = 0; t < S; t++) {
          const int c = st * S + t;
          if (t + 1 < S) {
#pragma unroll
            for (int j = 0; j < NB; j++) qv[(t + 1) & 1][j] = qs[buf][t + 1][j][lane];
          }
          if (c < nch) {
#pragma unroll
            for (int g = 0; g < G; g++)
#pragma unroll
              for (int j = 0; j < NB; j++)
                acc[g][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ring[t][g], qv[t & 1][j],
                                                                    acc[g][j], 0, 0, 0);
          }
          // slot t's MFMAs are issued: it takes chunk c + S, a whole stage
          // ahead of its use
          __builtin_amdgcn_sched_barrier(0);
          bload(t, c + S);
          __builtin_amdgcn_sched_barrier(0);
        }
        qstore(buf ^ 1);
        qload(st + 2);
        __syncthreads();
      }
#pragma unroll
      for (int g = 0; g < G; g++) {
        const long bg = bw + g * a.bstep;
        if (bg < bend) sel.block(a, acc[g], bg, bend, h, g & 1);
      }
    }
    sel.finish(aw, qbase, split * W + wave, r, h);
  }  // work items
}

// rows [0, n) of a row-major BF16 copy (row stride ldh) -> fragment-major
// (block b, chunk c at ((b * nch + c) * 64 + lane) * 8, lane h*32 + r holding
// row 32b + r, columns 16c + 8h .. +7), rows [n, nrows) zero; one wave per
// (block, chunk)
__global__ __launch_bounds__(256) void yk_frag_from_rows_kernel(const uint16_t *__restrict__ src,
                                                                long n, long nrows, int ldh,
                                                                uint16_t *__restrict__ dst) {
  const int nch = ldh >> 4;
  const long w = blockIdx.x * 4L + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63, r = lane & 31, h = lane >> 5;
  if (w >= (nrows >> 5) * nch) return;
  const long b = w / nch;
  const int c = (int)(w - b * nch);
  const long row = b * 32 + r;
  uint4 v = make_uint4(0, 0, 0, 0);
  if (row < n) v = *(const uint4 *)(src + row * ldh + c * 16 + h * 8);
  *(uint4 *)(dst + (w * 64 + lane) * 8) = v;
}

// ---------------------------------------------------------------------------
// per-query sorts (same contracts as in yk_select.hip)

__device__ void yk_bitonic_b(uint64_t *s, int p2) {
  for (int size = 2; size <= p2; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int t = threadIdx.x; t < p2 / 2; t += blockDim.x) {
        int lo = 2 * t - (t & (stride - 1)), hi = lo + stride;
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

// sort n keys per query; write the kp smallest keys (candidates) or the
// kp-th smallest as threshold
__global__ __launch_bounds__(256) void yk_bf_sortsel_kernel(const uint64_t *__restrict__ src, int n,
                                                            int kp, int p2, uint64_t *cand_out,
                                                            uint64_t *thr) {
  extern __shared__ uint64_t s[];
  const long q = blockIdx.x;
  const uint64_t *p = src + q * (long)n;
  for (int i = threadIdx.x; i < p2; i += blockDim.x) s[i] = i < n ? p[i] : YK_KEY_NONE;
  __syncthreads();
  yk_bitonic_b(s, p2);
  if (thr) {
    if (threadIdx.x == 0) thr[q] = kp <= p2 ? s[kp - 1] : YK_KEY_NONE;
    return;
  }
  for (int i = threadIdx.x; i < kp; i += blockDim.x) cand_out[q * kp + i] = i < p2 ? s[i] : YK_KEY_NONE;
}

// ---------------------------------------------------------------------------
// One wave per query, keys in registers: E keys per lane (64 E keys), the
// lane's keys being elements lane * E .. lane * E + E - 1 of the sequence.
// Bitonic network (ascending): strides below E are compare-exchanges inside
// the lane, strides of E and more exchange with lane ^ (stride / E) by a
// shuffle, each lane keeping the min or the max of its pair.  No LDS and no
// barrier: the per-query sorts of the filtered pass's candidates (a few
// hundred keys) measured ~4x faster than the workgroup LDS network.

__device__ __forceinline__ uint64_t yk_shfl_xor64(uint64_t x, int m) {
  const uint32_t lo = __shfl_xor((uint32_t)x, m), hi = __shfl_xor((uint32_t)(x >> 32), m);
  return ((uint64_t)hi << 32) | lo;
}

template <int E>
__device__ __forceinline__ void yk_wave_sort(uint64_t (&v)[E], int lane) {
#pragma unroll
  for (int size = 2; size <= 64 * E; size <<= 1) {
#pragma unroll
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      if (stride >= E) {
#pragma unroll
        for (int i = 0; i < E; i++) {
          const int e = lane * E + i;
          const uint64_t o = yk_shfl_xor64(v[i], stride / E);
          const bool keep_min = ((e & stride) == 0) == ((e & size) == 0);
          v[i] = keep_min ? (v[i] < o ? v[i] : o) : (v[i] < o ? o : v[i]);
        }
      } else {
#pragma unroll
        for (int i = 0; i < E; i++) {
          if (i & stride) continue;
          const int j = i | stride;
          const bool up = ((lane * E + i) & size) == 0;
          const uint64_t x = v[i], y = v[j];
          if ((x > y) == up) {
            v[i] = y;
            v[j] = x;
          }
        }
      }
    }
  }
}

// the m (<= 64 E) keys at src, sorted; the kp smallest to dst (kp <= 64 E),
// or only the kp-th smallest to *kth
template <int E>
__device__ __forceinline__ void yk_wave_select(const uint64_t *src, int m, int kp, int lane,
                                               uint64_t *dst, uint64_t *kth) {
  uint64_t v[E];
#pragma unroll
  for (int i = 0; i < E; i++) {
    const int e = lane * E + i;
    v[i] = e < m ? src[e] : YK_KEY_NONE;
  }
  yk_wave_sort<E>(v, lane);
#pragma unroll
  for (int i = 0; i < E; i++) {
    const int e = lane * E + i;
    if (dst && e < kp) dst[e] = v[i];
    if (kth && e == kp - 1) *kth = v[i];
  }
}

// ---------------------------------------------------------------------------
// Selection without a full sort: one wave per query holds the query's m keys
// E per lane (element i * 64 + lane, so the loads are coalesced; padding
// YK_KEY_NONE) and finds the kp-th smallest key bit by bit, top bit first:
// with X the bits fixed so far, the candidate C = X with the current bit 0
// and every lower bit 1 counts the keys <= C by one compare per held key
// and a ballot popcount; fewer than kp means the bit is 1.  After the last
// bit X is the least value with at least kp keys <= it: the kp-th smallest
// key itself.  With lo_const (the sketch's keys, whose low word is always
// 0xffffffff) only the 32 high bits are searched.  The search starts below
// the bits every real key shares (the highest bit where the least and the
// greatest real key differ: distances of one magnitude share sign and
// exponent), and once the high word is fixed, a high word held by a single
// key makes that key the answer without searching the index bits.

template <int E>
__device__ __forceinline__ void yk_wave_load(const uint64_t *src, int m, int lane,
                                             uint64_t (&v)[E]) {
#pragma unroll
  for (int i = 0; i < E; i++) {
    const int e = i * 64 + lane;
    v[i] = e < m ? src[e] : YK_KEY_NONE;
  }
}

__device__ __forceinline__ uint64_t yk_wave_min64(uint64_t a) {
  for (int o = 32; o > 0; o >>= 1) {
    const uint32_t lo = __shfl_xor((uint32_t)a, o), hi = __shfl_xor((uint32_t)(a >> 32), o);
    const uint64_t b = ((uint64_t)hi << 32) | lo;
    a = b < a ? b : a;
  }
  return a;
}

// keys <= c among the held keys
template <int E>
__device__ __forceinline__ int yk_wave_count_le(const uint64_t (&v)[E], uint64_t c) {
  int n = 0;
#pragma unroll
  for (int i = 0; i < E; i++) n += __popcll(__builtin_amdgcn_ballot_w64(v[i] <= c));
  return n;
}

template <int E>
__device__ __forceinline__ uint64_t yk_wave_kth(const uint64_t (&v)[E], int kp, bool lo_const) {
  uint64_t x = 0;
  int top = 63;
  if (E <= 16) {  // (with more keys per lane the bookkeeping costs occupancy)
    // the least and the greatest real key (the greatest as the least of the
    // complements), and the real key count
    uint64_t mn = YK_KEY_NONE, cmx = YK_KEY_NONE;
    int nreal = 0;
#pragma unroll
    for (int i = 0; i < E; i++) {
      const bool real = v[i] != YK_KEY_NONE;
      nreal += __popcll(__builtin_amdgcn_ballot_w64(real));
      mn = v[i] < mn ? v[i] : mn;
      cmx = real && ~v[i] < cmx ? ~v[i] : cmx;
    }
    if (nreal < kp) return YK_KEY_NONE;
    mn = yk_wave_min64(mn);
    const uint64_t diff = mn ^ ~yk_wave_min64(cmx);
    if (diff == 0) return mn;
    top = 63 - __builtin_clzll(diff);  // the bits above it: every real key's
    x = top == 63 ? 0 : (mn >> (top + 1)) << (top + 1);
  }
  // the high word
  for (int bit = top; bit >= 32; bit--)
    if (yk_wave_count_le<E>(v, x | ((1ull << bit) - 1)) < kp) x |= 1ull << bit;
  if (lo_const) return x | 0xffffffffull;
  if (E <= 16 && top >= 32) {
    // the high word is fixed: if one key holds it, that key is the answer
    const uint64_t xh = x >> 32;
    int eq = 0;
    uint64_t keq = YK_KEY_NONE;
#pragma unroll
    for (int i = 0; i < E; i++) {
      const bool in = (v[i] >> 32) == xh && v[i] != YK_KEY_NONE;
      eq += __popcll(__builtin_amdgcn_ballot_w64(in));
      keq = in && v[i] < keq ? v[i] : keq;
    }
    if (eq == 1) return yk_wave_min64(keq);
  }
  // the index bits
  for (int bit = top < 31 ? top : 31; bit >= 0; bit--)
    if (yk_wave_count_le<E>(v, x | ((1ull << bit) - 1)) < kp) x |= 1ull << bit;
  return x;
}

// the kp (<= 64 E2) smallest of the keys held in v (keys are distinct;
// fewer than kp real keys: padded with YK_KEY_NONE), ascending, to dst:
// the keys <= the kp-th are compacted into the wave's LDS slice (ballot
// ranks), then sorted in registers
template <int E, int E2>
__device__ __forceinline__ void yk_wave_topk(const uint64_t (&v)[E], int kp, int lane,
                                             uint64_t *lds, uint64_t *dst) {
  const uint64_t kth = yk_wave_kth<E>(v, kp, false);
  int base = 0;
#pragma unroll
  for (int i = 0; i < E; i++) {
    const bool take = v[i] <= kth && v[i] != YK_KEY_NONE;
    const uint64_t m = __builtin_amdgcn_ballot_w64(take);
    const int rank = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                               __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
    if (take) lds[base + rank] = v[i];
    base += __popcll(m);
  }
  // (base <= kp <= 64 E2: the LDS slice holds 64 E2 keys)
  for (int e = base + lane; e < 64 * E2; e += 64) lds[e] = YK_KEY_NONE;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  uint64_t w[E2];
#pragma unroll
  for (int i = 0; i < E2; i++) w[i] = lds[lane * E2 + i];
  yk_wave_sort<E2>(w, lane);
#pragma unroll
  for (int i = 0; i < E2; i++) {
    const int e = lane * E2 + i;
    if (e < kp) dst[e] = w[i];
  }
}

template <int E>
__device__ __forceinline__ void yk_rsel_thr(const uint64_t *p, int n, int m, int lane,
                                            uint64_t *mth) {
  uint64_t v[E];
  yk_wave_load<E>(p, n, lane, v);
  const uint64_t x = yk_wave_kth<E>(v, m, true);
  if (lane == 0) *mth = x;
}

// the sketch's bound: per query the m-th smallest of its n partial keys (low
// words 0xffffffff), m <= kp, the bound the filtered pass starts with (for
// m = kp an upper bound of the kp-th smallest first-pass key);
// YK_KEY_NONE when n < kp.  One wave per query.
__global__ __launch_bounds__(256) void yk_bf_rsel_thr_kernel(const uint64_t *__restrict__ src,
                                                             long nq, int n, int kp, int m,
                                                             uint64_t *thr) {
  const long q = blockIdx.x * 4L + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (q >= nq) return;
  const uint64_t *p = src + q * (long)n;
  if (kp > n) {
    if (lane == 0) thr[q] = YK_KEY_NONE;
    return;
  }
  if (n <= 256)
    yk_rsel_thr<4>(p, n, m, lane, thr + q);
  else if (n <= 512)
    yk_rsel_thr<8>(p, n, m, lane, thr + q);
  else if (n <= 1024)
    yk_rsel_thr<16>(p, n, m, lane, thr + q);
  else if (n <= 2048)
    yk_rsel_thr<32>(p, n, m, lane, thr + q);
  else
    yk_rsel_thr<64>(p, n, m, lane, thr + q);
}

template <int E, int E2>
__device__ __forceinline__ void yk_rsel_cand(const uint64_t *p, int m, int n, int kp, int lane,
                                             uint64_t *lds, uint64_t *dst, uint64_t *thr,
                                             int *cnt, int *flag) {
  if (n < kp && *thr != YK_KEY_NONE) {
    // fewer than kp keys within the starting bound (an estimated one: under
    // a proven bound, or under an overflow's, kp keys or more are within
    // it, the data being finite): the query's next pass takes no bound; its
    // list then overflows and the overflow's bound below is a proven one
    for (int i = lane; i < kp; i += 64) dst[i] = YK_KEY_NONE;
    if (lane == 0) {
      *thr = YK_KEY_NONE;
      *cnt = 0;
      atomicOr(flag, 1);
    }
    return;
  }
  uint64_t v[E];
  yk_wave_load<E>(p, m, lane, v);
  if (n > YK_CAND_CAP) {
    // the list overflowed: the kp-th smallest captured key is a tighter
    // bound for the query's next pass; the query's list is written empty
    // (every reader of it stays in bounds until that pass replaces it)
    const uint64_t x = yk_wave_kth<E>(v, kp, false);
    for (int i = lane; i < kp; i += 64) dst[i] = YK_KEY_NONE;
    if (lane == 0) {
      *thr = x;
      *cnt = 0;
      atomicOr(flag, 1);
    }
    return;
  }
  yk_wave_topk<E, E2>(v, kp, lane, lds, dst);
  if (lane == 0) *thr = 0;
}

// The filtered pass's candidate lists: per query (thr != 0) the kp smallest
// of its captured keys, sorted, to cand_out, and thr cleared.  A list that
// overflowed YK_CAND_CAP gets a tighter bound and the flag; one with fewer
// than kp keys (under an estimated starting bound) gets no bound and the
// flag.  Either way the query's list is written empty until its next pass.
// One wave per query; kp <= 64 E2 (E2 = 16 covers YK_KMAX).
template <int E2>
__global__ __launch_bounds__(256) void yk_bf_rsel_cand_kernel(const uint64_t *__restrict__ cand,
                                                              long nq, int *cnt, uint64_t *thr,
                                                              int kp, uint64_t *cand_out,
                                                              int *flag) {
  __shared__ uint64_t lds[4][64 * E2];
  const long q = blockIdx.x * 4L + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (q >= nq || thr[q] == 0) return;
  const int n = cnt[q];
  const int m = n < YK_CAND_CAP ? n : YK_CAND_CAP;
  const uint64_t *p = cand + q * (long)YK_CAND_CAP;
  uint64_t *dst = cand_out + q * kp;
  if (m <= 256)
    yk_rsel_cand<4, E2>(p, m, n, kp, lane, lds[wave], dst, thr + q, cnt + q, flag);
  else if (m <= 512)
    yk_rsel_cand<8, E2>(p, m, n, kp, lane, lds[wave], dst, thr + q, cnt + q, flag);
  else if (m <= 1024)
    yk_rsel_cand<16, E2>(p, m, n, kp, lane, lds[wave], dst, thr + q, cnt + q, flag);
  else if (m <= 2048)
    yk_rsel_cand<32, E2>(p, m, n, kp, lane, lds[wave], dst, thr + q, cnt + q, flag);
  else
    yk_rsel_cand<64, E2>(p, m, n, kp, lane, lds[wave], dst, thr + q, cnt + q, flag);
}

// yk_bf_sortsel_kernel for n <= 1024 keys per query: four queries per
// workgroup, one per wave
__global__ __launch_bounds__(256) void yk_bf_wsortsel_kernel(const uint64_t *__restrict__ src,
                                                             long nq, int n, int kp,
                                                             uint64_t *cand_out, uint64_t *thr) {
  const long q = blockIdx.x * 4L + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (q >= nq) return;
  const uint64_t *p = src + q * (long)n;
  uint64_t *dst = thr ? nullptr : cand_out + q * kp;
  uint64_t *kth = thr ? thr + q : nullptr;
  if (thr && kp > n) {  // fewer keys than kp: no bound
    if (lane == 0) *kth = YK_KEY_NONE;
    return;
  }
  if (n <= 256)
    yk_wave_select<4>(p, n, kp, lane, dst, kth);
  else if (n <= 512)
    yk_wave_select<8>(p, n, kp, lane, dst, kth);
  else
    yk_wave_select<16>(p, n, kp, lane, dst, kth);
}

__global__ void yk_bf_zero_kernel(int *p, long n) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------------------
// FP32 re-rank (bit-identical to the FP32 form) + certificate.  MFMA rows =
// 32 candidates, every column the query.

// The re-rank is split over waves: yk_rerank_tile_kernel, one wave per
// (query, 32-candidate tile) -- the tiles of all queries run at once --
// writes each candidate's FP32-form key to ex[q * kp + candidate] (the value
// is the FP32 form's instruction sequence for the pair: the same MFMA chain
// in the same order, then float(qn + bn) - 2 dot); yk_rerank_fin_kernel, one
// wave per query, sorts the re-ranked keys in registers, writes the k
// results and applies the certificate.
//
// E = E_bf + E_32 bounds |FP32-form value - first-pass value| for any pair
// of the query (also the certificate's margin).  Only candidates whose
// first-pass value is <= v_k + 2E (v_k: the k-th first-pass value) can be
// among the FP32 form's k nearest: the k first-pass-smallest have FP32
// values <= v_k + E, any candidate above the bound one > v_k + E.  The
// candidates are in ascending first-pass order, so the re-ranked ones are
// the leading tiles that start within the bound.

// E = E_bf + E_32 for query stats qs and base maxima bs, and the re-rank
// bound v_k + 2E (INFINITY when the k-th key is missing)
__device__ __forceinline__ void yk_rerank_bound(bool dot, float4 qs, float4 bs, float gam_bf,
                                                float gam_32, uint64_t kkey, double *ebf,
                                                double *e32, double *bound) {
  if (dot) {
    *ebf = (double)qs.z * bs.x + (double)qs.y * bs.z + gam_bf * (double)qs.y * bs.y;
    *e32 = gam_32 * (double)qs.x * bs.x;
  } else {
    *ebf = 2.0 * ((double)qs.z * bs.x + (double)qs.y * bs.z) +
           gam_bf * (2.0 * qs.y * bs.y + qs.w + bs.w);
    *e32 = gam_32 * (2.0 * qs.x * bs.x + qs.w + bs.w);
  }
  *bound = kkey == YK_KEY_NONE ? INFINITY
                               : (double)yk_fkey_inv((uint32_t)(kkey >> 32)) + 2.0 * (*ebf + *e32);
}

__device__ __forceinline__ bool yk_rerank_needed(const uint64_t *cq, int kp, int t, double bound) {
  if (t >= kp) return false;
  const uint64_t kk = cq[t];
  return kk != YK_KEY_NONE && (double)yk_fkey_inv((uint32_t)(kk >> 32)) <= bound;
}

template <int NCH>
__global__ __launch_bounds__(64) void yk_rerank_tile_kernel(
    const float *__restrict__ q, const double *__restrict__ qn, const float *__restrict__ b,
    const float *__restrict__ bn, int ld, int type, const uint64_t *__restrict__ cand, int kp,
    int k, const float4 *__restrict__ qstats, const float4 *__restrict__ bmax, float gam_bf,
    float gam_32, uint64_t *__restrict__ ex) {
  const long qi = blockIdx.y;
  const int t0 = blockIdx.x * 32;
  const int lane = threadIdx.x, r = lane & 31, h = lane >> 5;
  const bool dot = type == YK_DOT;
  const uint64_t *cq = cand + qi * kp;
  double ebf, e32, bound;
  yk_rerank_bound(dot, qstats[qi], *bmax, gam_bf, gam_32, cq[k - 1], &ebf, &e32, &bound);
  if (!yk_rerank_needed(cq, kp, t0, bound)) return;
  // A candidate whose first-pass value is above the bound cannot be among
  // the k nearest (its FP32-form value exceeds the k-th's, as for a whole
  // tile above): its row is not gathered (the lane reads row 0, shared) and
  // its value is not written.
  auto beyond = [&](uint64_t kk) {
    return kk == YK_KEY_NONE || (double)yk_fkey_inv((uint32_t)(kk >> 32)) > bound;
  };
  const int ci = t0 + r;
  const uint64_t key = ci < kp ? cq[ci] : YK_KEY_NONE;
  const long row = beyond(key) ? 0 : (long)(uint32_t)key;
  const float *bp = b + row * ld + h * 8;
  const float *qp = q + qi * ld + h * 8;
  const int nch = NCH > 0 ? NCH : ld >> 4;
  f32x16 acc;
#pragma unroll
  for (int e = 0; e < 16; e++) acc[e] = 0.f;
  auto chunk = [&](int c) {
    const float4 a0 = *(const float4 *)(bp + c * 16), a1 = *(const float4 *)(bp + c * 16 + 4);
    const float4 b0 = *(const float4 *)(qp + c * 16), b1 = *(const float4 *)(qp + c * 16 + 4);
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.x, b0.x, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.y, b0.y, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.z, b0.z, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.w, b0.w, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.x, b1.x, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.y, b1.y, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.z, b1.z, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.w, b1.w, acc, 0, 0, 0);
  };
  if (NCH > 0) {
    // every row load issued before the chain starts (one memory round trip
    // per wave instead of one per chunk)
    float4 av[2 * (NCH > 0 ? NCH : 1)];
#pragma unroll
    for (int c = 0; c < NCH; c++) {
      av[2 * c] = *(const float4 *)(bp + c * 16);
      av[2 * c + 1] = *(const float4 *)(bp + c * 16 + 4);
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int c = 0; c < NCH; c++) {
      const float4 a0 = av[2 * c], a1 = av[2 * c + 1];
      // the query's chunk: the same address for every lane of a half (an
      // L1 hit); hoisting it too costs two waves per SIMD and measured slower
      const float4 b0 = *(const float4 *)(qp + c * 16), b1 = *(const float4 *)(qp + c * 16 + 4);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.x, b0.x, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.y, b0.y, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.z, b0.z, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.w, b0.w, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.x, b1.x, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.y, b1.y, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.z, b1.z, acc, 0, 0, 0);
      acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.w, b1.w, acc, 0, 0, 0);
    }
  } else {
    for (int c = 0; c < nch; c++) chunk(c);
  }
  // column 0 (lanes 0 and 32) holds the tile's candidate rows 8g + 4h + e
  if (r == 0) {
    const double qnd = dot ? 0.0 : qn[qi];
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int cidx = t0 + 8 * (i >> 2) + 4 * h + (i & 3);
      if (cidx >= kp) continue;
      const uint64_t ck = cq[cidx];
      uint64_t xk = YK_KEY_NONE;
      if (!beyond(ck)) {
        const int bi = (int)(uint32_t)ck;
        float v = dot ? acc[i] : (float)(qnd + (double)bn[bi]) - 2.0f * acc[i];
        // k = 1 follows nn.c:404-440: only values below 1e30 are taken
        if (v == v && (k > 1 || v < 1e30f)) xk = yk_key(v, bi);
      }
      ex[qi * kp + cidx] = xk;
    }
  }
}

template <int E>
__device__ __forceinline__ void yk_rerank_finish(const uint64_t *exq, int nproc, int k, int lane,
                                                 int *oi, float *od, uint64_t *kth) {
  uint64_t v[E];
#pragma unroll
  for (int i = 0; i < E; i++) {
    const int e = lane * E + i;
    v[i] = e < nproc ? exq[e] : YK_KEY_NONE;
  }
  yk_wave_sort<E>(v, lane);
  // no result: (-1, 1e30) for k = 1 (nn.c:404), (-1, NaN bits) padding for
  // k > 1 (nn.c:515-518)
  const float none_d = k == 1 ? 1e30f : __uint_as_float(0xffffffffu);
#pragma unroll
  for (int i = 0; i < E; i++) {
    const int e = lane * E + i;
    if (e < k) {
      oi[e] = v[i] == YK_KEY_NONE ? -1 : (int)(uint32_t)v[i];
      od[e] = v[i] == YK_KEY_NONE ? none_d : yk_fkey_inv((uint32_t)(v[i] >> 32));
    }
    if (e == k - 1) *kth = v[i];
  }
}

__global__ __launch_bounds__(256) void yk_rerank_fin_kernel(
    const uint64_t *__restrict__ cand, long n, long nb, int type, int kp, int k,
    const float4 *__restrict__ qstats, const float4 *__restrict__ bmax, float gam_bf,
    float gam_32, const uint64_t *__restrict__ ex, int *out_i, float *out_d, int *unc_list,
    int *n_unc) {
  __shared__ uint64_t kth_s[4];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long qi = blockIdx.x * 4L + wave;
  if (qi >= n) return;
  const uint64_t *cq = cand + qi * kp;
  double ebf, e32, bound;
  yk_rerank_bound(type == YK_DOT, qstats[qi], *bmax, gam_bf, gam_32, cq[k - 1], &ebf, &e32,
                  &bound);
  int nproc = 0;  // candidates re-ranked: the leading tiles that start within the bound
  for (int t0 = 0; yk_rerank_needed(cq, kp, t0, bound); t0 += 32) nproc = min(kp, t0 + 32);
  const int m = max(nproc, k);
  const uint64_t *exq = ex + qi * kp;
  int *oi = out_i + qi * k;
  float *od = out_d + qi * k;
  uint64_t *kth = kth_s + wave;
  if (m <= 256)
    yk_rerank_finish<4>(exq, nproc, k, lane, oi, od, kth);
  else if (m <= 512)
    yk_rerank_finish<8>(exq, nproc, k, lane, oi, od, kth);
  else
    yk_rerank_finish<16>(exq, nproc, k, lane, oi, od, kth);
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  if (lane == 0) {
    // the first pass kept the kp smallest keys; fewer than kp real ones means
    // every base vector was a candidate
    const uint64_t last = cq[kp - 1], fk = *kth;
    bool ok;
    if (last == YK_KEY_NONE || kp >= nb) {
      ok = true;
    } else if (fk == YK_KEY_NONE) {
      ok = false;
    } else {
      const float Tb = yk_fkey_inv((uint32_t)(last >> 32));
      const float Fk = yk_fkey_inv((uint32_t)(fk >> 32));
      ok = (double)Fk < (double)Tb - ebf - e32;
    }
    if (!ok) unc_list[atomicAdd(n_unc, 1)] = (int)qi;
  }
}

// k' <= 8 (the k-means assign): four queries per wave.  Row r of the MFMA
// tile is candidate r % 8 of query r / 8, column s < 4 is query s, so the
// useful outputs (query s's candidates x column s) sit in lanes s and s + 32
// (values i = 4s .. 4s + 3): a quarter of the MFMA chains of one query per
// wave.  Each output element is still the FP32 form's instruction sequence
// for its pair.  Results, conventions and certificate as in
// yk_rerank_tile_kernel + yk_rerank_fin_kernel.
template <int NCH>
__global__ __launch_bounds__(64) void yk_rerank4_kernel(
    const float *__restrict__ q, const double *__restrict__ qn, const float *__restrict__ b,
    const float *__restrict__ bn, long nb, int ld, int type, const uint64_t *__restrict__ cand,
    long nq, int kp, int k, const float4 *__restrict__ qstats, const float4 *__restrict__ bmax,
    float gam_bf, float gam_32, int *out_i, float *out_d, int *unc_list, int *n_unc) {
  __shared__ uint64_t ex[4][8];
  const long q0 = blockIdx.x * 4L;
  const int lane = threadIdx.x, r = lane & 31, h = lane >> 5;
  const int nch = NCH > 0 ? NCH : ld >> 4;
  const bool dot = type == YK_DOT;
  // A operand: row r = candidate r % 8 of query q0 + r / 8
  const long qa = q0 + (r >> 3);
  const int ca = r & 7;
  const uint64_t akey = (qa < nq && ca < kp) ? cand[qa * kp + ca] : YK_KEY_NONE;
  const long arow = akey == YK_KEY_NONE ? 0 : (long)(uint32_t)akey;
  const float *bp = b + arow * ld + h * 8;
  // B operand: column r = query q0 + r for r < 4 (others repeat query q0)
  const long qb = (r < 4 && q0 + r < nq) ? q0 + r : q0;
  const float *qp = q + qb * ld + h * 8;
  f32x16 acc;
#pragma unroll
  for (int e = 0; e < 16; e++) acc[e] = 0.f;
  for (int c = 0; c < nch; c++) {
    const float4 a0 = *(const float4 *)(bp + c * 16), a1 = *(const float4 *)(bp + c * 16 + 4);
    const float4 b0 = *(const float4 *)(qp + c * 16), b1 = *(const float4 *)(qp + c * 16 + 4);
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.x, b0.x, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.y, b0.y, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.z, b0.z, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.w, b0.w, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.x, b1.x, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.y, b1.y, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.z, b1.z, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.w, b1.w, acc, 0, 0, 0);
  }
  // lane (column s, half h) holds rows 8s + 4h + e, e < 4, in acc[4s + e]:
  // candidates 4h + e of query q0 + s
  if (r < 4 && q0 + r < nq) {
    const long qi = q0 + r;
    const double qnd = dot ? 0.0 : qn[qi];
#pragma unroll
    for (int s4 = 0; s4 < 4; s4++) {
      if (s4 != r) continue;
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int ci = 4 * h + e;
        uint64_t xk = YK_KEY_NONE;
        const uint64_t ck = ci < kp ? cand[qi * kp + ci] : YK_KEY_NONE;
        if (ck != YK_KEY_NONE) {
          const int bi = (int)(uint32_t)ck;
          const float dv = acc[4 * s4 + e];
          float v = dot ? dv : (float)(qnd + (double)bn[bi]) - 2.0f * dv;
          // k = 1 follows nn.c:404-440: only values below 1e30 are taken
          if (v == v && (k > 1 || v < 1e30f)) xk = yk_key(v, bi);
        }
        ex[r][ci] = xk;
      }
    }
  }
  __syncthreads();
  if (lane >= 4 || q0 + lane >= nq) return;
  const int s4 = lane;
  const long qi = q0 + s4;
  uint64_t t[8];
#pragma unroll
  for (int i = 0; i < 8; i++) t[i] = ex[s4][i];
  // sort the 8 keys (insertion network)
#pragma unroll
  for (int i = 1; i < 8; i++)
#pragma unroll
    for (int j2 = i; j2 > 0; j2--) {
      const uint64_t x = t[j2 - 1], y = t[j2];
      t[j2 - 1] = x < y ? x : y;
      t[j2] = x < y ? y : x;
    }
  const float none_d = k == 1 ? 1e30f : __uint_as_float(0xffffffffu);
#pragma unroll
  for (int i = 0; i < 8; i++)
    if (i < k) {
      out_i[qi * k + i] = t[i] == YK_KEY_NONE ? -1 : (int)(uint32_t)t[i];
      out_d[qi * k + i] = t[i] == YK_KEY_NONE ? none_d : yk_fkey_inv((uint32_t)(t[i] >> 32));
    }
  const uint64_t last = cand[qi * kp + kp - 1];
  uint64_t kth = YK_KEY_NONE;
#pragma unroll
  for (int i = 0; i < 8; i++)
    if (i == k - 1) kth = t[i];
  bool ok;
  if (last == YK_KEY_NONE || kp >= nb) {
    ok = true;
  } else if (kth == YK_KEY_NONE) {
    ok = false;
  } else {
    const float Tb = yk_fkey_inv((uint32_t)(last >> 32));
    const float Fk = yk_fkey_inv((uint32_t)(kth >> 32));
    const float4 qs = qstats[qi], bs = *bmax;  // {|x|, |bf x|, |e|, |x|^2}
    double ebf, e32;
    if (dot) {
      ebf = (double)qs.z * bs.x + (double)qs.y * bs.z + gam_bf * (double)qs.y * bs.y;
      e32 = gam_32 * (double)qs.x * bs.x;
    } else {
      ebf = 2.0 * ((double)qs.z * bs.x + (double)qs.y * bs.z) +
            gam_bf * (2.0 * qs.y * bs.y + qs.w + bs.w);
      e32 = gam_32 * (2.0 * qs.x * bs.x + qs.w + bs.w);
    }
    ok = (double)Fk < (double)Tb - ebf - e32;
  }
  if (!ok) unc_list[atomicAdd(n_unc, 1)] = (int)qi;
}

// scatter FP32-fallback rows back: out[list[i]] = src[i]
__global__ void yk_scatter_rows_kernel(const int *__restrict__ list, int n, int k,
                                       const int *__restrict__ si, const float *__restrict__ sd,
                                       int *oi, float *od) {
  long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)n * k) return;
  long i = t / k, j = t % k;
  long dst = (long)list[i] * k + j;
  oi[dst] = si[t];
  od[dst] = sd[t];
}

// dst row i = src row idx[i], rows of `units` 16-byte units
__global__ void yk_gather16_kernel(const uint4 *__restrict__ src, const int *__restrict__ idx,
                                   long n, int units, uint4 *__restrict__ dst) {
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= n * units) return;
  const long i = t / units;
  const int u = (int)(t - i * units);
  dst[t] = src[(long)idx[i] * units + u];
}

__global__ void yk_gather_f64_kernel(const double *__restrict__ src, const int *__restrict__ idx,
                                     long n, double *dst) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}

// ---------------------------------------------------------------------------
// host orchestration

// Scratch buffers kept across calls, slots 0..9 per nesting level (the
// escalation of uncertified queries runs the precision form again, one level
// down, while the outer level still holds its own buffers).
// bytes rounded up to a whole number of 256-byte units (scratch pieces
// carved from one buffer each start on such a boundary)
static size_t yk_b256(size_t b) { return (b + 255) & ~(size_t)255; }

static void *g_bs[20];
static size_t g_bs_bytes[20];
static int g_bs_level = 0;

static int yk_bscratch(int slot, size_t bytes, void **out) {
  slot += 10 * g_bs_level;
  if (bytes > g_bs_bytes[slot]) {
    if (g_bs[slot] && hipFree(g_bs[slot]) != hipSuccess) (void)hipGetLastError();
    g_bs[slot] = nullptr;
    g_bs_bytes[slot] = 0;
    hipError_t e = hipMalloc(&g_bs[slot], bytes);
    if (e != hipSuccess) return yk_hip_fail(e, "hipMalloc(bf16 scratch)");
    g_bs_bytes[slot] = bytes;
  }
  *out = g_bs[slot];
  return YK_OK;
}

struct BPlan {
  int tiles, nsplit;
  long brange, bstep;
};

// Launch-plan knobs the autotuner chooses among (every choice gives the
// same bits): the work-item target of the base split and a scale of the
// sketch's sampling stride; cfg = target index + 3 x stride index, 0 the
// default plan.  The plan of the search in progress (calls are serialised
// by the library lock; the nested escalation search keeps its caller's).
static const long yk_bitem_targets[3] = {4096, 2048, 8192};
static const int yk_bstride_scale[4][2] = {{1, 1}, {8, 3}, {16, 3}, {3, 2}};
#define YK_BNCFG 12
static int g_bcfg = 0;

// the sketch's sampling stride for k' > 16 (same reasoning as the FP32
// form in yk_select.hip), scaled by the current plan
static long yk_bsketch_stride(long nb, int kp) {
  const int *sc = yk_bstride_scale[(g_bcfg / 3) % 4];
  long stride = 600L * sc[0] / sc[1] / kp;
  const long cap = sc[0] == 16 ? 16 : 8;  // the 16/3 scale may sample 1 in 16
  if (stride > cap) stride = cap;
  if (stride < 1) stride = 1;
  const long nblocks = (nb + 31) / 32;
  while (stride > 1 && nblocks / stride < 2L * kp) stride >>= 1;
  return stride;
}

// bbytes: bytes of one 32-row block of the base copy (32 rows x ldh BF16)
static BPlan yk_bplan(long nq, long nb, int nblk, long sample, long min_splits, long max_split,
                      long bbytes) {
  BPlan p;
  const long qt = 32L * nblk * 4;
  p.tiles = (int)((nq + qt - 1) / qt);
  p.bstep = 32 * sample;
  const long nblocks = (nb + p.bstep - 1) / p.bstep;
  const long target = yk_bitem_targets[g_bcfg % 3];
  long want = (target + p.tiles - 1) / p.tiles;  // ~8 items per resident workgroup
  long maxs = (nblocks + 15) / 16;
  if (want > maxs) want = maxs;
  if (want < min_splits) want = min_splits;
  if (want > max_split) want = max_split;
  if (want > nblocks) want = nblocks;
  if (want < 1) want = 1;
  long per = (nblocks + want - 1) / want;
  // an item's fragments stay within 2^30 bytes (buffer-load offsets are
  // 32-bit signed; a 32-row block is 64 * ldh bytes, 61 KB at d' = 976)
  long max_per = (1L << 30) / (bbytes * sample);
  if (max_per < 1) max_per = 1;
  if (per > max_per) per = max_per;
  p.brange = per * p.bstep;
  p.nsplit = (int)((nb + p.brange - 1) / p.brange);
  return p;
}

// Query tiles (32 queries each) per wave: three where the kernel still fits
// two waves per SIMD (256 VGPRs: query fragments 12 * nch, accumulators 48,
// base buffer 4 * nch, selection state), in the long-vector kernel and in the
// register top-8 scan (the k-means assign; yk_scan_bf16_w2, measured 5.5%
// faster than two tiles at k = 65536); else two.  Each base fragment loaded
// then feeds three MFMAs instead of two.
static int yk_bnb(int mode, int nch, int r = YK_REG_R) {
  if (mode != 2 && nch > 16) return 1;  // yk_scan_bf16_tiled: 128 queries per workgroup
  if (mode == 2 && r <= 8 && nch >= 1 && nch <= 9) return 3;  // yk_scan_bf16_w2
  return mode != 2 && nch >= 1 && nch <= 9 ? 3 : 2;
}

// streaming kernel with exactly nch register-resident chunks (nch <= 16),
// else the one streaming both operands
template <int MODE, int R>
static void (*yk_bscan_pick(int nch))(Bf16Args) {
  constexpr int N3 = MODE != 2 ? 3 : 2;
  switch (nch) {
    case 1: return yk_scan_bf16<N3, 1, MODE, R>;
    case 2: return yk_scan_bf16<N3, 2, MODE, R>;
    case 3: return yk_scan_bf16<N3, 3, MODE, R>;
    case 4: return yk_scan_bf16<N3, 4, MODE, R>;
    case 5: return yk_scan_bf16<N3, 5, MODE, R>;
    case 6: return yk_scan_bf16<N3, 6, MODE, R>;
    case 7: return yk_scan_bf16<N3, 7, MODE, R>;
    case 8: return yk_scan_bf16<N3, 8, MODE, R>;
    case 9: return yk_scan_bf16<N3, 9, MODE, R>;
    case 10: return yk_scan_bf16<2, 10, MODE, R>;
    case 11: return yk_scan_bf16<2, 11, MODE, R>;
    case 12: return yk_scan_bf16<2, 12, MODE, R>;
    case 13: return yk_scan_bf16<2, 13, MODE, R>;
    case 14: return yk_scan_bf16<2, 14, MODE, R>;
    case 15: return yk_scan_bf16<2, 15, MODE, R>;
    case 16: return yk_scan_bf16<2, 16, MODE, R>;
    default: return yk_scan_bf16<2, 0, MODE, R>;
  }
}

template <int MODE, int R>
static int yk_launch_bscan(const BPlan &p, Bf16Args a) {
  // 64 or 96 queries per wave at two waves per SIMD (measured faster than
  // 128 queries per wave at one wave per SIMD)
  const int nbq = yk_bnb(MODE, a.ldh >> 4, R);
  void (*kern)(Bf16Args) = yk_bscan_pick<MODE, R>(a.ldh >> 4);
  if constexpr (MODE == 2 && R <= 8) {
    if (nbq == 3) {
      switch (a.ldh >> 4) {
        case 1: kern = yk_scan_bf16_w2<3, 1, MODE, R>; break;
        case 2: kern = yk_scan_bf16_w2<3, 2, MODE, R>; break;
        case 3: kern = yk_scan_bf16_w2<3, 3, MODE, R>; break;
        case 4: kern = yk_scan_bf16_w2<3, 4, MODE, R>; break;
        case 5: kern = yk_scan_bf16_w2<3, 5, MODE, R>; break;
        case 6: kern = yk_scan_bf16_w2<3, 6, MODE, R>; break;
        case 7: kern = yk_scan_bf16_w2<3, 7, MODE, R>; break;
        case 8: kern = yk_scan_bf16_w2<3, 8, MODE, R>; break;
        default: kern = yk_scan_bf16_w2<3, 9, MODE, R>; break;
      }
    }
  }
  if constexpr (MODE != 2) {
    if ((a.ldh >> 4) > 16) kern = yk_scan_bf16_tiled<MODE, YK_TW>;
  }
  a.brange = p.brange;
  a.bstep = p.bstep;
  a.nsplit = p.nsplit;
  a.tiles_x = (int)((a.nq + 128L * nbq - 1) / (128L * nbq));
  a.total = a.tiles_x * p.nsplit;
  a.per_xcd = (a.total + 7) / 8;
  const int nthr = (MODE != 2 && (a.ldh >> 4) > 16) ? 64 * YK_TW : 256;
  const int per_res = yk_resident_wgs((const void *)kern, nthr) / 8;
  const dim3 grid(8 * (a.per_xcd < per_res ? a.per_xcd : per_res)), blk(nthr);
  // item tickets for the scans built on yk_scan_bf16_body
  a.tick = nullptr;
  if (MODE == 2 || (a.ldh >> 4) <= 16) YK_TRY(yk_item_tickets(&a.tick));
  int tok = yk_prof_begin(MODE == 1 ? YK_PROF_BF16_SKETCH : YK_PROF_BF16);
  hipLaunchKernelGGL(kern, grid, blk, 0, yk_stream(), a);
  yk_prof_end(tok);
  YK_HIP(hipGetLastError());
  return YK_OK;
}

// per query, the n keys at src[q * n ..]: the kp smallest to cand_out, or
// the m-th smallest (m <= kp; the kp-th where the keys are too many for the
// select) to thr
static int yk_bf_sortsel(const uint64_t *src, long nq, int n, int kp, uint64_t *cand_out,
                         uint64_t *thr, int m = 0) {
  if (thr && n <= 4096) {  // the sketch's bound: a select, no sort
    hipLaunchKernelGGL(yk_bf_rsel_thr_kernel, dim3((unsigned)((nq + 3) / 4)), dim3(256), 0,
                       yk_stream(), src, nq, n, kp, m > 0 && m < kp ? m : kp, thr);
    YK_HIP(hipGetLastError());
    return YK_OK;
  }
  if (n <= 1024) {
    hipLaunchKernelGGL(yk_bf_wsortsel_kernel, dim3((unsigned)((nq + 3) / 4)), dim3(256), 0,
                       yk_stream(), src, nq, n, kp, cand_out, thr);
  } else {
    int p2 = 1;
    while (p2 < n) p2 <<= 1;
    hipLaunchKernelGGL(yk_bf_sortsel_kernel, dim3((unsigned)nq), dim3(256), sizeof(uint64_t) * p2,
                       yk_stream(), src, n, kp, p2, cand_out, thr);
  }
  YK_HIP(hipGetLastError());
  return YK_OK;
}


// kp smallest first-pass keys per query -> cand [nq][kp].  For kp > YK_REG_R
// the filtered pass captures at most YK_CAND_CAP keys per query; a query
// whose capture overflowed gets an empty list, a tighter bound and the
// device flag `ovf` set (the caller zeroes it first).  Without resume the
// call enqueues one filtered pass and returns without waiting; with resume
// (after the caller read the flag set) it repeats the filtered pass for the
// overflowed queries with their tighter bounds, waiting on the flag each time.
static int yk_bf16_candidates(Bf16Args a, int kp, uint64_t *cand_out, int *ovf, bool resume) {
  const long nq = a.nq;
  void *vpart;
  if (kp <= YK_REG_R) {
    const long bbytes = 64L * a.ldh;
    const int R = kp <= 8 ? 8 : YK_REG_R;
    const int nbq = yk_bnb(2, a.ldh >> 4, R);
    BPlan p = yk_bplan(nq, a.nb, nbq, 1, 1, 256, bbytes);
    // enough query tiles to fill the chip: no base split, and the lane pairs
    // write their merged lists straight to cand_out
    if (p.tiles * nbq >= 4096) p = yk_bplan(nq, a.nb, nbq, 1, 1, 1, bbytes);
    if (p.nsplit == 1) {
      a.cand = cand_out;
      a.kp_direct = kp;
      if (R == 8)
        YK_TRY((yk_launch_bscan<2, 8>(p, a)));
      else
        YK_TRY((yk_launch_bscan<2, YK_REG_R>(p, a)));
      return YK_OK;
    }
    YK_TRY(yk_bscratch(0, sizeof(uint64_t) * nq * p.nsplit * 2 * R, &vpart));
    a.part = (uint64_t *)vpart;
    if (R == 8)
      YK_TRY((yk_launch_bscan<2, 8>(p, a)));
    else
      YK_TRY((yk_launch_bscan<2, YK_REG_R>(p, a)));
    YK_TRY(yk_bf_sortsel(a.part, nq, p.nsplit * 2 * R, kp, cand_out, nullptr));
    return YK_OK;
  }
  void *vthr, *vcnt, *vcand;
  YK_TRY(yk_bscratch(1, sizeof(uint64_t) * nq, &vthr));
  YK_TRY(yk_bscratch(2, sizeof(int) * (nq + 64), &vcnt));
  YK_TRY(yk_bscratch(3, sizeof(uint64_t) * nq * YK_CAND_CAP, &vcand));
  uint64_t *thr = (uint64_t *)vthr;
  int *cnt = (int *)vcnt;
  if (!resume) {
    // sketch: per-lane minima over a 1-in-stride sample, >= kp splits (same
    // reasoning as the FP32 form in yk_select.hip)
    const long stride = yk_bsketch_stride(a.nb, kp);
    BPlan pa = yk_bplan(nq, a.nb, yk_bnb(1, a.ldh >> 4), stride, kp, kp > 256 ? kp : 256,
                        64L * a.ldh);
    // per query and split: 2 lane halves x 2 block parities of minima, and
    // for long vectors that per wave of the tiled kernel's workgroup
    const int nsub = (a.ldh >> 4) > 16 ? 4 * YK_TW : 4;
    YK_TRY(yk_bscratch(0, sizeof(uint64_t) * nq * pa.nsplit * nsub, &vpart));
    a.part = (uint64_t *)vpart;
    YK_TRY((yk_launch_bscan<1, 1>(pa, a)));
    YK_TRY(yk_bf_sortsel(a.part, nq, pa.nsplit * nsub, kp, nullptr, thr,
                         yk_bound_rank(kp, stride)));
    hipLaunchKernelGGL(yk_bf_zero_kernel, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0,
                       yk_stream(), cnt, nq);
  }
  BPlan pb = yk_bplan(nq, a.nb, yk_bnb(3, a.ldh >> 4), 1, 1, 4096, 64L * a.ldh);
  a.thr = thr;
  a.cnt = cnt;
  a.cand = (uint64_t *)vcand;
  for (int tries = resume ? 1 : 0;; tries++) {
    if (resume) hipLaunchKernelGGL(yk_bf_zero_kernel, dim3(1), dim3(64), 0, yk_stream(), ovf, 1L);
    YK_TRY((yk_launch_bscan<3, 1>(pb, a)));
    const dim3 g4((unsigned)((nq + 3) / 4));
    if (kp <= 256)
      hipLaunchKernelGGL(yk_bf_rsel_cand_kernel<4>, g4, dim3(256), 0, yk_stream(), a.cand, nq, cnt,
                         thr, kp, cand_out, ovf);
    else if (kp <= 512)
      hipLaunchKernelGGL(yk_bf_rsel_cand_kernel<8>, g4, dim3(256), 0, yk_stream(), a.cand, nq, cnt,
                         thr, kp, cand_out, ovf);
    else
      hipLaunchKernelGGL(yk_bf_rsel_cand_kernel<16>, g4, dim3(256), 0, yk_stream(), a.cand, nq,
                         cnt, thr, kp, cand_out, ovf);
    YK_HIP(hipGetLastError());
    if (!resume) break;
    int hflag = 0;
    YK_TRY(yk_read_small(&hflag, ovf, sizeof(int)));
    if (!hflag) break;
    if (tries == 5)
      return yk_set_error(YK_EINVAL, "k'=%d: too many tied first-pass distances", kp);
  }
  return YK_OK;
}

static int yk_knn_bf16_run(int type, long nq, long nb, int d, int k, int kp, const float *q,
                           const double *qn, const uint16_t *qh, const float *qstats,
                           const float *b, const float *bn, const uint16_t *bh,
                           const float *bstats_max, int *assign, float *dis, int *n_fallback);

// a top-level search through yk_time_ms, with plan cfg
struct YkBf16Call {
  int type;
  long nq, nb;
  int d, k, kp;
  const float *q;
  const double *qn;
  const uint16_t *qh;
  const float *qstats, *b, *bn;
  const uint16_t *bh;
  const float *bstats_max;
  int *assign;
  float *dis;
  int *n_fallback;
  int cfg;
};
static int yk_bf16_call(void *p) {
  const YkBf16Call &c = *(const YkBf16Call *)p;
  g_bcfg = c.cfg;
  const int rc = yk_knn_bf16_run(c.type, c.nq, c.nb, c.d, c.k, c.kp, c.q, c.qn, c.qh, c.qstats,
                                 c.b, c.bn, c.bh, c.bstats_max, c.assign, c.dis, c.n_fallback);
  g_bcfg = 0;
  return rc;
}

static int yk_blg(long x) {
  int l = 0;
  while (x > 1) {
    x >>= 1;
    l++;
  }
  return l;
}

extern "C" int yk_knn_bf16(int type, long nq, long nb, int d, int k, int kp, const float *q,
                           const double *qn, const uint16_t *qh, const float *qstats,
                           const float *b, const float *bn, const uint16_t *bh,
                           const float *bstats_max, int *assign, float *dis, int *n_fallback) {
  YK_TRY(yk_init());
  YkGuard guard;  // process-wide scratch (g_bs, the escalation level) is shared
  YkBf16Call call = {type, nq, nb, d, k, kp, q, qn, qh, qstats, b, bn, bh, bstats_max,
                     assign, dis, n_fallback, 0};
  // The plan: the tuning cache's choice for the shape class; on a miss (a
  // top-level search of at least 2^32 flop, where one timing is
  // meaningful) every candidate is timed once after an untimed run of the
  // default and the fastest recorded; the outputs then hold the last
  // candidate's results, the same bits as any plan's.
  if (g_bs_level == 0 && yk_tune_enabled() && nq > 0 &&
      (double)nq * nb * d >= (double)(1L << 31)) {
    const int kpc = kp <= 8 ? 8 : kp <= YK_REG_R ? 16 : 1 << yk_blg(kp);
    char key[128];
    snprintf(key, sizeof key, "bf16:t%d:kp%d:ldh%d:q%d:b%d", type == 12 ? YK_L2 : type, kpc,
             yk_bf16_ld(d, type == 12 ? YK_L2 : type), yk_blg(nq), yk_blg(nb));
    int cfg = 0;
    if (yk_tune_get(key, &cfg)) {
      call.cfg = cfg >= 0 && cfg < YK_BNCFG ? cfg : 0;
      return yk_bf16_call(&call);
    }
    const int ncand = kp > YK_REG_R ? YK_BNCFG : 3;  // the stride matters for k' > 16 only
    // candidates whose plans come out the same as an earlier one's are not
    // timed again
    bool dup[YK_BNCFG] = {};
    long sig[YK_BNCFG][3];
    int distinct = 0;
    const int ldh = yk_bf16_ld(d, type == 12 ? YK_L2 : type), nch = ldh >> 4;
    const long nqb = kp <= YK_REG_R ? nq : (nq < 16384 ? nq : 16384);
    for (int i = 0; i < ncand; i++) {
      g_bcfg = i;
      if (kp <= YK_REG_R) {
        const int R = kp <= 8 ? 8 : YK_REG_R, nbq = yk_bnb(2, nch, R);
        BPlan p = yk_bplan(nqb, nb, nbq, 1, 1, 256, 64L * ldh);
        if (p.tiles * nbq >= 4096) p = yk_bplan(nqb, nb, nbq, 1, 1, 1, 64L * ldh);
        sig[i][0] = p.nsplit, sig[i][1] = p.brange, sig[i][2] = 0;
      } else {
        const long st = yk_bsketch_stride(nb, kp);
        const BPlan pa =
            yk_bplan(nqb, nb, yk_bnb(1, nch), st, kp, kp > 256 ? kp : 256, 64L * ldh);
        const BPlan pb = yk_bplan(nqb, nb, yk_bnb(3, nch), 1, 1, 4096, 64L * ldh);
        sig[i][0] = st, sig[i][1] = pa.nsplit, sig[i][2] = pb.nsplit;
      }
      for (int j = 0; j < i && !dup[i]; j++)
        dup[i] = !dup[j] && sig[j][0] == sig[i][0] && sig[j][1] == sig[i][1] &&
                 sig[j][2] == sig[i][2];
      distinct += !dup[i];
    }
    g_bcfg = 0;
    if (distinct < 2)  // one plan: nothing to choose
      return yk_knn_bf16_run(type, nq, nb, d, k, kp, q, qn, qh, qstats, b, bn, bh, bstats_max,
                             assign, dis, n_fallback);
    YK_TRY(yk_bf16_call(&call));  // warm-up, default plan
    float best = 1e30f;
    for (int i = 0; i < ncand; i++) {
      if (dup[i]) continue;
      call.cfg = i;
      float ms = 0;
      YK_TRY(yk_time_ms(yk_bf16_call, &call, &ms));
      if (ms < best) {
        best = ms;
        cfg = i;
      }
    }
    yk_tune_put(key, cfg);
    return YK_OK;
  }
  return yk_knn_bf16_run(type, nq, nb, d, k, kp, q, qn, qh, qstats, b, bn, bh, bstats_max,
                         assign, dis, n_fallback);
}

static int yk_knn_bf16_run(int type, long nq, long nb, int d, int k, int kp, const float *q,
                           const double *qn, const uint16_t *qh, const float *qstats,
                           const float *b, const float *bn, const uint16_t *bh,
                           const float *bstats_max, int *assign, float *dis, int *n_fallback) {
  if (type == 12) type = YK_L2;
  if (type != YK_L2 && type != YK_DOT)
    return yk_set_error(YK_EINVAL, "distance type %d has no GPU path (2, 12, 16 only)", type);
  if (kp > nb) kp = (int)nb;
  if (kp < k) kp = k;
  if (k < 1 || k > nb || kp > YK_KMAX)
    return yk_set_error(YK_EINVAL, "k=%d k'=%d out of range (nb=%ld, max %d)", k, kp, nb, YK_KMAX);
  if (nq <= 0) return YK_OK;
  // query batches bound the candidate memory: kp <= 16 keeps kp keys per
  // query, larger kp a YK_CAND_CAP list per query during the filtered pass
  const long QB = kp <= YK_REG_R ? (1L << 20) : 16384;
  void *vcand, *vunc;
  const long qb = nq < QB ? nq : QB;
  YK_TRY(yk_bscratch(4, sizeof(uint64_t) * qb * kp, &vcand));
  YK_TRY(yk_bscratch(5, sizeof(int) * (qb + 64), &vunc));
  // the uncertified count and the candidate-overflow flag side by side: one
  // read after the re-rank brings both back
  int *unc = (int *)vunc, *nunc = unc + qb + 32, *ovf = nunc + 1;
  const int ld = yk_ld(d), ldh = yk_bf16_ld(d, type);
  const float gam_bf = 2.0f * (ldh + 8) * 0x1p-23f, gam_32 = 2.0f * (ld + 8) * 0x1p-23f;
  int total_fb = 0;
  for (long q0 = 0; q0 < nq; q0 += QB) {
    const long n = nq - q0 < QB ? nq - q0 : QB;
    Bf16Args a = {};
    a.q = qh + q0 * ldh;
    a.nq = n;
    a.b = bh;
    a.nb = nb;
    a.ldh = ldh;
    if ((ldh >> 4) > 16 && kp > YK_REG_R) {
      // long vectors: the batch's queries fragment-major for the long scan,
      // rows padded to its workgroup span (4 waves x 32 x query tiles)
      const long span = 128L * yk_bnb(3, ldh >> 4);
      const long rows = (n + span - 1) / span * span, nw = (rows >> 5) * (ldh >> 4);
      void *vqf;
      YK_TRY(yk_bscratch(9, sizeof(uint16_t) * (size_t)rows * ldh, &vqf));
      hipLaunchKernelGGL(yk_frag_from_rows_kernel, dim3((unsigned)((nw + 3) / 4)), dim3(256), 0,
                         yk_stream(), a.q, n, rows, ldh, (uint16_t *)vqf);
      YK_HIP(hipGetLastError());
      a.qf = (const uint16_t *)vqf;
    }
    hipLaunchKernelGGL(yk_bf_zero_kernel, dim3(1), dim3(64), 0, yk_stream(), nunc, 2L);
    YK_TRY(yk_bf16_candidates(a, kp, (uint64_t *)vcand, ovf, false));
    auto rerank = [&]() -> int {
      int tok = yk_prof_begin(YK_PROF_RERANK);
      if (kp <= 8) {
        void (*rk4)(const float *, const double *, const float *, const float *, long, int, int,
                    const uint64_t *, long, int, int, const float4 *, const float4 *, float, float,
                    int *, float *, int *, int *);
        switch (ld >> 4) {
          case 1: rk4 = yk_rerank4_kernel<1>; break;
          case 2: rk4 = yk_rerank4_kernel<2>; break;
          case 3: rk4 = yk_rerank4_kernel<3>; break;
          case 4: rk4 = yk_rerank4_kernel<4>; break;
          case 5: rk4 = yk_rerank4_kernel<5>; break;
          case 6: rk4 = yk_rerank4_kernel<6>; break;
          case 7: rk4 = yk_rerank4_kernel<7>; break;
          case 8: rk4 = yk_rerank4_kernel<8>; break;
          default: rk4 = yk_rerank4_kernel<0>; break;
        }
        hipLaunchKernelGGL(rk4, dim3((unsigned)((n + 3) / 4)), dim3(64), 0, yk_stream(),
                           q + q0 * ld, qn ? qn + q0 : nullptr, b, bn, nb, ld, type,
                           (const uint64_t *)vcand, n, kp, k, (const float4 *)qstats + q0,
                           (const float4 *)bstats_max, gam_bf, gam_32, assign + q0 * k,
                           dis + q0 * k, unc, nunc);
      } else {
        void (*rt)(const float *, const double *, const float *, const float *, int, int,
                   const uint64_t *, int, int, const float4 *, const float4 *, float, float,
                   uint64_t *);
        switch (ld >> 4) {
          case 1: rt = yk_rerank_tile_kernel<1>; break;
          case 2: rt = yk_rerank_tile_kernel<2>; break;
          case 3: rt = yk_rerank_tile_kernel<3>; break;
          case 4: rt = yk_rerank_tile_kernel<4>; break;
          case 5: rt = yk_rerank_tile_kernel<5>; break;
          case 6: rt = yk_rerank_tile_kernel<6>; break;
          case 7: rt = yk_rerank_tile_kernel<7>; break;
          case 8: rt = yk_rerank_tile_kernel<8>; break;
          default: rt = yk_rerank_tile_kernel<0>; break;
        }
        void *vex;
        YK_TRY(yk_bscratch(8, sizeof(uint64_t) * (size_t)n * kp, &vex));
        hipLaunchKernelGGL(rt, dim3((unsigned)((kp + 31) / 32), (unsigned)n), dim3(64), 0,
                           yk_stream(), q + q0 * ld, qn ? qn + q0 : nullptr, b, bn, ld, type,
                           (const uint64_t *)vcand, kp, k, (const float4 *)qstats + q0,
                           (const float4 *)bstats_max, gam_bf, gam_32, (uint64_t *)vex);
        hipLaunchKernelGGL(yk_rerank_fin_kernel, dim3((unsigned)((n + 3) / 4)), dim3(256), 0,
                           yk_stream(), (const uint64_t *)vcand, n, nb, type, kp, k,
                           (const float4 *)qstats + q0, (const float4 *)bstats_max, gam_bf,
                           gam_32, (const uint64_t *)vex, assign + q0 * k, dis + q0 * k, unc,
                           nunc);
      }
      yk_prof_end(tok);
      YK_HIP(hipGetLastError());
      return YK_OK;
    };
    YK_TRY(rerank());
    int hv[2] = {0, 0};
    YK_TRY(yk_read_small(hv, nunc, sizeof(hv)));
    if (hv[1]) {
      // a query's capture overflowed (massive ties): its list was empty in
      // this re-rank; the filtered pass again for such queries, then the
      // re-rank of the batch again
      YK_TRY(yk_bf16_candidates(a, kp, (uint64_t *)vcand, ovf, true));
      hipLaunchKernelGGL(yk_bf_zero_kernel, dim3(1), dim3(64), 0, yk_stream(), nunc, 1L);
      YK_TRY(rerank());
      YK_TRY(yk_read_small(hv, nunc, sizeof(int)));
    }
    const int hn = hv[0];
    // Uncertified queries, first level: the precision form again on just
    // those queries with 8x the candidates (the certificate's margin grows
    // with the k'-th first-pass value); what that level cannot certify goes
    // to the FP32 form there.  The result is the FP32 form's either way.
    const int kp2 = (int)min((long)min(8 * kp, YK_KMAX), nb);
    if (hn > 0 && g_bs_level == 0 && kp2 > kp) {
      // each piece starts on a 256-byte boundary: the gathers write and the
      // nested search reads them with 16-byte vector accesses
      const size_t szq = yk_b256(sizeof(float) * (size_t)hn * ld), szn = yk_b256(sizeof(double) * hn),
                   szh = yk_b256(sizeof(uint16_t) * (size_t)hn * ldh),
                   szs = yk_b256(sizeof(float4) * hn);
      void *vg, *voi;
      YK_TRY(yk_bscratch(6, szq + szn + szh + szs, &vg));
      YK_TRY(yk_bscratch(7, yk_b256(sizeof(int) * (size_t)hn * k) + sizeof(float) * (size_t)hn * k,
                         &voi));
      float *gq = (float *)vg;
      double *gqn = (double *)((char *)vg + szq);
      uint16_t *gqh = (uint16_t *)((char *)gqn + szn);
      float4 *gqs = (float4 *)((char *)gqh + szh);
      int *foi = (int *)voi;
      float *fod = (float *)((char *)voi + yk_b256(sizeof(int) * (size_t)hn * k));
      YK_TRY(yk_gather_rows(q + q0 * ld, unc, hn, d, gq));
      if (type != YK_DOT)
        hipLaunchKernelGGL(yk_gather_f64_kernel, dim3((hn + 255) / 256), dim3(256), 0, yk_stream(),
                           qn + q0, unc, (long)hn, gqn);
      const int hu = ldh / 8;  // 16-byte units of a BF16 row
      hipLaunchKernelGGL(yk_gather16_kernel, dim3((unsigned)(((long)hn * hu + 255) / 256)),
                         dim3(256), 0, yk_stream(), (const uint4 *)(qh + q0 * ldh), unc, (long)hn,
                         hu, (uint4 *)gqh);
      hipLaunchKernelGGL(yk_gather16_kernel, dim3((unsigned)((hn + 255) / 256)), dim3(256), 0,
                         yk_stream(), (const uint4 *)(qstats + 4 * q0), unc, (long)hn, 1,
                         (uint4 *)gqs);
      YK_HIP(hipGetLastError());
      int fb2 = 0;
      g_bs_level = 1;
      const int rc2 = yk_knn_bf16(type, hn, nb, d, k, kp2, gq, type == YK_DOT ? nullptr : gqn, gqh,
                                  (const float *)gqs, b, bn, bh, bstats_max, foi, fod, &fb2);
      g_bs_level = 0;
      YK_TRY(rc2);
      long tot = (long)hn * k;
      hipLaunchKernelGGL(yk_scatter_rows_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                         yk_stream(), unc, hn, k, foi, fod, assign + q0 * k, dis + q0 * k);
      YK_HIP(hipGetLastError());
      total_fb += fb2;
    } else if (hn > 0) {
      // uncertified queries: the FP32 form on their rows
      void *vq, *vqn, *voi, *vod;
      const size_t szq = yk_b256(sizeof(float) * (size_t)hn * ld);
      YK_TRY(yk_bscratch(6, szq + sizeof(double) * hn, &vq));
      YK_TRY(yk_bscratch(7, yk_b256(sizeof(int) * (size_t)hn * k) + sizeof(float) * (size_t)hn * k,
                         &voi));
      float *fq = (float *)vq;
      double *fqn = (double *)((char *)vq + szq);
      vqn = fqn;
      int *foi = (int *)voi;
      float *fod = (float *)((char *)voi + yk_b256(sizeof(int) * (size_t)hn * k));
      vod = fod;
      YK_TRY(yk_gather_rows(q + q0 * ld, unc, hn, d, fq));
      if (type != YK_DOT)
        hipLaunchKernelGGL(yk_gather_f64_kernel, dim3((hn + 255) / 256), dim3(256), 0, yk_stream(),
                           qn + q0, unc, (long)hn, (double *)vqn);
      YK_TRY(yk_knn_f32(type, hn, nb, d, k, fq, type == YK_DOT ? nullptr : (double *)vqn, b, bn,
                        nullptr, foi, (float *)vod));
      long tot = (long)hn * k;
      hipLaunchKernelGGL(yk_scatter_rows_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                         yk_stream(), unc, hn, k, foi, fod, assign + q0 * k, dis + q0 * k);
      YK_HIP(hipGetLastError());
      total_fb += hn;
    }
  }
  if (n_fallback) *n_fallback = total_fb;
  return YK_OK;
}

// yk_select.hip -- fused FP32 distance + k-smallest selection on MFMA.
//
// Reference semantics (yael nn.c):
//   dist(q,b) = float( double(sum_j q_j*q_j) [float products, double sum]
//                    + double(float-accumulated sum_j b_j*b_j) )  - 2 * q.b
//   (nn.c:108-126), optionally times w[b] (nn.c:497-500); type 16 is the
//   plain dot product (nn.c:266-306).  k = 1: strict '<' from (-1, 1e30)
//   (nn.c:404-440).  k > 1: the k smallest non-NaN values, sorted, padded
//   with (-1, 0xffffffff) (nn.c:463-525).  Equal values are ordered by base
//   index (the reference's ascending scan gives the same for k = 1).
//
// Every selection works on 64-bit keys (order-preserving float bits << 32 |
// base index): key order is (value, index) order, so the result does not
// depend on how the base is split across lanes, waves and workgroups.
//
// Scan kernel (gfx950, wave64, v_mfma_f32_32x32x2_f32):
//   * a workgroup of W waves; each wave owns 32*NB queries (the MFMA
//     columns, held in registers when d <= 128) and streams its workgroup's
//     base range in blocks of 32 rows (the MFMA rows) from global memory;
//   * per 16-wide k-chunk a lane loads 8 contiguous floats (two dwordx4) of
//     its row: A and B use the same k permutation, each 64-byte row segment
//     is read once per block;
//   * the 32x32 accumulator gives each lane one query and 16 base rows, so
//     the epilogue is lane-local:
//       MODE 0  running (value, index) minimum           -> k = 1
//       MODE 1  each lane's own minimum over its rows     -> the threshold
//               sketch for k > 16
//       MODE 2  register top-R of keys (R <= 16)          -> 2 <= k <= 16
//       MODE 3  append every key <= the query's threshold T_q to a per-query
//               candidate list (staged in LDS, one global
//               atomic per YK_CSTAGE keys)                -> k > 16
//   * a persistent grid (resident workgroups only) walks the work items;
//     each XCD owns a contiguous run of items so workgroups that share a
//     base range share that XCD's L2;
//   * the base is split over work items when the query tiles alone cannot
//     fill 256 CUs; split results are merged by small sort kernels.
//
// k > 16 is exact in two passes: (A) MODE 1 over a 1-in-stride sample of
// base blocks with >= k splits gives, per query, >= 2k minima of disjoint
// row subsets -- distinct real keys -- whose k-th smallest T_q bounds the
// true k-th smallest key from above.  (B) MODE 3 over the whole base
// collects every key <= T_q, then a per-query sort keeps the k smallest.  A
// query whose candidates overflow the list gets a tighter T_q (the k-th
// smallest of the keys it did capture, again an upper bound) and is scanned
// again.
#include <hip/hip_runtime.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>

#include "../../include/yael_mi355x.h"
#include "yk.h"
#include "yk_dev.h"

typedef float f32x16 __attribute__((ext_vector_type(16)));
// word 3 of a raw (untyped, stride 0) buffer descriptor on gfx9-family GPUs
#define YK_BUF_DWORD3 0x00020000

#define YK_CAND_CAP 4096  // candidate keys kept per query (MODE 3)
#define YK_CSTAGE 8       // MODE 3: keys staged in LDS per (lane, query tile)
#define YK_REG_R 16       // register top-R of MODE 2

struct ScanArgs {
  const float *q;
  const double *qn;
  long nq;
  const float *b;
  const float *bn;  // >= round_up(nb, 32) entries
  const float *w;   // optional, >= round_up(nb, 32) entries
  long nb;
  int ld;
  int type;
  long brange;  // base rows per split, multiple of bstep
  long bstep;   // rows between processed 32-row blocks (32 * sampling stride)
  int nsplit;
  int tiles_x;
  int total;
  int per_xcd;
  int *out_i;      // MODE 0, nsplit == 1
  float *out_d;
  uint64_t *part;  // MODE 0: [nq][nsplit]; MODE 2: [nq][nsplit][2][R]
  const uint64_t *thr;  // MODE 3: per-query threshold key (0 = inactive)
  int *cnt;             // MODE 3: per-query candidate count
  uint64_t *cand;       // MODE 3: [nq][YK_CAND_CAP]
  // batched form (item_inst != NULL): work item -> (instance, point tile);
  // q/qn/out_* are the concatenated points of all instances, b/bn the
  // instances' centroid sets (kpad rows apart), nb is unused
  const int *item_inst;
  const int *item_tile;
  const long *inst_off;  // ninst + 1 point offsets
  int kc, kpad;
  const int *done;  // instances already finished (skipped)
  const int *qidx;  // optional: point j of the concatenation is row qidx[j] of q / qn
  int *tick;        // per XCD, zero at launch: item tickets (nullptr: static order)
};

// The hardware hands consecutive block ids to the 8 XCDs in turn; give each
// XCD a contiguous run of work items (split-major), so the workgroups that
// share a base range also share that XCD's L2.
__device__ __forceinline__ int yk_xcd_item(int per_xcd) {
  int id = blockIdx.x;
  return (id & 7) * per_xcd + (id >> 3);
}

// sorted (ascending) register list: insert x, drop the largest
template <int R>
__device__ __forceinline__ void yk_reg_insert(uint64_t (&t)[R], uint64_t x) {
#pragma unroll
  for (int i = R - 1; i > 0; i--) {
    uint64_t lo = t[i - 1], cur = t[i];
    t[i] = x < lo ? lo : (x < cur ? x : cur);
  }
  t[0] = x < t[0] ? x : t[0];
}

// One full block's epilogue for k = 1 and the sketch (each lane keeps the best
// of its own rows), unweighted.  The lane's 16 values are tested first on
// the float approximation qn + (bn - 2 acc), which differs from the exact
// expression by far less than slack = 2^-19 (|q|^2 + max|b|^2)
// (|2 q.b| <= |q|^2 + |b|^2); x -> fl(qn + x) is monotone, so the query
// norm is added once, to the minimum of bn - 2 acc.  A block that may hold
// a better value takes the exact FP64-normed expression for all 16 values,
// branch-free, in ascending row order with a strict '<' (the smallest index
// wins a tie).  The test is skipped (test = false) for a range's first
// blocks: a wave runs the exact pass when any of its 64 lanes may improve,
// which at block j of a range happens with probability about
// 1 - (1 - 1/(j+1))^64, so early blocks would pay for both.
template <bool DOT>
__device__ __forceinline__ void yk_best1_block(const f32x16 &acc, const float (&nn)[16],
                                               double qnd, long bb, int h, bool test, float &bv,
                                               int &bi) {
  // Empty asm statements on the operands keep each pass inside its branch
  // (the compiler would otherwise hoist parts it shares with the general
  // path, executing them for every block).
  if (test) {
    const float qf = (float)qnd;
    float mm = INFINITY, nmax = 0.f;
#pragma unroll
    for (int i = 0; i < 16; i++) {
      float ax = acc[i], nx = nn[i];
      asm volatile("" : "+v"(ax), "+v"(nx));
      mm = fminf(mm, DOT ? ax : __builtin_fmaf(-2.0f, ax, nx));
      nmax = fmaxf(nmax, nx);
    }
    const float m = DOT ? mm : qf + mm;
    const float slack = DOT ? 0.f : 0x1p-19f * (qf + nmax);
    if (!(m - slack <= bv) && m == m) return;
  }
  double qd = qnd;
  asm volatile("" : "+v"(qd));
  // value i sits at row bb + 4 h + 8 (i >> 2) + (i & 3)
  int bo = -1;
#pragma unroll
  for (int i = 0; i < 16; i++) {
    float ax = acc[i], nx = nn[i];
    asm volatile("" : "+v"(ax), "+v"(nx));
    const float v = DOT ? ax : (float)(qd + (double)nx) - 2.0f * ax;
    const bool take = v < bv;
    bv = take ? v : bv;
    bo = take ? 8 * (i >> 2) + (i & 3) : bo;
  }
  if (bo >= 0) bi = (int)bb + 4 * h + bo;
}

// NCH > 0: queries register-resident, exactly NCH 16-wide chunks (ld = 16
// NCH, so every chunk loop is unrolled without guards); NCH = 0: both
// operands streamed from memory.
template <int NB, int NCH, int MODE, int R>
__global__ __launch_bounds__(256) void yk_scan_f32(ScanArgs a0) {
  constexpr bool QRES = NCH > 0;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  const int r = lane & 31, h = lane >> 5;
  // Persistent grid: the hardware deals block ids to the 8 XCDs in turn;
  // XCD x owns the contiguous run of work items [x*per_xcd, (x+1)*per_xcd)
  // (split-major, so concurrently running workgroups of one XCD share a base
  // range in its L2), and its workgroups stride through that run.
  const int xcd = blockIdx.x & 7, local = blockIdx.x >> 3, wpx = gridDim.x >> 3;
  const int ibeg = xcd * a0.per_xcd, iend = min(a0.total, ibeg + a0.per_xcd);
  // MODE 3: candidate keys are staged in LDS, YK_CSTAGE per (query tile,
  // lane), and appended to the query's global list a region at a time (one
  // atomic per region instead of one per key; the count keeps growing past
  // YK_CAND_CAP so the select kernel sees an overflow)
  __shared__ uint64_t cstage[MODE == 3 ? 4 * NB * 64 * YK_CSTAGE : 1];
  uint64_t *const cs = cstage + (MODE == 3 ? wave * NB * 64 * YK_CSTAGE : 0);
  // a fixed first item, then tickets from the XCD's counter: the XCD's
  // workgroups keep to a window of consecutive items (one or two base ranges
  // in its L2) instead of drifting apart
  __shared__ int s_item;
  auto next_item = [&](int it) {
    if (!a0.tick) return it + wpx;
    __syncthreads();  // every wave has read the previous ticket
    if (threadIdx.x == 0) s_item = ibeg + wpx + atomicAdd(&a0.tick[xcd], 1);
    __syncthreads();
    return s_item;
  };
  for (int item = ibeg + local; item < iend; item = next_item(item)) {
  ScanArgs a = a0;
  int tile, split;
  if (a0.item_inst) {
    const int inst = a0.item_inst[item];
    if (a0.done && a0.done[inst]) continue;
    tile = a0.item_tile[item];
    split = 0;
    const long off = a0.inst_off[inst];
    if (a0.qidx) {
      a.qidx = a0.qidx + off;  // rows through the index table, q and qn as given
    } else {
      a.q = a0.q + off * a0.ld;
      a.qn = a0.qn + off;
    }
    a.nq = a0.inst_off[inst + 1] - off;
    a.b = a0.b + (long)inst * a0.kpad * a0.ld;
    a.bn = a0.bn + (long)inst * a0.kpad;
    a.nb = a0.kc;
    a.out_i = a0.out_i + off;
    a.out_d = a0.out_d + off;
  } else {
    tile = item % a0.tiles_x;
    split = item / a0.tiles_x;
  }
  const long qbase = ((long)tile * nwave + wave) * (32 * NB);
  const long bbeg = (long)split * a.brange;
  const long bend = min(a.nb, bbeg + a.brange);
  const int nch = a.ld >> 4;
  const int ld = a.ld;
  const bool dot = (a.type == YK_DOT);

  long qrow[NB];  // the query's position (outputs)
  long qsrc[NB];  // its row of q / qn
  bool qok[NB];
  double qnd[NB];
#pragma unroll
  for (int j = 0; j < NB; j++) {
    long qi = qbase + j * 32 + r;
    qok[j] = qi < a.nq;
    qrow[j] = qok[j] ? qi : a.nq - 1;
    qsrc[j] = a.qidx ? (long)a.qidx[qrow[j]] : qrow[j];
    qnd[j] = dot ? 0.0 : a.qn[qsrc[j]];
  }

  // selection state
  float bv[NB];
  int bi[NB];
  int ncs[NB];
  auto flush = [&](int j, int n) {
    const int base = atomicAdd(&a.cnt[qrow[j]], n);
    const uint64_t *src = cs + (j * 64 + lane) * YK_CSTAGE;
    for (int s_ = 0; s_ < n; s_++)
      if (base + s_ < YK_CAND_CAP) a.cand[qrow[j] * YK_CAND_CAP + base + s_] = src[s_];
  };
  uint64_t top[NB][MODE == 2 ? R : 1];
  uint64_t T[NB];
  float Tv[NB];
#pragma unroll
  for (int j = 0; j < NB; j++) {
    bv[j] = 1e30f;
    bi[j] = -1;
    ncs[j] = 0;
#pragma unroll
    for (int s = 0; s < (MODE == 2 ? R : 1); s++) top[j][s] = YK_KEY_NONE;
    T[j] = MODE == 3 ? a.thr[qrow[j]] : 0;
    Tv[j] = (T[j] == 0) ? -INFINITY
                        : (T[j] == YK_KEY_NONE ? INFINITY : yk_fkey_inv((uint32_t)(T[j] >> 32)));
  }
  if (MODE == 3) {
    bool act = false;
#pragma unroll
    for (int j = 0; j < NB; j++) act |= qok[j] && T[j] != 0;
    if (!__any(act)) continue;  // every query of this wave is already final
  }

  float4 qreg[NB][QRES ? 2 * NCH : 1];
  if (QRES) {
#pragma unroll
    for (int j = 0; j < NB; j++)
#pragma unroll
      for (int c = 0; c < NCH; c++) {
        const float *p = a.q + qsrc[j] * ld + c * 16 + h * 8;
        qreg[j][2 * c] = *(const float4 *)p;
        qreg[j][2 * c + 1] = *(const float4 *)(p + 4);
      }
  }

  // the next block's base operand is loaded while the current one is
  // multiplied (register double buffer, QRES only).  Buffer loads: a
  // descriptor based at the item's first row, the block's row offset in a
  // scalar register, the lane's (row, half) offset in a vector register
  // (rows past nb clamped to the last row, masked by the epilogue) -- no
  // 64-bit vector address arithmetic per block.  The planner keeps an
  // item's rows within 2^31 bytes.
  const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
      (void *)(a.b + bbeg * ld), 0, (int)min((a.nb - bbeg) * ld * 4, 0x7fffffffL), YK_BUF_DWORD3);
  // With one or two chunks a block is only 8-16 MFMAs: its base norms
  // (rows 8g + 4h .. +3) are loaded at the end of the previous block, into
  // the registers the epilogue reads, so that the epilogue does not wait for
  // them -- nor, with them, for the next block's operands issued meanwhile
  // (those are younger).  Dot-product searches read no norms.
  // Weighted searches are served by the other variants (the launcher's
  // choice), so these kernels compile without the weights' loads.
  constexpr bool NPF = QRES && NCH <= 2;
  const float *const wts = NPF ? nullptr : a.w;
  float4 anext[QRES ? 2 * NCH : 1];
  // Buffer loads based at the item's first row (bn holds round_up(nb, 32)
  // entries); for a dot-product search the descriptor has no records, so
  // the loads are unconditional and return zeros without touching memory.
  float4 nn4[NPF ? 4 : 1];
  const __amdgpu_buffer_rsrc_t rsn = __builtin_amdgcn_make_buffer_rsrc(
      (void *)(dot ? a.b : a.bn + bbeg), 0,
      dot ? 0 : (int)min((((a.nb + 31) & ~31L) - bbeg) * 4, 0x7fffffffL), YK_BUF_DWORD3);
  auto loadn = [&](long bb) {
#pragma unroll
    for (int g = 0; g < 4; g++)
      nn4[g] = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(
                                              rsn, 16 * h, (int)((bb - bbeg) * 4) + 32 * g, 0));
  };
  auto loadb = [&](long bb) {
    const int soff = (int)((bb - bbeg) * ld * 4);
    const int voff = ((int)min((long)r, a.nb - 1 - bb) * ld + h * 8) * 4;
#pragma unroll
    for (int c = 0; c < NCH; c++) {
      anext[2 * c] = __builtin_bit_cast(
          float4, __builtin_amdgcn_raw_buffer_load_b128(rs, voff, soff + c * 64, 0));
      anext[2 * c + 1] = __builtin_bit_cast(
          float4, __builtin_amdgcn_raw_buffer_load_b128(rs, voff + 16, soff + c * 64, 0));
    }
  };
  if (QRES && bbeg < bend) loadb(bbeg);
  if (NPF && bbeg < bend) loadn(bbeg);
  for (long bb = bbeg; bb < bend; bb += a.bstep) {
    const long brow = min(bb + r, a.nb - 1);
    const float *bp = a.b + brow * ld + h * 8;
    f32x16 acc[NB];
#pragma unroll
    for (int j = 0; j < NB; j++)
#pragma unroll
      for (int e = 0; e < 16; e++) acc[j][e] = 0.f;

    if (QRES) {
      float4 av[QRES ? 2 * NCH : 1];
#pragma unroll
      for (int c = 0; c < 2 * NCH; c++) av[c] = anext[c];
      {
        // unconditional (past the range it reloads the current block): a
        // conditional prefetch joins paths with different outstanding-load
        // counts, and the wait before the MFMAs would then also wait for the
        // loads just issued
        loadb(bb + a.bstep < bend ? bb + a.bstep : bb);
      }
#pragma unroll
      for (int c = 0; c < NCH; c++)
        {
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
          const float *qp = a.q + qsrc[j] * ld + c * 16 + h * 8;
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
    float nn[16], ww[16];
#pragma unroll
    for (int g = 0; g < 4; g++) {
      const long b4 = bb + 8 * g + 4 * h;
      float4 n4 = NPF   ? nn4[NPF ? g : 0]
                  : dot ? make_float4(0.f, 0.f, 0.f, 0.f)
                        : *(const float4 *)(a.bn + b4);
      float4 w4 = wts ? *(const float4 *)(wts + b4) : make_float4(1.f, 1.f, 1.f, 1.f);
      nn[4 * g] = n4.x; nn[4 * g + 1] = n4.y; nn[4 * g + 2] = n4.z; nn[4 * g + 3] = n4.w;
      ww[4 * g] = w4.x; ww[4 * g + 1] = w4.y; ww[4 * g + 2] = w4.z; ww[4 * g + 3] = w4.w;
    }
#pragma unroll
    for (int j = 0; j < NB; j++) {
      if constexpr (MODE == 0 || MODE == 1) {
        // full blocks of unweighted searches (the last, partial block of a
        // range and weighted searches take the general path below)
        if (!wts && bb + 32 <= bend) {
          const bool test = bb - bbeg >= 48L * a.bstep;
          if (dot)
            yk_best1_block<true>(acc[j], nn, qnd[j], bb, h, test, bv[j], bi[j]);
          else
            yk_best1_block<false>(acc[j], nn, qnd[j], bb, h, test, bv[j], bi[j]);
          continue;
        }
      }
      // Prefilter on float approximations qn + (bn - 2 acc), which differ
      // from the exact expression by far less than 2^-19 (|q|^2 + max|b|^2)
      // (|2 q.b| <= |q|^2 + |b|^2): a min over the lane's 16 values, then a
      // per-value test; only values that may matter take the exact
      // FP64-normed expression.  Weighted searches evaluate every value.
      float nmax = 0.f;
#pragma unroll
      for (int i = 0; i < 16; i++) nmax = fmaxf(nmax, nn[i]);
      const float qf = (float)qnd[j];
      const float slack = dot ? 0.f : 0x1p-19f * (qf + nmax);
      float ap[16];
#pragma unroll
      for (int i = 0; i < 16; i++)
        ap[i] = dot ? acc[j][i] : qf + __builtin_fmaf(-2.0f, acc[j][i], nn[i]);
      const float lim = MODE == 3 ? Tv[j] : (MODE == 2 ? INFINITY : bv[j]);
      if (!wts && MODE != 2) {
        float m = ap[0];
#pragma unroll
        for (int i = 1; i < 16; i++) m = fminf(m, ap[i]);
        if (!(m - slack <= lim) && m == m) continue;
      }
#pragma unroll
      for (int i = 0; i < 16; i++) {
        if (!wts && MODE != 2 && !(ap[i] - slack <= (MODE == 3 ? Tv[j] : bv[j])) && ap[i] == ap[i])
          continue;
        const long bidx = bb + 8 * (i >> 2) + 4 * h + (i & 3);
        const bool bok = bidx < bend;
        float v = dot ? acc[j][i] : (float)(qnd[j] + (double)nn[i]) - 2.0f * acc[j][i];
        if (wts) v *= ww[i];
        if (MODE == 0 || MODE == 1) {
          if (bok && v < bv[j]) {  // strict '<' in ascending index order
            bv[j] = v;
            bi[j] = (int)bidx;
          }
        } else if (MODE == 2) {
          uint64_t key = (bok && v == v) ? yk_key(v, (int)bidx) : YK_KEY_NONE;
          if (key < top[j][R - 1]) yk_reg_insert<R>(top[j], key);
        } else if (bok && qok[j] && v == v) {
          uint64_t key = yk_key(v, (int)bidx);
          if (key <= T[j]) {
            const int n = ncs[j];
            cs[(j * 64 + lane) * YK_CSTAGE + n] = key;
            if (n + 1 == YK_CSTAGE) {
              flush(j, YK_CSTAGE);
              ncs[j] = 0;
            } else {
              ncs[j] = n + 1;
            }
          }
        }
      }
    }
    // the next block's norms (unconditional, as the operands' prefetch)
    if constexpr (NPF) loadn(bb + a.bstep < bend ? bb + a.bstep : bb);
  }

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
      if (h == 0 && qok[j]) {
        if (a.nsplit == 1) {
          a.out_d[qi] = bv[j];
          a.out_i[qi] = bi[j];
        } else {
          a.part[qi * a.nsplit + split] = bi[j] < 0 ? YK_KEY_NONE : yk_key(bv[j], bi[j]);
        }
      }
    }
  } else if (MODE == 1) {
    // sketch: each lane's own minimum (distinct rows per lane -> distinct keys)
#pragma unroll
    for (int j = 0; j < NB; j++) {
      const long qi = qbase + j * 32 + r;
      if (qok[j])
        a.part[(qi * a.nsplit + split) * 2 + h] = bi[j] < 0 ? YK_KEY_NONE : yk_key(bv[j], bi[j]);
    }
  } else if (MODE == 2) {
#pragma unroll
    for (int j = 0; j < NB; j++) {
      const long qi = qbase + j * 32 + r;
      if (!qok[j]) continue;
      uint64_t *dst = a.part + ((qi * a.nsplit + split) * 2 + h) * R;
#pragma unroll
      for (int s = 0; s < R; s++) dst[s] = top[j][s];
    }
  } else {
#pragma unroll
    for (int j = 0; j < NB; j++)
      if (ncs[j] > 0) flush(j, ncs[j]);
  }
  }  // work items
}

// ---------------------------------------------------------------------------
// small per-query sorts

__device__ __forceinline__ void yk_store_key(uint64_t key, int *oi, float *od) {
  *oi = key == YK_KEY_NONE ? -1 : (int)(uint32_t)key;
  *od = key == YK_KEY_NONE ? __uint_as_float(0xffffffffu) : yk_fkey_inv((uint32_t)(key >> 32));
}

// k = 1 merge of per-split minima (nothing below 1e30 -> (-1, 1e30))
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

// ascending bitonic sort of s[0..p2) by the block (p2 a power of two)
__device__ void yk_block_bitonic(uint64_t *s, int p2) {
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
}

// Sort the n keys of each query (src + q*n); then either write the k
// smallest (out_i/out_d) or, when thr != NULL, the k-th smallest as the
// query's threshold (YK_KEY_NONE when fewer than k real keys: admit all).
__global__ __launch_bounds__(256) void yk_sortsel_kernel(const uint64_t *__restrict__ src,
                                                         int n, int k, int p2, int *out_i,
                                                         float *out_d, uint64_t *thr) {
  extern __shared__ uint64_t s[];
  const long q = blockIdx.x;
  const uint64_t *p = src + q * (long)n;
  for (int i = threadIdx.x; i < p2; i += blockDim.x) s[i] = i < n ? p[i] : YK_KEY_NONE;
  __syncthreads();
  yk_block_bitonic(s, p2);
  if (thr) {
    if (threadIdx.x == 0) thr[q] = k <= p2 ? s[k - 1] : YK_KEY_NONE;
    return;
  }
  for (int i = threadIdx.x; i < k; i += blockDim.x)
    yk_store_key(i < p2 ? s[i] : YK_KEY_NONE, out_i + q * k + i, out_d + q * k + i);
}

// After a MODE 3 pass: queries whose list fitted get their k smallest and
// become inactive (thr = 0); overflowing queries get the k-th smallest
// captured key as a tighter threshold, their count is reset, *flag is set.
// A query with fewer than k keys under a bound (the estimated starting
// bound, or with NaN distances, which no bound admits, any bound) gets no
// bound, its count reset and the flag: its next pass admits every value
// (and overflows, the overflow's bound then being a proven one); a query
// with fewer than k keys and no bound has no more to find.
__global__ __launch_bounds__(256) void yk_select_kernel(const uint64_t *__restrict__ cand,
                                                        int *cnt, uint64_t *thr, int k,
                                                        int *out_i, float *out_d, int *flag) {
  __shared__ uint64_t s[YK_CAND_CAP];
  const long q = blockIdx.x;
  if (thr[q] == 0) return;
  const int n = cnt[q];
  if (n < k && thr[q] != YK_KEY_NONE) {
    if (threadIdx.x == 0) {
      thr[q] = YK_KEY_NONE;
      cnt[q] = 0;
      atomicOr(flag, 1);
    }
    return;
  }
  const int m = n < YK_CAND_CAP ? n : YK_CAND_CAP;
  int p2 = 1;
  while (p2 < m) p2 <<= 1;
  const uint64_t *p = cand + q * (long)YK_CAND_CAP;
  for (int i = threadIdx.x; i < p2; i += blockDim.x) s[i] = i < m ? p[i] : YK_KEY_NONE;
  __syncthreads();
  yk_block_bitonic(s, p2);
  if (n > YK_CAND_CAP) {
    if (threadIdx.x == 0) {
      thr[q] = s[k - 1];
      cnt[q] = 0;
      atomicOr(flag, 1);
    }
    return;
  }
  for (int i = threadIdx.x; i < k; i += blockDim.x)
    yk_store_key(i < p2 ? s[i] : YK_KEY_NONE, out_i + q * k + i, out_d + q * k + i);
  __syncthreads();
  if (threadIdx.x == 0) thr[q] = 0;
}

__global__ void yk_zero_cnt_kernel(int *cnt, long n) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i < n) cnt[i] = 0;
}

// ---------------------------------------------------------------------------
// planning and launch

struct YkPlan {
  int nblk;    // query blocks of 32 per wave (NB)
  int waves;   // waves per workgroup
  int tiles;   // query tiles
  int nsplit;  // base splits
  long brange;
  long bstep;
};

// Launch-plan knobs the autotuner chooses among (every choice gives the
// same bits): the number of work items the base split aims at, and a
// scale of the k > 16 sketch's sampling stride.  cfg = target index + 3 x
// stride index; 0 is the default plan.
static const long yk_item_targets[3] = {4096, 2048, 8192};
static const int yk_stride_scale[4][2] = {{1, 1}, {1, 2}, {2, 1}, {16, 3}};  // num., denom.
#define YK_NCFG 12

// Split the base so that tiles * nsplit reaches the item target (~8 per
// resident workgroup by default), with at least min_rows rows per split and
// at most max_split splits.
static YkPlan yk_plan(long nq, long nb, long sample, long min_splits, long max_split,
                      int cfg = 0) {
  YkPlan p;
  p.nblk = 1;
  p.waves = 4;
  const long qt = 32L * p.nblk * p.waves;
  p.tiles = (int)((nq + qt - 1) / qt);
  p.bstep = 32 * sample;
  const long nblocks = (nb + p.bstep - 1) / p.bstep;  // processed blocks
  // ~8 work items per resident workgroup (512 of them: 256 CUs x 2) keeps
  // the tail of the persistent grid short
  const long target = yk_item_targets[cfg % 3];
  long want = (target + p.tiles - 1) / p.tiles;
  long maxs = (nblocks + 15) / 16;  // >= 16 processed blocks per split
  if (want > maxs) want = maxs;
  if (want < min_splits) want = min_splits;
  if (want > max_split) want = max_split;
  if (want > nblocks) want = nblocks;
  if (want < 1) want = 1;
  long per = (nblocks + want - 1) / want;  // processed blocks per split
  // an item's rows stay within 2^30 bytes (buffer-load offsets are 32-bit)
  const long max_per = (1L << 30) / (32L * 128 * 4 * sample);
  if (per > max_per) per = max_per;
  p.brange = per * p.bstep;
  p.nsplit = (int)((nb + p.brange - 1) / p.brange);
  return p;
}

extern "C" int yk_knn_plan(long nq, long nb, int d, int k, int *nsplit, int *waves, int *nblk) {
  (void)d;
  YkPlan p = yk_plan(nq, nb, 1, 1, k > 1 ? 256 : 4096);
  *nsplit = p.nsplit;
  *waves = p.waves;
  *nblk = p.nblk;
  return YK_OK;
}

// Device scratch of the selection pipeline.  Two sources: the grow-only
// buffers the library allocates and keeps (arena == nullptr: slot i is one
// buffer, reallocated when a call needs more), or a caller-given region
// carved front to back in 256-byte-aligned pieces (the no-allocation entry
// point's static array).  A request that does not fit the region is an
// error, never an allocation.
static void *g_scratch[4];
static size_t g_scratch_bytes[4];

struct YkArena {
  char *base;
  size_t cap, used;
};

static size_t yk_align256(size_t b) { return (b + 255) & ~(size_t)255; }

static int yk_scratch(YkArena *ar, int slot, size_t bytes, void **out) {
  if (ar) {
    const size_t at = yk_align256(ar->used);
    if (at > ar->cap || bytes > ar->cap - at)
      return yk_set_error(YK_EINVAL, "scratch request of %zu bytes exceeds the %zu-byte region",
                          bytes, ar->cap);
    *out = ar->base + at;
    ar->used = at + bytes;
    return YK_OK;
  }
  if (bytes > g_scratch_bytes[slot]) {
    if (g_scratch[slot] && hipFree(g_scratch[slot]) != hipSuccess) (void)hipGetLastError();
    g_scratch[slot] = nullptr;
    g_scratch_bytes[slot] = 0;
    hipError_t e = hipMalloc(&g_scratch[slot], bytes);
    if (e != hipSuccess) return yk_hip_fail(e, "hipMalloc(selection scratch)");
    g_scratch_bytes[slot] = bytes;
  }
  *out = g_scratch[slot];
  return YK_OK;
}

// the kernel with exactly ld/16 register-resident chunks (ld <= 128), else
// the one streaming both operands
template <int MODE, int R>
static void (*yk_f32_pick(int ld, bool weighted))(ScanArgs) {
  if (weighted && ld <= 32) return yk_scan_f32<1, 0, MODE, R>;  // no weights in NCH <= 2
  switch (ld >> 4) {
    case 1: return yk_scan_f32<1, 1, MODE, R>;
    case 2: return yk_scan_f32<1, 2, MODE, R>;
    case 3: return yk_scan_f32<1, 3, MODE, R>;
    case 4: return yk_scan_f32<1, 4, MODE, R>;
    case 5: return yk_scan_f32<1, 5, MODE, R>;
    case 6: return yk_scan_f32<1, 6, MODE, R>;
    case 7: return yk_scan_f32<1, 7, MODE, R>;
    case 8: return yk_scan_f32<1, 8, MODE, R>;
    default: return yk_scan_f32<1, 0, MODE, R>;
  }
}

template <int MODE, int R>
static int yk_launch_scan(const YkPlan &p, ScanArgs a, bool qres) {
  a.brange = p.brange;
  a.bstep = p.bstep;
  a.nsplit = p.nsplit;
  a.tiles_x = p.tiles;
  a.total = p.tiles * p.nsplit;
  a.per_xcd = (a.total + 7) / 8;
  (void)qres;
  auto kern = yk_f32_pick<MODE, R>(a.ld, a.w != nullptr);
  YK_TRY(yk_item_tickets(&a.tick));
  const int per_xcd_res = yk_resident_wgs((const void *)kern, 64 * p.waves) / 8;
  dim3 grid(8 * (a.per_xcd < per_xcd_res ? a.per_xcd : per_xcd_res)), block(64 * p.waves);
  int tok = yk_prof_begin(MODE == 1 ? YK_PROF_F32_SKETCH : YK_PROF_FUSED);
  hipLaunchKernelGGL(kern, grid, block, 0, yk_stream(), a);
  yk_prof_end(tok);
  YK_HIP(hipGetLastError());
  return YK_OK;
}

static int yk_sortsel(const uint64_t *src, long nq, int n, int k, int *oi, float *od,
                      uint64_t *thr) {
  int p2 = 1;
  while (p2 < n) p2 <<= 1;
  int tok = yk_prof_begin(YK_PROF_MERGE);
  hipLaunchKernelGGL(yk_sortsel_kernel, dim3((unsigned)nq), dim3(256), sizeof(uint64_t) * p2,
                     yk_stream(), src, n, k, p2, oi, od, thr);
  yk_prof_end(tok);
  YK_HIP(hipGetLastError());
  return YK_OK;
}

// Sketch stride of the k > 16 path: every lane's minimum over its rows of
// a 1-in-stride sample of base blocks, with >= k splits -> >= 2k distinct
// real keys per query.  The k-th smallest of 2k subset minima sits near
// global rank 1.4 * stride * k, the expected candidate count; ~800 keeps
// the filtered pass's append path rare and the list far from overflow.
static long yk_sketch_stride(long nb, int k, int cfg = 0) {
  const int *sc = yk_stride_scale[(cfg / 3) % 4];
  long stride = 600L * sc[0] / sc[1] / k;
  const long cap = sc[0] == 16 ? 16 : 8;  // the last scale may sample 1 in 16
  if (stride > cap) stride = cap;
  if (stride < 1) stride = 1;
  const long nblocks = (nb + 31) / 32;
  while (stride > 1 && nblocks / stride < 2L * k) stride >>= 1;
  return stride;
}

// queries per pass of the k > 16 path (bounds the candidate memory)
#define YK_LARGE_K_QB 16384L

// k > 16: sketch threshold, filtered scan, per-query sort (queries in batches
// that bound the candidate memory).
static int yk_knn_large_k(ScanArgs base, int k, bool qres, int *assign, float *dis,
                          YkArena *ar, int cfg) {
  const long QB = YK_LARGE_K_QB;
  const long nq_all = base.nq;
  void *vthr, *vcnt, *vcand, *vpart, *vflag;
  const long qb = nq_all < QB ? nq_all : QB;
  YK_TRY(yk_scratch(ar, 0, sizeof(uint64_t) * qb, &vthr));
  YK_TRY(yk_scratch(ar, 1, sizeof(int) * (qb + 64), &vcnt));
  YK_TRY(yk_scratch(ar, 2, sizeof(uint64_t) * qb * YK_CAND_CAP, &vcand));
  const size_t fixed_used = ar ? ar->used : 0;
  uint64_t *thr = (uint64_t *)vthr;
  int *cnt = (int *)vcnt;
  int *flag = cnt + qb + 32;
  vflag = flag;
  const long stride = yk_sketch_stride(base.nb, k, cfg);
  const long min_splits = k;

  for (long q0 = 0; q0 < nq_all; q0 += QB) {
    const long nq = nq_all - q0 < QB ? nq_all - q0 : QB;
    ScanArgs a = base;
    a.q = base.q + q0 * base.ld;
    a.qn = base.qn ? base.qn + q0 : nullptr;
    a.nq = nq;
    // (A) sketch: per-lane minima over a strided sample
    YkPlan pa = yk_plan(nq, a.nb, stride, min_splits, min_splits > 256 ? min_splits : 256, cfg);
    if (ar) ar->used = fixed_used;  // each pass reuses the region's partial-key piece
    YK_TRY(yk_scratch(ar, 3, sizeof(uint64_t) * nq * pa.nsplit * 2, &vpart));
    a.part = (uint64_t *)vpart;
    YK_TRY((yk_launch_scan<1, 1>(pa, a, qres)));
    // the starting bound: the m-th smallest minimum (yk_bound_rank)
    YK_TRY(yk_sortsel(a.part, nq, pa.nsplit * 2, yk_bound_rank(k, stride), nullptr, nullptr,
                      thr));
    // (B) filtered scan + per-query sort, retried for overflowing queries
    YkPlan pb = yk_plan(nq, a.nb, 1, 1, 4096, cfg);
    a.thr = thr;
    a.cnt = cnt;
    a.cand = (uint64_t *)vcand;
    hipLaunchKernelGGL(yk_zero_cnt_kernel, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0,
                       yk_stream(), cnt, nq);
    hipLaunchKernelGGL(yk_zero_cnt_kernel, dim3(1), dim3(64), 0, yk_stream(), flag, 1L);
    int tries = 0;
    for (;;) {
      YK_TRY((yk_launch_scan<3, 1>(pb, a, qres)));
      int tok = yk_prof_begin(YK_PROF_MERGE);
      hipLaunchKernelGGL(yk_select_kernel, dim3((unsigned)nq), dim3(256), 0, yk_stream(), a.cand,
                         cnt, thr, k, assign + q0 * k, dis + q0 * k, flag);
      yk_prof_end(tok);
      YK_HIP(hipGetLastError());
      int hflag = 0;
      YK_TRY(yk_read_small(&hflag, vflag, sizeof(int)));
      if (!hflag) break;
      if (++tries == 6)
        return yk_set_error(YK_EINVAL,
                            "k=%d: more than %d keys tie below the k-th distance of a query", k,
                            YK_CAND_CAP);
      hipLaunchKernelGGL(yk_zero_cnt_kernel, dim3(1), dim3(64), 0, yk_stream(), flag, 1L);
    }
  }
  return YK_OK;
}

// Checks shared by the FP32 entry points; 1 when there is nothing to do.
static int yk_knn_f32_check(int *type, long nq, long nb, int k) {
  if (*type == 12) *type = YK_L2;
  if (*type != YK_L2 && *type != YK_DOT)
    return yk_set_error(YK_EINVAL, "distance type %d has no GPU path (2, 12, 16 only)", *type);
  if (k < 1 || k > nb)
    return yk_set_error(YK_EINVAL, "k=%d out of range (nb=%ld)", k, nb);
  if (nb > 0x7fffffffL || nq > 0x7fffffffL) return yk_set_error(YK_EINVAL, "sizes exceed int32");
  return nq <= 0 ? 1 : YK_OK;
}

// The FP32 search of nq queries, scratch from ar (nullptr: the grow-only
// buffers).
static int yk_knn_f32_run(int type, long nq, long nb, int d, int k, const float *q,
                          const double *qn, const float *b, const float *bn, const float *w,
                          int *assign, float *dis, YkArena *ar, int cfg = 0) {
  if (k > YK_KMAX) {  // beyond the on-chip selection: distance rows + radix select
    if (ar) return yk_set_error(YK_EINVAL, "k=%d > %d needs allocated scratch", k, YK_KMAX);
    return yk_knn_bigk(type, nq, nb, d, k, q, qn, b, bn, w, assign, dis);
  }
  ScanArgs a = {};
  a.q = q;
  a.qn = qn;
  a.nq = nq;
  a.b = b;
  a.bn = bn;
  a.w = w;
  a.nb = nb;
  a.ld = yk_ld(d);
  a.type = type;
  a.out_i = assign;
  a.out_d = dis;
  const bool qres = a.ld <= 128;
  if (k == 1) {
    YkPlan p = yk_plan(nq, nb, 1, 1, 4096, cfg);
    if (p.nsplit > 1) {
      void *buf;
      YK_TRY(yk_scratch(ar, 3, sizeof(uint64_t) * nq * p.nsplit, &buf));
      a.part = (uint64_t *)buf;
    }
    YK_TRY((yk_launch_scan<0, 1>(p, a, qres)));
    if (p.nsplit > 1) {
      int tok = yk_prof_begin(YK_PROF_MERGE);
      hipLaunchKernelGGL(yk_merge1_kernel, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0,
                         yk_stream(), a.part, nq, p.nsplit, assign, dis);
      yk_prof_end(tok);
      YK_HIP(hipGetLastError());
    }
    return YK_OK;
  }
  if (k <= YK_REG_R) {
    YkPlan p = yk_plan(nq, nb, 1, 1, 256, cfg);
    void *buf;
    YK_TRY(yk_scratch(ar, 3, sizeof(uint64_t) * nq * p.nsplit * 2 * YK_REG_R, &buf));
    a.part = (uint64_t *)buf;
    YK_TRY((yk_launch_scan<2, YK_REG_R>(p, a, qres)));
    return yk_sortsel(a.part, nq, p.nsplit * 2 * YK_REG_R, k, assign, dis, nullptr);
  }
  return yk_knn_large_k(a, k, qres, assign, dis, ar, cfg);
}

// a search through yk_time_ms
struct YkF32Call {
  int type;
  long nq, nb;
  int d, k;
  const float *q;
  const double *qn;
  const float *b, *bn, *w;
  int *assign;
  float *dis;
  int cfg;
};
static int yk_f32_call(void *p) {
  const YkF32Call &c = *(const YkF32Call *)p;
  return yk_knn_f32_run(c.type, c.nq, c.nb, c.d, c.k, c.q, c.qn, c.b, c.bn, c.w, c.assign, c.dis,
                        nullptr, c.cfg);
}

// log2 bucket of a size (shape classes of the tuning cache)
static int yk_lg(long x) {
  int l = 0;
  while (x > 1) {
    x >>= 1;
    l++;
  }
  return l;
}

// The plan: the tuning cache's choice for the shape class; on a miss (and
// a search of at least 2^30 flop, where one timing is meaningful), every
// candidate plan is timed once after an untimed run of the default, the
// fastest recorded.  The outputs then hold the last candidate's results --
// the same bits as any other plan's.  *ran = 1 when the outputs are already
// written.
static int yk_f32_plan(const YkF32Call &call, int *cfg, int *ran) {
  *cfg = 0;
  *ran = 0;
  if (call.k > YK_KMAX || !yk_tune_enabled()) return YK_OK;
  if ((double)call.nq * call.nb * call.d < (double)(1 << 29)) return YK_OK;
  const int kc = call.k == 1 ? 1 : call.k <= YK_REG_R ? 16 : 1 << yk_lg(call.k);
  char key[128];
  snprintf(key, sizeof key, "f32:t%d:k%d:ld%d:q%d:b%d", call.type, kc, yk_ld(call.d),
           yk_lg(call.nq), yk_lg(call.nb));
  if (yk_tune_get(key, cfg)) {
    if (*cfg < 0 || *cfg >= YK_NCFG) *cfg = 0;
    return YK_OK;
  }
  const int ncand = call.k > YK_REG_R ? YK_NCFG : 3;  // the stride matters for k > 16 only
  // candidates whose plans come out the same as an earlier one's are not
  // timed again (e.g. queries enough to fill the chip without any split)
  long sig[YK_NCFG][3];
  int distinct = 0;
  bool dup[YK_NCFG] = {};
  for (int i = 0; i < ncand; i++) {
    const long nqb = call.nq < YK_LARGE_K_QB ? call.nq : YK_LARGE_K_QB;
    if (call.k == 1) {
      const YkPlan p = yk_plan(call.nq, call.nb, 1, 1, 4096, i);
      sig[i][0] = p.nsplit, sig[i][1] = p.brange, sig[i][2] = 0;
    } else if (call.k <= YK_REG_R) {
      const YkPlan p = yk_plan(call.nq, call.nb, 1, 1, 256, i);
      sig[i][0] = p.nsplit, sig[i][1] = p.brange, sig[i][2] = 0;
    } else {
      const long st = yk_sketch_stride(call.nb, call.k, i);
      const YkPlan pa = yk_plan(nqb, call.nb, st, call.k, call.k > 256 ? call.k : 256, i);
      const YkPlan pb = yk_plan(nqb, call.nb, 1, 1, 4096, i);
      sig[i][0] = st, sig[i][1] = pa.nsplit, sig[i][2] = pb.nsplit;
    }
    for (int j = 0; j < i && !dup[i]; j++)
      dup[i] = !dup[j] && sig[j][0] == sig[i][0] && sig[j][1] == sig[i][1] &&
               sig[j][2] == sig[i][2];
    distinct += !dup[i];
  }
  if (distinct < 2) return YK_OK;  // one plan: nothing to choose
  YkF32Call c = call;
  c.cfg = 0;
  YK_TRY(yk_f32_call(&c));  // warm-up
  float best = 1e30f;
  for (int i = 0; i < ncand; i++) {
    if (dup[i]) continue;
    c.cfg = i;
    float ms = 0;
    YK_TRY(yk_time_ms(yk_f32_call, &c, &ms));
    if (ms < best) {
      best = ms;
      *cfg = i;
    }
  }
  yk_tune_put(key, *cfg);
  *ran = 1;
  return YK_OK;
}

extern "C" int yk_knn_f32(int type, long nq, long nb, int d, int k, const float *q,
                          const double *qn, const float *b, const float *bn, const float *w,
                          int *assign, float *dis) {
  YK_TRY(yk_init());
  YkGuard guard;
  const int c = yk_knn_f32_check(&type, nq, nb, k);
  if (c != YK_OK) return c > 0 ? YK_OK : c;
  const YkF32Call call = {type, nq, nb, d, k, q, qn, b, bn, w, assign, dis, 0};
  int cfg = 0, ran = 0;
  YK_TRY(yk_f32_plan(call, &cfg, &ran));
  if (ran) return YK_OK;
  return yk_knn_f32_run(type, nq, nb, d, k, q, qn, b, bn, w, assign, dis, nullptr, cfg);
}

// ---------------------------------------------------------------------------
// the FP32 search without allocation: scratch in a static device array

// Every GPU the library runs on holds one copy of this array for the life
// of the process.  Host code never passes its name to a kernel (on the host
// it is only a stand-in for the symbol): the device address of each GPU's
// copy comes from hipGetSymbolAddress, taken once per device.
static __device__ __align__(256) unsigned char scratch[YK_NOALLOC_SCRATCH_BYTES];

#define YK_NOALLOC_MAX_DEV 64
static void *g_noalloc_ptr[YK_NOALLOC_MAX_DEV];
static pthread_mutex_t g_noalloc_mu = PTHREAD_MUTEX_INITIALIZER;

// the current device's copy of the array (set-up on the first call there)
static int yk_noalloc_region(char **out) {
  int dev = -1;
  YK_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= YK_NOALLOC_MAX_DEV)
    return yk_set_error(YK_EINVAL, "device %d beyond the %d the scratch table covers", dev,
                        YK_NOALLOC_MAX_DEV);
  pthread_mutex_lock(&g_noalloc_mu);
  int rc = YK_OK;
  if (!g_noalloc_ptr[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess)
      rc = yk_hip_fail(e, "hipGetSymbolAddress(scratch)");
    else if (!ptr)
      rc = yk_set_error(YK_EHIP, "hipGetSymbolAddress(scratch) gave a null address");
    else
      g_noalloc_ptr[dev] = ptr;
  }
  *out = (char *)g_noalloc_ptr[dev];
  pthread_mutex_unlock(&g_noalloc_mu);
  return rc;
}

// Scratch bytes one yk_knn_f32_run over nq queries takes from a region (the
// same plans and pieces as the run itself, each rounded up to 256 bytes).
static size_t yk_knn_f32_need(long nq, long nb, int k) {
  if (k == 1) {
    const YkPlan p = yk_plan(nq, nb, 1, 1, 4096);
    return p.nsplit > 1 ? yk_align256(sizeof(uint64_t) * nq * p.nsplit) : 0;
  }
  if (k <= YK_REG_R) {
    const YkPlan p = yk_plan(nq, nb, 1, 1, 256);
    return yk_align256(sizeof(uint64_t) * nq * p.nsplit * 2 * YK_REG_R);
  }
  const long qb = nq < YK_LARGE_K_QB ? nq : YK_LARGE_K_QB;
  const long stride = yk_sketch_stride(nb, k);
  const YkPlan pa = yk_plan(qb, nb, stride, k, k > 256 ? k : 256);
  return yk_align256(sizeof(uint64_t) * qb) + yk_align256(sizeof(int) * (qb + 64)) +
         yk_align256(sizeof(uint64_t) * qb * YK_CAND_CAP) +
         yk_align256(sizeof(uint64_t) * qb * pa.nsplit * 2);
}

extern "C" int yk_knn_f32_noalloc(int type, long nq, long nb, int d, int k, const float *q,
                                  const double *qn, const float *b, const float *bn,
                                  const float *w, int *assign, float *dis) {
  YK_TRY(yk_init());
  YkGuard guard;
  const int c = yk_knn_f32_check(&type, nq, nb, k);
  if (c != YK_OK) return c > 0 ? YK_OK : c;
  char *region = nullptr;
  YK_TRY(yk_noalloc_region(&region));
  // Queries go in passes small enough for the region: the largest
  // power-of-two pass (at most 16384 queries) whose scratch fits.  Each
  // query's result depends only on that query, so the passes give the
  // single call's results.
  long pass = 16384;
  while (pass > 32 && yk_knn_f32_need(pass < nq ? pass : nq, nb, k) > YK_NOALLOC_SCRATCH_BYTES)
    pass >>= 1;
  const size_t need = yk_knn_f32_need(pass < nq ? pass : nq, nb, k);
  if (need > YK_NOALLOC_SCRATCH_BYTES)
    return yk_set_error(YK_EINVAL,
                        "k=%d, nb=%ld needs %zu scratch bytes even for %ld queries at a time; "
                        "the no-allocation search holds %zu",
                        k, nb, need, pass, (size_t)YK_NOALLOC_SCRATCH_BYTES);
  const int ld = yk_ld(d);
  for (long q0 = 0; q0 < nq; q0 += pass) {
    const long n = nq - q0 < pass ? nq - q0 : pass;
    YkArena ar = {region, (size_t)YK_NOALLOC_SCRATCH_BYTES, 0};
    YK_TRY(yk_knn_f32_run(type, n, nb, d, k, q + q0 * ld, qn ? qn + q0 : nullptr, b, bn, w,
                          assign + q0 * k, dis + q0 * k, &ar));
  }
  return YK_OK;
}

// ---------------------------------------------------------------------------
// batched form: the assign step of many k-means instances in one launch

// Small centroid sets (NBLK 32-row blocks, d <= 16 NCH): one workgroup per
// update chunk (up to `chunk` consecutive points of one instance) stages the
// instance's centroid fragments in LDS and streams the chunk's points
// through in 32-point blocks, wave w taking blocks w, w + W, ...; a block's
// fragments are reloaded chunk by chunk right after the centroid blocks'
// MFMAs have read them, so every load has the next block's MFMAs to
// arrive in.  (The general scan holds the points and streams the
// centroids, which for k <= 64 leaves each 32-point tile two centroid blocks
// of work against the full latency of loading its points.)  Same MFMA
// operands, same order and the same epilogue arithmetic as the general
// scan's MODE 0: the outputs are its bits.
struct BassignArgs {
  const float *v;
  const double *vn;
  const float *cent, *cn;
  int ld, kc, kpad, chunk;
  const long *inst_off;
  const int *uitem_inst, *uitem_first;
  int nuitems, per_xcd;
  const int *done;
  int *assign;
  float *dis;
  int *tick;
};

template <int NCH, int NBLK>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void
yk_bassign_res_kernel(BassignArgs a) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwave = blockDim.x >> 6;
  const int r = lane & 31, h = lane >> 5;
  const int xcd = blockIdx.x & 7, local = blockIdx.x >> 3, wpx = gridDim.x >> 3;
  const int ibeg = xcd * a.per_xcd, iend = min(a.nuitems, ibeg + a.per_xcd);
  const int ld = a.ld;
  __shared__ float4 cfs[NBLK * 2 * NCH * 64];
  __shared__ float4 nns[NBLK * 4 * 64];
  __shared__ int s_item;
  auto next_item = [&](int it) {
    __syncthreads();  // also: every wave is done with the staged centroids
    if (threadIdx.x == 0) s_item = ibeg + wpx + atomicAdd(&a.tick[xcd], 1);
    __syncthreads();
    return s_item;
  };
  for (int u = ibeg + local; u < iend; u = next_item(u)) {
    const int inst = a.uitem_inst[u];
    if (a.done[inst]) continue;
    const long off = a.inst_off[inst], n = a.inst_off[inst + 1] - off;
    const long p0 = (long)(u - a.uitem_first[inst]) * a.chunk;
    const long p1 = p0 + a.chunk < n ? p0 + a.chunk : n;
    const float *cb0 = a.cent + (long)inst * a.kpad * ld;
    const float *cnb = a.cn + (long)inst * a.kpad;
    // centroid fragments and norms, staged in LDS once for the workgroup
    // in the lanes' order: fragment entry (cb, half-chunk q, lane) holds row
    // 32 cb + r, dimensions 16 (q / 2) + 8 h + 4 (q % 2) .. + 3 (rows past kc
    // clamped to the last, masked below); norm entry (cb, g, lane) rows
    // 32 cb + 8 g + 4 h .. + 3
    for (int i = threadIdx.x; i < NBLK * 2 * NCH * 64; i += blockDim.x) {
      const int ln = i & 63, q = (i >> 6) % (2 * NCH), cb = (i >> 6) / (2 * NCH);
      const float *row = cb0 + (long)min(32 * cb + (ln & 31), a.kc - 1) * ld + (ln >> 5) * 8;
      cfs[i] = *(const float4 *)(row + 16 * (q >> 1) + 4 * (q & 1));
    }
    for (int i = threadIdx.x; i < NBLK * 4 * 64; i += blockDim.x) {
      const int ln = i & 63, g = (i >> 6) & 3, cb = i >> 8;
      nns[i] = *(const float4 *)(cnb + 32 * cb + 8 * g + 4 * (ln >> 5));
    }
    __syncthreads();
    const long step = 32L * nwave;
    long pb = p0 + 32L * wave;
    if (pb >= p1) continue;
    // the block's point fragments and norm (rows past the chunk clamped to
    // its last point, not written)
    float4 bq[2 * NCH];
    auto loadq = [&](long blk, int c) {
      const float *pr = a.v + (off + min(blk + r, p1 - 1)) * ld + h * 8 + 16 * c;
      bq[2 * c] = *(const float4 *)pr;
      bq[2 * c + 1] = *(const float4 *)(pr + 4);
    };
#pragma unroll
    for (int c = 0; c < NCH; c++) loadq(pb, c);
    double qd = a.vn[off + min(pb + r, p1 - 1)];
    for (; pb < p1; pb += step) {
      const long pn = pb + step < p1 ? pb + step : pb;  // past the chunk: reload this block
      const double qdn = a.vn[off + min(pn + r, p1 - 1)];
      // the centroid blocks' MFMA chains interleaved chunk by chunk (two
      // independent accumulators); chunk c of the next point block loads
      // once every chain has read it, a whole block of MFMAs ahead of use
      f32x16 acc[NBLK];
#pragma unroll
      for (int cb = 0; cb < NBLK; cb++)
#pragma unroll
        for (int e = 0; e < 16; e++) acc[cb][e] = 0.f;
      // the centroid operand of chunk c + 1 is read from LDS while chunk
      // c's MFMAs run
      float4 av[2][NBLK][2];
#pragma unroll
      for (int cb = 0; cb < NBLK; cb++) {
        av[0][cb][0] = cfs[(cb * 2 * NCH) * 64 + lane];
        av[0][cb][1] = cfs[(cb * 2 * NCH + 1) * 64 + lane];
      }
#pragma unroll
      for (int c = 0; c < NCH; c++) {
        const float4 b0 = bq[2 * c], b1 = bq[2 * c + 1];
        if (c + 1 < NCH) {
#pragma unroll
          for (int cb = 0; cb < NBLK; cb++) {
            av[(c + 1) & 1][cb][0] = cfs[(cb * 2 * NCH + 2 * c + 2) * 64 + lane];
            av[(c + 1) & 1][cb][1] = cfs[(cb * 2 * NCH + 2 * c + 3) * 64 + lane];
          }
        }
#pragma unroll
        for (int cb = 0; cb < NBLK; cb++) {
          const float4 a0 = av[c & 1][cb][0], a1 = av[c & 1][cb][1];
          acc[cb] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.x, b0.x, acc[cb], 0, 0, 0);
          acc[cb] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.y, b0.y, acc[cb], 0, 0, 0);
          acc[cb] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.z, b0.z, acc[cb], 0, 0, 0);
          acc[cb] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0.w, b0.w, acc[cb], 0, 0, 0);
          acc[cb] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.x, b1.x, acc[cb], 0, 0, 0);
          acc[cb] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.y, b1.y, acc[cb], 0, 0, 0);
          acc[cb] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.z, b1.z, acc[cb], 0, 0, 0);
          acc[cb] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1.w, b1.w, acc[cb], 0, 0, 0);
        }
        __builtin_amdgcn_sched_barrier(0);
        loadq(pn, c);
        __builtin_amdgcn_sched_barrier(0);
      }
      float bv = 1e30f;
      int bi = -1;
#pragma unroll
      for (int cb = 0; cb < NBLK; cb++) {
        float nn[16];
#pragma unroll
        for (int g = 0; g < 4; g++) {
          const float4 n4 = nns[(cb * 4 + g) * 64 + lane];
          nn[4 * g] = n4.x; nn[4 * g + 1] = n4.y; nn[4 * g + 2] = n4.z; nn[4 * g + 3] = n4.w;
        }
        if (32 * cb + 32 <= a.kc) {
          yk_best1_block<false>(acc[cb], nn, qd, 32L * cb, h, false, bv, bi);
        } else {
#pragma unroll
          for (int i = 0; i < 16; i++) {
            const int bidx = 32 * cb + 8 * (i >> 2) + 4 * h + (i & 3);
            const float v = (float)(qd + (double)nn[i]) - 2.0f * acc[cb][i];
            if (bidx < a.kc && v < bv) {
              bv = v;
              bi = bidx;
            }
          }
        }
      }
      const float ov = __shfl_xor(bv, 32);
      const int oi = __shfl_xor(bi, 32);
      if (ov < bv || (ov == bv && oi < bi)) {
        bv = ov;
        bi = oi;
      }
      if (h == 0 && pb + r < p1) {
        a.dis[off + pb + r] = bv;
        a.assign[off + pb + r] = bi;
      }
      qd = qdn;
    }
  }
}

template <int NCH, int NBLK>
static void (*yk_bassign_pick2())(BassignArgs) {
  return yk_bassign_res_kernel<NCH, NBLK>;
}

// The kernel for (nch, nblk), or null when the centroid set does not fit
// (centroid fragments of nblk x nch x 2 KB of LDS, at most 32 KB).
static void (*yk_bassign_pick(int nch, int nblk))(BassignArgs) {
  if (nblk == 1) {
    switch (nch) {
      case 1: return yk_bassign_pick2<1, 1>();
      case 2: return yk_bassign_pick2<2, 1>();
      case 3: return yk_bassign_pick2<3, 1>();
      case 4: return yk_bassign_pick2<4, 1>();
      case 5: return yk_bassign_pick2<5, 1>();
      case 6: return yk_bassign_pick2<6, 1>();
      case 7: return yk_bassign_pick2<7, 1>();
      case 8: return yk_bassign_pick2<8, 1>();
    }
  } else if (nblk == 2) {
    switch (nch) {
      case 1: return yk_bassign_pick2<1, 2>();
      case 2: return yk_bassign_pick2<2, 2>();
      case 3: return yk_bassign_pick2<3, 2>();
      case 4: return yk_bassign_pick2<4, 2>();
      case 5: return yk_bassign_pick2<5, 2>();
      case 6: return yk_bassign_pick2<6, 2>();
      case 7: return yk_bassign_pick2<7, 2>();
      case 8: return yk_bassign_pick2<8, 2>();
    }
  } else if (nblk <= 4 && nch <= 4) {
    switch (nch * 8 + nblk) {
      case 1 * 8 + 3: return yk_bassign_pick2<1, 3>();
      case 2 * 8 + 3: return yk_bassign_pick2<2, 3>();
      case 3 * 8 + 3: return yk_bassign_pick2<3, 3>();
      case 4 * 8 + 3: return yk_bassign_pick2<4, 3>();
      case 1 * 8 + 4: return yk_bassign_pick2<1, 4>();
      case 2 * 8 + 4: return yk_bassign_pick2<2, 4>();
      case 3 * 8 + 4: return yk_bassign_pick2<3, 4>();
      case 4 * 8 + 4: return yk_bassign_pick2<4, 4>();
    }
  }
  return nullptr;
}

// The batched assign over the update chunks with the centroids
// register-resident; returns 1 when the shape has no such kernel (the
// caller then uses yk_batch_assign).
int yk_batch_assign_res(int ninst, int d, int kc, int kpad, int chunk, const float *v,
                        const double *vn, const float *cent, const float *cn,
                        const long *inst_off, long nuitems, const int *uitem_inst,
                        const int *uitem_first, const int *done, int *assign, float *dis) {
  YK_TRY(yk_init());
  (void)ninst;
  const int ld = yk_ld(d);
  auto kern = yk_bassign_pick(ld >> 4, (kc + 31) / 32);
  if (!kern || nuitems <= 0 || nuitems > 0x7fffffffL) return 1;
  BassignArgs a = {};
  a.v = v;
  a.vn = vn;
  a.cent = cent;
  a.cn = cn;
  a.ld = ld;
  a.kc = kc;
  a.kpad = kpad;
  a.chunk = chunk;
  a.inst_off = inst_off;
  a.uitem_inst = uitem_inst;
  a.uitem_first = uitem_first;
  a.nuitems = (int)nuitems;
  a.per_xcd = (a.nuitems + 7) / 8;
  a.done = done;
  a.assign = assign;
  a.dis = dis;
  YK_TRY(yk_item_tickets(&a.tick));
  const int per_res = yk_resident_wgs((const void *)kern, 256) / 8;
  const dim3 grid(8 * (a.per_xcd < per_res ? a.per_xcd : per_res)), block(256);
  hipLaunchKernelGGL(kern, grid, block, 0, yk_stream(), a);
  YK_HIP(hipGetLastError());
  return YK_OK;
}

extern "C" int yk_batch_assign(long nitems, const int *item_inst, const int *item_tile,
                               const long *inst_off, int d, int kc, int kpad, const float *v,
                               const double *vn, const float *cent, const float *cn,
                               const int *done, int *assign, float *dis) {
  return yk_batch_assign_idx(nitems, item_inst, item_tile, inst_off, d, kc, kpad, v, vn, nullptr,
                             cent, cn, done, assign, dis);
}

int yk_batch_assign_idx(long nitems, const int *item_inst, const int *item_tile,
                        const long *inst_off, int d, int kc, int kpad, const float *v,
                        const double *vn, const int *qidx, const float *cent, const float *cn,
                        const int *done, int *assign, float *dis) {
  YK_TRY(yk_init());
  if (nitems <= 0) return YK_OK;
  ScanArgs a = {};
  a.q = v;
  a.qn = vn;
  a.b = cent;
  a.bn = cn;
  a.ld = yk_ld(d);
  a.type = YK_L2;
  a.brange = kpad;
  a.bstep = 32;
  a.nsplit = 1;
  a.tiles_x = 1;
  a.out_i = assign;
  a.out_d = dis;
  a.item_inst = item_inst;
  a.item_tile = item_tile;
  a.inst_off = inst_off;
  a.kc = kc;
  a.kpad = kpad;
  a.done = done;
  a.qidx = qidx;
  a.total = (int)nitems;
  a.per_xcd = (a.total + 7) / 8;
  auto kern = yk_f32_pick<0, 1>(a.ld, false);
  YK_TRY(yk_item_tickets(&a.tick));
  const int per_res = yk_resident_wgs((const void *)kern, 256) / 8;
  dim3 grid(8 * (a.per_xcd < per_res ? a.per_xcd : per_res)), block(256);
  hipLaunchKernelGGL(kern, grid, block, 0, yk_stream(), a);
  YK_HIP(hipGetLastError());
  return YK_OK;
}

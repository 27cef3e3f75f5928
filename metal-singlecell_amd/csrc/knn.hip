// Exact Euclidean kNN on gfx950: MFMA distance tiles with a fused top-L selection, an exact
// float32 re-rank of the admitted candidates, a per-row certificate, and an exact
// brute-force pass for the rows the certificate cannot clear.
//
// Replaces the reference's `neighbors._knn` (src/metalsinglecell/neighbors.py:698-724):
// `_knn_gpu` (fp16 Gram + `topk_rows`, :25-132) below 30k cells and NNDescent above
// (:677-695).  Both are approximate there; this path is exact at every size.
//
// Order key (identical to oracle/knn.py): (j != i, d2, j) with
//   d2 = float32 sum_t (x_it - x_jt)^2, t ascending, each op separately rounded.
//
// Pipeline (all on one stream):
//   1. colsum/colmean  float64 column means (deterministic two-level sum)
//   2. prep            centred rows in the MFMA operand format, norms and rounding
//                      residuals per point (float64 math)
//   3. select          per 128 queries: stream every candidate tile through LDS, MFMA
//                      acc = |c~|^2 - 2 q^.c~ (the |c~|^2 term rides in the K padding),
//                      keep each lane's L smallest (acc, j) keys in registers
//   4. finalize        merge the lane lists (and candidate splits), exact d2 for the 64
//                      admitted candidates, sort by the exact key, certify that no
//                      rejected candidate can beat the k-th, write idx/dist
//   5. second pass     rows that fail the certificate are selected again with a wide
//                      list (16 queries per wave, 8 lists of L per query) and re-certified
//   6. fallback        exact brute force over all n for rows still uncertified
//
// Representations.  q^ is the query side of the dot product, c~ the candidate side:
//   FP32 form:  q^ = c~ = float32(x - mu)
//   BF16 form:  q^ = hi(x - mu), c~ = hi + lo, where hi = bf16(x - mu) and
//               lo = bf16((x - mu) - hi); the candidate is carried to ~16 bits, so the
//               ranking error is set by the query's own rounding r_q = |q^ - (q - mu)|.
//
// Certificate: a rejected candidate c has acc(q,c) >= tau (tau = the smallest list-tail
// value over the lists that saw q).  With E bounding the float32 accumulation error of the
// MFMA dot product, |q^ - c~|^2 >= tau + |q^|^2 - E, and by the triangle inequality
// |q - c| >= sqrt(tau + |q^|^2 - E) - r_q - r_c,max with r_c = |c~ - (c - mu)|.  A row is
// certified when its k-th exact d2 is below that bound squared (less the float32 rounding
// of the exact sum).
#include <hip/hip_bf16.h>

#include <cmath>
#include <vector>

#include "common.h"
#include "msc_hip.h"
#include "prims.h"
#include "part.h"
#include "wide.h"

namespace msc {
namespace knn {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

constexpr int kThreads = 256;       // 4 waves
constexpr int kQueriesPerBlock = 128; // 4 waves x 2 N-blocks of 16 queries (the wide form: 64)
constexpr int kTile = 64;           // candidates per LDS tile = 4 M-blocks of 16
constexpr int kListStride = 64;     // keys written per query per split
constexpr int kHitCapMax = 12;      // per-lane LDS hit-buffer slots per list (at most)
constexpr int kLdsHalfCU = 81920;   // LDS bytes per block that still lets two blocks share a CU
constexpr uint64_t kEmpty = ~0ull;

// ------------------------------------------------------------------------------------------
// operand formats
// ------------------------------------------------------------------------------------------
// BF16: row of KB = 32*KS bf16: [hi_0 .. hi_{d-1}, lo_0 .. lo_{d-1}, cn_hi, cn_mid, cn_lo, 0...]
//       cn = |hi + lo|^2 split into three bf16 terms; padding rows carry cn_hi = +inf.
//       The query operand of the same row is [-2 hi, -2 hi, 1, 1, 1, 0...].
// FP32: row of KF = 8*KS floats in fragment order: logical k = 4s + g (MFMA step s, lane
//       group g) sits at position 8*(s/2) + 2g + (s%2), so a lane reads its two steps with
//       one 8-byte load.  Logical layout [x~_0 .. x~_{d-1}, cn, 0...].

__host__ __device__ inline int bf16_ks(int d) { return (2 * d + 3 + 31) / 32; }
__host__ __device__ inline int f32_ks(int d) { return (d + 1 + 7) / 8; }
__host__ __device__ inline int f32_pos(int k) {
  int s = k >> 2, g = k & 3;
  return 8 * (s >> 1) + 2 * g + (s & 1);
}

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7F800000u) == 0x7F800000u) return static_cast<uint16_t>(u >> 16);  // inf/nan
  u += 0x7FFFu + ((u >> 16) & 1u);
  return static_cast<uint16_t>(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t h) {
  return __uint_as_float(static_cast<uint32_t>(h) << 16);
}

// ------------------------------------------------------------------------------------------
// 1. column means (float64, fixed summation order)
// ------------------------------------------------------------------------------------------
__global__ void colsum_partial(const float* __restrict__ X, int64_t n, int d, int64_t rows_per,
                               double* __restrict__ partial) {
  const int64_t r0 = blockIdx.x * rows_per;
  const int64_t r1 = min(n, r0 + rows_per);
  for (int t = threadIdx.x; t < d; t += blockDim.x) {
    double s = 0.0;
#pragma unroll 8
    for (int64_t r = r0; r < r1; ++r) s += static_cast<double>(X[r * d + t]);
    partial[static_cast<int64_t>(blockIdx.x) * d + t] = s;
  }
}

__global__ void colmean_final(const double* __restrict__ partial, int nb, int d, int64_t n,
                              double* __restrict__ mean) {
  for (int t = threadIdx.x; t < d; t += blockDim.x) {
    double s = 0.0;
    for (int b = 0; b < nb; ++b) s += partial[static_cast<int64_t>(b) * d + t];
    mean[t] = s / static_cast<double>(n);
  }
}

// ------------------------------------------------------------------------------------------
// 2. operand preparation
// ------------------------------------------------------------------------------------------
// nrm2/res: candidate side |c~|^2 and |c~ - x_c| (for the global maxima);
// qn2/qres: query side |hi|^2 and |hi - x_c| (per-row certificate terms)
__global__ void prep_bf16(const float* __restrict__ X, int64_t n, int64_t n_pad, int d, int KB,
                          const double* __restrict__ mean, uint16_t* __restrict__ Xb,
                          float* __restrict__ nrm2, float* __restrict__ res, float* __restrict__ qn2,
                          float* __restrict__ qres) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= n_pad) return;
  uint16_t* row = Xb + i * KB;
  if (i >= n) {
    for (int t = 0; t < KB; ++t) row[t] = 0;
    row[2 * d] = 0x7F80u;  // +inf: padding candidates never beat a threshold
    nrm2[i] = 0.0f;
    res[i] = 0.0f;
    qn2[i] = 0.0f;
    qres[i] = 0.0f;
    return;
  }
  double cn = 0.0, r2 = 0.0, hn = 0.0, hr2 = 0.0;
#pragma unroll 8
  for (int t = 0; t < d; ++t) {
    const double xc = static_cast<double>(X[i * d + t]) - mean[t];
    const uint16_t h = f32_to_bf16_rne(static_cast<float>(xc));
    const double xh = static_cast<double>(bf16_to_f32(h));
    const uint16_t lo = f32_to_bf16_rne(static_cast<float>(xc - xh));
    const double xt = xh + static_cast<double>(bf16_to_f32(lo));
    cn += xt * xt;
    r2 += (xt - xc) * (xt - xc);
    hn += xh * xh;
    hr2 += (xh - xc) * (xh - xc);
    row[t] = h;
    row[d + t] = lo;
  }
  const uint16_t hi = f32_to_bf16_rne(static_cast<float>(cn));
  const double rem1 = cn - static_cast<double>(bf16_to_f32(hi));
  const uint16_t mid = f32_to_bf16_rne(static_cast<float>(rem1));
  const double rem2 = rem1 - static_cast<double>(bf16_to_f32(mid));
  const uint16_t lo3 = f32_to_bf16_rne(static_cast<float>(rem2));
  row[2 * d] = hi;
  row[2 * d + 1] = mid;
  row[2 * d + 2] = lo3;
  for (int t = 2 * d + 3; t < KB; ++t) row[t] = 0;
  nrm2[i] = static_cast<float>(cn);
  res[i] = static_cast<float>(sqrt(r2));
  qn2[i] = static_cast<float>(hn);
  qres[i] = static_cast<float>(sqrt(hr2));
}

__global__ void prep_f32(const float* __restrict__ X, int64_t n, int64_t n_pad, int d, int KF,
                         const double* __restrict__ mean, float* __restrict__ Xf,
                         float* __restrict__ nrm2, float* __restrict__ res) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= n_pad) return;
  float* row = Xf + i * KF;
  if (i >= n) {
    for (int t = 0; t < KF; ++t) row[t] = 0.0f;
    row[f32_pos(d)] = __uint_as_float(0x7F800000u);
    nrm2[i] = 0.0f;
    res[i] = 0.0f;
    return;
  }
  double cn = 0.0, r2 = 0.0;
#pragma unroll 8
  for (int t = 0; t < d; ++t) {
    const double xc = static_cast<double>(X[i * d + t]) - mean[t];
    const float xf = static_cast<float>(xc);
    cn += static_cast<double>(xf) * static_cast<double>(xf);
    r2 += (static_cast<double>(xf) - xc) * (static_cast<double>(xf) - xc);
    row[f32_pos(t)] = xf;
  }
  row[f32_pos(d)] = static_cast<float>(cn);
  for (int t = d + 1; t < KF; ++t) row[f32_pos(t)] = 0.0f;
  nrm2[i] = static_cast<float>(cn);
  res[i] = static_cast<float>(sqrt(r2));
}

// max over points of |x~|^2 and r (max is order independent); one block
__global__ void point_max(const float* __restrict__ nrm2, const float* __restrict__ res,
                          int64_t n, float* __restrict__ gstat) {
  float a = 0.0f, b = 0.0f;
#pragma unroll 8
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    a = fmaxf(a, nrm2[i]);
    b = fmaxf(b, res[i]);
  }
  a = wave_max_f32(a);
  b = wave_max_f32(b);
  __shared__ float sa[kThreads / kWave], sb[kThreads / kWave];
  const int w = threadIdx.x / kWave;
  if (lane_id() == 0) { sa[w] = a; sb[w] = b; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int j = 1; j < static_cast<int>(blockDim.x) / kWave; ++j) {
      a = fmaxf(a, sa[j]);
      b = fmaxf(b, sb[j]);
    }
    gstat[0] = a;
    gstat[1] = b;
  }
}

// ------------------------------------------------------------------------------------------
// 3. select: MFMA tiles + lane-private top-L
// ------------------------------------------------------------------------------------------
template <int L>
__device__ __forceinline__ void list_insert(uint64_t (&lst)[L], uint64_t key) {
#pragma unroll
  for (int i = L - 1; i > 0; --i) {
    const uint64_t a = lst[i - 1];
    const uint64_t b = lst[i];
    lst[i] = (key < a) ? a : ((key < b) ? key : b);
  }
  lst[0] = (key < lst[0]) ? key : lst[0];
}

__device__ __forceinline__ float tail_value(uint64_t key) {
  return key == kEmpty ? __uint_as_float(0x7F800000u) : ord_float(static_cast<uint32_t>(key >> 32));
}

template <bool BF16, int KS>
struct SelCfg {
  static constexpr int K = BF16 ? 32 * KS : 8 * KS;         // elements per row
  static constexpr int ROWB = BF16 ? 2 * K : 4 * K;          // bytes per row
  // LDS row stride = 8 (mod 16) dwords.  gfx950 serves 512 B of a wave's LDS read per cycle
  // over 128 banks: 32 lanes of a 16-byte read (16 rows x 2 column groups, BF16) or 64 lanes
  // of an 8-byte read (16 rows x 4 groups, FP32), i.e. 8 consecutive dwords of each of 16
  // rows.  Rows 8 * odd dwords apart start on 16 distinct multiples of 8 (mod 128), so the
  // 128 dwords fall in 128 distinct banks.  (A stride of 4 (mod 32) dwords costs 4 conflict
  // cycles per 16-byte read on gfx950.)
  static constexpr int rdw = ROWB / 4;
  static constexpr int SDW = rdw + ((8 - (rdw % 16) + 16) % 16);
  static constexpr int STRIDE = 4 * SDW;
  static_assert(STRIDE % 32 == 0 && SDW % 16 == 8, "LDS row stride");
  static constexpr int CHUNK = BF16 ? 16 : 8;                // bytes per global->LDS move
  static constexpr int NCHUNK = kTile * ROWB / CHUNK;
  static constexpr int PER_THREAD = (NCHUNK + kThreads - 1) / kThreads;
  static constexpr int NB_FRAG = BF16 ? KS : KS;             // B fragments per N-block
};

// a block-uniform value, declared uniform (kept in a scalar register)
__device__ __forceinline__ int32_t uniform_i32(int64_t v) {
  return static_cast<int32_t>(__builtin_amdgcn_readfirstlane(static_cast<uint32_t>(v)));
}

// Candidate tiles of one block: one range [c_begin, c_end) (a split of [0, n), or a sample's
// own rows), or -- with rng4 -- up to four row ranges of the partition-pruned layout
// (rng4[8 b + 2 r], rng4[8 b + 2 r + 1] = begin, end of range r of block b), walked in order as
// one stream of tiles.  Tile T starts at row MSC_TROW(T); rows of its range end at MSC_TEND(T).
// Everything is block-uniform and 32-bit (row ids and tile counts stay below 2^31), so the
// per-tile arithmetic runs on the scalar unit.
#define MSC_TILE_MAP(RNG4, BLK, CB, CE)                                                              \
  const int64_t* tm_r = (RNG4) ? (RNG4) + static_cast<int64_t>(BLK) * 8 : nullptr;                  \
  const int32_t tm_b0 = static_cast<int32_t>(tm_r ? tm_r[0] : (CB));                               \
  const int32_t tm_e0 = static_cast<int32_t>(tm_r ? tm_r[1] : (CE));                               \
  const int32_t tm_b1 = static_cast<int32_t>(tm_r ? tm_r[2] : 0), tm_e1 = static_cast<int32_t>(tm_r ? tm_r[3] : 0); \
  const int32_t tm_b2 = static_cast<int32_t>(tm_r ? tm_r[4] : 0), tm_e2 = static_cast<int32_t>(tm_r ? tm_r[5] : 0); \
  const int32_t tm_b3 = static_cast<int32_t>(tm_r ? tm_r[6] : 0), tm_e3 = static_cast<int32_t>(tm_r ? tm_r[7] : 0); \
  const int32_t tm_t1 = tm_e0 > tm_b0 ? (tm_e0 - tm_b0 + kTile - 1) / kTile : 0;                    \
  const int32_t tm_t2 = tm_t1 + (tm_e1 > tm_b1 ? (tm_e1 - tm_b1 + kTile - 1) / kTile : 0);          \
  const int32_t tm_t3 = tm_t2 + (tm_e2 > tm_b2 ? (tm_e2 - tm_b2 + kTile - 1) / kTile : 0);          \
  const int64_t tm_n = tm_t3 + (tm_e3 > tm_b3 ? (tm_e3 - tm_b3 + kTile - 1) / kTile : 0);           \
  /* row(T) = b_r + 64 (T - t_r) in range r, as increments over range 0 */                         \
  const int32_t tm_d1 = uniform_i32((tm_b1 - tm_t1 * kTile) - tm_b0);                              \
  const int32_t tm_d2 = uniform_i32((tm_b2 - tm_t2 * kTile) - (tm_b1 - tm_t1 * kTile));            \
  const int32_t tm_d3 = uniform_i32((tm_b3 - tm_t3 * kTile) - (tm_b2 - tm_t2 * kTile));            \
  const int32_t tm_f1 = uniform_i32(tm_e1 - tm_e0), tm_f2 = uniform_i32(tm_e2 - tm_e1),             \
                tm_f3 = uniform_i32(tm_e3 - tm_e2);                                                 \
  const int32_t tm_u0 = uniform_i32(tm_b0), tm_v0 = uniform_i32(tm_e0), tm_u1 = uniform_i32(tm_t1), \
                tm_u2 = uniform_i32(tm_t2), tm_u3 = uniform_i32(tm_t3);
// (sums of independent selects: a chain of selects over four values is turned into a lookup
// in a private array, i.e. scratch memory)
#define MSC_TROW(T)                                                                                 \
  (tm_u0 + static_cast<int32_t>(T) * kTile + (static_cast<int32_t>(T) >= tm_u1 ? tm_d1 : 0) +       \
   (static_cast<int32_t>(T) >= tm_u2 ? tm_d2 : 0) + (static_cast<int32_t>(T) >= tm_u3 ? tm_d3 : 0))
#define MSC_TEND(T)                                                                                 \
  (tm_v0 + (static_cast<int32_t>(T) >= tm_u1 ? tm_f1 : 0) + (static_cast<int32_t>(T) >= tm_u2 ? tm_f2 : 0) + \
   (static_cast<int32_t>(T) >= tm_u3 ? tm_f3 : 0))

// WIDE = false: each wave holds 32 queries (2 N-blocks), one list of L keys per lane and
//   query; a query's 64 candidates of every tile spread over 4 lanes (lists) -> 4L keys.
// WIDE = true: each wave holds 16 queries (1 N-block); a lane keeps two lists for its
//   query, one for M-blocks 0-1 and one for 2-3 -> 8 lists of L per query, written as two
//   virtual splits (2 * split + list), so the finalize step merges them like splits.
template <bool BF16, int KS, int L, bool WIDE>
__global__ __launch_bounds__(kThreads, 2) void select_kernel(
    const void* __restrict__ Xp, int64_t n, int64_t n_slots, int d, int64_t split_len,
    const int32_t* __restrict__ qrows, const int64_t* __restrict__ blk_range,
    const int64_t* __restrict__ rng4, uint64_t* __restrict__ keys_out, float* __restrict__ tau_out) {
  using C = SelCfg<BF16, KS>;
  constexpr int NB = WIDE ? 1 : 2;         // N-blocks (of 16 queries) per wave
  constexpr int QPB = 4 * 16 * NB;         // queries per block
  // hit-buffer slots per lane and list: as many as keep two blocks per CU (>= 8), else the max
  constexpr int kFit = (kLdsHalfCU - 2 * kTile * C::STRIDE) / ((kThreads / kWave) * 2 * kWave * 8);
  constexpr int kHitCap = (kFit >= 8) ? (kFit < kHitCapMax ? kFit : kHitCapMax) : kHitCapMax;
  __shared__ __attribute__((aligned(16))) char tile[2][kTile * C::STRIDE];
  // per-lane hit buffers: [wave][nb][slot][lane], lane-minor (conflict-free 8-byte access)
  __shared__ uint64_t hitbuf[kThreads / kWave][2][kHitCap][kWave];

  const int l = lane_id();
  const int wv = threadIdx.x / kWave;
  const int g = l >> 4;     // lane group 0..3
  const int lm = l & 15;    // row within an MFMA block
  const int split = blockIdx.y;
  // candidate range: one split of [0, n), or the block's own range (batched samples)
  // candidate range: one split of [0, n), or this block's split of its own sample (batched
  // samples: blk_range holds [begin, end) per (split, block))
  const int64_t br = 2 * (static_cast<int64_t>(split) * gridDim.x + blockIdx.x);
  const int64_t c_begin = blk_range ? blk_range[br] : static_cast<int64_t>(split) * split_len;
  const int64_t c_end = blk_range ? blk_range[br + 1] : min(n, c_begin + split_len);
  MSC_TILE_MAP(rng4, blockIdx.x, c_begin, c_end)
  // query slots: slot s is row s, or row qrows[s] when a row list is given (a padding slot, -1,
  // computes row 0's lists, which nothing reads)
  const int64_t sbase = static_cast<int64_t>(blockIdx.x) * QPB + wv * (16 * NB);
  const char* Xb = static_cast<const char*>(Xp);

  // ---- query operand fragments, resident in registers for the whole stream ----
  bf16x8 qb[NB][BF16 ? KS : 1];
  float qf[NB][BF16 ? 1 : 2 * KS];
#pragma unroll
  for (int nb = 0; nb < NB; ++nb) {
    const int64_t slot = sbase + 16 * nb + lm;
    const int64_t q = qrows ? max(static_cast<int64_t>(qrows[slot]), int64_t(0)) : slot;
    const char* qrow = Xb + q * C::ROWB;
    if constexpr (BF16) {
      // [-2 hi, -2 hi, 1, 1, 1, 0 ...] against the candidate row [hi, lo, cn_hi, cn_mid, cn_lo, 0 ...]
      const uint16_t* src = reinterpret_cast<const uint16_t*>(qrow);
#pragma unroll
      for (int s = 0; s < KS; ++s) {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int k = 32 * s + 8 * g + e;
          float v;
          if (k < d) v = -2.0f * bf16_to_f32(src[k]);
          else if (k < 2 * d) v = -2.0f * bf16_to_f32(src[k - d]);
          else if (k < 2 * d + 3) v = 1.0f;
          else v = 0.0f;
          qb[nb][s][e] = static_cast<__bf16>(v);
        }
      }
    } else {
      const float* row = reinterpret_cast<const float*>(qrow);
#pragma unroll
      for (int s = 0; s < 2 * KS; ++s) {
        const int k = 4 * s + g;
        float v;
        if (k < d) v = -2.0f * row[f32_pos(k)];
        else if (k == d) v = 1.0f;
        else v = 0.0f;
        qf[nb][s] = v;
      }
    }
  }

  uint64_t lst[2][L];
  float tau[2];  // each list's own tail (the certificate threshold)
  float tq[2];   // admission threshold: the smallest tail over the lanes (and lists) of the query
  int cnt[2] = {0, 0};
#pragma unroll
  for (int nb = 0; nb < 2; ++nb) {
#pragma unroll
    for (int i = 0; i < L; ++i) lst[nb][i] = kEmpty;
    tau[nb] = __uint_as_float(0x7F800000u);
    tq[nb] = tau[nb];
  }

  const int64_t ntiles = tm_n;

  // Prefetch registers for one tile: up to eight named 16-byte registers chosen at compile
  // time (an array here gets demoted to LDS or scratch by the compiler).  Global reads and
  // LDS stores are 16-byte chunks for both precisions (the row stride is a multiple of 32 B).
  constexpr int kRowChunks = C::ROWB / 16;
  constexpr int kNChunk = kTile * kRowChunks;
  constexpr int kPer = (kNChunk + kThreads - 1) / kThreads;
  static_assert(kPer <= 8, "tile prefetch needs more than 8 registers");
  // Two register sets (A: odd tiles, B: even tiles) so that a tile's global load has two
  // tiles of compute to complete before its LDS store.
  uint4 pa0{}, pa1{}, pa2{}, pa3{}, pa4{}, pa5{}, pa6{}, pa7{};
  uint4 pb0{}, pb1{}, pb2{}, pb3{}, pb4{}, pb5{}, pb6{}, pb7{};
#define MSC_PF(S, i) (i == 0 ? S##0 : i == 1 ? S##1 : i == 2 ? S##2 : i == 3 ? S##3 : i == 4 ? S##4 : i == 5 ? S##5 : i == 6 ? S##6 : S##7)
#define MSC_LOAD1(S, i, T)                                                                      \
  if constexpr (kPer > i) {                                                                     \
    const int c_ = threadIdx.x + i * kThreads;                                                  \
    if (kNChunk % kThreads == 0 || c_ < kNChunk)                                                \
      MSC_PF(S, i) = *reinterpret_cast<const uint4*>(Xb + static_cast<int64_t>(MSC_TROW(T)) * C::ROWB + static_cast<int64_t>(c_) * 16); \
  }
#define MSC_STORE1(S, i, B)                                                                     \
  if constexpr (kPer > i) {                                                                     \
    const int c_ = threadIdx.x + i * kThreads;                                                  \
    if (kNChunk % kThreads == 0 || c_ < kNChunk) {                                              \
      char* dst_ = &tile[(B)][(c_ / kRowChunks) * C::STRIDE + (c_ % kRowChunks) * 16];           \
      *reinterpret_cast<uint4*>(dst_) = MSC_PF(S, i);                                           \
    }                                                                                           \
  }
#define MSC_LOAD_TILE(S, T) { MSC_LOAD1(S, 0, T) MSC_LOAD1(S, 1, T) MSC_LOAD1(S, 2, T) MSC_LOAD1(S, 3, T) \
                              MSC_LOAD1(S, 4, T) MSC_LOAD1(S, 5, T) MSC_LOAD1(S, 6, T) MSC_LOAD1(S, 7, T) }
#define MSC_STORE_TILE(S, B) { MSC_STORE1(S, 0, B) MSC_STORE1(S, 1, B) MSC_STORE1(S, 2, B) MSC_STORE1(S, 3, B) \
                               MSC_STORE1(S, 4, B) MSC_STORE1(S, 5, B) MSC_STORE1(S, 6, B) MSC_STORE1(S, 7, B) }


  // BF16 with up to 4 K-steps: the next M-block's fragments are read from LDS while the
  // current M-block's MFMAs run (two fragment sets, 8 * KS VGPRs), so the LDS latency hides
  // behind 2 * KS MFMAs instead of two.  A tile's first M-block is read right after the barrier
  // that publishes it, before the previous tile's selection.  Same operands in the same order
  // as the plain loop below.
  constexpr bool kAhead = BF16 && KS <= 4;
  bf16x8 af0[kAhead ? KS : 1];
  auto read_first = [&](const int buf) __attribute__((always_inline)) {
    if constexpr (kAhead) {
#pragma unroll
      for (int s = 0; s < KS; ++s)
        af0[s] = *reinterpret_cast<const bf16x8*>(&tile[buf][lm * C::STRIDE] + 64 * s + 16 * g);
    }
  };

  // one candidate tile's MFMA distance block from LDS buffer buf
  auto tile_mfma = [&](const int buf, f32x4 (&acc)[4][NB]) __attribute__((always_inline)) {
    if constexpr (kAhead) {
      bf16x8 af[2][KS];
#pragma unroll
      for (int s = 0; s < KS; ++s) af[0][s] = af0[s];
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) {
        if (mb + 1 < 4) {
          const char* nrow = &tile[buf][(16 * (mb + 1) + lm) * C::STRIDE];
#pragma unroll
          for (int s = 0; s < KS; ++s)
            af[(mb + 1) & 1][s] = *reinterpret_cast<const bf16x8*>(nrow + 64 * s + 16 * g);
        }
        // keep the next M-block's reads issued ahead of this M-block's MFMAs (the scheduler
        // would otherwise sink each read next to its first use)
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s = 0; s < KS; ++s)
#pragma unroll
          for (int nb = 0; nb < NB; ++nb)
            acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[mb & 1][s], qb[nb][s], acc[mb][nb], 0, 0, 0);
      }
    }
#pragma unroll
    for (int mb = 0; mb < (kAhead ? 0 : 4); ++mb) {
#pragma unroll
      for (int nb = 0; nb < NB; ++nb) acc[mb][nb] = f32x4{0.f, 0.f, 0.f, 0.f};
      const char* arow = &tile[buf][(16 * mb + lm) * C::STRIDE];
      if constexpr (BF16) {
#pragma unroll
        for (int s = 0; s < KS; ++s) {
          const bf16x8 a = *reinterpret_cast<const bf16x8*>(arow + 64 * s + 16 * g);
#pragma unroll
          for (int nb = 0; nb < NB; ++nb)
            acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, qb[nb][s], acc[mb][nb], 0, 0, 0);
        }
      } else {
#pragma unroll
        for (int h = 0; h < KS; ++h) {
          const float2 a2 = *reinterpret_cast<const float2*>(arow + 32 * h + 8 * g);
#pragma unroll
          for (int nb = 0; nb < NB; ++nb) {
            acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a2.x, qf[nb][2 * h], acc[mb][nb], 0, 0, 0);
            acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a2.y, qf[nb][2 * h + 1], acc[mb][nb], 0, 0, 0);
          }
        }
      }
    }

  };

  // the fused selection over one tile's distance block
  auto tile_select = [&](const f32x4 (&acc)[4][NB], const int64_t t) __attribute__((always_inline)) {
    // Hits (acc <= the lane's list tail) are appended to the lane's LDS buffer; buffers are
    // merged into the sorted register lists when any lane's buffer could overflow on the
    // next M-block, so a rare hit costs one LDS store instead of a wave-wide insertion.
    // Each (M-block, N-block) group is gated by a wave-uniform test: after the first few
    // hundred tiles almost every group has no hit and costs four compares.
    const int32_t c0 = MSC_TROW(t);
    const int32_t cend = MSC_TEND(t);
    const bool tail_tile = c0 + kTile > cend;  // the tile's range ends inside it
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) {
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) {
        const int li = WIDE ? (mb >> 1) : nb;  // list receiving this group
        const f32x4 a4 = acc[mb][nb];
        const bool h = fminf(fminf(a4[0], a4[1]), fminf(a4[2], a4[3])) <= tq[li];
        if (__builtin_amdgcn_ballot_w64(h) == 0ull) continue;
        const uint32_t cg = static_cast<uint32_t>(c0) + 16 * mb + 4 * g;  // candidate ids < 2^31
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const float v = a4[i];
          const uint32_t c = cg + i;
          if (v <= tq[li] && (!tail_tile || c < static_cast<uint32_t>(cend))) {
            hitbuf[wv][li][cnt[li]][l] = (static_cast<uint64_t>(float_ord(v)) << 32) | c;
            ++cnt[li];
          }
        }
        // an M-block adds at most 4 hits per lane: flush before a buffer could overflow
        if (__builtin_amdgcn_ballot_w64(cnt[li] > kHitCap - 4) != 0ull) {
          for (int j = 0; j < cnt[li]; ++j) {
            const uint64_t key = hitbuf[wv][li][j][l];
            if (key < lst[li][L - 1]) list_insert<L>(lst[li], key);
          }
          cnt[li] = 0;
          tau[li] = tail_value(lst[li][L - 1]);
          // A candidate above the smallest tail of its query's lists can never be needed:
          // the certificate bounds every rejected candidate by that smallest tail.
          float t = WIDE ? fminf(tau[0], tau[1]) : tau[li];
          t = fminf(t, __shfl_xor(t, 16, kWave));
          t = fminf(t, __shfl_xor(t, 32, kWave));
          if (WIDE) {
            tq[0] = t;
            tq[1] = t;
          } else {
            tq[li] = t;
          }
        }
      }
    }

  };

  // Tiles in pairs: tile t's MFMAs read buffer 0, then tile t+1 (set A, loaded two tiles
  // earlier) is stored to buffer 1, whose last reader (tile t-1's MFMAs) finished before the
  // previous barrier; then the same with the roles swapped.  One barrier per tile publishes the
  // stored tile; tile t's selection runs after it, behind the first reads of tile t+1, so the
  // MFMA drain overlaps the barrier and the LDS latency overlaps the selection.  The compiler
  // fences keep each store after its tile's MFMAs and each read-ahead before the selection, so
  // the global load latency hides behind two tiles of work.
  // Every load and store is unconditional: a tile index past the end is clamped to the last
  // tile (already read in full, so the repeat stays in bounds) and its store goes to the buffer
  // that no later tile reads.  Conditional loads would make the compiler's wait counts assume
  // the worst at every join, so a tile's store would also wait for the next tile's loads.
#define MSC_CLAMP(T) ((T) < ntiles ? (T) : ntiles - 1)
  if (ntiles > 0) {  // block-uniform
    MSC_LOAD_TILE(pb, 0);
    MSC_STORE_TILE(pb, 0);
    __syncthreads();
    MSC_LOAD_TILE(pa, MSC_CLAMP(1));
    asm volatile("" ::: "memory");  // set A's loads before set B's, the order the loop keeps
    MSC_LOAD_TILE(pb, MSC_CLAMP(2));
    read_first(0);
    f32x4 acc[4][NB];
    // The pair loop has its only exit at the bottom (an exit between the two halves would give
    // the loop header a path on which set A's loads are the newest, and the wait counts would
    // follow it).  An odd last tile is in buffer 0 with its first fragments read.
    int64_t t = 0;
    for (; t + 1 < ntiles; t += 2) {
      tile_mfma(0, acc);
      asm volatile("" ::: "memory");
      MSC_STORE_TILE(pa, 1);
      __syncthreads();
      read_first(1);
      MSC_LOAD_TILE(pa, MSC_CLAMP(t + 3));
      asm volatile("" ::: "memory");
      tile_select(acc, t);
      tile_mfma(1, acc);
      asm volatile("" ::: "memory");
      MSC_STORE_TILE(pb, 0);
      __syncthreads();
      read_first(0);
      MSC_LOAD_TILE(pb, MSC_CLAMP(t + 4));
      asm volatile("" ::: "memory");
      tile_select(acc, t + 1);
    }
    if (t < ntiles) {
      tile_mfma(0, acc);
      tile_select(acc, t);
    }
  }
#undef MSC_CLAMP
#undef MSC_LOAD_TILE
#undef MSC_STORE_TILE
#undef MSC_LOAD1
#undef MSC_STORE1
#undef MSC_PF

  // ---- drain the hit buffers, write the lane lists and the per-query threshold ----
#pragma unroll
  for (int nb = 0; nb < 2; ++nb) {
    for (int j = 0; j < cnt[nb]; ++j) {
      const uint64_t key = hitbuf[wv][nb][j][l];
      if (key < lst[nb][L - 1]) list_insert<L>(lst[nb], key);
    }
    tau[nb] = tail_value(lst[nb][L - 1]);
  }
#pragma unroll
  for (int li = 0; li < 2; ++li) {
    const int64_t slot = sbase + (WIDE ? 0 : 16 * li) + lm;
    const int64_t vsplit = WIDE ? 2 * static_cast<int64_t>(split) + li : split;
    uint64_t* out = keys_out + (vsplit * n_slots + slot) * kListStride;
#pragma unroll
    for (int i = 0; i < L; ++i) out[g * L + i] = lst[li][i];
    if (4 * L < kListStride) {
      for (int i = 4 * L + g; i < kListStride; i += 4) out[i] = kEmpty;
    }
    float tv = tau[li];
    tv = fminf(tv, __shfl_xor(tv, 16, kWave));
    tv = fminf(tv, __shfl_xor(tv, 32, kWave));
    if (g == 0) tau_out[vsplit * n_slots + slot] = tv;
  }
}

// ------------------------------------------------------------------------------------------
// 3b. select, BF16 main pass on v_mfma_f32_32x32x16_bf16
// ------------------------------------------------------------------------------------------
// Each wave holds 32 queries as ONE N-block of the 32x32x16 MFMA and streams 64-candidate
// tiles as two M-blocks of 32.  K runs in steps of 16, so d = 50 (2d + 3 = 103 operand
// terms) takes 7 steps (K = 112) instead of the 4 x 32 = 128 of the 16x16x32 form.
//   A (candidates, LDS): lane l reads row (l mod 32) of the M-block, 8 terms at 16 s + 8 (l div 32)
//   B (queries, registers): lane l holds query (l mod 32), the same 8 terms
//   C: lane l holds query (l mod 32) against rows 8 (i div 4) + 4 (l div 32) + (i mod 4), i < 16
// Two lanes share a query; each keeps a sorted register list of its L = 32 smallest keys, so 64
// candidates are admitted per query, as in the 16x16x32 form, and a query's admission threshold
// is the smaller of its two lanes' tails (about its 60th candidate, against about the 40th
// with four lanes of 16: far fewer rows fail the certificate).  Hits go to a per-lane LDS
// buffer merged into the list when a lane's buffer could overflow.  The M-blocks are
// software pipelined: the selection of one M-block runs while the next one's MFMAs execute.
//
// Tiles are copied global -> LDS by the load unit itself (global_load_lds_dwordx4: a wave's
// 64 lanes fill 1 KB of LDS), three tile buffers in flight, no staging registers.  A tile is
// SC = CH | 1 wave loads: row r of the tile sits at r * 16 SC bytes, its 16-byte chunk c at
// + 16 c; the one pad chunk of a row (SC > CH) loads a copy of the row's last chunk.  The odd
// row stride in 16-byte groups puts the 32 rows of a half-wave's fragment read on 32 distinct
// bank groups.

// one 16-byte global -> LDS copy per lane through the load unit (LDS address: the wave-uniform
// base + 16 x lane); the builtin exists only in device compilation
__device__ __forceinline__ void lds_load16(const void* src, void* lds_base) {
#if defined(__HIP_DEVICE_COMPILE__)
  __builtin_amdgcn_global_load_lds(src, (__attribute__((address_space(3))) void*)lds_base, 16, 0, 0);
#endif
}

// block shape of the 32x32x16 form: kSelWaves waves of 32 queries share every tile
constexpr int kSelWaves = 4;
constexpr int kSelThreads = kSelWaves * kWave;
constexpr int kSelQueries = 32 * kSelWaves;
constexpr int kSelBlocksPerCU = 8 / kSelWaves;           // two waves per SIMD
constexpr int kSelLds = 163840 / kSelBlocksPerCU;          // LDS per block (160 KB per CU)

template <int KT>
struct Sel32Cfg {
  static constexpr int CH = 2 * KT;     // 16-byte chunks of a row used (K = 16 KT bf16)
  static constexpr int SC = CH | 1;     // chunks per LDS row: an odd number of 16 B groups
  static constexpr int STRIDE = 16 * SC;
  static constexpr int TB = kTile * STRIDE;  // = 1 KB x SC
  static constexpr int NI = SC;              // wave loads per tile
  static constexpr int PERW = (NI + kSelWaves - 1) / kSelWaves;  // per wave (the last one repeated)
  static constexpr int FIT = (kSelLds - 3 * TB) / (kSelWaves * kWave * 8);
  static constexpr int CAP = FIT < 24 ? FIT : 24;
  static_assert(CAP >= 12, "hit buffer");
};

template <int KT>
__global__ __launch_bounds__(kSelThreads, kSelBlocksPerCU) void select32_kernel(
    const void* __restrict__ Xp, int64_t n, int64_t n_slots, int d, int rowb, int64_t split_len,
    const int32_t* __restrict__ qrows, const int64_t* __restrict__ blk_range, const int64_t* __restrict__ rng4,
    uint64_t* __restrict__ keys_out, float* __restrict__ tau_out) {
  using C = Sel32Cfg<KT>;
  typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;
  constexpr int L = 32;
  __shared__ __attribute__((aligned(1024))) char tA[C::TB];
  __shared__ __attribute__((aligned(1024))) char tB[C::TB];
  __shared__ __attribute__((aligned(1024))) char tC[C::TB];
  __shared__ uint64_t hitbuf[kSelWaves][C::CAP][kWave];

  const int l = lane_id();
  const int wv = threadIdx.x / kWave;
  const int h = l >> 5;
  const int lq = l & 31;
  const int split = blockIdx.y;
  const int64_t br = 2 * (static_cast<int64_t>(split) * gridDim.x + blockIdx.x);
  const int64_t c_begin = blk_range ? blk_range[br] : static_cast<int64_t>(split) * split_len;
  const int64_t c_end = blk_range ? blk_range[br + 1] : min(n, c_begin + split_len);
  MSC_TILE_MAP(rng4, blockIdx.x, c_begin, c_end)
  const int64_t slot = static_cast<int64_t>(blockIdx.x) * kSelQueries + wv * 32 + lq;
  // slot s is row s, or row qrows[s] (-1: a padding slot); a lane without a query only helps
  // load the tiles
  const int64_t qrow = slot < n_slots ? (qrows ? static_cast<int64_t>(qrows[slot]) : slot) : -1;
  const bool own = qrow >= 0;
  const char* Xb = static_cast<const char*>(Xp);

  bf16x8 qb[KT];
  {
    const uint16_t* src = reinterpret_cast<const uint16_t*>(Xb + (own ? qrow : 0) * rowb);
#pragma unroll
    for (int s = 0; s < KT; ++s) {
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int k = 16 * s + 8 * h + e;
        float v;
        if (k < d) v = -2.0f * bf16_to_f32(src[k]);
        else if (k < 2 * d) v = -2.0f * bf16_to_f32(src[k - d]);
        else if (k < 2 * d + 3) v = 1.0f;
        else v = 0.0f;
        qb[s][e] = static_cast<__bf16>(v);
      }
    }
  }

  uint64_t lst[L];
#pragma unroll
  for (int i = 0; i < L; ++i) lst[i] = kEmpty;
  float tq = __uint_as_float(0x7F800000u);
  int cnt = 0;
  const int64_t ntiles = tm_n;
  // With partition-pruned ranges the walk starts at the tile that holds the block's first query
  // (its own partition, where its nearest candidates lie) and wraps around, so the admission
  // threshold tightens early and fewer M-blocks take the hit path.  The tiles scanned and each
  // candidate's place in its tile (hence its lane) do not change.
  int32_t tstart = 0;
  if (rng4 && qrows) {
    const int32_t q0 = qrows[static_cast<int64_t>(blockIdx.x) * kSelQueries];
    if (q0 >= tm_b0 && q0 < tm_e0) tstart = (q0 - tm_b0) / kTile;
    else if (q0 >= tm_b1 && q0 < tm_e1) tstart = tm_t1 + (q0 - tm_b1) / kTile;
    else if (q0 >= tm_b2 && q0 < tm_e2) tstart = tm_t2 + (q0 - tm_b2) / kTile;
    else if (q0 >= tm_b3 && q0 < tm_e3) tstart = tm_t3 + (q0 - tm_b3) / kTile;
  }
  tstart = uniform_i32(tstart);
  // the tile of the walk's step T (0 <= T < ntiles)
  auto rot = [&](const int64_t T) __attribute__((always_inline)) -> int32_t {
    const int32_t r = static_cast<int32_t>(T) + tstart;
    return r >= static_cast<int32_t>(ntiles) ? r - static_cast<int32_t>(ntiles) : r;
  };

  // this wave's loads of a tile: wave load u = wv + kSelWaves j (clamped), lane l fills LDS chunk
  // 64 u + l = (row, chunk) of the tile
  int lds_off[C::PERW], src_off[C::PERW];
#pragma unroll
  for (int j = 0; j < C::PERW; ++j) {
    const int u = min(wv + kSelWaves * j, C::NI - 1);
    const int pch = 64 * u + l;
    const int row = pch / C::SC, col = pch % C::SC;
    lds_off[j] = 1024 * u;
    src_off[j] = row * rowb + min(col, C::CH - 1) * 16;
  }
  auto load_tile = [&](char* buf, const int64_t T) __attribute__((always_inline)) {
    const int32_t tr = rot(T < ntiles ? T : ntiles - 1);
    const char* base = Xb + static_cast<int64_t>(MSC_TROW(tr)) * rowb;
#pragma unroll
    for (int j = 0; j < C::PERW; ++j)
      lds_load16(base + src_off[j], buf + lds_off[j]);
  };
  // wait until at most `later` of this wave's tile loads are outstanding (vmcnt; expcnt and
  // lgkmcnt left at their maxima)
  constexpr int kWaitNext = (C::PERW & 15) | (7 << 4) | (15 << 8) | (((C::PERW >> 4) & 3) << 14);

  bf16x8 af[2][KT];
  auto read_mb = [&](const char* buf, const int mb, bf16x8 (&a)[KT]) __attribute__((always_inline)) {
    const char* row = buf + (32 * mb + lq) * C::STRIDE + 16 * h;
#pragma unroll
    for (int s = 0; s < KT; ++s) a[s] = *reinterpret_cast<const bf16x8*>(row + 32 * s);
  };
  auto mfma_mb = [&](const bf16x8 (&a)[KT], f32x16& acc) __attribute__((always_inline)) {
    acc = f32x16{};
#pragma unroll
    for (int s = 0; s < KT; ++s) acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[s], qb[s], acc, 0, 0, 0);
  };
  auto flush = [&]() __attribute__((always_inline)) {
    for (int j = 0; j < cnt; ++j) {
      const uint64_t key = hitbuf[wv][j][l];
      if (key < lst[L - 1]) list_insert<L>(lst, key);
    }
    cnt = 0;
    const float t = tail_value(lst[L - 1]);
    tq = fminf(t, __shfl_xor(t, 32, kWave));
  };
  // gate + hits of one M-block: element i = 4 g + j holds candidate c0 + 8 g + 4 h + j; a
  // wave-uniform test per M-block, then per group of four
  auto select_mb = [&](const f32x16& acc, const int32_t c0, const int32_t cend) __attribute__((always_inline)) {
    float gm[4];
#pragma unroll
    for (int g = 0; g < 4; ++g) gm[g] = fminf(fminf(acc[4 * g], acc[4 * g + 1]), fminf(acc[4 * g + 2], acc[4 * g + 3]));
    if (__builtin_amdgcn_ballot_w64(fminf(fminf(gm[0], gm[1]), fminf(gm[2], gm[3])) <= tq) == 0ull) return;
    const uint32_t cg = static_cast<uint32_t>(c0) + 4 * h;  // candidate ids < 2^31
#pragma unroll
    for (int g = 0; g < 4; ++g) {
      if (__builtin_amdgcn_ballot_w64(gm[g] <= tq) == 0ull) continue;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const float v = acc[4 * g + j];
        const uint32_t c = cg + 8 * g + j;
        if (v <= tq && c < static_cast<uint32_t>(cend)) {  // rows past the range end are not candidates
          hitbuf[wv][cnt][l] = (static_cast<uint64_t>(float_ord(v)) << 32) | c;
          ++cnt;
        }
      }
      if (__builtin_amdgcn_ballot_w64(cnt > C::CAP - 4) != 0ull) flush();
    }
  };

  if (ntiles > 0) {  // block-uniform
    load_tile(tA, 0);
    load_tile(tB, 1);
    __builtin_amdgcn_s_waitcnt(kWaitNext);
    __syncthreads();
    read_mb(tA, 0, af[0]);
    f32x16 acc0, acc1;
    bool have1 = false;
    // tile t in X, t+1 in Y (loading), Z free: start t+2 into Z; M-block 1 of t read, M-block 0
    // multiplied, the previous M-block 1 selected, M-block 1 multiplied; wait for t+1, barrier,
    // M-block 0 of t+1 read; M-block 0 of t selected
    auto step = [&](char* X, char* Y, char* Z, const int64_t t) __attribute__((always_inline)) {
      load_tile(Z, t + 2);
      read_mb(X, 1, af[1]);
      mfma_mb(af[0], acc0);
      if (have1) {
        const int32_t tp = rot(t - 1);
        select_mb(acc1, MSC_TROW(tp) + 32, MSC_TEND(tp));
      }
      mfma_mb(af[1], acc1);
      // this wave's loads of tile t+1 are done; the barrier publishes everyone's.  A plain
      // barrier: __syncthreads' release fence would also wait for tile t+2's loads.
      __builtin_amdgcn_s_waitcnt(kWaitNext);
      asm volatile("" ::: "memory");
      __builtin_amdgcn_s_barrier();
      asm volatile("" ::: "memory");
      read_mb(Y, 0, af[0]);
      const int32_t tc = rot(t);
      select_mb(acc0, MSC_TROW(tc), MSC_TEND(tc));
      have1 = true;
    };
    int64_t t = 0;
    for (; t + 3 <= ntiles; t += 3) {
      step(tA, tB, tC, t);
      step(tB, tC, tA, t + 1);
      step(tC, tA, tB, t + 2);
    }
    if (t < ntiles) step(tA, tB, tC, t);
    if (t + 1 < ntiles) step(tB, tC, tA, t + 1);
    const int32_t tl = rot(ntiles - 1);
    select_mb(acc1, MSC_TROW(tl) + 32, MSC_TEND(tl));
  }

  flush();
  if (!own) return;
  uint64_t* out = keys_out + (static_cast<int64_t>(split) * n_slots + slot) * kListStride;
#pragma unroll
  for (int i = 0; i < L; ++i) out[h * L + i] = lst[i];
  if (h == 0) tau_out[static_cast<int64_t>(split) * n_slots + slot] = tq;
}

// ------------------------------------------------------------------------------------------
// 4./5. exact re-rank, certificate, fallback
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ float exact_d2(const float* __restrict__ X, int d, int64_t i, int64_t j) {
  const float* a = X + i * d;
  const float* b = X + j * d;
  float acc = 0.0f;
  if ((d & 1) == 0 && (reinterpret_cast<uintptr_t>(X) & 7) == 0) {  // 8-byte aligned rows: pairwise
                                                                      // loads, same summation order
    const float2* a2 = reinterpret_cast<const float2*>(a);
    const float2* b2 = reinterpret_cast<const float2*>(b);
    // unrolled so that several iterations' loads are in flight before the (unchanged,
    // sequential) chain of adds consumes them
#pragma unroll 8
    for (int t = 0; t < d / 2; ++t) {
      const float2 x = a2[t], y = b2[t];
      const float d0 = __fsub_rn(x.x, y.x);
      acc = __fadd_rn(acc, __fmul_rn(d0, d0));
      const float d1 = __fsub_rn(x.y, y.y);
      acc = __fadd_rn(acc, __fmul_rn(d1, d1));
    }
    return acc;
  }
#pragma unroll 8
  for (int t = 0; t < d; ++t) {
    const float diff = __fsub_rn(a[t], b[t]);
    acc = __fadd_rn(acc, __fmul_rn(diff, diff));
  }
  return acc;
}

// exact_d2 of row i against rows j[0..3], each with exact_d2's loads and summation order (so
// bit-identical), the row-i loads shared and the four rows' loads in flight together
__device__ __forceinline__ void exact_d2_x4(const float* __restrict__ X, int d, int64_t i, const int64_t (&j)[4],
                                            float (&out)[4]) {
  const float* a = X + i * d;
  float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  if ((d & 1) == 0 && (reinterpret_cast<uintptr_t>(X) & 7) == 0) {
    const float2* a2 = reinterpret_cast<const float2*>(a);
    const float2* b2[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) b2[u] = reinterpret_cast<const float2*>(X + j[u] * d);
    for (int t = 0; t < d / 2; ++t) {
      const float2 x = a2[t];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float2 y = b2[u][t];
        const float d0 = __fsub_rn(x.x, y.x);
        acc[u] = __fadd_rn(acc[u], __fmul_rn(d0, d0));
        const float d1 = __fsub_rn(x.y, y.y);
        acc[u] = __fadd_rn(acc[u], __fmul_rn(d1, d1));
      }
    }
  } else {
    for (int t = 0; t < d; ++t) {
      const float x = a[t];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float diff = __fsub_rn(x, X[j[u] * d + t]);
        acc[u] = __fadd_rn(acc[u], __fmul_rn(diff, diff));
      }
    }
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) out[u] = acc[u];
}

// exact key: self -> 0, others ((d2 bits << 32) | j) + 1, invalid -> kEmpty
__device__ __forceinline__ uint64_t exact_key(int64_t q, int64_t j, float d2) {
  if (j == q) return 0ull;
  return ((static_cast<uint64_t>(__float_as_uint(d2)) << 32) | static_cast<uint64_t>(j)) + 1ull;
}
__device__ __forceinline__ void decode_exact(uint64_t key, int64_t q, int64_t* j, float* d2) {
  if (key == 0ull) { *j = q; *d2 = 0.0f; return; }
  key -= 1ull;
  *j = static_cast<int64_t>(key & 0xFFFFFFFFull);
  *d2 = __uint_as_float(static_cast<uint32_t>(key >> 32));
}

struct Bound {
  double gam;       // relative error bound of the float32 MFMA accumulation
  double split_rel; // residual of the |c~|^2 split, relative to |c~|^2
};

// The certificate above: true when every candidate that the list (tail tau) did not keep is
// provably farther than the list's k-th exact distance (key kth).
__device__ __forceinline__ bool certified(float tau, uint64_t kth, int64_t q, int d, const float* __restrict__ nrm2,
                                          const float* __restrict__ res, const float* __restrict__ gstat, Bound bnd) {
  int64_t jk;
  float d2k;
  decode_exact(kth, q, &jk, &d2k);
  const double nq2 = static_cast<double>(nrm2[q]);
  const double nmax2 = static_cast<double>(gstat[0]);
  const double nq = sqrt(nq2), nmax = sqrt(nmax2);
  // sum |terms| <= |c~|^2 + 2 |q^| (|hi_c| + |lo_c|) <= nmax^2 + 2.02 |q^| nmax
  const double E = bnd.gam * (2.02 * nq * nmax + nmax2) + bnd.split_rel * nmax2 + 4.0e-7 * (nq2 + nmax2);
  const double lb2 = static_cast<double>(tau) + nq2 - E;
  const double lb = (lb2 > 0.0 ? sqrt(lb2) : 0.0) -
                    (static_cast<double>(res[q]) + static_cast<double>(gstat[1])) * (1.0 + 1e-6) - 1e-12 * nmax;
  const double slack = 1.0 - 8.0 * (d + 2) * 5.9604644775390625e-08;
  return lb > 0.0 && static_cast<double>(d2k) < lb * lb * slack;
}

__global__ __launch_bounds__(kThreads) void finalize_kernel(
    const float* __restrict__ X, int64_t nq, int d, int k, int S, int64_t n_slots,
    const int32_t* __restrict__ qrows, const uint64_t* __restrict__ keys, const float* __restrict__ taus,
    const float* __restrict__ nrm2, const float* __restrict__ res, const float* __restrict__ gstat,
    Bound bnd, const int64_t* __restrict__ row_range, const int32_t* __restrict__ perm,
    int32_t* __restrict__ out_idx, float* __restrict__ out_dist, int32_t* __restrict__ flags) {
  // Rows of X (and of keys, flags) are in the order the select used; with perm, row r of X is
  // row perm[r] of the caller's data: exact keys and outputs use the caller's row ids.
  const int l = lane_id();
  const int64_t slot = static_cast<int64_t>(blockIdx.x) * (kThreads / kWave) + threadIdx.x / kWave;
  if (slot >= nq) return;
  const int64_t q = qrows ? static_cast<int64_t>(qrows[slot]) : slot;
  if (q < 0) return;  // padding slot of a query-block layout
  const int64_t qo = perm ? static_cast<int64_t>(perm[q]) : q;
  if (row_range && row_range[2 * q + 1] == 0) {  // padding row of a batched layout
    if (l == 0) flags[q] = 0;
    return;
  }

  uint64_t cur = wave_sort_u64(keys[slot * kListStride + l]);
  float tau = taus[slot];
  for (int s = 1; s < S; ++s) {
    const uint64_t nk = wave_sort_u64(keys[(static_cast<int64_t>(s) * n_slots + slot) * kListStride + l]);
    cur = wave_merge_min_u64(cur, nk);
    tau = fminf(tau, taus[static_cast<int64_t>(s) * n_slots + slot]);
  }
  if (S > 1) tau = fminf(tau, tail_value(shfl_u64(cur, kWave - 1)));

  const bool valid = cur != kEmpty;
  const int64_t j = static_cast<int64_t>(cur & 0xFFFFFFFFull);
  uint64_t ek = kEmpty;
  if (valid) ek = exact_key(qo, perm ? static_cast<int64_t>(perm[j]) : j, exact_d2(X, d, q, j));
  ek = wave_sort_u64(ek);

  const uint64_t kth = shfl_u64(ek, k - 1);
  const bool self_first = shfl_u64(ek, 0) == 0ull;
  bool ok = self_first && kth != kEmpty;
  if (ok && !isinf(tau)) ok = certified(tau, kth, q, d, nrm2, res, gstat, bnd);
  if (ok) {
    if (l < k) {
      int64_t jj;
      float d2;
      decode_exact(ek, qo, &jj, &d2);
      out_idx[qo * k + l] = static_cast<int32_t>(jj);
      out_dist[qo * k + l] = sqrtf(d2);
    }
    if (l == 0) flags[q] = 0;
  } else if (l == 0) {
    flags[q] = 1;
  }
}

// The 64 smallest exact keys of row q over candidates [cb, ce), ascending in lane order (one
// wave): chunks of 64 candidates merged in candidate order, four chunks' rows loaded at once.
__device__ __forceinline__ uint64_t brute_top64(const float* __restrict__ X, int d, int64_t q, int64_t cb,
                                                int64_t ce, const int32_t* __restrict__ perm = nullptr) {
  const int64_t qo = perm ? static_cast<int64_t>(perm[q]) : q;
  const int l = lane_id();
  uint64_t cur = kEmpty, tail = kEmpty;
  for (int64_t base = cb; base < ce; base += 4 * kWave) {
    int64_t j[4];
    bool v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t jj = base + u * kWave + l;
      v[u] = jj < ce;
      j[u] = v[u] ? jj : cb;  // an in-range row for the load; its key is dropped
    }
    float d2[4];
    exact_d2_x4(X, d, q, j, d2);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint64_t ek = v[u] ? exact_key(qo, perm ? static_cast<int64_t>(perm[j[u]]) : j[u], d2[u]) : kEmpty;
      if (__any(ek < tail)) {
        cur = wave_merge_min_u64(cur, wave_sort_u64(ek));
        tail = shfl_u64(cur, kWave - 1);
      }
    }
  }
  return cur;
}

__global__ __launch_bounds__(kThreads) void fallback_kernel(
    const float* __restrict__ X, int64_t n, int d, int k, const int32_t* __restrict__ flags,
    const int64_t* __restrict__ row_range, const int32_t* __restrict__ perm, int32_t* __restrict__ out_idx,
    float* __restrict__ out_dist) {
  const int l = lane_id();
  const int64_t q = static_cast<int64_t>(blockIdx.x) * (kThreads / kWave) + threadIdx.x / kWave;
  if (q >= n || flags[q] == 0) return;
  const int64_t cb = row_range ? row_range[2 * q] : 0;
  const int64_t ce = row_range ? row_range[2 * q + 1] : n;
  const uint64_t cur = brute_top64(X, d, q, cb, ce, perm);
  const int64_t qo = perm ? static_cast<int64_t>(perm[q]) : q;
  if (l < k) {
    int64_t jj;
    float d2;
    decode_exact(cur, qo, &jj, &d2);
    out_idx[qo * k + l] = static_cast<int32_t>(jj);
    out_dist[qo * k + l] = sqrtf(d2);
  }
}

// Fallback for the few rows no pass certifies, split across waves: wave (r, s) brute-forces
// chunk s of row rows[r]'s candidates into a sorted list of 64 exact keys; the merge kernel then
// keeps the 64 smallest over the chunks (lists sorted ascending, merged in chunk order: the
// same keys as one scan of all candidates, so the same result).
__global__ __launch_bounds__(kThreads) void fallback_chunk_kernel(
    const float* __restrict__ X, int64_t n, int d, const int32_t* __restrict__ rows, int64_t m, int S,
    int64_t chunk, const int32_t* __restrict__ perm, uint64_t* __restrict__ lists) {
  const int l = lane_id();
  const int64_t w = static_cast<int64_t>(blockIdx.x) * (kThreads / kWave) + threadIdx.x / kWave;
  if (w >= m * S) return;
  const int64_t r = w / S;
  const int s = static_cast<int>(w % S);
  const int64_t cb = static_cast<int64_t>(s) * chunk;
  const int64_t ce = min(n, cb + chunk);
  lists[w * kWave + l] = cb < ce ? brute_top64(X, d, rows[r], cb, ce, perm) : kEmpty;
}

__global__ __launch_bounds__(kThreads) void fallback_merge_kernel(
    const int32_t* __restrict__ rows, int64_t m, int S, int k, const int32_t* __restrict__ perm,
    const uint64_t* __restrict__ lists, int32_t* __restrict__ out_idx, float* __restrict__ out_dist) {
  const int l = lane_id();
  const int64_t r = static_cast<int64_t>(blockIdx.x) * (kThreads / kWave) + threadIdx.x / kWave;
  if (r >= m) return;
  uint64_t cur = lists[r * S * kWave + l];
  for (int s = 1; s < S; ++s) cur = wave_merge_min_u64(cur, lists[(r * S + s) * kWave + l]);
  const int64_t q = rows[r];
  const int64_t qo = perm ? static_cast<int64_t>(perm[q]) : q;
  if (l < k) {
    int64_t jj;
    float d2;
    decode_exact(cur, qo, &jj, &d2);
    out_idx[qo * k + l] = static_cast<int32_t>(jj);
    out_dist[qo * k + l] = sqrtf(d2);
  }
}

// distances CSR: row i holds the k-1 non-self neighbours sorted by column (canonical CSR,
// explicit zeros kept, as scipy's coo->csr in neighbors.py:755-760)
__global__ __launch_bounds__(kThreads) void distances_csr_kernel(
    const int32_t* __restrict__ idx, const float* __restrict__ dist, int64_t n, int k,
    int64_t* __restrict__ indptr, int32_t* __restrict__ indices, float* __restrict__ data) {
  const int l = lane_id();
  const int64_t q = static_cast<int64_t>(blockIdx.x) * (kThreads / kWave) + threadIdx.x / kWave;
  if (q > n) return;
  if (l == 0) indptr[q] = q * (k - 1);
  if (q == n) return;
  uint64_t key = kEmpty;
  if (l < k - 1) {
    key = (static_cast<uint64_t>(static_cast<uint32_t>(idx[q * k + 1 + l])) << 32) |
          __float_as_uint(dist[q * k + 1 + l]);
  }
  key = wave_sort_u64(key);
  if (l < k - 1) {
    indices[q * (k - 1) + l] = static_cast<int32_t>(key >> 32);
    data[q * (k - 1) + l] = __uint_as_float(static_cast<uint32_t>(key));
  }
}

// the same rows for k > 65: the k - 1 keys sorted in the wave's LDS region (KP >= k - 1)
__global__ __launch_bounds__(kThreads) void distances_csr_wide_kernel(
    const int32_t* __restrict__ idx, const float* __restrict__ dist, int64_t n, int k, int KP,
    int64_t* __restrict__ indptr, int32_t* __restrict__ indices, float* __restrict__ data) {
  extern __shared__ uint64_t skeys[];
  const int w = threadIdx.x / kWave, l = lane_id();
  const int64_t q = static_cast<int64_t>(blockIdx.x) * (kThreads / kWave) + w;
  if (q > n) return;
  if (l == 0) indptr[q] = q * (k - 1);
  if (q == n) return;
  uint64_t* a = skeys + static_cast<size_t>(w) * KP;
  for (int p = l; p < KP; p += kWave)
    a[p] = p < k - 1 ? (static_cast<uint64_t>(static_cast<uint32_t>(idx[q * k + 1 + p])) << 32) |
                           __float_as_uint(dist[q * k + 1 + p])
                     : kEmpty;
  wave_lds_sync();
  wave_lds_sort_u64(a, KP);
  for (int p = l; p < k - 1; p += kWave) {
    indices[q * (k - 1) + p] = static_cast<int32_t>(a[p] >> 32);
    data[q * (k - 1) + p] = __uint_as_float(static_cast<uint32_t>(a[p]));
  }
}

// ------------------------------------------------------------------------------------------
// host orchestration
// ------------------------------------------------------------------------------------------
constexpr int kL = 16;              // per-lane list length: 4 lanes x 16 = 64 admitted candidates
constexpr int kLShort = 8;          // the short-list form (batch-balanced kNN with k <= kShortK)
constexpr int kShortK = 4;

// compact flagged rows into a list (deterministic order: ascending row)
__global__ void k_flag_rows(const int32_t* __restrict__ flags, const int32_t* __restrict__ pos, int64_t n,
                            int32_t* __restrict__ rows) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i < n && flags[i]) rows[pos[i]] = static_cast<int32_t>(i);
}
// pad the row list to a whole number of query blocks by repeating its last row
__global__ void k_pad_rows(int32_t* __restrict__ rows, int64_t m, int64_t m_pad) {
  const int64_t i = m + blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i < m_pad) rows[i] = rows[m - 1];
}

struct Plan {
  int64_t n, n_pad, split_len;
  int64_t n_ops;  // operand rows: n_pad plus one tile of padding, so that a tile starting at any
                  // row below n stays inside the array
  int d, k, S, precision;
  int ks_bf16, ks_f32;
};

static Plan make_plan(int64_t n, int d, int k, int precision) {
  Plan p{};
  p.n = n;
  p.d = d;
  p.k = k;
  p.precision = precision;
  p.n_pad = ceil_div(n, kQueriesPerBlock) * kQueriesPerBlock;
  p.n_ops = p.n_pad + kTile;
  p.ks_bf16 = bf16_ks(d);
  p.ks_f32 = f32_ks(d);
  const int64_t qblocks = p.n_pad / kQueriesPerBlock;
  // split candidates so that the grid holds >= ~4 blocks per CU, each split >= 4096 rows
  int64_t S = ceil_div(1024, qblocks);
  S = std::min<int64_t>(S, std::max<int64_t>(1, n / 4096));
  S = std::max<int64_t>(1, std::min<int64_t>(S, 16));
  p.split_len = ceil_div(ceil_div(n, S), kTile) * kTile;
  p.S = static_cast<int>(ceil_div(n, p.split_len));
  return p;
}

struct Buffers {
  double* partial;
  double* mean;
  void* Xb;        // BF16 operands (BF16 form only)
  void* Xf;        // FP32 operands
  float* nrm2b;    // BF16 candidate side |c~|^2, |c~ - x_c|
  float* resb;
  float* qn2b;     // BF16 query side |hi|^2, |hi - x_c|
  float* qresb;
  float* nrm2f;
  float* resf;
  float* gstatb;
  float* gstatf;
  uint64_t* keys;
  float* taus;
  int32_t* flags;
  int32_t* pos;
  int32_t* rows;
  int64_t* nfall;
  void* scan_ws;
  size_t scan_ws_bytes;
};

static constexpr int kColBlocks = 256;

static size_t second_pass_slots(int64_t n_pad) {
  return 2 * static_cast<size_t>(n_pad) + 2048 * static_cast<size_t>(kQueriesPerBlock);
}

static Buffers carve(Carve& cv, const Plan& p) {
  Buffers b{};
  const bool bf = p.precision == MSC_PRECISION_BF16_RERANK;
  b.partial = cv.take<double>(static_cast<size_t>(kColBlocks) * p.d);
  b.mean = cv.take<double>(p.d);
  b.Xb = bf ? cv.take<char>(static_cast<size_t>(p.n_ops) * 64 * p.ks_bf16) : nullptr;
  b.Xf = bf ? nullptr : cv.take<char>(static_cast<size_t>(p.n_ops) * 32 * p.ks_f32);
  b.nrm2b = cv.take<float>(p.n_ops);
  b.resb = cv.take<float>(p.n_ops);
  b.qn2b = cv.take<float>(p.n_ops);
  b.qresb = cv.take<float>(p.n_ops);
  b.nrm2f = cv.take<float>(p.n_ops);
  b.resf = cv.take<float>(p.n_ops);
  b.gstatb = cv.take<float>(4);
  b.gstatf = cv.take<float>(4);
  // the wide second pass over m rows writes 2 * S2 * m_pad <= second_pass_slots(n_pad) slots
  const size_t slots = std::max<size_t>(static_cast<size_t>(p.S) * p.n_pad, second_pass_slots(p.n_pad));
  b.keys = cv.take<uint64_t>(slots * kListStride);
  b.taus = cv.take<float>(slots);
  b.flags = cv.take<int32_t>(p.n_pad + 1);
  b.pos = cv.take<int32_t>(p.n_pad + 1);
  b.rows = cv.take<int32_t>(p.n_pad + kQueriesPerBlock);
  b.nfall = cv.take<int64_t>(1);
  b.scan_ws_bytes = prims::scan_workspace<int32_t>(p.n_pad + 1);
  b.scan_ws = cv.take<char>(b.scan_ws_bytes);
  return b;
}

template <bool BF16, int KS, bool WIDE, int L>
static void launch_select_t(const void* Xp, int64_t n, int64_t n_slots, int d, int64_t split_len, int S,
                            const int32_t* qrows, const int64_t* blk_range, const int64_t* rng4, uint64_t* keys,
                            float* taus, hipStream_t st) {
  const int64_t qpb = WIDE ? kQueriesPerBlock / 2 : kQueriesPerBlock;
  dim3 grid(static_cast<unsigned>(n_slots / qpb), static_cast<unsigned>(S));
  select_kernel<BF16, KS, L, WIDE><<<grid, kThreads, 0, st>>>(Xp, n, n_slots, d, split_len, qrows, blk_range, rng4,
                                                              keys, taus);
}

template <int KT>
static void launch_select32_t(const void* Xp, int64_t n, int64_t n_slots, int d, int rowb, int64_t split_len, int S,
                              const int32_t* qrows, const int64_t* blk_range, const int64_t* rng4, uint64_t* keys,
                              float* taus, hipStream_t st) {
  dim3 grid(static_cast<unsigned>(ceil_div(n_slots, kSelQueries)), static_cast<unsigned>(S));
  select32_kernel<KT><<<grid, kSelThreads, 0, st>>>(Xp, n, n_slots, d, rowb, split_len, qrows, blk_range, rng4, keys,
                                                     taus);
}

// K-steps of 16 of the 32x32x16 form: the 2d + 3 operand terms, rounded up
static inline int bf16_kt(int d) { return (2 * d + 3 + 15) / 16; }

// n_slots must be a multiple of kQueriesPerBlock; the wide form writes 2 * S key sets.
// short_list keeps kLShort keys per lane (4 kLShort per query) instead of kL: for small k,
// where many short candidate ranges would otherwise spend their time filling long lists.
static int launch_select(bool bf16, bool wide, int ks, const void* Xp, int64_t n, int64_t n_slots, int d,
                         int64_t split_len, int S, const int32_t* qrows, const int64_t* blk_range, uint64_t* keys,
                         float* taus, hipStream_t st, bool short_list = false, const int64_t* rng4 = nullptr) {
  // the BF16 main pass (64-key lists) runs on the 32x32x16 form up to 8 K-steps
  if (bf16 && !wide && !short_list && bf16_kt(d) <= 8) {
    const int rowb = 64 * ks;
    switch (bf16_kt(d)) {
#define MSC_SEL32(KT) \
  launch_select32_t<KT>(Xp, n, n_slots, d, rowb, split_len, S, qrows, blk_range, rng4, keys, taus, st)
      case 1: MSC_SEL32(1); break;
      case 2: MSC_SEL32(2); break;
      case 3: MSC_SEL32(3); break;
      case 4: MSC_SEL32(4); break;
      case 5: MSC_SEL32(5); break;
      case 6: MSC_SEL32(6); break;
      case 7: MSC_SEL32(7); break;
      case 8: MSC_SEL32(8); break;
#undef MSC_SEL32
    }
    MSC_LAUNCH_CHECK();
    return kOk;
  }
#define MSC_SEL(B, K)                                                                                             \
  (wide ? launch_select_t<B, K, true, kL>(Xp, n, n_slots, d, split_len, S, qrows, blk_range, rng4, keys, taus, st) \
        : (short_list                                                                                               \
               ? launch_select_t<B, K, false, kLShort>(Xp, n, n_slots, d, split_len, S, qrows, blk_range, rng4,     \
                                                       keys, taus, st)                                              \
               : launch_select_t<B, K, false, kL>(Xp, n, n_slots, d, split_len, S, qrows, blk_range, rng4, keys,    \
                                                  taus, st)))
  if (bf16) {
    switch (ks) {
      case 1: MSC_SEL(true, 1); break;
      case 2: MSC_SEL(true, 2); break;
      case 3: MSC_SEL(true, 3); break;
      case 4: MSC_SEL(true, 4); break;
      case 5: MSC_SEL(true, 5); break;
      case 6: MSC_SEL(true, 6); break;
      case 7: MSC_SEL(true, 7); break;
      case 8: MSC_SEL(true, 8); break;
      default: set_error("knn: d=%d too large for the BF16 form (max 125)", d); return kUnsupported;
    }
  } else {
    switch (ks) {
      case 1: MSC_SEL(false, 1); break;
      case 2: MSC_SEL(false, 2); break;
      case 3: MSC_SEL(false, 3); break;
      case 4: MSC_SEL(false, 4); break;
      case 5: MSC_SEL(false, 5); break;
      case 6: MSC_SEL(false, 6); break;
      case 7: MSC_SEL(false, 7); break;
      case 8: MSC_SEL(false, 8); break;
      case 9: MSC_SEL(false, 9); break;
      case 10: MSC_SEL(false, 10); break;
      case 11: MSC_SEL(false, 11); break;
      case 12: MSC_SEL(false, 12); break;
      case 13: MSC_SEL(false, 13); break;
      case 14: MSC_SEL(false, 14); break;
      case 15: MSC_SEL(false, 15); break;
      case 16: MSC_SEL(false, 16); break;
      default: set_error("knn: d=%d too large for the FP32 form (max 127)", d); return kUnsupported;
    }
  }
#undef MSC_SEL
  MSC_LAUNCH_CHECK();
  return kOk;
}

static Bound make_bound(bool bf16, int ks) {
  Bound bnd;
  const double u = 5.9604644775390625e-08;  // 2^-24
  const int terms = bf16 ? 32 * ks : 2 * 8 * ks;
  bnd.gam = terms * u / (1.0 - terms * u);
  bnd.split_rel = bf16 ? 1.0 / (1 << 26) : 2.0 * u;
  return bnd;
}

// Shapes the MFMA select takes: at most 64 neighbours (its 64 admitted candidates per query)
// and d within its K padding (4 K-steps of 32 bf16 / 8 floats, after the norm terms).
static bool select_shape(int d, int k, int precision) {
  return k <= 64 && (precision == MSC_PRECISION_FP32 ? d <= 127 : d <= 125);
}

static int validate(int64_t n, int d, int k, int precision) {
  MSC_REQUIRE(n >= 1 && n < (int64_t(1) << 31) - kQueriesPerBlock, "knn: n=%lld out of range", (long long)n);
  MSC_REQUIRE(d >= 1, "knn: d must be >= 1");
  MSC_REQUIRE(k >= 1 && k <= wide::kMaxK, "knn: k=%d must be in [1, %d]", k, wide::kMaxK);
  MSC_REQUIRE(k <= n, "knn: k=%d exceeds n=%lld", k, (long long)n);
  MSC_REQUIRE(precision == MSC_PRECISION_FP32 || precision == MSC_PRECISION_BF16_RERANK,
              "knn: unknown precision %d", precision);
  return kOk;
}

// flags -> compacted row list padded to whole query blocks; returns the row count
static int compact_flags(const Plan& p, const Buffers& b, int64_t* m_out, hipStream_t st) {
  MSC_TRY(prims::exclusive_scan<int32_t>(b.flags, b.pos, p.n + 1, b.scan_ws, b.scan_ws_bytes, st));
  int32_t m = 0;
  MSC_HIP(hipMemcpyAsync(&m, b.pos + p.n, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  MSC_HIP(hipStreamSynchronize(st));
  if (m > 0) {
    k_flag_rows<<<static_cast<unsigned>(ceil_div(p.n, kThreads)), kThreads, 0, st>>>(b.flags, b.pos, p.n, b.rows);
    const int64_t m_pad = ceil_div(m, kQueriesPerBlock) * kQueriesPerBlock;
    if (m_pad > m)
      k_pad_rows<<<static_cast<unsigned>(ceil_div(m_pad - m, kThreads)), kThreads, 0, st>>>(b.rows, m, m_pad);
    MSC_LAUNCH_CHECK();
  }
  *m_out = m;
  return kOk;
}

// column means, operand rows, certificate terms for one precision
struct Operands {
  const void* X;
  const float* qn2;
  const float* qres;
  const float* gstat;
  int ks;
  Bound bnd;
};

static Operands prepare(const Plan& p, const Buffers& b, const float* X, int64_t n, bool bf16, hipStream_t st) {
  const int d = p.d;
  const int64_t rows_per = ceil_div(n, kColBlocks);
  colsum_partial<<<kColBlocks, 64, 0, st>>>(X, n, d, rows_per, b.partial);
  colmean_final<<<1, 64, 0, st>>>(b.partial, kColBlocks, d, n, b.mean);
  const unsigned pb = static_cast<unsigned>(ceil_div(p.n_ops, kThreads));
  Operands op{};
  if (bf16) {
    prep_bf16<<<pb, kThreads, 0, st>>>(X, n, p.n_ops, d, 32 * p.ks_bf16, b.mean, static_cast<uint16_t*>(b.Xb),
                                       b.nrm2b, b.resb, b.qn2b, b.qresb);
    point_max<<<1, kThreads, 0, st>>>(b.nrm2b, b.resb, n, b.gstatb);
    op = Operands{b.Xb, b.qn2b, b.qresb, b.gstatb, p.ks_bf16, make_bound(true, p.ks_bf16)};
  } else {
    prep_f32<<<pb, kThreads, 0, st>>>(X, n, p.n_ops, d, 8 * p.ks_f32, b.mean, static_cast<float*>(b.Xf), b.nrm2f,
                                      b.resf);
    point_max<<<1, kThreads, 0, st>>>(b.nrm2f, b.resf, n, b.gstatf);
    op = Operands{b.Xf, b.nrm2f, b.resf, b.gstatf, p.ks_f32, make_bound(false, p.ks_f32)};
  }
  return op;
}

}  // namespace knn
}  // namespace msc

using namespace msc;
using namespace msc::knn;

namespace msc {
namespace knn {
// The partition-pruned candidate ranges (part.h) apply to the MFMA select shapes with d <= 64;
// they are used from kPruneMinRows rows on unless msc_knn_info.prune asks otherwise.
constexpr int64_t kPruneMinRows = 65536;

static bool prune_shape(int64_t n, int d, int k, int precision) {
  return select_shape(d, k, precision) && d <= part::kDMax && n >= 2 * kQueriesPerBlock;
}

}  // namespace knn
}  // namespace msc

extern "C" int msc_knn_workspace_size(int64_t n, int d, int k, int precision, size_t* bytes) {
  MSC_TRY(validate(n, d, k, precision));
  MSC_REQUIRE(bytes != nullptr, "knn: bytes is null");
  if (!select_shape(d, k, precision)) {  // the wide path keeps its lists in LDS
    *bytes = 256;
    return kOk;
  }
  Plan p = make_plan(n, d, k, precision);
  Carve cv(nullptr, 0);
  carve(cv, p);
  if (prune_shape(n, d, k, precision)) cv.take<char>(part::workspace(n, d));
  *bytes = cv.used + 256;
  return kOk;
}

extern "C" int msc_knn_exact(const float* X, int64_t n, int d, int k, int precision,
                             int32_t* out_idx, float* out_dist, void* workspace,
                             size_t workspace_bytes, void* stream, msc_knn_info* info) {
  MSC_TRY(validate(n, d, k, precision));
  MSC_REQUIRE(X && out_idx && out_dist && workspace, "knn: null pointer argument");
  const int mode = info ? info->prune : MSC_PRUNE_AUTO;
  MSC_REQUIRE(mode == MSC_PRUNE_AUTO || mode == MSC_PRUNE_OFF || mode == MSC_PRUNE_ON, "knn: unknown prune mode %d",
              mode);
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (info) {
    info->n_partitions = info->n_groups = info->n_ranges = 0;
    info->scanned = 1.0;
    info->partition_ms = 0.0;
  }
  if (!select_shape(d, k, precision)) {
    // every candidate's exact distance (both precisions: the result is the same exact kNN)
    hipEvent_t e0 = nullptr, e1 = nullptr;
    const bool timing = info && info->timing;
    if (timing) {
      MSC_HIP(hipEventCreate(&e0));
      MSC_HIP(hipEventCreate(&e1));
      MSC_HIP(hipEventRecord(e0, st));
    }
    MSC_TRY(wide::knn(X, n, d, k, false, 0, out_idx, out_dist, st));
    if (info) {
      info->n_splits = 0;
      info->list_len = k;
      info->n_second_pass = 0;
      info->n_fallback = n;
      info->prep_ms = info->select_ms = info->finalize_ms = info->second_pass_ms = 0.0;
      info->fallback_ms = 0.0;
    }
    if (timing) {
      MSC_HIP(hipEventRecord(e1, st));
      MSC_HIP(hipEventSynchronize(e1));
      float ms = 0.0f;
      MSC_HIP(hipEventElapsedTime(&ms, e0, e1));
      info->fallback_ms = ms;
      MSC_HIP(hipEventDestroy(e0));
      MSC_HIP(hipEventDestroy(e1));
    }
    return kOk;
  }
  Plan p = make_plan(n, d, k, precision);
  const bool prune_ok = prune_shape(n, d, k, precision);
  MSC_REQUIRE(mode != MSC_PRUNE_ON || prune_ok,
              "knn: the pruned candidate ranges need d <= %d, k <= 64 and n >= %d (d=%d, k=%d, n=%lld)", part::kDMax,
              2 * kQueriesPerBlock, d, k, (long long)n);
  const bool pruned = prune_ok && (mode == MSC_PRUNE_ON || (mode == MSC_PRUNE_AUTO && n >= kPruneMinRows));
  Carve cv(workspace, workspace_bytes);
  Buffers b = carve(cv, p);
  const size_t pbytes = prune_ok ? part::workspace(n, d) : 0;
  char* pws = prune_ok ? cv.take<char>(pbytes) : nullptr;
  if (!cv.ok()) {
    set_error("knn: workspace too small (%zu < %zu)", workspace_bytes, cv.used);
    return kWorkspaceTooSmall;
  }
  const bool bf16 = precision == MSC_PRECISION_BF16_RERANK;
  const bool timing = info && info->timing;
  hipEvent_t ev[7];
  if (timing)
    for (auto& e : ev) MSC_HIP(hipEventCreate(&e));
  if (timing) MSC_HIP(hipEventRecord(ev[0], st));

  // Rows in select order: X itself, or its rows in partition order (Xw[r] = X[perm[r]]) with
  // every query block given its candidate ranges.
  part::Layout L{};
  const float* Xw = X;
  const int32_t* perm = nullptr;
  if (pruned) {
    MSC_TRY(part::build(X, n, d, k, pws, pbytes, &L, st));
    Xw = L.Xs;
    perm = L.perm;
  }
  if (timing) MSC_HIP(hipEventRecord(ev[1], st));

  Operands op = prepare(p, b, Xw, n, bf16, st);
  MSC_LAUNCH_CHECK();
  if (timing) MSC_HIP(hipEventRecord(ev[2], st));

  if (pruned) {
    MSC_TRY(launch_select(bf16, false, op.ks, op.X, n, L.n_slots, d, n, 1, L.qrows, nullptr, b.keys, b.taus, st,
                          false, L.blk_range));
  } else {
    MSC_TRY(launch_select(bf16, false, op.ks, op.X, n, p.n_pad, d, p.split_len, p.S, nullptr, nullptr, b.keys,
                          b.taus, st));
  }
  if (timing) MSC_HIP(hipEventRecord(ev[3], st));

  const unsigned wb = static_cast<unsigned>(ceil_div(n, kThreads / kWave));
  if (pruned) {
    finalize_kernel<<<static_cast<unsigned>(ceil_div(L.n_slots, kThreads / kWave)), kThreads, 0, st>>>(
        Xw, L.n_slots, d, k, 1, L.n_slots, L.qrows, b.keys, b.taus, op.qn2, op.qres, op.gstat, op.bnd, nullptr, perm,
        out_idx, out_dist, b.flags);
  } else {
    finalize_kernel<<<wb, kThreads, 0, st>>>(X, n, d, k, p.S, p.n_pad, nullptr, b.keys, b.taus, op.qn2, op.qres,
                                             op.gstat, op.bnd, nullptr, nullptr, out_idx, out_dist, b.flags);
  }
  MSC_LAUNCH_CHECK();
  if (timing) MSC_HIP(hipEventRecord(ev[4], st));

  // second pass: rows the certificate could not clear are selected again, over all candidates,
  // with the wide list
  int64_t m1 = 0, m2 = 0;
  MSC_TRY(compact_flags(p, b, &m1, st));
  if (m1 > 0) {
    const int64_t m_pad = ceil_div(m1, kQueriesPerBlock) * kQueriesPerBlock;
    const int64_t wblocks = m_pad / (kQueriesPerBlock / 2);
    // split the candidates of the few rows finely enough that the grid still fills the device
    int64_t S2 = std::max<int64_t>(1, std::min<int64_t>(ceil_div(2048, wblocks), 16));
    S2 = std::min<int64_t>(S2, std::max<int64_t>(1, n / 4096));
    S2 = std::max<int64_t>(1, std::min<int64_t>(S2, static_cast<int64_t>(second_pass_slots(p.n_pad)) / (2 * m_pad)));
    const int64_t split2 = ceil_div(ceil_div(n, S2), kTile) * kTile;
    S2 = ceil_div(n, split2);
    MSC_REQUIRE(2 * S2 * m_pad <= static_cast<int64_t>(second_pass_slots(p.n_pad)), "knn: internal second-pass sizing");
    MSC_TRY(launch_select(bf16, true, op.ks, op.X, n, m_pad, d, split2, static_cast<int>(S2), b.rows, nullptr, b.keys,
                          b.taus, st));
    finalize_kernel<<<static_cast<unsigned>(ceil_div(m1, kThreads / kWave)), kThreads, 0, st>>>(
        Xw, m1, d, k, static_cast<int>(2 * S2), m_pad, b.rows, b.keys, b.taus, op.qn2, op.qres, op.gstat, op.bnd,
        nullptr, perm, out_idx, out_dist, b.flags);
    MSC_LAUNCH_CHECK();
  }
  if (timing) MSC_HIP(hipEventRecord(ev[5], st));
  if (m1 > 0) {
    // rows still uncertified: each row's candidates in S chunks of >= 16,384 rows, one wave per
    // (row, chunk), then a merge per row
    MSC_TRY(compact_flags(p, b, &m2, st));
    if (m2 > 0) {
      const int64_t slots = static_cast<int64_t>(second_pass_slots(p.n_pad));
      int64_t S3 = std::max<int64_t>(1, std::min<int64_t>({ceil_div(n, 16384), 256, slots / m2}));
      const int64_t chunk = ceil_div(n, S3);
      S3 = ceil_div(n, chunk);
      fallback_chunk_kernel<<<static_cast<unsigned>(ceil_div(m2 * S3, kThreads / kWave)), kThreads, 0, st>>>(
          Xw, n, d, b.rows, m2, static_cast<int>(S3), chunk, perm, b.keys);
      fallback_merge_kernel<<<static_cast<unsigned>(ceil_div(m2, kThreads / kWave)), kThreads, 0, st>>>(
          b.rows, m2, static_cast<int>(S3), k, perm, b.keys, out_idx, out_dist);
      MSC_LAUNCH_CHECK();
    }
  }
  if (timing) MSC_HIP(hipEventRecord(ev[6], st));
  if (info) {
    info->n_fallback = m2;
    info->n_second_pass = m1;
    info->n_splits = pruned ? 1 : p.S;
    info->list_len = 4 * kL;
    if (pruned) {
      info->n_partitions = L.n_parts;
      info->n_groups = L.n_groups;
      info->n_ranges = L.S;
      info->scanned = L.pairs / (static_cast<double>(n) * static_cast<double>(n));
    }
    if (timing) {
      MSC_HIP(hipEventSynchronize(ev[6]));
      float ms;
      MSC_HIP(hipEventElapsedTime(&ms, ev[0], ev[1])); info->partition_ms = ms;
      MSC_HIP(hipEventElapsedTime(&ms, ev[1], ev[2])); info->prep_ms = ms;
      MSC_HIP(hipEventElapsedTime(&ms, ev[2], ev[3])); info->select_ms = ms;
      MSC_HIP(hipEventElapsedTime(&ms, ev[3], ev[4])); info->finalize_ms = ms;
      MSC_HIP(hipEventElapsedTime(&ms, ev[4], ev[5])); info->second_pass_ms = ms;
      MSC_HIP(hipEventElapsedTime(&ms, ev[5], ev[6])); info->fallback_ms = ms;
    }
  }
  if (timing)
    for (auto& e : ev) MSC_HIP(hipEventDestroy(e));
  return kOk;
}

extern "C" int msc_knn_distances_csr(const int32_t* idx, const float* dist, int64_t n, int k,
                                     int64_t* out_indptr, int32_t* out_indices, float* out_data,
                                     void* stream) {
  MSC_REQUIRE(idx && dist && out_indptr && (k == 1 || (out_indices && out_data)),
              "distances_csr: null pointer argument");
  MSC_REQUIRE(k >= 1 && k <= wide::kMaxK + 1 && n >= 1, "distances_csr: bad n/k");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const unsigned wb = static_cast<unsigned>(ceil_div(n + 1, kThreads / kWave));
  if (k <= kWave + 1) {
    distances_csr_kernel<<<wb, kThreads, 0, st>>>(idx, dist, n, k, out_indptr, out_indices, out_data);
  } else {
    const int KP = pow2_at_least(k - 1, kWave);
    distances_csr_wide_kernel<<<wb, kThreads, (kThreads / kWave) * KP * sizeof(uint64_t), st>>>(
        idx, dist, n, k, KP, out_indptr, out_indices, out_data);
  }
  MSC_LAUNCH_CHECK();
  return kOk;
}

// ------------------------------------------------------------------------------------------
// batched samples: one launch over many independent samples
// ------------------------------------------------------------------------------------------
namespace msc {
namespace knn {
// padded layout: sample s occupies rows [P_s, P_s + n_s) of an n_tot x d copy, P_s % 128 == 0
__global__ void k_pad_copy(const float* __restrict__ X, const int32_t* __restrict__ prow2row, int64_t n_tot, int d,
                           float* __restrict__ Xpad) {
  const int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (e >= n_tot * d) return;
  const int64_t r = e / d;
  const int32_t src = prow2row[r];
  Xpad[e] = src >= 0 ? X[static_cast<int64_t>(src) * d + (e % d)] : 0.0f;
}
// results back to input rows with unpadded global column ids
__global__ void k_unpad(const int32_t* __restrict__ idx_pad, const float* __restrict__ dist_pad,
                        const int32_t* __restrict__ prow2row, const int64_t* __restrict__ row_range,
                        const int64_t* __restrict__ in_off, const int32_t* __restrict__ prow_sample, int64_t n_tot, int k,
                        int32_t* __restrict__ out_idx, float* __restrict__ out_dist) {
  const int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (e >= n_tot * k) return;
  const int64_t r = e / k;
  const int32_t dst = prow2row[r];
  if (dst < 0) return;
  const int64_t t = e % k;
  const int32_t s = prow_sample[r];
  out_idx[static_cast<int64_t>(dst) * k + t] = static_cast<int32_t>(idx_pad[e] - row_range[2 * r] + in_off[s]);
  out_dist[static_cast<int64_t>(dst) * k + t] = dist_pad[e];
}

// ---- batch-balanced kNN (one list per query and batch) ------------------------------------
// Rows sorted by batch and padded per batch as in the batched layout; select split t of every
// query block scans batch t.  Wave w = t * n_slots + q finalises query slot q against batch t:
// exact re-rank, the certificate over batch t's range alone (the query itself is required
// first only in its own batch), kk[t] results at columns coff[t].. of the K-wide padded row.
__global__ __launch_bounds__(kThreads) void finalize_sep_kernel(
    const float* __restrict__ X, int64_t n_slots, int d, int nb, const uint64_t* __restrict__ keys,
    const float* __restrict__ taus, const float* __restrict__ nrm2, const float* __restrict__ res,
    const float* __restrict__ gstat, Bound bnd, const int32_t* __restrict__ prow2row,
    const int32_t* __restrict__ prow_batch, const int32_t* __restrict__ kk, const int32_t* __restrict__ coff, int K,
    int32_t* __restrict__ idx_pad, float* __restrict__ dist_pad, int32_t* __restrict__ flags) {
  const int l = lane_id();
  const int64_t w = static_cast<int64_t>(blockIdx.x) * (kThreads / kWave) + threadIdx.x / kWave;
  const int64_t M = static_cast<int64_t>(nb) * n_slots;
  if (w >= M) {
    if (w == M && l == 0) flags[M] = 0;  // the flag scan's last input
    return;
  }
  const int t = static_cast<int>(w / n_slots);
  const int64_t q = w - static_cast<int64_t>(t) * n_slots;
  if (prow2row[q] < 0) {  // padding row
    if (l == 0) flags[w] = 0;
    return;
  }
  const int k = kk[t];
  const uint64_t cur = keys[w * kListStride + l];
  const float tau = taus[w];
  uint64_t ek = kEmpty;
  if (cur != kEmpty) {
    const int64_t j = static_cast<int64_t>(cur & 0xFFFFFFFFull);
    ek = exact_key(q, j, exact_d2(X, d, q, j));
  }
  ek = wave_sort_u64(ek);
  const uint64_t kth = shfl_u64(ek, k - 1);
  const bool own = prow_batch[q] == t;
  bool ok = kth != kEmpty && (!own || shfl_u64(ek, 0) == 0ull);
  if (ok && !isinf(tau)) ok = certified(tau, kth, q, d, nrm2, res, gstat, bnd);
  if (ok) {
    if (l < k) {
      int64_t jj;
      float d2;
      decode_exact(ek, q, &jj, &d2);
      idx_pad[q * K + coff[t] + l] = static_cast<int32_t>(jj);
      dist_pad[q * K + coff[t] + l] = sqrtf(d2);
    }
    if (l == 0) flags[w] = 0;
  } else if (l == 0) {
    flags[w] = 1;
  }
}

// exact brute force of each flagged (batch t, slot q) — the compacted list lists[0, *count)
// of w = t * n_slots + q — over batch t's rows [brange[2t], brange[2t+1]); a fixed grid of
// waves strides over the list
constexpr int kFallbackBlocks = 1024;
__global__ __launch_bounds__(kThreads) void fallback_sep_kernel(
    const float* __restrict__ X, int64_t n_slots, int d, const int32_t* __restrict__ lists,
    const int32_t* __restrict__ count, const int64_t* __restrict__ brange, const int32_t* __restrict__ kk,
    const int32_t* __restrict__ coff, int K, int32_t* __restrict__ idx_pad, float* __restrict__ dist_pad) {
  const int l = lane_id();
  const int64_t m = *count;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * (kThreads / kWave);
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * (kThreads / kWave) + threadIdx.x / kWave; i < m; i += stride) {
    const int64_t w = lists[i];
    const int t = static_cast<int>(w / n_slots);
    const int64_t q = w - static_cast<int64_t>(t) * n_slots;
    const int64_t cb = brange[2 * t], ce = brange[2 * t + 1];
    const uint64_t cur = brute_top64(X, d, q, cb, ce);
    if (l < kk[t]) {
      int64_t jj;
      float d2;
      decode_exact(cur, q, &jj, &d2);
      idx_pad[q * K + coff[t] + l] = static_cast<int32_t>(jj);
      dist_pad[q * K + coff[t] + l] = sqrtf(d2);
    }
  }
}

// the K per-batch results of a row sorted by (float32 distance, column): the stable sort of
// the batch-ordered concatenation; written at the input row with input column ids
__global__ __launch_bounds__(kThreads) void k_bbknn_combine(
    const int32_t* __restrict__ idx_pad, const float* __restrict__ dist_pad, int64_t n_slots, int K,
    const int32_t* __restrict__ prow2row, int32_t* __restrict__ out_idx, float* __restrict__ out_dist) {
  const int l = lane_id();
  const int64_t q = static_cast<int64_t>(blockIdx.x) * (kThreads / kWave) + threadIdx.x / kWave;
  if (q >= n_slots) return;
  const int32_t row = prow2row[q];
  if (row < 0) return;
  uint64_t key = kEmpty;
  if (l < K) key = (static_cast<uint64_t>(__float_as_uint(dist_pad[q * K + l])) << 32) | static_cast<uint32_t>(l);
  key = wave_sort_u64(key);
  if (l < K) {
    const int64_t e = q * K + static_cast<int64_t>(key & 0xFFFFFFFFull);
    out_idx[static_cast<int64_t>(row) * K + l] = prow2row[idx_pad[e]];
    out_dist[static_cast<int64_t>(row) * K + l] = dist_pad[e];
  }
}

// The same combine for K > 64: the row's K keys sorted in the wave's LDS region (KP >= K)
__global__ __launch_bounds__(kThreads) void k_bbknn_combine_wide(
    const int32_t* __restrict__ idx_pad, const float* __restrict__ dist_pad, int64_t n_slots, int K, int KP,
    const int32_t* __restrict__ prow2row, int32_t* __restrict__ out_idx, float* __restrict__ out_dist) {
  extern __shared__ uint64_t skeys[];
  const int w = threadIdx.x / kWave, l = lane_id();
  const int64_t q = static_cast<int64_t>(blockIdx.x) * (kThreads / kWave) + w;
  if (q >= n_slots) return;
  const int32_t row = prow2row[q];
  if (row < 0) return;
  uint64_t* a = skeys + static_cast<size_t>(w) * KP;
  for (int p = l; p < KP; p += kWave)
    a[p] = p < K ? (static_cast<uint64_t>(__float_as_uint(dist_pad[q * K + p])) << 32) | static_cast<uint32_t>(p)
                 : kEmpty;
  wave_lds_sync();
  wave_lds_sort_u64(a, KP);
  for (int p = l; p < K; p += kWave) {
    const int64_t e = q * K + static_cast<int64_t>(a[p] & 0xFFFFFFFFull);
    out_idx[static_cast<int64_t>(row) * K + p] = prow2row[idx_pad[e]];
    out_dist[static_cast<int64_t>(row) * K + p] = dist_pad[e];
  }
}

// obsp['distances'] of the combined lists: the self column and zero distances dropped,
// columns ascending (scipy's coo->csr, setdiag(0), eliminate_zeros)
__global__ __launch_bounds__(kThreads) void k_bbknn_dcount(const int32_t* __restrict__ idx,
                                                           const float* __restrict__ dist, int64_t n, int K,
                                                           int64_t* __restrict__ cnt) {
  const int l = lane_id();
  const int64_t q = static_cast<int64_t>(blockIdx.x) * (kThreads / kWave) + threadIdx.x / kWave;
  if (q > n) return;
  if (q == n) {
    if (l == 0) cnt[n] = 0;
    return;
  }
  int64_t c = 0;
  for (int p0 = 0; p0 < K; p0 += kWave) {
    const int p = p0 + l;
    c += __popcll(__ballot(p < K && idx[q * K + p] != static_cast<int32_t>(q) && dist[q * K + p] != 0.0f));
  }
  if (l == 0) cnt[q] = c;
}
__global__ __launch_bounds__(kThreads) void k_bbknn_dfill(const int32_t* __restrict__ idx,
                                                          const float* __restrict__ dist, int64_t n, int K,
                                                          const int64_t* __restrict__ indptr,
                                                          int32_t* __restrict__ indices, float* __restrict__ data) {
  const int l = lane_id();
  const int64_t q = static_cast<int64_t>(blockIdx.x) * (kThreads / kWave) + threadIdx.x / kWave;
  if (q >= n) return;
  uint64_t key = kEmpty;
  if (l < K && idx[q * K + l] != static_cast<int32_t>(q) && dist[q * K + l] != 0.0f)
    key = (static_cast<uint64_t>(static_cast<uint32_t>(idx[q * K + l])) << 32) | __float_as_uint(dist[q * K + l]);
  key = wave_sort_u64(key);
  const int64_t b = indptr[q], m = indptr[q + 1] - b;
  if (l < m) {
    indices[b + l] = static_cast<int32_t>(key >> 32);
    data[b + l] = __uint_as_float(static_cast<uint32_t>(key));
  }
}

// the same fill for K > 64: the kept (column, distance) keys sorted in the wave's LDS region
__global__ __launch_bounds__(kThreads) void k_bbknn_dfill_wide(const int32_t* __restrict__ idx,
                                                               const float* __restrict__ dist, int64_t n, int K,
                                                               int KP, const int64_t* __restrict__ indptr,
                                                               int32_t* __restrict__ indices,
                                                               float* __restrict__ data) {
  extern __shared__ uint64_t skeys[];
  const int w = threadIdx.x / kWave, l = lane_id();
  const int64_t q = static_cast<int64_t>(blockIdx.x) * (kThreads / kWave) + w;
  if (q >= n) return;
  uint64_t* a = skeys + static_cast<size_t>(w) * KP;
  for (int p = l; p < KP; p += kWave) {
    uint64_t key = kEmpty;
    if (p < K && idx[q * K + p] != static_cast<int32_t>(q) && dist[q * K + p] != 0.0f)
      key = (static_cast<uint64_t>(static_cast<uint32_t>(idx[q * K + p])) << 32) | __float_as_uint(dist[q * K + p]);
    a[p] = key;
  }
  wave_lds_sync();
  wave_lds_sort_u64(a, KP);
  const int64_t b = indptr[q], m = indptr[q + 1] - b;
  for (int64_t p = l; p < m; p += kWave) {
    indices[b + p] = static_cast<int32_t>(a[p] >> 32);
    data[b + p] = __uint_as_float(static_cast<uint32_t>(a[p]));
  }
}

}  // namespace knn
}  // namespace msc

namespace {
// Padded layout: sample s occupies rows [P_s, P_s + n_s) of an n_tot x d copy, P_s % 128 == 0.
// Each query block of a sample scans its own sample, split into S candidate ranges of whole
// 64-row tiles (ranges past the sample's end are empty).
struct BatchLayout {
  int64_t n_tot = 0, nblk = 0;
  int S = 1;
  std::vector<int32_t> prow2row, prow_sample;
  std::vector<int64_t> row_range, blk_range;
};

int make_layout(const int64_t* offsets, int n_samples, int k, int S, BatchLayout* L) {
  using msc::knn::kQueriesPerBlock;
  using msc::knn::kTile;
  MSC_REQUIRE(S >= 1 && S <= 16, "knn_batched: splits must be in [1, 16]");
  int64_t P = 0;
  for (int s = 0; s < n_samples; ++s) {
    const int64_t ns = offsets[s + 1] - offsets[s];
    MSC_REQUIRE(ns >= k, "knn_batched: sample %d has %lld cells < k=%d", s, (long long)ns, k);
    P += msc::ceil_div(ns, kQueriesPerBlock) * kQueriesPerBlock;
  }
  MSC_REQUIRE(P < (int64_t(1) << 31) - 256, "knn_batched: total padded rows too large");
  L->n_tot = P;
  L->nblk = P / kQueriesPerBlock;
  L->S = S;
  L->prow2row.assign(P, -1);
  L->prow_sample.assign(P, 0);
  L->row_range.assign(2 * P, 0);
  L->blk_range.assign(2 * L->nblk * S, 0);
  int64_t base = 0;
  for (int s = 0; s < n_samples; ++s) {
    const int64_t ns = offsets[s + 1] - offsets[s];
    const int64_t padded = msc::ceil_div(ns, kQueriesPerBlock) * kQueriesPerBlock;
    for (int64_t i = 0; i < padded; ++i) {
      const int64_t r = base + i;
      L->prow_sample[r] = s;
      if (i < ns) {
        L->prow2row[r] = static_cast<int32_t>(offsets[s] + i);
        L->row_range[2 * r] = base;
        L->row_range[2 * r + 1] = base + ns;
      }
    }
    const int64_t slen = msc::ceil_div(msc::ceil_div(ns, S), kTile) * kTile;
    for (int64_t b = base / kQueriesPerBlock; b < (base + padded) / kQueriesPerBlock; ++b)
      for (int t = 0; t < S; ++t) {
        const int64_t lo = std::min(base + ns, base + t * slen);
        const int64_t hi = std::min(base + ns, lo + slen);
        L->blk_range[2 * (t * L->nblk + b)] = lo;
        L->blk_range[2 * (t * L->nblk + b) + 1] = hi;
      }
    base += padded;
  }
  return msc::kOk;
}

// device copies of the layout tables and the per-call buffers, carved identically by the
// size query, prepare and launch
struct BatchBufs {
  msc::knn::Plan p;
  msc::knn::Buffers b;
  float* Xpad;
  int32_t *prow2row, *prow_sample;
  int64_t *row_range, *blk_range, *in_off;
  int32_t* idx_pad;
  float* dist_pad;
};

BatchBufs carve_batched(msc::Carve& cv, const BatchLayout& L, int n_samples, int d, int k, int precision) {
  using namespace msc::knn;
  BatchBufs t{};
  t.p = make_plan(L.n_tot, d, k, precision);
  t.p.S = L.S;  // keys for S splits of every slot
  t.b = carve(cv, t.p);
  t.Xpad = cv.take<float>(static_cast<size_t>(L.n_tot) * d);
  t.prow2row = cv.take<int32_t>(L.n_tot);
  t.prow_sample = cv.take<int32_t>(L.n_tot);
  t.row_range = cv.take<int64_t>(2 * L.n_tot);
  t.blk_range = cv.take<int64_t>(2 * L.nblk * L.S);
  t.in_off = cv.take<int64_t>(n_samples + 1);
  t.idx_pad = cv.take<int32_t>(static_cast<size_t>(L.n_tot) * k);
  t.dist_pad = cv.take<float>(static_cast<size_t>(L.n_tot) * k);
  return t;
}

int batched_common(const int64_t* offsets, int n_samples, int d, int k, int precision, int splits, int list_len,
                   BatchLayout* L) {
  MSC_REQUIRE(list_len == 0 || list_len == kL || (list_len == kLShort && k <= 4 * kLShort),
              "knn_batched: list_len must be 0, %d or %d (the latter for k <= %d)", kL, kLShort, 4 * kLShort);
  MSC_REQUIRE(offsets && n_samples >= 1, "knn_batched: bad arguments");
  MSC_REQUIRE(precision == MSC_PRECISION_FP32 || precision == MSC_PRECISION_BF16_RERANK, "knn_batched: precision");
  MSC_REQUIRE(k >= 1 && k <= 64 && d >= 1 && d <= 125, "knn_batched: bad d/k");
  return make_layout(offsets, n_samples, k, splits <= 0 ? 1 : splits, L);
}
}  // namespace

extern "C" int msc_knn_batched_workspace_size(const int64_t* offsets, int n_samples, int d, int k, int precision,
                                              int splits, int list_len, size_t* bytes) {
  MSC_REQUIRE(bytes, "knn_batched: bytes is null");
  BatchLayout L;
  MSC_TRY(batched_common(offsets, n_samples, d, k, precision, splits, list_len, &L));
  Carve cv(nullptr, 0);
  carve_batched(cv, L, n_samples, d, k, precision);
  *bytes = cv.used + 256;
  return kOk;
}

extern "C" int msc_knn_batched_prepare(const int64_t* offsets, int n_samples, int d, int k, int precision, int splits,
                                       int list_len,
                                       void* workspace, size_t workspace_bytes, void* stream) {
  BatchLayout L;
  MSC_TRY(batched_common(offsets, n_samples, d, k, precision, splits, list_len, &L));
  MSC_REQUIRE(workspace, "knn_batched: null workspace");
  hipStream_t st = static_cast<hipStream_t>(stream);
  Carve cv(workspace, workspace_bytes);
  BatchBufs t = carve_batched(cv, L, n_samples, d, k, precision);
  if (!cv.ok()) {
    set_error("knn_batched: workspace too small");
    return kWorkspaceTooSmall;
  }
  MSC_HIP(hipMemcpyAsync(t.prow2row, L.prow2row.data(), L.n_tot * sizeof(int32_t), hipMemcpyHostToDevice, st));
  MSC_HIP(hipMemcpyAsync(t.prow_sample, L.prow_sample.data(), L.n_tot * sizeof(int32_t), hipMemcpyHostToDevice, st));
  MSC_HIP(hipMemcpyAsync(t.row_range, L.row_range.data(), 2 * L.n_tot * sizeof(int64_t), hipMemcpyHostToDevice, st));
  MSC_HIP(hipMemcpyAsync(t.blk_range, L.blk_range.data(), 2 * L.nblk * L.S * sizeof(int64_t), hipMemcpyHostToDevice,
                         st));
  MSC_HIP(hipMemcpyAsync(t.in_off, offsets, (n_samples + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
  MSC_HIP(hipStreamSynchronize(st));  // the host tables are pageable and local to this call
  return kOk;
}

extern "C" int msc_knn_batched_launch(const float* X, const int64_t* offsets, int n_samples, int d, int k,
                                      int precision, int splits, int list_len, int32_t* out_idx, float* out_dist,
                                      void* workspace,
                                      size_t workspace_bytes, void* stream) {
  BatchLayout L;
  MSC_TRY(batched_common(offsets, n_samples, d, k, precision, splits, list_len, &L));
  MSC_REQUIRE(X && out_idx && out_dist && workspace, "knn_batched: null pointer argument");
  hipStream_t st = static_cast<hipStream_t>(stream);
  Carve cv(workspace, workspace_bytes);
  BatchBufs t = carve_batched(cv, L, n_samples, d, k, precision);
  if (!cv.ok()) {
    set_error("knn_batched: workspace too small");
    return kWorkspaceTooSmall;
  }
  const Plan& p = t.p;
  const Buffers& b = t.b;
  const int64_t n = L.n_tot;
  k_pad_copy<<<static_cast<unsigned>(ceil_div(n * d, kThreads)), kThreads, 0, st>>>(X, t.prow2row, n, d, t.Xpad);
  MSC_LAUNCH_CHECK();
  const bool bf16 = precision == MSC_PRECISION_BF16_RERANK;
  Operands op = prepare(p, b, t.Xpad, n, bf16, st);
  MSC_LAUNCH_CHECK();
  MSC_TRY(launch_select(bf16, false, op.ks, op.X, n, n, d, n, L.S, nullptr, t.blk_range, b.keys, b.taus, st,
                        list_len == kLShort));
  const unsigned wb = static_cast<unsigned>(ceil_div(n, kThreads / kWave));
  finalize_kernel<<<wb, kThreads, 0, st>>>(t.Xpad, n, d, k, L.S, n, nullptr, b.keys, b.taus, op.qn2, op.qres,
                                           op.gstat, op.bnd, t.row_range, nullptr, t.idx_pad, t.dist_pad, b.flags);
  MSC_LAUNCH_CHECK();
  fallback_kernel<<<wb, kThreads, 0, st>>>(t.Xpad, n, d, k, b.flags, t.row_range, nullptr, t.idx_pad, t.dist_pad);
  k_unpad<<<static_cast<unsigned>(ceil_div(n * k, kThreads)), kThreads, 0, st>>>(
      t.idx_pad, t.dist_pad, t.prow2row, t.row_range, t.in_off, t.prow_sample, n, k, out_idx, out_dist);
  MSC_LAUNCH_CHECK();
  return kOk;
}

extern "C" int msc_knn_batched(const float* X, const int64_t* offsets, int n_samples, int d, int k, int precision,
                               int splits, int list_len, int32_t* out_idx, float* out_dist, void* workspace,
                               size_t workspace_bytes,
                               void* stream, msc_knn_info* info) {
  MSC_TRY(msc_knn_batched_prepare(offsets, n_samples, d, k, precision, splits, list_len, workspace, workspace_bytes,
                                  stream));
  MSC_TRY(msc_knn_batched_launch(X, offsets, n_samples, d, k, precision, splits, list_len, out_idx, out_dist, workspace,
                                 workspace_bytes, stream));
  if (info) {
    BatchLayout L;
    MSC_TRY(batched_common(offsets, n_samples, d, k, precision, splits, list_len, &L));
    Carve cv(workspace, workspace_bytes);
    BatchBufs t = carve_batched(cv, L, n_samples, d, k, precision);
    int64_t m = 0;
    MSC_TRY(compact_flags(t.p, t.b, &m, static_cast<hipStream_t>(stream)));
    info->n_fallback = m;
    info->n_second_pass = 0;
    info->n_splits = L.S;
    info->list_len = 4 * (list_len == kLShort ? kLShort : kL);
  }
  return kOk;
}

// ------------------------------------------------------------------------------------------
// batch-balanced kNN
// ------------------------------------------------------------------------------------------
namespace {
// Batch t occupies padded rows [P_t, P_t + n_t) (its input rows in ascending order),
// P_t % 128 == 0; every query block of every batch runs one select split per batch t, over
// exactly batch t's rows.  Results: kk_t = min(k, n_t) columns per batch at offset coff_t of
// a K = sum kk_t wide row.
// combined neighbours per row (the K-wide rows are sorted in LDS beyond 64)
constexpr int kBbMaxWidth = 1024;

struct BbLayout {
  int64_t n_tot = 0, nblk = 0;
  int nb = 0, K = 0;
  std::vector<int32_t> prow2row, prow_batch, kk, coff;
  std::vector<int64_t> brange, blk_range;
};

int make_bb_layout(const int32_t* batch, int64_t n, int n_batches, int k, BbLayout* L) {
  using msc::knn::kQueriesPerBlock;
  MSC_REQUIRE(batch && n >= 1 && n < (int64_t(1) << 31) - 8192, "bbknn: bad n or null batch");
  MSC_REQUIRE(n_batches >= 1 && n_batches <= kBbMaxWidth, "bbknn: n_batches must be in [1, %d]", kBbMaxWidth);
  MSC_REQUIRE(k >= 1 && k <= 64, "bbknn: neighbors_within_batch must be in [1, 64]");
  std::vector<int64_t> cnt(n_batches, 0);
  for (int64_t i = 0; i < n; ++i) {
    const int32_t b = batch[i];
    MSC_REQUIRE(b >= 0 && b < n_batches, "bbknn: batch code %d of row %lld outside [0, %d)", b, (long long)i,
                n_batches);
    ++cnt[b];
  }
  L->nb = n_batches;
  L->kk.resize(n_batches);
  L->coff.resize(n_batches);
  L->brange.resize(2 * n_batches);
  std::vector<int64_t> P(n_batches);
  int64_t base = 0;
  int K = 0;
  for (int t = 0; t < n_batches; ++t) {
    MSC_REQUIRE(cnt[t] >= 1, "bbknn: batch %d has no cells", t);
    L->kk[t] = static_cast<int32_t>(std::min<int64_t>(k, cnt[t]));
    L->coff[t] = K;
    K += L->kk[t];
    P[t] = base;
    L->brange[2 * t] = base;
    L->brange[2 * t + 1] = base + cnt[t];
    base += msc::ceil_div(cnt[t], kQueriesPerBlock) * kQueriesPerBlock;
  }
  MSC_REQUIRE(K <= kBbMaxWidth, "bbknn: %d combined neighbours (sum over batches of min(k, batch size)) exceed %d",
              K, kBbMaxWidth);
  L->K = K;
  MSC_REQUIRE(base * n_batches < (int64_t(1) << 31) - 1, "bbknn: n_batches x padded cells must stay below 2^31");
  L->n_tot = base;
  L->nblk = base / kQueriesPerBlock;
  L->prow2row.assign(base, -1);
  L->prow_batch.assign(base, 0);
  std::vector<int64_t> pos(P);
  for (int64_t i = 0; i < n; ++i) L->prow2row[pos[batch[i]]++] = static_cast<int32_t>(i);
  for (int t = 0; t < n_batches; ++t)
    for (int64_t r = P[t]; r < P[t] + msc::ceil_div(cnt[t], kQueriesPerBlock) * kQueriesPerBlock; ++r)
      L->prow_batch[r] = t;
  L->blk_range.resize(2 * L->nblk * n_batches);
  for (int t = 0; t < n_batches; ++t)
    for (int64_t b = 0; b < L->nblk; ++b) {
      L->blk_range[2 * (t * L->nblk + b)] = L->brange[2 * t];
      L->blk_range[2 * (t * L->nblk + b) + 1] = L->brange[2 * t + 1];
    }
  return msc::kOk;
}

struct BbBufs {
  msc::knn::Plan p;
  msc::knn::Buffers b;
  float* Xpad;
  int32_t *prow2row, *prow_batch, *kk, *coff, *flags, *fpos, *flist;
  int64_t *brange, *blk_range;
  void* fscan;
  size_t fscan_bytes;
  int32_t* idx_pad;
  float* dist_pad;
};

BbBufs carve_bb(msc::Carve& cv, const BbLayout& L, int d, int k, int precision) {
  using namespace msc::knn;
  BbBufs t{};
  t.p = make_plan(L.n_tot, d, k, precision);
  t.p.S = L.nb;  // one key set per (batch, slot)
  t.b = carve(cv, t.p);
  t.Xpad = cv.take<float>(static_cast<size_t>(L.n_tot) * d);
  t.prow2row = cv.take<int32_t>(L.n_tot);
  t.prow_batch = cv.take<int32_t>(L.n_tot);
  t.kk = cv.take<int32_t>(L.nb);
  t.coff = cv.take<int32_t>(L.nb);
  const int64_t M = static_cast<int64_t>(L.nb) * L.n_tot;
  t.flags = cv.take<int32_t>(M + 1);
  t.fpos = cv.take<int32_t>(M + 1);
  t.flist = cv.take<int32_t>(M);
  t.fscan_bytes = msc::prims::scan_workspace<int32_t>(M + 1);
  t.fscan = cv.take<char>(t.fscan_bytes);
  t.brange = cv.take<int64_t>(2 * L.nb);
  t.blk_range = cv.take<int64_t>(2 * L.nblk * L.nb);
  t.idx_pad = cv.take<int32_t>(static_cast<size_t>(L.n_tot) * L.K);
  t.dist_pad = cv.take<float>(static_cast<size_t>(L.n_tot) * L.K);
  return t;
}

int bb_common(const int32_t* batch, int64_t n, int n_batches, int d, int k, int precision, BbLayout* L) {
  MSC_REQUIRE(precision == MSC_PRECISION_FP32 || precision == MSC_PRECISION_BF16_RERANK, "bbknn: precision");
  MSC_REQUIRE(d >= 1 && (precision == MSC_PRECISION_FP32 ? d <= 127 : d <= 125), "bbknn: d=%d unsupported", d);
  return make_bb_layout(batch, n, n_batches, k, L);
}
}  // namespace

extern "C" int msc_bbknn_workspace_size(const int32_t* batch, int64_t n, int n_batches, int d, int k, int precision,
                                        size_t* bytes) {
  MSC_REQUIRE(bytes, "bbknn: bytes is null");
  BbLayout L;
  MSC_TRY(bb_common(batch, n, n_batches, d, k, precision, &L));
  Carve cv(nullptr, 0);
  carve_bb(cv, L, d, k, precision);
  *bytes = cv.used + 256;
  return kOk;
}

extern "C" int msc_bbknn(const float* X, int64_t n, int d, const int32_t* batch, int n_batches, int k,
                         int precision, int32_t* out_idx, float* out_dist, int* out_width, void* workspace,
                         size_t workspace_bytes, void* stream, msc_knn_info* info) {
  BbLayout L;
  MSC_TRY(bb_common(batch, n, n_batches, d, k, precision, &L));
  MSC_REQUIRE(X && out_idx && out_dist && workspace, "bbknn: null pointer argument");
  hipStream_t st = static_cast<hipStream_t>(stream);
  Carve cv(workspace, workspace_bytes);
  BbBufs t = carve_bb(cv, L, d, k, precision);
  if (!cv.ok()) {
    set_error("bbknn: workspace too small (%zu < %zu)", workspace_bytes, cv.used);
    return kWorkspaceTooSmall;
  }
  const int64_t nt = L.n_tot;
  const int nb = L.nb;
  MSC_HIP(hipMemcpyAsync(t.prow2row, L.prow2row.data(), nt * sizeof(int32_t), hipMemcpyHostToDevice, st));
  MSC_HIP(hipMemcpyAsync(t.prow_batch, L.prow_batch.data(), nt * sizeof(int32_t), hipMemcpyHostToDevice, st));
  MSC_HIP(hipMemcpyAsync(t.kk, L.kk.data(), nb * sizeof(int32_t), hipMemcpyHostToDevice, st));
  MSC_HIP(hipMemcpyAsync(t.coff, L.coff.data(), nb * sizeof(int32_t), hipMemcpyHostToDevice, st));
  MSC_HIP(hipMemcpyAsync(t.brange, L.brange.data(), 2 * nb * sizeof(int64_t), hipMemcpyHostToDevice, st));
  MSC_HIP(hipMemcpyAsync(t.blk_range, L.blk_range.data(), L.blk_range.size() * sizeof(int64_t),
                         hipMemcpyHostToDevice, st));
  MSC_HIP(hipStreamSynchronize(st));  // the host tables are pageable and local to this call
  k_pad_copy<<<static_cast<unsigned>(ceil_div(nt * d, kThreads)), kThreads, 0, st>>>(X, t.prow2row, nt, d, t.Xpad);
  MSC_LAUNCH_CHECK();
  const bool bf16 = precision == MSC_PRECISION_BF16_RERANK;
  Operands op = prepare(t.p, t.b, t.Xpad, nt, bf16, st);
  MSC_LAUNCH_CHECK();
  // the certificate compares the k-th exact distance with the smallest lane tail, which is at
  // least the (per-lane list length)-th approximate distance: 8 per lane keeps that margin
  // wide for k <= 4 at half the list upkeep of the 16-per-lane lists
  MSC_TRY(launch_select(bf16, false, op.ks, op.X, nt, nt, d, nt, nb, nullptr, t.blk_range, t.b.keys, t.b.taus, st,
                        k <= kShortK));
  const int64_t M = static_cast<int64_t>(nb) * nt;
  const unsigned wb = static_cast<unsigned>(ceil_div(M + 1, kThreads / kWave));
  finalize_sep_kernel<<<wb, kThreads, 0, st>>>(t.Xpad, nt, d, nb, t.b.keys, t.b.taus, op.qn2, op.qres, op.gstat,
                                               op.bnd, t.prow2row, t.prow_batch, t.kk, t.coff, L.K, t.idx_pad,
                                               t.dist_pad, t.flags);
  MSC_LAUNCH_CHECK();
  // compact the uncertified (batch, slot) lists; the count stays on the device
  MSC_TRY(prims::exclusive_scan<int32_t>(t.flags, t.fpos, M + 1, t.fscan, t.fscan_bytes, st));
  k_flag_rows<<<static_cast<unsigned>(ceil_div(M, kThreads)), kThreads, 0, st>>>(t.flags, t.fpos, M, t.flist);
  fallback_sep_kernel<<<kFallbackBlocks, kThreads, 0, st>>>(t.Xpad, nt, d, t.flist, t.fpos + M, t.brange, t.kk,
                                                            t.coff, L.K, t.idx_pad, t.dist_pad);
  MSC_LAUNCH_CHECK();
  const unsigned cb = static_cast<unsigned>(ceil_div(nt, kThreads / kWave));
  if (L.K <= kWave) {
    k_bbknn_combine<<<cb, kThreads, 0, st>>>(t.idx_pad, t.dist_pad, nt, L.K, t.prow2row, out_idx, out_dist);
  } else {
    const int KP = pow2_at_least(L.K, 128);
    k_bbknn_combine_wide<<<cb, kThreads, (kThreads / kWave) * KP * sizeof(uint64_t), st>>>(
        t.idx_pad, t.dist_pad, nt, L.K, KP, t.prow2row, out_idx, out_dist);
  }
  MSC_LAUNCH_CHECK();
  if (out_width) *out_width = L.K;
  if (info) {
    int32_t m = 0;
    MSC_HIP(hipMemcpyAsync(&m, t.fpos + M, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    MSC_HIP(hipStreamSynchronize(st));
    info->n_fallback = m;
    info->n_second_pass = 0;
    info->n_splits = nb;
    info->list_len = 4 * (k <= kShortK ? kLShort : kL);
  }
  return kOk;
}

extern "C" int msc_bbknn_distances_workspace_size(int64_t n, size_t* bytes) {
  MSC_REQUIRE(bytes && n >= 1, "bbknn_distances: bad arguments");
  *bytes = prims::scan_workspace<int64_t>(n + 1) + 256;
  return kOk;
}

extern "C" int msc_bbknn_distances_csr(const int32_t* idx, const float* dist, int64_t n, int K, int64_t* out_indptr,
                                       int32_t* out_indices, float* out_data, void* workspace,
                                       size_t workspace_bytes, void* stream) {
  MSC_REQUIRE(idx && dist && out_indptr && out_indices && out_data && workspace, "bbknn_distances: null pointer");
  MSC_REQUIRE(n >= 1 && K >= 1 && K <= kBbMaxWidth, "bbknn_distances: need n >= 1 and 1 <= K <= %d", kBbMaxWidth);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const unsigned wb = static_cast<unsigned>(ceil_div(n + 1, kThreads / kWave));
  k_bbknn_dcount<<<wb, kThreads, 0, st>>>(idx, dist, n, K, out_indptr);
  MSC_LAUNCH_CHECK();
  MSC_TRY(prims::exclusive_scan<int64_t>(out_indptr, out_indptr, n + 1, workspace, workspace_bytes, st));
  if (K <= kWave) {
    k_bbknn_dfill<<<wb, kThreads, 0, st>>>(idx, dist, n, K, out_indptr, out_indices, out_data);
  } else {
    const int KP = pow2_at_least(K, 128);
    k_bbknn_dfill_wide<<<wb, kThreads, (kThreads / kWave) * KP * sizeof(uint64_t), st>>>(
        idx, dist, n, K, KP, out_indptr, out_indices, out_data);
  }
  MSC_LAUNCH_CHECK();
  return kOk;
}

// Exact kNN for wide shapes (k > 64, or d beyond the MFMA select's K padding; the spatial kNN
// for k > 32 or d outside {2, 3}).  Replaces the reference's fallbacks for these shapes:
// argpartition over the Gram distances when K > 32 (src/metalsinglecell/neighbors.py:66-68)
// and exact brute force for the grid kNN when k > 32 or d not in {2, 3} (neighbors.py:350-355).
//
// Every candidate's exact distance is computed (float32 direct differences, summed in
// coordinate order with separately rounded operations — the definition of neighbors.py:
// 420-462), so no certificate is needed.  Work layout:
//   * a workgroup (4 waves) owns 4*QW query rows; each wave owns QW of them;
//   * candidates stream through LDS 64 rows at a time (a lane per candidate row), in chunks
//     of 64 coordinates (row stride 65 floats: lane l's reads hit bank (l + t) mod 64, so a
//     wave's 64 reads are conflict-free); the block's query rows sit beside them and are read
//     as 16-byte broadcasts;
//   * per query the wave keeps its KP best keys sorted in LDS plus a buffer of admitted keys;
//     a key is admitted only below the current k-th best, so after the first tiles almost
//     every candidate costs one compare; a full buffer is sorted and merged into the list.
#include "wide.h"

#include "msc_hip.h"

namespace msc {
namespace wide {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / kWave;
constexpr int kTile = 64;    // candidate rows per tile (one per lane)
constexpr int kChunk = 64;   // coordinates per LDS chunk
constexpr int kStride = kChunk + 1;
constexpr uint64_t kEmpty = ~0ull;

// (j != q, value, j): self -> 0, others ((value bits << 32) | j) + 1
__device__ __forceinline__ uint64_t make_key(int64_t q, int64_t j, float v) {
  if (j == q) return 0ull;
  return ((static_cast<uint64_t>(__float_as_uint(v)) << 32) | static_cast<uint64_t>(j)) + 1ull;
}

template <int QW, bool SQRT_KEY>
__global__ __launch_bounds__(kThreads) void wide_select_kernel(const float* __restrict__ X, int64_t n, int d, int k,
                                                               int KP, int skip, int32_t* __restrict__ out_idx,
                                                               float* __restrict__ out_dist) {
  constexpr int QB = QW * kWaves;
  extern __shared__ __align__(16) unsigned char smem[];
  uint64_t* lists = reinterpret_cast<uint64_t*>(smem);                 // QB x {list, buffer} x KP
  float* tile = reinterpret_cast<float*>(lists + static_cast<size_t>(QB) * 2 * KP);  // kTile x kStride
  float* qv = tile + kTile * kStride;                                   // QB x kChunk
  const int w = threadIdx.x / kWave, l = lane_id();
  const int64_t q0 = static_cast<int64_t>(blockIdx.x) * QB;
  const int kout = k - skip;

  uint64_t tau[QW];
  int nbuf[QW];
#pragma unroll
  for (int u = 0; u < QW; ++u) {
    uint64_t* lst = lists + static_cast<size_t>(w * QW + u) * 2 * KP;
    for (int p = l; p < 2 * KP; p += kWave) lst[p] = kEmpty;
    tau[u] = kEmpty;
    nbuf[u] = 0;
  }

  for (int64_t c0 = 0; c0 < n; c0 += kTile) {
    float acc[QW];
#pragma unroll
    for (int u = 0; u < QW; ++u) acc[u] = 0.0f;
    for (int t0 = 0; t0 < d; t0 += kChunk) {
      const int dc = min(kChunk, d - t0);
      __syncthreads();  // the previous chunk's readers are done with tile and qv
      for (int r = w; r < kTile; r += kWaves) {
        const int64_t row = c0 + r;
        if (l < dc) tile[r * kStride + l] = row < n ? X[row * d + t0 + l] : 0.0f;
      }
      for (int r = w; r < QB; r += kWaves) {
        const int64_t row = q0 + r;
        if (l < dc) qv[r * kChunk + l] = row < n ? X[row * d + t0 + l] : 0.0f;
      }
      __syncthreads();
      const float* crow = tile + l * kStride;
      const float* qw = qv + w * QW * kChunk;
      int t = 0;
      for (; t + 4 <= dc; t += 4) {
        const float c0v = crow[t], c1v = crow[t + 1], c2v = crow[t + 2], c3v = crow[t + 3];
#pragma unroll
        for (int u = 0; u < QW; ++u) {
          const float4 qq = *reinterpret_cast<const float4*>(qw + u * kChunk + t);
          float df = __fsub_rn(qq.x, c0v);
          acc[u] = __fadd_rn(acc[u], __fmul_rn(df, df));
          df = __fsub_rn(qq.y, c1v);
          acc[u] = __fadd_rn(acc[u], __fmul_rn(df, df));
          df = __fsub_rn(qq.z, c2v);
          acc[u] = __fadd_rn(acc[u], __fmul_rn(df, df));
          df = __fsub_rn(qq.w, c3v);
          acc[u] = __fadd_rn(acc[u], __fmul_rn(df, df));
        }
      }
      for (; t < dc; ++t) {
        const float cv = crow[t];
#pragma unroll
        for (int u = 0; u < QW; ++u) {
          const float df = __fsub_rn(qw[u * kChunk + t], cv);
          acc[u] = __fadd_rn(acc[u], __fmul_rn(df, df));
        }
      }
    }
    const int64_t j = c0 + l;
#pragma unroll
    for (int u = 0; u < QW; ++u) {
      const int64_t q = q0 + w * QW + u;
      uint64_t key = kEmpty;
      if (j < n && q < n) key = make_key(q, j, SQRT_KEY ? sqrtf(acc[u]) : acc[u]);
      const bool admit = key < tau[u];
      const uint64_t m = __ballot(admit);
      if (m == 0) continue;
      uint64_t* lst = lists + static_cast<size_t>(w * QW + u) * 2 * KP;
      uint64_t* buf = lst + KP;
      if (admit) buf[nbuf[u] + __popcll(m & ((1ull << l) - 1ull))] = key;
      nbuf[u] += __popcll(m);
      if (nbuf[u] > KP - kWave) {
        wave_lds_sync();
        for (int p = nbuf[u] + l; p < KP; p += kWave) buf[p] = kEmpty;
        wave_lds_sync();
        wave_lds_sort_u64(buf, KP);
        wave_lds_merge_min_u64(lst, buf, KP);
        nbuf[u] = 0;
        tau[u] = lst[k - 1];
      }
    }
  }

#pragma unroll
  for (int u = 0; u < QW; ++u) {
    const int64_t q = q0 + w * QW + u;
    uint64_t* lst = lists + static_cast<size_t>(w * QW + u) * 2 * KP;
    uint64_t* buf = lst + KP;
    if (nbuf[u] > 0) {
      wave_lds_sync();
      for (int p = nbuf[u] + l; p < KP; p += kWave) buf[p] = kEmpty;
      wave_lds_sync();
      wave_lds_sort_u64(buf, KP);
      wave_lds_merge_min_u64(lst, buf, KP);
    }
    if (q >= n) continue;
    for (int r = skip + l; r < k; r += kWave) {
      uint64_t key = lst[r];
      int64_t jj = q;
      float v = 0.0f;
      if (key != 0ull) {
        key -= 1ull;
        jj = static_cast<int64_t>(key & 0xFFFFFFFFull);
        v = __uint_as_float(static_cast<uint32_t>(key >> 32));
      }
      out_idx[q * kout + (r - skip)] = static_cast<int32_t>(jj);
      out_dist[q * kout + (r - skip)] = SQRT_KEY ? v : sqrtf(v);
    }
  }
}

template <int QW, bool SQRT_KEY>
static int launch(const float* X, int64_t n, int d, int k, int KP, int skip, int32_t* out_idx, float* out_dist,
                  hipStream_t st) {
  constexpr int QB = QW * kWaves;
  const size_t lds = static_cast<size_t>(QB) * 2 * KP * sizeof(uint64_t) +
                     (static_cast<size_t>(kTile) * kStride + static_cast<size_t>(QB) * kChunk) * sizeof(float);
  MSC_REQUIRE(lds <= 64 * 1024, "wide kNN: internal LDS sizing (%zu bytes)", lds);
  const unsigned grid = static_cast<unsigned>(ceil_div(n, QB));
  wide_select_kernel<QW, SQRT_KEY><<<grid, kThreads, lds, st>>>(X, n, d, k, KP, skip, out_idx, out_dist);
  MSC_LAUNCH_CHECK();
  return kOk;
}

int knn(const float* X, int64_t n, int d, int k, bool sqrt_key, int skip, int32_t* out_idx, float* out_dist,
        hipStream_t st) {
  MSC_REQUIRE(n >= 1 && n < (int64_t(1) << 31) - 1, "wide kNN: n=%lld out of range", (long long)n);
  MSC_REQUIRE(d >= 1, "wide kNN: d must be >= 1");
  MSC_REQUIRE(k >= 1 && k <= kMaxK && k <= n, "wide kNN: k=%d must be in [1, min(n, %d)]", k, kMaxK);
  MSC_REQUIRE(skip >= 0 && skip < k, "wide kNN: bad skip");
  const int KP = pow2_at_least(k, 128);
  // lists of 4*QW queries x 2 x KP keys stay within 32 KB (64 KB at KP = 512)
  if (KP <= 128)
    return sqrt_key ? launch<4, true>(X, n, d, k, KP, skip, out_idx, out_dist, st)
                    : launch<4, false>(X, n, d, k, KP, skip, out_idx, out_dist, st);
  if (KP <= 256)
    return sqrt_key ? launch<2, true>(X, n, d, k, KP, skip, out_idx, out_dist, st)
                    : launch<2, false>(X, n, d, k, KP, skip, out_idx, out_dist, st);
  return sqrt_key ? launch<1, true>(X, n, d, k, KP, skip, out_idx, out_dist, st)
                  : launch<1, false>(X, n, d, k, KP, skip, out_idx, out_dist, st);
}

}  // namespace wide
}  // namespace msc

// Deterministic scans, a stable atomics-free LSD radix sort and segment boundaries.
#include "prims.h"

namespace msc {
namespace prims {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

template <typename T>
__device__ __forceinline__ T shfl_up_t(T v, int d) {
  if constexpr (sizeof(T) == 8) {
    uint64_t u = static_cast<uint64_t>(v);
    uint32_t lo = __shfl_up(static_cast<uint32_t>(u), d, kWave);
    uint32_t hi = __shfl_up(static_cast<uint32_t>(u >> 32), d, kWave);
    return static_cast<T>((static_cast<uint64_t>(hi) << 32) | lo);
  } else {
    return __shfl_up(v, d, kWave);
  }
}

// inclusive scan within a wave
template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v) {
  const int l = lane_id();
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    T o = shfl_up_t(v, d);
    if (l >= d) v += o;
  }
  return v;
}

// exclusive scan of one value per thread across the block; returns the block total in *tot
template <typename T>
__device__ __forceinline__ T block_excl_scan(T v, T* tot) {
  __shared__ T wsum[kScanThreads / kWave];
  const int w = threadIdx.x / kWave;
  T inc = wave_incl_scan(v);
  if (lane_id() == kWave - 1) wsum[w] = inc;
  __syncthreads();
  T off = 0, all = 0;
#pragma unroll
  for (int i = 0; i < kScanThreads / kWave; ++i) {
    if (i < w) off += wsum[i];
    all += wsum[i];
  }
  __syncthreads();
  *tot = all;
  return off + inc - v;
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_block_reduce(const T* __restrict__ in, int64_t n,
                                                               T* __restrict__ sums) {
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
  T s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i)
    if (base + i < n) s += in[base + i];
  T tot;
  block_excl_scan(s, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_block_scan(const T* in, T* out, int64_t n,
                                                             const T* __restrict__ offsets) {
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
  T v[kScanItems];
  T s = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    v[i] = (base + i < n) ? in[base + i] : T(0);
    s += v[i];
  }
  T tot;
  T pre = block_excl_scan(s, &tot);
  if (offsets) pre += offsets[blockIdx.x];
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    if (base + i < n) out[base + i] = pre;
    pre += v[i];
  }
}

template <typename T>
size_t scan_workspace(int64_t n) {
  size_t bytes = 0;
  while (n > kScanTile) {
    n = ceil_div(n, kScanTile);
    bytes += ((n * sizeof(T) + 255) / 256) * 256;
  }
  return bytes + 256;
}

template <typename T>
int exclusive_scan(const T* in, T* out, int64_t n, void* ws, size_t ws_bytes, hipStream_t st) {
  if (n <= 0) return kOk;
  if (n <= kScanTile) {
    k_block_scan<T><<<1, kScanThreads, 0, st>>>(in, out, n, nullptr);
    MSC_LAUNCH_CHECK();
    return kOk;
  }
  const int64_t nt = ceil_div(n, kScanTile);
  const size_t mine = ((nt * sizeof(T) + 255) / 256) * 256;
  if (mine > ws_bytes) {
    set_error("scan: workspace too small");
    return kWorkspaceTooSmall;
  }
  T* sums = static_cast<T*>(ws);
  k_block_reduce<T><<<static_cast<unsigned>(nt), kScanThreads, 0, st>>>(in, n, sums);
  MSC_LAUNCH_CHECK();
  MSC_TRY(exclusive_scan<T>(sums, sums, nt, static_cast<char*>(ws) + mine, ws_bytes - mine, st));
  k_block_scan<T><<<static_cast<unsigned>(nt), kScanThreads, 0, st>>>(in, out, n, sums);
  MSC_LAUNCH_CHECK();
  return kOk;
}

template size_t scan_workspace<int32_t>(int64_t);
template size_t scan_workspace<int64_t>(int64_t);
template int exclusive_scan<int32_t>(const int32_t*, int32_t*, int64_t, void*, size_t, hipStream_t);
template int exclusive_scan<int64_t>(const int64_t*, int64_t*, int64_t, void*, size_t, hipStream_t);

// ------------------------------------------------------------------------------------------
// radix sort
// ------------------------------------------------------------------------------------------
constexpr int kRsThreads = 256;
constexpr int kRsRounds = 16;
constexpr int kRsTile = kRsThreads * kRsRounds;
constexpr int kRsWaves = kRsThreads / kWave;

// lanes of the wave holding the same 8-bit digit (valid lanes only)
__device__ __forceinline__ uint64_t peer_mask(uint32_t digit, bool valid) {
  uint64_t m = __ballot(valid);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const bool bit = (digit >> b) & 1u;
    const uint64_t bm = __ballot(bit);
    m &= bit ? bm : ~bm;
  }
  return m;
}

__global__ __launch_bounds__(kRsThreads) void k_rs_hist(const uint32_t* __restrict__ keys, int64_t n,
                                                        int shift, int nblocks,
                                                        int32_t* __restrict__ counts) {
  __shared__ int32_t h[kRsWaves][256];
  const int tid = threadIdx.x, w = tid / kWave, l = lane_id();
  for (int i = tid; i < kRsWaves * 256; i += kRsThreads) (&h[0][0])[i] = 0;
  __syncthreads();
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kRsTile;
  for (int r = 0; r < kRsRounds; ++r) {
    const int64_t pos = base + r * kRsThreads + tid;
    const bool valid = pos < n;
    const uint32_t digit = valid ? (keys[pos] >> shift) & 255u : 0u;
    const uint64_t peers = peer_mask(digit, valid);
    if (valid && (__ffsll(static_cast<long long>(peers)) - 1) == l)
      h[w][digit] += __popcll(peers);
  }
  __syncthreads();
  int32_t s = 0;
#pragma unroll
  for (int i = 0; i < kRsWaves; ++i) s += h[i][tid];
  counts[static_cast<int64_t>(tid) * nblocks + blockIdx.x] = s;
}

__global__ __launch_bounds__(kRsThreads) void k_rs_scatter(
    const uint32_t* __restrict__ kin, const uint32_t* __restrict__ vin, uint32_t* __restrict__ kout,
    uint32_t* __restrict__ vout, int64_t n, int shift, int nblocks, const int32_t* __restrict__ offsets) {
  __shared__ int32_t basep[256];
  __shared__ int32_t wc[kRsWaves][256];
  const int tid = threadIdx.x, w = tid / kWave, l = lane_id();
  basep[tid] = offsets[static_cast<int64_t>(tid) * nblocks + blockIdx.x];
#pragma unroll
  for (int i = 0; i < kRsWaves; ++i) wc[i][tid] = 0;
  __syncthreads();
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kRsTile;
  const uint64_t lt = (1ull << l) - 1ull;
  for (int r = 0; r < kRsRounds; ++r) {
    const int64_t pos = base + r * kRsThreads + tid;
    const bool valid = pos < n;
    const uint32_t key = valid ? kin[pos] : 0u;
    const uint32_t digit = (key >> shift) & 255u;
    const uint64_t peers = peer_mask(digit, valid);
    const int rank = __popcll(peers & lt);
    if (valid && (__ffsll(static_cast<long long>(peers)) - 1) == l) wc[w][digit] = __popcll(peers);
    __syncthreads();
    if (valid) {
      int32_t off = basep[digit] + rank;
      for (int i = 0; i < w; ++i) off += wc[i][digit];
      kout[off] = key;
      vout[off] = vin[pos];
    }
    __syncthreads();
    int32_t add = 0;
#pragma unroll
    for (int i = 0; i < kRsWaves; ++i) {
      add += wc[i][tid];
      wc[i][tid] = 0;
    }
    basep[tid] += add;
    __syncthreads();
  }
}

size_t radix_workspace(int64_t n) {
  const int64_t nb = ceil_div(std::max<int64_t>(n, 1), kRsTile);
  const int64_t nc = 256 * nb;
  size_t b = 0;
  b += ((nc * sizeof(int32_t) + 255) / 256) * 256 * 2;    // counts + offsets
  b += ((n * sizeof(uint32_t) + 255) / 256) * 256 * 2;    // ping-pong keys/vals
  b += scan_workspace<int32_t>(nc);
  return b + 1024;
}

int radix_sort_pairs(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
                     uint32_t* vals_out, int64_t n, int key_bits, void* ws, size_t ws_bytes,
                     hipStream_t st) {
  if (n <= 0) return kOk;
  MSC_REQUIRE(n < (int64_t(1) << 31), "radix sort: n too large");
  const int nb = static_cast<int>(ceil_div(n, kRsTile));
  const int64_t nc = 256 * static_cast<int64_t>(nb);
  Carve cv(ws, ws_bytes);
  int32_t* counts = cv.take<int32_t>(nc);
  int32_t* offs = cv.take<int32_t>(nc);
  uint32_t* tk = cv.take<uint32_t>(n);
  uint32_t* tv = cv.take<uint32_t>(n);
  const size_t sw = scan_workspace<int32_t>(nc);
  void* sws = cv.take<char>(sw);
  if (!cv.ok()) {
    set_error("radix sort: workspace too small");
    return kWorkspaceTooSmall;
  }
  const int passes = std::max(1, (key_bits + 7) / 8);
  // pass p reads src, writes dst; the last pass writes (keys_out, vals_out)
  const uint32_t* sk = keys_in;
  const uint32_t* sv = vals_in;
  for (int p = 0; p < passes; ++p) {
    const bool last = p == passes - 1;
    // alternate temporaries so that the final pass lands in the outputs
    uint32_t* dk = last ? keys_out : (((passes - 1 - p) & 1) ? tk : keys_out);
    uint32_t* dv = last ? vals_out : (((passes - 1 - p) & 1) ? tv : vals_out);
    k_rs_hist<<<nb, kRsThreads, 0, st>>>(sk, n, 8 * p, nb, counts);
    MSC_LAUNCH_CHECK();
    MSC_TRY(exclusive_scan<int32_t>(counts, offs, nc, sws, sw, st));
    k_rs_scatter<<<nb, kRsThreads, 0, st>>>(sk, sv, dk, dv, n, 8 * p, nb, offs);
    MSC_LAUNCH_CHECK();
    sk = dk;
    sv = dv;
  }
  return kOk;
}

__global__ void k_segment_starts(const uint32_t* __restrict__ keys, int64_t m, int64_t nseg,
                                 int64_t* __restrict__ seg_start) {
  const int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (p > m) return;
  const int64_t prev = (p == 0) ? -1 : std::min<int64_t>(keys[p - 1], nseg);
  const int64_t cur = (p == m) ? nseg : std::min<int64_t>(keys[p], nseg);
  for (int64_t j = prev + 1; j <= cur; ++j) seg_start[j] = p;
}

int segment_starts(const uint32_t* keys, int64_t m, int64_t nseg, int64_t* seg_start, hipStream_t st) {
  const unsigned g = static_cast<unsigned>(ceil_div(m + 1, 256));
  k_segment_starts<<<g, 256, 0, st>>>(keys, m, nseg, seg_start);
  MSC_LAUNCH_CHECK();
  return kOk;
}

int bits_for(int64_t v) {
  int b = 1;
  while (b < 63 && (int64_t(1) << b) <= v) ++b;
  return b;
}

}  // namespace prims
}  // namespace msc

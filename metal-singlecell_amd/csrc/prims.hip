// Deterministic scans, a stable atomics-free LSD radix sort and segment boundaries.
#include "prims.h"

namespace msc {
namespace prims {

constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;
constexpr int64_t kScanLookback = 2048;  // tiles up to which a block sums its predecessors

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

// tile_sums != nullptr: this block's offset is the sum of tile_sums[0, blockIdx.x), summed here
template <typename T>
__global__ __launch_bounds__(kScanThreads) void k_block_scan(const T* in, T* out, int64_t n,
                                                             const T* __restrict__ offsets,
                                                             const T* __restrict__ tile_sums = nullptr) {
  __shared__ T shoff;
  if (tile_sums) {
    T s = 0;
    for (int i = threadIdx.x; i < static_cast<int>(blockIdx.x); i += kScanThreads) s += tile_sums[i];
    T tot;
    block_excl_scan(s, &tot);
    if (threadIdx.x == 0) shoff = tot;
    __syncthreads();
  }
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
  if (tile_sums) pre += shoff;
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
  if (nt <= kScanLookback) {  // each block adds up the preceding tile sums itself
    k_block_scan<T><<<static_cast<unsigned>(nt), kScanThreads, 0, st>>>(in, out, n, nullptr, sums);
    MSC_LAUNCH_CHECK();
    return kOk;
  }
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
// One pass = histogram (per block, per digit) -> exclusive scan over (digit, block) ->
// scatter.  A tile of kRsTile keys is split into kRsWaves contiguous wave segments; lane l
// of wave w holds items seg(w) + 64 i + l, i < kRsItems, all loaded up front.  Within a
// wave, keys of equal digit are ranked by ballot peer masks and a wave-private LDS counter
// (written by one lane per peer group), so ranks follow input order: the sort is stable
// and uses no atomics.  The scatter stages the tile in LDS in digit order so each digit's
// run leaves the block as one contiguous, coalesced write.
constexpr int kRsThreads = 256;
constexpr int kRsItems = 8;
constexpr int kRsWaves = kRsThreads / kWave;
constexpr int kRsSeg = kRsItems * kWave;         // keys per wave segment
constexpr int kRsTile = kRsThreads * kRsItems;   // keys per block

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
#pragma unroll
  for (int i = 0; i < kRsWaves; ++i) h[i][tid] = 0;
  __syncthreads();
  const int64_t seg = static_cast<int64_t>(blockIdx.x) * kRsTile + w * kRsSeg;
  uint32_t dg[kRsItems];
#pragma unroll
  for (int i = 0; i < kRsItems; ++i) {
    const int64_t pos = seg + i * kWave + l;
    dg[i] = pos < n ? (keys[pos] >> shift) & 255u : 256u;
  }
#pragma unroll
  for (int i = 0; i < kRsItems; ++i) {
    const bool valid = dg[i] < 256u;
    const uint64_t peers = peer_mask(dg[i], valid);
    if (valid && (__ffsll(static_cast<long long>(peers)) - 1) == l) h[w][dg[i]] += __popcll(peers);
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
  __shared__ int32_t wc[kRsWaves][256];   // per-wave digit counts, then per-wave digit offsets
  __shared__ int32_t dstart[256];         // first tile slot of each digit
  __shared__ int32_t gbase[256];          // global position of each digit's first key
  __shared__ uint32_t sk[kRsTile];
  __shared__ uint32_t sv[kRsTile];
  const int tid = threadIdx.x, w = tid / kWave, l = lane_id();
#pragma unroll
  for (int i = 0; i < kRsWaves; ++i) wc[i][tid] = 0;
  gbase[tid] = offsets[static_cast<int64_t>(tid) * nblocks + blockIdx.x];
  const int64_t tile0 = static_cast<int64_t>(blockIdx.x) * kRsTile;
  const int64_t seg = tile0 + w * kRsSeg;
  uint32_t key[kRsItems], val[kRsItems];
  int rank[kRsItems];
#pragma unroll
  for (int i = 0; i < kRsItems; ++i) {
    const int64_t pos = seg + i * kWave + l;
    const bool valid = pos < n;
    key[i] = valid ? kin[pos] : 0u;
    val[i] = valid ? vin[pos] : 0u;
  }
  __syncthreads();
  const uint64_t lt = (1ull << l) - 1ull;
  const int64_t nvalid = n - seg;  // keys of this wave segment that exist (may be <= 0)
#pragma unroll
  for (int i = 0; i < kRsItems; ++i) {
    const bool valid = i * kWave + l < nvalid;
    const uint32_t digit = (key[i] >> shift) & 255u;
    const uint64_t peers = peer_mask(digit, valid);
    const int before = valid ? wc[w][digit] : 0;   // all reads of this round precede the update
    rank[i] = before + __popcll(peers & lt);
    if (valid && (__ffsll(static_cast<long long>(peers)) - 1) == l) wc[w][digit] = before + __popcll(peers);
  }
  __syncthreads();
  // thread tid owns digit tid: per-wave offsets within the digit and the digit's tile start
  int32_t run = 0;
#pragma unroll
  for (int i = 0; i < kRsWaves; ++i) {
    const int32_t c = wc[i][tid];
    wc[i][tid] = run;
    run += c;
  }
  // exclusive scan of the digit totals across the block (one digit per thread)
  __shared__ int32_t wsum[kRsWaves];
  int32_t inc = run;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const int32_t o = __shfl_up(inc, d, kWave);
    if (l >= d) inc += o;
  }
  if (l == kWave - 1) wsum[w] = inc;
  __syncthreads();
  int32_t off = 0;
#pragma unroll
  for (int i = 0; i < kRsWaves; ++i)
    if (i < w) off += wsum[i];
  dstart[tid] = off + inc - run;
  __syncthreads();
  // stage in digit order
#pragma unroll
  for (int i = 0; i < kRsItems; ++i) {
    if (i * kWave + l < nvalid) {
      const uint32_t digit = (key[i] >> shift) & 255u;
      const int slot = dstart[digit] + wc[w][digit] + rank[i];
      sk[slot] = key[i];
      sv[slot] = val[i];
    }
  }
  __syncthreads();
  // coalesced write-out: slot p holds the (p - dstart[d])-th key of digit d in this tile
  const int64_t m = n - tile0 < kRsTile ? n - tile0 : kRsTile;
  for (int p = tid; p < m; p += kRsThreads) {
    const uint32_t k = sk[p];
    const uint32_t digit = (k >> shift) & 255u;
    const int64_t dst = static_cast<int64_t>(gbase[digit]) + (p - dstart[digit]);
    kout[dst] = k;
    vout[dst] = sv[p];
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

// seg_start[j] = first p with min(keys[p], nseg) >= j, for j = 0..nseg (keys ascending)
__global__ void k_segment_starts(const uint32_t* __restrict__ keys, int64_t m, int64_t nseg,
                                 int64_t* __restrict__ seg_start) {
  const int64_t j = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (j > nseg) return;
  int64_t lo = 0, hi = m;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (static_cast<int64_t>(keys[mid]) < j) lo = mid + 1;
    else hi = mid;
  }
  seg_start[j] = lo;
}

int segment_starts(const uint32_t* keys, int64_t m, int64_t nseg, int64_t* seg_start, hipStream_t st) {
  const unsigned g = static_cast<unsigned>(ceil_div(nseg + 1, 256));
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

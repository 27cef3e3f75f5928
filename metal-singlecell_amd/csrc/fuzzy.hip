// UMAP fuzzy-simplicial-set connectivities on gfx950.
//
// Replaces the reference's CPU call of umap-learn's fuzzy_simplicial_set
// (src/metalsinglecell/neighbors.py:742-753; set_op_mix_ratio=1, local_connectivity=1,
// precomputed kNN).  Stages, all deterministic and free of global atomics:
//   1. mean of all kNN distances (float64, fixed reduction order)
//   2. per row: rho (first positive distance), sigma (<= 64 float32 bisection steps toward
//      log2 k), the 1e-3 * mean floors, membership strengths v_ij
//   3. transpose: stable radix sort of the nonzero off-diagonal edges by destination
//   4. per row: union of the out-list (sorted by column) and the in-list (sources ascending),
//      P_ij = (a + b) - a*b in float32; count, scan, fill -> canonical CSR
// exp is the project's float32 expf32 (common.h), identical to the oracle's.
#include <mutex>

#include "common.h"
#include "msc_hip.h"
#include "prims.h"

namespace msc {
namespace fuzzy {

constexpr int kThreads = 256;

// Mean of the kNN distances, per group of rows (batched samples) or over all rows (goff null):
// group g (blockIdx.y) splits its m_g = rows_g * k distances into gridDim.x fixed ranges, so
// a group's sum is formed exactly as for a call on that group alone.
__device__ __forceinline__ void group_span(const int64_t* goff, int64_t m, int k, int g, int64_t* base, int64_t* mg) {
  if (goff) {
    *base = goff[g] * k;
    *mg = (goff[g + 1] - goff[g]) * k;
  } else {
    *base = 0;
    *mg = m;
  }
}
__global__ void k_dist_partial(const float* __restrict__ dist, int64_t m, int k, const int64_t* __restrict__ goff,
                               double* __restrict__ partial) {
  int64_t base, mg;
  group_span(goff, m, k, blockIdx.y, &base, &mg);
  dist += base;
  partial += static_cast<int64_t>(blockIdx.y) * gridDim.x;
  const int64_t per = (mg + gridDim.x - 1) / gridDim.x;
  const int64_t b0 = blockIdx.x * per, b1 = min(mg, b0 + per);
  double s = 0.0;
#pragma unroll 8
  for (int64_t i = b0 + threadIdx.x; i < b1; i += blockDim.x) s += static_cast<double>(dist[i]);
  s = wave_sum_f64(s);
  __shared__ double sh[kThreads / kWave];
  if (lane_id() == 0) sh[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kThreads / kWave; ++w) t += sh[w];
    partial[blockIdx.x] = t;
  }
}

__global__ void k_dist_mean(const double* __restrict__ partial, int nb, int64_t m, int k,
                            const int64_t* __restrict__ goff, double* __restrict__ mean) {
  if (threadIdx.x == 0) {
    int64_t base, mg;
    group_span(goff, m, k, blockIdx.x, &base, &mg);
    double t = 0.0;
    for (int b = 0; b < nb; ++b) t += partial[static_cast<int64_t>(blockIdx.x) * nb + b];
    mean[blockIdx.x] = t / static_cast<double>(mg);
  }
}

// one thread per row; STAGED: the row's distances staged in LDS (k <= 64), else read from
// global memory (k > 64; the row stays in the caches over the bisection)
template <bool STAGED>
__global__ __launch_bounds__(64) void k_smooth_membership(
    const int32_t* __restrict__ idx, const float* __restrict__ dist, int64_t n, int k,
    double target, const double* __restrict__ mean_all, const int64_t* __restrict__ goff, int n_groups,
    float* __restrict__ sigma_out,
    float* __restrict__ rho_out, float* __restrict__ vals) {
  __shared__ float sd[STAGED ? 64 * 65 : 1];
  const int64_t r0 = static_cast<int64_t>(blockIdx.x) * 64;
  const int rows = static_cast<int>(min<int64_t>(64, n - r0));
  if constexpr (STAGED) {
    for (int e = threadIdx.x; e < rows * k; e += 64) sd[(e / k) * 65 + (e % k)] = dist[r0 * k + e];
    __syncthreads();
  }
  const int t = threadIdx.x;
  if (t >= rows) return;
  const int64_t i = r0 + t;
  const float* D = STAGED ? &sd[t * 65] : dist + i * k;

  float rho = 0.0f;
  for (int j = 0; j < k; ++j)
    if (D[j] > 0.0f) { rho = D[j]; break; }
  double rowsum = 0.0;
  for (int j = 0; j < k; ++j) rowsum += static_cast<double>(D[j]);

  float lo = 0.0f, hi = __uint_as_float(0x7F800000u), mid = 1.0f;
  for (int it = 0; it < 64; ++it) {
    float psum = 0.0f;
    for (int j = 1; j < k; ++j) {
      const float dd = __fsub_rn(D[j], rho);
      const float e = dd > 0.0f ? expf32(-(dd / mid)) : 1.0f;
      psum = __fadd_rn(psum, e);
    }
    const double ps = static_cast<double>(psum);
    if (fabs(ps - target) < 1e-5) break;
    if (ps > target) {
      hi = mid;
      mid = __fadd_rn(lo, hi) / 2.0f;
    } else {
      lo = mid;
      mid = isinf(hi) ? mid * 2.0f : __fadd_rn(lo, hi) / 2.0f;
    }
  }
  float sigma = mid;
  int grp = 0;  // group (sample) of row i: last g with goff[g] <= i
  if (goff) {
    int lo = 0, hi = n_groups - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (goff[mid] <= i) lo = mid;
      else hi = mid - 1;
    }
    grp = lo;
  }
  const double floor_v = (rho > 0.0f) ? 1e-3 * (rowsum / static_cast<double>(k)) : 1e-3 * mean_all[grp];
  if (static_cast<double>(sigma) < floor_v) sigma = static_cast<float>(floor_v);
  sigma_out[i] = sigma;
  rho_out[i] = rho;

  for (int j = 0; j < k; ++j) {
    const int32_t c = idx[i * k + j];
    float v;
    if (c == static_cast<int32_t>(i)) v = 0.0f;
    else {
      const float dd = __fsub_rn(D[j], rho);
      v = (dd <= 0.0f || sigma == 0.0f) ? 1.0f : expf32(-(dd / sigma));
    }
    vals[i * k + j] = v;
  }
}

// transposed edge list: key = destination for kept edges (off-diagonal, v != 0), n otherwise
__global__ void k_edge_keys(const int32_t* __restrict__ idx, const float* __restrict__ vals, int64_t n,
                            int k, uint32_t* __restrict__ key, uint32_t* __restrict__ eid) {
  const int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (e >= n * k) return;
  const int64_t i = e / k;
  const int32_t c = idx[e];
  const bool keep = c >= 0 && c != static_cast<int32_t>(i) && vals[e] != 0.0f;
  key[e] = keep ? static_cast<uint32_t>(c) : static_cast<uint32_t>(n);
  eid[e] = static_cast<uint32_t>(e);
}

// lower_bound of x in the source ids of the in-list [lo, hi) (sources ascending)
__device__ __forceinline__ int64_t in_lower_bound(const uint32_t* __restrict__ in_eid, int64_t lo,
                                                  int64_t hi, int k, int64_t x) {
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (static_cast<int64_t>(in_eid[mid] / k) < x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// one wave per row; FILL=false counts entries, FILL=true writes them
template <bool FILL>
__global__ __launch_bounds__(kThreads) void k_union(
    const int32_t* __restrict__ idx, const float* __restrict__ vals, int64_t n, int k,
    const int64_t* __restrict__ in_start, const uint32_t* __restrict__ in_eid,
    int64_t* __restrict__ row_nnz, const int64_t* __restrict__ indptr,
    int32_t* __restrict__ out_idx, float* __restrict__ out_val, int64_t cap) {
  __shared__ int64_t fcol[kThreads / kWave][64];
  __shared__ int32_t fpre[kThreads / kWave][65];
  const int w = threadIdx.x / kWave, l = lane_id();
  const int64_t i = static_cast<int64_t>(blockIdx.x) * (kThreads / kWave) + w;
  if (i >= n) return;

  // out-list F: kept entries of row i sorted by column (key = col << 32 | slot)
  uint64_t key = ~0ull;
  if (l < k) {
    const int32_t c = idx[i * k + l];
    if (c >= 0 && c != static_cast<int32_t>(i) && vals[i * k + l] != 0.0f)
      key = (static_cast<uint64_t>(static_cast<uint32_t>(c)) << 32) | static_cast<uint32_t>(l);
  }
  key = wave_sort_u64(key);
  const bool inF = key != ~0ull;
  const int nF = __popcll(__ballot(inF));
  const int64_t col = inF ? static_cast<int64_t>(key >> 32) : INT64_MAX;
  const int slot = static_cast<int>(key & 0xFFFFFFFFu);
  const int64_t s0 = in_start[i], s1 = in_start[i + 1];
  const int64_t nS = s1 - s0;

  // membership of each F column in the in-list
  int64_t lb = s0;
  bool found = false;
  if (inF) {
    lb = in_lower_bound(in_eid, s0, s1, k, col);
    found = lb < s1 && static_cast<int64_t>(in_eid[lb] / k) == col;
  }
  const uint64_t fmask = __ballot(found);
  const int nBoth = __popcll(fmask);
  if (!FILL) {
    if (l == 0) row_nnz[i] = nF + nS - nBoth;
    if (i == 0 && l == 0) row_nnz[n] = 0;
    return;
  }
  const int64_t base = indptr[i];
  fcol[w][l] = col;
  // fpre[j] = number of found flags among F[0..j)
  const int pre = __popcll(fmask & ((1ull << l) - 1ull));
  fpre[w][l] = pre;
  if (l == kWave - 1) fpre[w][kWave] = nBoth;
  // fcol/fpre are private to this wave: LDS ops of one wave complete in order, so only a
  // compiler-level ordering point is needed (no block barrier; other waves may have exited)
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

  if (inF) {
    const float a = vals[i * k + slot];
    float v = a;
    if (found) {
      const float b = vals[in_eid[lb]];
      v = __fsub_rn(__fadd_rn(a, b), __fmul_rn(a, b));
    }
    const int64_t pos = base + l + (lb - s0) - pre;
    if (pos < cap) {
      out_idx[pos] = static_cast<int32_t>(col);
      out_val[pos] = v;
    }
  }
  // in-only entries
  for (int64_t p = s0 + l; p < s1; p += kWave) {
    const uint32_t e = in_eid[p];
    const int64_t s = static_cast<int64_t>(e / k);
    // lower_bound of s in F
    int lo = 0, hi = nF;
    while (lo < hi) {
      const int m = (lo + hi) >> 1;
      if (fcol[w][m] < s) lo = m + 1;
      else hi = m;
    }
    const bool dup = lo < nF && fcol[w][lo] == s;
    if (!dup) {
      const int64_t pos = base + lo + ((p - s0) - fpre[w][lo]);
      if (pos < cap) {
        out_idx[pos] = static_cast<int32_t>(s);
        out_val[pos] = vals[e];
      }
    }
  }
}

// k > 64: the same union with the out-list F of up to k - 1 entries sorted in the wave's LDS
// region (KP >= k keys) and the found-prefix counts kept beside it
template <bool FILL>
__global__ __launch_bounds__(kThreads) void k_union_wide(
    const int32_t* __restrict__ idx, const float* __restrict__ vals, int64_t n, int k, int KP,
    const int64_t* __restrict__ in_start, const uint32_t* __restrict__ in_eid,
    int64_t* __restrict__ row_nnz, const int64_t* __restrict__ indptr,
    int32_t* __restrict__ out_idx, float* __restrict__ out_val, int64_t cap) {
  extern __shared__ uint64_t ushared[];
  const int w = threadIdx.x / kWave, l = lane_id();
  const int64_t i = static_cast<int64_t>(blockIdx.x) * (kThreads / kWave) + w;
  if (i >= n) return;
  uint64_t* F = ushared + static_cast<size_t>(w) * KP;
  int32_t* fpre = reinterpret_cast<int32_t*>(ushared + static_cast<size_t>(kThreads / kWave) * KP) + w * (KP + 1);

  int nF = 0;
  for (int p = l; p < KP; p += kWave) {
    uint64_t key = ~0ull;
    if (p < k) {
      const int32_t c = idx[i * k + p];
      if (c >= 0 && c != static_cast<int32_t>(i) && vals[i * k + p] != 0.0f)
        key = (static_cast<uint64_t>(static_cast<uint32_t>(c)) << 32) | static_cast<uint32_t>(p);
    }
    nF += __popcll(__ballot(key != ~0ull));
    F[p] = key;
  }
  wave_lds_sync();
  wave_lds_sort_u64(F, KP);
  const int64_t s0 = in_start[i], s1 = in_start[i + 1];
  const int64_t nS = s1 - s0;
  const int64_t base = FILL ? indptr[i] : 0;

  // F entries in column order: found in the in-list or not, and their output positions
  int nBoth = 0;
  for (int f0 = 0; f0 < KP; f0 += kWave) {
    const int f = f0 + l;
    const bool inF = f < nF;
    const uint64_t key = F[f];
    const int64_t col = inF ? static_cast<int64_t>(key >> 32) : INT64_MAX;
    int64_t lb = s0;
    bool found = false;
    if (inF) {
      lb = in_lower_bound(in_eid, s0, s1, k, col);
      found = lb < s1 && static_cast<int64_t>(in_eid[lb] / k) == col;
    }
    const uint64_t fmask = __ballot(found);
    const int pre = nBoth + __popcll(fmask & ((1ull << l) - 1ull));
    fpre[f] = pre;
    if (FILL && inF) {
      const int slot = static_cast<int>(key & 0xFFFFFFFFu);
      const float a = vals[i * k + slot];
      float v = a;
      if (found) {
        const float b = vals[in_eid[lb]];
        v = __fsub_rn(__fadd_rn(a, b), __fmul_rn(a, b));
      }
      const int64_t pos = base + f + (lb - s0) - pre;
      if (pos < cap) {
        out_idx[pos] = static_cast<int32_t>(col);
        out_val[pos] = v;
      }
    }
    nBoth += __popcll(fmask);
  }
  if (l == 0) fpre[KP] = nBoth;
  if (!FILL) {
    if (l == 0) row_nnz[i] = nF + nS - nBoth;
    if (i == 0 && l == 0) row_nnz[n] = 0;
    return;
  }
  wave_lds_sync();
  // in-only entries
  for (int64_t p = s0 + l; p < s1; p += kWave) {
    const uint32_t e = in_eid[p];
    const int64_t src = static_cast<int64_t>(e / k);
    int lo = 0, hi = nF;
    while (lo < hi) {
      const int m = (lo + hi) >> 1;
      if (static_cast<int64_t>(F[m] >> 32) < src) lo = m + 1;
      else hi = m;
    }
    const bool dup = lo < nF && static_cast<int64_t>(F[lo] >> 32) == src;
    if (!dup) {
      const int64_t pos = base + lo + ((p - s0) - fpre[lo]);
      if (pos < cap) {
        out_idx[pos] = static_cast<int32_t>(src);
        out_val[pos] = vals[e];
      }
    }
  }
}

struct Bufs {
  double* partial;
  double* mean;
  float* vals;
  uint32_t* key;
  uint32_t* eid;
  uint32_t* skey;
  uint32_t* seid;
  int64_t* in_start;
  int64_t* row_nnz;
  void* sort_ws;
  size_t sort_ws_bytes;
  void* scan_ws;
  size_t scan_ws_bytes;
};

constexpr int kMeanBlocks = 256;

static Bufs carve(Carve& cv, int64_t n, int k, int n_groups = 1) {
  Bufs b{};
  const int64_t m = n * k;
  b.partial = cv.take<double>(static_cast<size_t>(kMeanBlocks) * n_groups);
  b.mean = cv.take<double>(n_groups);
  b.vals = cv.take<float>(m);
  b.key = cv.take<uint32_t>(m);
  b.eid = cv.take<uint32_t>(m);
  b.skey = cv.take<uint32_t>(m);
  b.seid = cv.take<uint32_t>(m);
  b.in_start = cv.take<int64_t>(n + 1);
  b.row_nnz = cv.take<int64_t>(n + 1);
  b.sort_ws_bytes = prims::radix_workspace(m);
  b.sort_ws = cv.take<char>(b.sort_ws_bytes);
  b.scan_ws_bytes = prims::scan_workspace<int64_t>(n + 1);
  b.scan_ws = cv.take<char>(b.scan_ws_bytes);
  return b;
}

static int validate(int64_t n, int k) {
  MSC_REQUIRE(n >= 1 && k >= 1 && k <= 1024, "fuzzy: need n >= 1 and 1 <= k <= 1024");
  MSC_REQUIRE(n * k < (int64_t(1) << 31), "fuzzy: n*k must be < 2^31");
  return kOk;
}

// ---- trimming (bbknn): keep each row's strongest entries ----------------------------------
// Row i's threshold is its trim-th largest value when it holds more than trim entries, else
// 0; entry (i, j) survives when it is nonzero and >= both row thresholds (the row pass, the
// transpose and the column pass of neighbors._trim_graph collapse to that test).
// The trim-th largest is found by a bitwise search over the order-preserving uint32 image:
// the largest T with #{v : ord(v) >= T} >= trim is exactly ord of the trim-th largest value.
__global__ __launch_bounds__(kThreads) void k_trim_thresh(const int64_t* __restrict__ indptr,
                                                          const float* __restrict__ data, int64_t n, int64_t trim,
                                                          float* __restrict__ thr) {
  const int l = lane_id();
  const int64_t q = static_cast<int64_t>(blockIdx.x) * (kThreads / kWave) + threadIdx.x / kWave;
  if (q >= n) return;
  const int64_t b = indptr[q], e = indptr[q + 1];
  if (e - b <= trim) {
    if (l == 0) thr[q] = 0.0f;
    return;
  }
  uint32_t T = 0;
  for (int bit = 31; bit >= 0; --bit) {
    const uint32_t cand = T | (1u << bit);
    int64_t c = 0;
    for (int64_t p = b + l; p < e; p += kWave) c += float_ord(data[p]) >= cand ? 1 : 0;
    c = wave_sum_i64(c);
    if (c >= trim) T = cand;
  }
  if (l == 0) thr[q] = ord_float(T);
}

__device__ __forceinline__ bool trim_keep(float v, float ti, float tj) { return v != 0.0f && v >= ti && v >= tj; }

__global__ __launch_bounds__(kThreads) void k_trim_count(const int64_t* __restrict__ indptr,
                                                         const int32_t* __restrict__ indices,
                                                         const float* __restrict__ data, int64_t n,
                                                         const float* __restrict__ thr, int64_t* __restrict__ cnt) {
  const int l = lane_id();
  const int64_t q = static_cast<int64_t>(blockIdx.x) * (kThreads / kWave) + threadIdx.x / kWave;
  if (q > n) return;
  if (q == n) {
    if (l == 0) cnt[n] = 0;
    return;
  }
  const int64_t b = indptr[q], e = indptr[q + 1];
  const float ti = thr[q];
  int64_t c = 0;
  for (int64_t p = b + l; p < e; p += kWave) c += trim_keep(data[p], ti, thr[indices[p]]) ? 1 : 0;
  c = wave_sum_i64(c);
  if (l == 0) cnt[q] = c;
}

__global__ __launch_bounds__(kThreads) void k_trim_fill(const int64_t* __restrict__ indptr,
                                                        const int32_t* __restrict__ indices,
                                                        const float* __restrict__ data, int64_t n,
                                                        const float* __restrict__ thr,
                                                        const int64_t* __restrict__ out_indptr,
                                                        int32_t* __restrict__ out_indices,
                                                        float* __restrict__ out_data) {
  const int l = lane_id();
  const int64_t q = static_cast<int64_t>(blockIdx.x) * (kThreads / kWave) + threadIdx.x / kWave;
  if (q >= n) return;
  const int64_t b = indptr[q], e = indptr[q + 1];
  const float ti = thr[q];
  int64_t o = out_indptr[q];
  for (int64_t base = b; base < e; base += kWave) {
    const int64_t p = base + l;
    bool keep = false;
    int32_t j = 0;
    float v = 0.0f;
    if (p < e) {
      j = indices[p];
      v = data[p];
      keep = trim_keep(v, ti, thr[j]);
    }
    const uint64_t m = __ballot(keep);
    if (keep) {
      const int64_t r = o + __popcll(m & ((1ull << l) - 1ull));
      out_indices[r] = j;
      out_data[r] = v;
    }
    o += __popcll(m);
  }
}

}  // namespace fuzzy
}  // namespace msc

using namespace msc;
using namespace msc::fuzzy;

extern "C" int msc_fuzzy_workspace_size(int64_t n, int k, size_t* bytes) {
  MSC_TRY(validate(n, k));
  MSC_REQUIRE(bytes, "fuzzy: bytes is null");
  Carve cv(nullptr, 0);
  carve(cv, n, k);
  *bytes = cv.used + 256;
  return kOk;
}

// all kernels of the connectivities pipeline; goff (device, n_groups + 1 row offsets) or null
static int fuzzy_kernels(const int32_t* knn_idx, const float* knn_dist, int64_t n, int k, const int64_t* goff,
                         int n_groups, int64_t* out_indptr, int32_t* out_indices, float* out_data, int64_t capacity,
                         float* out_sigma, float* out_rho, const Bufs& b, hipStream_t st, int64_t* host_nnz) {
  const int64_t m = n * k;
  k_dist_partial<<<dim3(kMeanBlocks, n_groups), kThreads, 0, st>>>(knn_dist, m, k, goff, b.partial);
  k_dist_mean<<<n_groups, 64, 0, st>>>(b.partial, kMeanBlocks, m, k, goff, b.mean);
  MSC_LAUNCH_CHECK();
  const double target = std::log2(static_cast<double>(k));
  if (k <= 64)
    k_smooth_membership<true><<<static_cast<unsigned>(ceil_div(n, 64)), 64, 0, st>>>(
        knn_idx, knn_dist, n, k, target, b.mean, goff, n_groups, out_sigma, out_rho, b.vals);
  else
    k_smooth_membership<false><<<static_cast<unsigned>(ceil_div(n, 64)), 64, 0, st>>>(
        knn_idx, knn_dist, n, k, target, b.mean, goff, n_groups, out_sigma, out_rho, b.vals);
  MSC_LAUNCH_CHECK();
  k_edge_keys<<<static_cast<unsigned>(ceil_div(m, kThreads)), kThreads, 0, st>>>(knn_idx, b.vals, n, k,
                                                                                 b.key, b.eid);
  MSC_LAUNCH_CHECK();
  MSC_TRY(prims::radix_sort_pairs(b.key, b.eid, b.skey, b.seid, m, prims::bits_for(n), b.sort_ws,
                                  b.sort_ws_bytes, st));
  MSC_TRY(prims::segment_starts(b.skey, m, n, b.in_start, st));
  const unsigned wb = static_cast<unsigned>(ceil_div(n, kThreads / kWave));
  const int KP = pow2_at_least(k, kWave);
  const size_t ulds = (kThreads / kWave) * (KP * sizeof(int64_t) + (KP + 1) * sizeof(int32_t));
  if (k <= kWave)
    k_union<false><<<wb, kThreads, 0, st>>>(knn_idx, b.vals, n, k, b.in_start, b.seid, b.row_nnz,
                                            nullptr, nullptr, nullptr, 0);
  else
    k_union_wide<false><<<wb, kThreads, ulds, st>>>(knn_idx, b.vals, n, k, KP, b.in_start, b.seid, b.row_nnz,
                                                    nullptr, nullptr, nullptr, 0);
  MSC_LAUNCH_CHECK();
  MSC_TRY(prims::exclusive_scan<int64_t>(b.row_nnz, out_indptr, n + 1, b.scan_ws, b.scan_ws_bytes, st));
  if (host_nnz) {
    int64_t nnz = 0;
    MSC_HIP(hipMemcpyAsync(&nnz, out_indptr + n, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    MSC_HIP(hipStreamSynchronize(st));
    if (nnz > capacity) {
      set_error("fuzzy: output capacity %lld < nnz %lld", (long long)capacity, (long long)nnz);
      return kInvalidArgument;
    }
    *host_nnz = nnz;
  }
  if (k <= kWave)
    k_union<true><<<wb, kThreads, 0, st>>>(knn_idx, b.vals, n, k, b.in_start, b.seid, nullptr,
                                           out_indptr, out_indices, out_data, capacity);
  else
    k_union_wide<true><<<wb, kThreads, ulds, st>>>(knn_idx, b.vals, n, k, KP, b.in_start, b.seid, nullptr,
                                                   out_indptr, out_indices, out_data, capacity);
  MSC_LAUNCH_CHECK();
  return kOk;
}

extern "C" int msc_fuzzy_connectivities(const int32_t* knn_idx, const float* knn_dist, int64_t n,
                                        int k, int64_t* out_indptr, int32_t* out_indices,
                                        float* out_data, int64_t capacity, float* out_sigma,
                                        float* out_rho, int64_t* out_nnz, void* workspace,
                                        size_t workspace_bytes, void* stream) {
  MSC_TRY(validate(n, k));
  MSC_REQUIRE(knn_idx && knn_dist && out_indptr && out_indices && out_data && out_sigma &&
                  out_rho && out_nnz && workspace,
              "fuzzy: null pointer argument");
  hipStream_t st = static_cast<hipStream_t>(stream);
  Carve cv(workspace, workspace_bytes);
  Bufs b = carve(cv, n, k);
  if (!cv.ok()) {
    set_error("fuzzy: workspace too small (%zu < %zu)", workspace_bytes, cv.used);
    return kWorkspaceTooSmall;
  }
  return fuzzy_kernels(knn_idx, knn_dist, n, k, nullptr, 1, out_indptr, out_indices, out_data, capacity, out_sigma,
                       out_rho, b, st, out_nnz);
}

extern "C" int msc_fuzzy_batched_workspace_size(int64_t n, int k, int n_groups, size_t* bytes) {
  MSC_TRY(validate(n, k));
  MSC_REQUIRE(bytes && n_groups >= 1, "fuzzy_batched: bad arguments");
  Carve cv(nullptr, 0);
  carve(cv, n, k, n_groups);
  *bytes = cv.used + 256;
  return kOk;
}

extern "C" int msc_fuzzy_connectivities_batched(const int32_t* knn_idx, const float* knn_dist, int64_t n, int k,
                                                const int64_t* group_offsets, int n_groups, int64_t* out_indptr,
                                                int32_t* out_indices, float* out_data, int64_t capacity,
                                                float* out_sigma, float* out_rho, void* workspace,
                                                size_t workspace_bytes, void* stream) {
  MSC_TRY(validate(n, k));
  MSC_REQUIRE(knn_idx && knn_dist && group_offsets && n_groups >= 1 && out_indptr && out_indices && out_data &&
                  out_sigma && out_rho && workspace,
              "fuzzy_batched: null pointer argument");
  MSC_REQUIRE(capacity >= 2 * n * (k - 1), "fuzzy_batched: capacity must be >= 2*n*(k-1)");
  hipStream_t st = static_cast<hipStream_t>(stream);
  Carve cv(workspace, workspace_bytes);
  Bufs b = carve(cv, n, k, n_groups);
  if (!cv.ok()) {
    set_error("fuzzy_batched: workspace too small (%zu < %zu)", workspace_bytes, cv.used);
    return kWorkspaceTooSmall;
  }
  return fuzzy_kernels(knn_idx, knn_dist, n, k, group_offsets, n_groups, out_indptr, out_indices, out_data, capacity,
                       out_sigma, out_rho, b, st, nullptr);
}

// ---- the batched connectivities without a workspace --------------------------------------
// Scratch for callers that cannot allocate: one static device array per GPU, held until the
// process ends.  Host code uses the device address from hipGetSymbolAddress (looked up once
// per device, on the first call there), never the array's name.  Calls that overlap on one
// GPU share the array and race; every byte a call reads is written by that call first.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

int scratch_on_current_device(void** out) {
  int dev = 0;
  MSC_HIP(hipGetDevice(&dev));
  MSC_REQUIRE(dev >= 0 && dev < kMaxDevices, "fuzzy_batched_noalloc: device %d beyond %d", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    MSC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    MSC_REQUIRE(p != nullptr, "fuzzy_batched_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return kOk;
}
}  // namespace

extern "C" int msc_fuzzy_connectivities_batched_noalloc(const int32_t* knn_idx, const float* knn_dist, int64_t n,
                                                        int k, const int64_t* group_offsets, int n_groups,
                                                        int64_t* out_indptr, int32_t* out_indices, float* out_data,
                                                        int64_t capacity, float* out_sigma, float* out_rho,
                                                        void* stream) {
  MSC_TRY(validate(n, k));
  MSC_REQUIRE(knn_idx && knn_dist && group_offsets && n_groups >= 1 && out_indptr && out_indices && out_data &&
                  out_sigma && out_rho,
              "fuzzy_batched_noalloc: null pointer argument");
  MSC_REQUIRE(capacity >= 2 * n * (k - 1), "fuzzy_batched_noalloc: capacity must be >= 2*n*(k-1)");
  Carve need(nullptr, 0);
  carve(need, n, k, n_groups);
  MSC_REQUIRE(need.used <= kNoAllocScratchBytes,
              "fuzzy_batched_noalloc: n=%lld k=%d needs %zu scratch bytes; the static scratch holds %zu "
              "(use msc_fuzzy_connectivities_batched with a workspace)",
              (long long)n, k, need.used, kNoAllocScratchBytes);
  void* ws = nullptr;
  MSC_TRY(scratch_on_current_device(&ws));
  Carve cv(ws, kNoAllocScratchBytes);
  Bufs b = carve(cv, n, k, n_groups);
  return fuzzy_kernels(knn_idx, knn_dist, n, k, group_offsets, n_groups, out_indptr, out_indices, out_data, capacity,
                       out_sigma, out_rho, b, static_cast<hipStream_t>(stream), nullptr);
}

extern "C" int msc_trim_graph_workspace_size(int64_t n, size_t* bytes) {
  MSC_REQUIRE(bytes && n >= 1, "trim: bad arguments");
  Carve cv(nullptr, 0);
  cv.take<float>(n);
  cv.take<char>(prims::scan_workspace<int64_t>(n + 1));
  *bytes = cv.used + 256;
  return kOk;
}

extern "C" int msc_trim_graph(const int64_t* indptr, const int32_t* indices, const float* data, int64_t n,
                              int64_t trim, int64_t* out_indptr, int32_t* out_indices, float* out_data,
                              int64_t* out_nnz, void* workspace, size_t workspace_bytes, void* stream) {
  MSC_REQUIRE(indptr && indices && data && out_indptr && out_indices && out_data && workspace,
              "trim: null pointer argument");
  MSC_REQUIRE(n >= 1 && trim >= 1, "trim: need n >= 1 and trim >= 1");
  hipStream_t st = static_cast<hipStream_t>(stream);
  Carve cv(workspace, workspace_bytes);
  float* thr = cv.take<float>(n);
  const size_t sb = prims::scan_workspace<int64_t>(n + 1);
  void* sws = cv.take<char>(sb);
  if (!cv.ok()) {
    set_error("trim: workspace too small (%zu < %zu)", workspace_bytes, cv.used);
    return kWorkspaceTooSmall;
  }
  const unsigned wb = static_cast<unsigned>(ceil_div(n + 1, kThreads / kWave));
  k_trim_thresh<<<wb, kThreads, 0, st>>>(indptr, data, n, trim, thr);
  k_trim_count<<<wb, kThreads, 0, st>>>(indptr, indices, data, n, thr, out_indptr);
  MSC_LAUNCH_CHECK();
  MSC_TRY(prims::exclusive_scan<int64_t>(out_indptr, out_indptr, n + 1, sws, sb, st));
  k_trim_fill<<<wb, kThreads, 0, st>>>(indptr, indices, data, n, thr, out_indptr, out_indices, out_data);
  MSC_LAUNCH_CHECK();
  if (out_nnz) {
    MSC_HIP(hipMemcpyAsync(out_nnz, out_indptr + n, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    MSC_HIP(hipStreamSynchronize(st));
  }
  return kOk;
}

// Cluster co-occurrence pair histogram on gfx950 (all ordered pairs, no global atomics).
//
// Replaces the fused histogram of spatial.co_occurrence (src/metalsinglecell/spatial.py:123-
// 171, kernel :123-146) as called by gr.co_occurrence (gr.py:514-530): for every ordered pair
// (i, j), i != j, the first threshold bin r with thr2[r] >= d2(i, j) (numpy searchsorted
// 'left'), counted into hist[code_i, code_j, r]; pairs beyond the last threshold are dropped.
// The caller turns the per-bin counts into the cumulative within-radius ratios.
//
// d2 is the separately rounded float32 sum of squared coordinate differences (the spatial kNN
// definition), not the reference's Gram form |x|^2 + |y|^2 - 2 x.y, which cancels at
// slide-scale coordinates.
//
// Layout: rows sorted by code (stable) into a device copy; query tiles of 256 rows never
// straddle two codes, so a block's histogram is the single K x L row of its tile's code and
// lives in LDS (per-workgroup LDS atomics).  Candidates stream through LDS 256 at a time;
// lane t takes candidate (t + s) mod 256 at step s so that the lanes of a wave hit different
// candidates (and mostly different bins) at once.  Tile histograms are summed per code in
// tile order, as 64-bit integers.
#include <cmath>
#include <vector>

#include "common.h"
#include "msc_hip.h"

namespace msc {
namespace cooccur {

constexpr int kThreads = 256;
constexpr int kMaxBins = 16384;  // K * L entries of the LDS histogram (64 KB)

// Xs[p] = X[perm[p]] (coordinates in code order)
__global__ void k_gather_rows(const float* __restrict__ X, const int32_t* __restrict__ perm, int64_t n, int dim,
                              float4* __restrict__ Xs, const int32_t* __restrict__ cs) {
  const int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (p >= n) return;
  const int64_t i = perm[p];
  float4 v;
  v.x = X[i * dim];
  v.y = X[i * dim + 1];
  v.z = dim == 3 ? X[i * dim + 2] : 0.0f;
  v.w = __int_as_float(cs[p]);
  Xs[p] = v;
}

__global__ __launch_bounds__(kThreads) void k_pair_hist(const float4* __restrict__ Xs, int64_t n, int dim,
                                                        const int64_t* __restrict__ tiles,
                                                        const float* __restrict__ thr2, int L, int K, int pow2L,
                                                        float ra, float rinv, uint32_t* __restrict__ partial) {
  extern __shared__ __align__(16) unsigned char smem[];
  float4* cand = reinterpret_cast<float4*>(smem);                          // kThreads
  float* thr = reinterpret_cast<float*>(smem + kThreads * sizeof(float4));  // L
  uint32_t* hist = reinterpret_cast<uint32_t*>(thr + L);                   // K * L
  const int t = threadIdx.x;
  const int KL = K * L;
  for (int e = t; e < KL; e += kThreads) hist[e] = 0u;
  for (int e = t; e < L; e += kThreads) thr[e] = thr2[e];
  const int64_t b0 = tiles[2 * blockIdx.x], b1 = tiles[2 * blockIdx.x + 1];
  const int64_t p = b0 + t;
  const bool active = p < b1;
  float4 q = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (active) q = Xs[p];
  for (int64_t base = 0; base < n; base += kThreads) {
    __syncthreads();  // thresholds and histogram zeroed / previous chunk consumed
    const int64_t j = base + t;
    cand[t] = j < n ? Xs[j] : make_float4(0.0f, 0.0f, 0.0f, __int_as_float(-1));
    __syncthreads();
    if (!active) continue;
    const float tmax = thr[L - 1];
    for (int s = 0; s < kThreads; ++s) {
      const int jj = (t + s) & (kThreads - 1);
      const float4 c = cand[jj];
      const int cj = __float_as_int(c.w);
      if (cj < 0 || base + jj == p) continue;
      const float dx = __fsub_rn(c.x, q.x);
      const float dy = __fsub_rn(c.y, q.y);
      float d2 = __fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy));
      if (dim == 3) {
        const float dz = __fsub_rn(c.z, q.z);
        d2 = __fadd_rn(d2, __fmul_rn(dz, dz));
      }
      if (!(d2 <= tmax)) continue;
      // number of thresholds below d2 = the first r with thr[r] >= d2 (thr[L - 1] >= d2 here)
      int pos;
      if (rinv > 0.0f) {
        // thresholds evenly spaced in distance: estimate ceil((sqrt(d2) - ra) * rinv), then walk
        // to the exact bin, so the result is the search's for any estimate
        const float est = ceilf((__builtin_sqrtf(d2) - ra) * rinv);
        pos = est <= 0.0f ? 0 : est >= static_cast<float>(L - 1) ? L - 1 : static_cast<int>(est);
        while (pos > 0 && thr[pos - 1] >= d2) --pos;
        while (thr[pos] < d2) ++pos;
      } else {
        pos = 0;  // branchless binary search
        for (int step = pow2L; step > 0; step >>= 1) {
          const int np = pos + step;
          if (np <= L && thr[np - 1] < d2) pos = np;
        }
      }
      atomicAdd(&hist[cj * L + pos], 1u);
    }
  }
  __syncthreads();
  uint32_t* out = partial + static_cast<int64_t>(blockIdx.x) * KL;
  for (int e = t; e < KL; e += kThreads) out[e] = hist[e];
}

// out[c][k][r] = sum over the tiles of code c (tiles [tc[c], tc[c+1]) ) of partial[tile][k][r]
__global__ void k_sum_tiles(const uint32_t* __restrict__ partial, const int32_t* __restrict__ tc, int K, int L,
                            int64_t* __restrict__ out) {
  const int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  const int64_t KL = static_cast<int64_t>(K) * L;
  if (e >= K * KL) return;
  const int c = static_cast<int>(e / KL);
  const int64_t f = e - c * KL;
  int64_t s = 0;
  for (int tile = tc[c]; tile < tc[c + 1]; ++tile) s += partial[tile * KL + f];
  out[e] = s;
}

struct Layout {
  std::vector<int32_t> perm, codes_sorted, tile_code_start;
  std::vector<int64_t> tiles;  // [begin, end) per tile
  int64_t ntiles = 0;
};

static int make_layout(const int32_t* codes, int64_t n, int K, Layout* Lo) {
  std::vector<int64_t> cnt(K, 0);
  for (int64_t i = 0; i < n; ++i) {
    MSC_REQUIRE(codes[i] >= 0 && codes[i] < K, "co_occurrence: code %d of row %lld outside [0, %d)", codes[i],
                (long long)i, K);
    ++cnt[codes[i]];
  }
  std::vector<int64_t> off(K + 1, 0);
  for (int c = 0; c < K; ++c) off[c + 1] = off[c] + cnt[c];
  Lo->perm.resize(n);
  Lo->codes_sorted.resize(n);
  std::vector<int64_t> pos(off.begin(), off.end() - 1);
  for (int64_t i = 0; i < n; ++i) {
    const int64_t p = pos[codes[i]]++;
    Lo->perm[p] = static_cast<int32_t>(i);
    Lo->codes_sorted[p] = codes[i];
  }
  Lo->tile_code_start.assign(K + 1, 0);
  Lo->tiles.clear();
  for (int c = 0; c < K; ++c) {
    Lo->tile_code_start[c] = static_cast<int32_t>(Lo->tiles.size() / 2);
    for (int64_t b = off[c]; b < off[c + 1]; b += kThreads) {
      Lo->tiles.push_back(b);
      Lo->tiles.push_back(std::min(off[c + 1], b + kThreads));
    }
  }
  Lo->tile_code_start[K] = static_cast<int32_t>(Lo->tiles.size() / 2);
  Lo->ntiles = static_cast<int64_t>(Lo->tiles.size() / 2);
  return kOk;
}

struct Bufs {
  int32_t *perm, *cs, *tc;
  int64_t* tiles;
  float* thr2;
  float4* Xs;
  uint32_t* partial;
};

static Bufs carve(Carve& cv, int64_t n, int K, int L, int64_t ntiles) {
  Bufs b{};
  b.perm = cv.take<int32_t>(n);
  b.cs = cv.take<int32_t>(n);
  b.tc = cv.take<int32_t>(K + 1);
  b.tiles = cv.take<int64_t>(2 * std::max<int64_t>(ntiles, 1));
  b.thr2 = cv.take<float>(L);
  b.Xs = cv.take<float4>(n);
  b.partial = cv.take<uint32_t>(static_cast<size_t>(std::max<int64_t>(ntiles, 1)) * K * L);
  return b;
}

static int validate(const int32_t* codes, int64_t n, int dim, int K, int L) {
  MSC_REQUIRE(codes && n >= 1 && n < (int64_t(1) << 24), "co_occurrence: need 1 <= n < 2^24 and host codes");
  MSC_REQUIRE(dim == 2 || dim == 3, "co_occurrence: dim must be 2 or 3");
  MSC_REQUIRE(K >= 1 && L >= 1 && static_cast<int64_t>(K) * L <= kMaxBins,
              "co_occurrence: need K >= 1, L >= 1 and K * L <= %d", kMaxBins);
  return kOk;
}

}  // namespace cooccur
}  // namespace msc

using namespace msc;
using namespace msc::cooccur;

extern "C" int msc_co_occurrence_workspace_size(const int32_t* codes, int64_t n, int dim, int n_codes, int n_bins,
                                                size_t* bytes) {
  MSC_TRY(validate(codes, n, dim, n_codes, n_bins));
  MSC_REQUIRE(bytes, "co_occurrence: bytes is null");
  Layout Lo;
  MSC_TRY(make_layout(codes, n, n_codes, &Lo));
  Carve cv(nullptr, 0);
  carve(cv, n, n_codes, n_bins, Lo.ntiles);
  *bytes = cv.used + 256;
  return kOk;
}

extern "C" int msc_co_occurrence(const float* coords, int64_t n, int dim, const int32_t* codes, int n_codes,
                                 const float* thr2, int n_bins, int64_t* out_hist, void* workspace,
                                 size_t workspace_bytes, void* stream) {
  MSC_TRY(validate(codes, n, dim, n_codes, n_bins));
  MSC_REQUIRE(coords && thr2 && out_hist && workspace, "co_occurrence: null pointer argument");
  for (int r = 0; r < n_bins; ++r) {
    MSC_REQUIRE(!std::isnan(thr2[r]), "co_occurrence: thresholds must not be NaN");
    MSC_REQUIRE(r == 0 || thr2[r - 1] <= thr2[r], "co_occurrence: thresholds must ascend");
  }
  const int K = n_codes, L = n_bins;
  Layout Lo;
  MSC_TRY(make_layout(codes, n, K, &Lo));
  hipStream_t st = static_cast<hipStream_t>(stream);
  Carve cv(workspace, workspace_bytes);
  Bufs b = carve(cv, n, K, L, Lo.ntiles);
  if (!cv.ok()) {
    set_error("co_occurrence: workspace too small (%zu < %zu)", workspace_bytes, cv.used);
    return kWorkspaceTooSmall;
  }
  MSC_HIP(hipMemcpyAsync(b.perm, Lo.perm.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, st));
  MSC_HIP(hipMemcpyAsync(b.cs, Lo.codes_sorted.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, st));
  MSC_HIP(hipMemcpyAsync(b.tc, Lo.tile_code_start.data(), (K + 1) * sizeof(int32_t), hipMemcpyHostToDevice, st));
  MSC_HIP(hipMemcpyAsync(b.tiles, Lo.tiles.data(), Lo.tiles.size() * sizeof(int64_t), hipMemcpyHostToDevice, st));
  MSC_HIP(hipMemcpyAsync(b.thr2, thr2, L * sizeof(float), hipMemcpyHostToDevice, st));
  MSC_HIP(hipStreamSynchronize(st));  // the host tables are pageable and local to this call
  k_gather_rows<<<static_cast<unsigned>(ceil_div(n, kThreads)), kThreads, 0, st>>>(coords, b.perm, n, dim, b.Xs,
                                                                                   b.cs);
  MSC_LAUNCH_CHECK();
  int pow2L = 1;
  while (pow2L < L) pow2L <<= 1;
  // Thresholds evenly spaced in distance (the default linspace intervals) let the kernel start
  // each bin lookup from an arithmetic estimate; any other spacing uses the binary search.
  float ra = 0.0f, rinv = 0.0f;
  if (L >= 2 && thr2[0] >= 0.0f) {
    const double a = std::sqrt(static_cast<double>(thr2[0]));
    const double h = (std::sqrt(static_cast<double>(thr2[L - 1])) - a) / (L - 1);
    bool even = h > 0.0;
    for (int r = 0; r < L && even; ++r)
      even = std::fabs(std::sqrt(static_cast<double>(thr2[r])) - (a + r * h)) <= 0.25 * h;
    if (even) {
      ra = static_cast<float>(a);  // t_r = a + r h: #{r : t_r < x} = ceil((x - a) / h)
      rinv = static_cast<float>(1.0 / h);
    }
  }
  const size_t shm = kThreads * sizeof(float4) + L * sizeof(float) + static_cast<size_t>(K) * L * sizeof(uint32_t);
  MSC_REQUIRE(shm <= 65536, "co_occurrence: %d codes x %d bins exceed the 64 KB workgroup histogram", K, L);
  k_pair_hist<<<static_cast<unsigned>(Lo.ntiles), kThreads, shm, st>>>(b.Xs, n, dim, b.tiles, b.thr2, L, K, pow2L,
                                                                       ra, rinv, b.partial);
  MSC_LAUNCH_CHECK();
  const int64_t m = static_cast<int64_t>(K) * K * L;
  k_sum_tiles<<<static_cast<unsigned>(ceil_div(m, kThreads)), kThreads, 0, st>>>(b.partial, b.tc, K, L, out_hist);
  MSC_LAUNCH_CHECK();
  return kOk;
}

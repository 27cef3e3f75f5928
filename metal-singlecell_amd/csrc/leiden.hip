// Synchronous parallel Leiden on gfx950, with exact fixed-point weights.
//
// Replaces graph.leiden.leiden / _leiden_pass (src/metalsinglecell/graph/leiden.py:352-455)
// with the move and refine kernels (graph/louvain.py:88-131, graph/leiden.py:100-144), the
// synchronous pass loop (louvain.py:282-394, leiden.py:251-349), contraction
// (louvain.py:440-453) and modularity (graph/primitives.py:94-111).
//
// Arithmetic (identical to oracle/leiden.py): weights quantised once to int64 fixed point
// (wq = rint(w * 2^s), s = 52 - ceil(log2 max w) - ceil(log2 E)), so degrees, sigma_tot,
// community weights and contracted weights are exact integer sums; gains are one fixed
// float64 expression; the coin is the counter RNG of common.h.  Every reduction is a
// sort-based segmented sum (stable radix sort + int64 prefix differences) or a per-vertex
// wave sum: no atomics anywhere, so labels are bit-identical to the oracle and independent
// of scheduling.
//
// Per-vertex best move, one launch per pass (k_move): vertices of degree <= 256 take one wave
// (up to four neighbours per lane, sorted inside the wave); degree 257..1024 and 1025..4096
// take one 256-thread block each from a list (4 / 16 neighbours per thread, sorted across the
// block).  Heavier vertices (and the list ranges whose community ids would not fit the sort
// key: 2^22 and 2^20 vertices and up) go through a sort path over their edges keyed by
// (vertex, community).
#include <cmath>
#include <vector>

#include "common.h"
#include "msc_hip.h"
#include "prims.h"

namespace msc {
namespace leiden {

constexpr int kThreads = 256;
constexpr int kWavesPerBlock = kThreads / kWave;
constexpr int kLightItems = 4;                  // neighbours per lane on the one-wave path
constexpr int kLightDeg = kWave * kLightItems;  // 256
constexpr int kMaxLevels = 20;
constexpr int kMaxPasses = 200;
constexpr int kSyncEvery = 4;
constexpr double kTol = 1e-9;

struct Graph {
  int64_t n = 0, E = 0;
  int64_t* indptr = nullptr;
  int32_t* indices = nullptr;
  int64_t* wq = nullptr;
  int64_t* k = nullptr;       // weighted degree (n)
  int64_t twom = 0;
  int64_t max_deg = 0;
};

// Per-vertex record for the light move kernel: a neighbour's community, refinement part
// and that community's sigma_tot in one 16-byte gather (instead of three dependent ones).
struct alignas(16) CommSig {
  int32_t comm;
  int32_t part;
  int64_t sig;
};

struct GainParams {
  double inv_s;   // 2^-s
  double inv;     // 1 / (twom * 2^-s)
  double gamma;
  // batched runs (many samples in one block-diagonal graph): each vertex takes its sample's
  // scale and total weight; null for a single graph
  const int32_t* vsid = nullptr;   // sample of each vertex
  const double* s_inv_s = nullptr; // per sample
  const double* s_inv = nullptr;   // per sample
};

__device__ __forceinline__ double gain_of(int64_t wc, int64_t sig, int64_t kv, const GainParams& gp) {
  const double W = static_cast<double>(wc) * gp.inv_s;
  const double S = static_cast<double>(sig) * gp.inv_s;
  const double K = static_cast<double>(kv) * gp.inv_s;
  return W - ((gp.gamma * S) * K) * gp.inv;
}

// the gain parameters of vertex v (its sample's, in a batched run)
__device__ __forceinline__ GainParams for_vertex(const GainParams& gp, int64_t v) {
  GainParams g = gp;
  if (gp.vsid) {
    const int32_t sid = gp.vsid[v];
    g.inv_s = gp.s_inv_s[sid];
    g.inv = gp.s_inv[sid];
  }
  return g;
}

// better(g1, c1) than (g2, c2): larger gain, ties to the smaller community
__device__ __forceinline__ bool better(double g1, int32_t c1, double g2, int32_t c2) {
  return g1 > g2 || (g1 == g2 && c1 < c2);
}

// ------------------------------------------------------------------------------------------
// simple element-wise kernels
// ------------------------------------------------------------------------------------------
__global__ void k_fill_i32(int32_t* p, int64_t n, int32_t v) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_iota_i32(int32_t* p, int64_t n) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i < n) p[i] = static_cast<int32_t>(i);
}
__global__ void k_fill_u8(uint8_t* p, int64_t n, uint8_t v) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_gather_i32(const int32_t* __restrict__ src, const int32_t* __restrict__ idx, int32_t* __restrict__ dst, int64_t n) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}
__global__ void k_copy_i32(const int32_t* __restrict__ src, int32_t* __restrict__ dst, int64_t n) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

static unsigned grid_for(int64_t n) { return static_cast<unsigned>(std::max<int64_t>(1, ceil_div(n, kThreads))); }

// ------------------------------------------------------------------------------------------
// graph set-up
// ------------------------------------------------------------------------------------------
// max weight: per-block maxima (grid-stride) into out[block], then one block over them;
// a maximum does not depend on the order of evaluation.  Non-finite and negative weights
// map to +inf so the caller's finiteness check rejects them.
constexpr int kMaxBlocks = 1024;
__global__ void k_max_weight(const float* __restrict__ w, int64_t E, float* __restrict__ out) {
  float m = 0.0f;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < E;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
  {
    const float x = w[i];
    m = fmaxf(m, (isfinite(x) && x >= 0.0f) ? x : __uint_as_float(0x7F800000u));
  }
  m = wave_max_f32(m);
  __shared__ float sh[kWavesPerBlock];
  if (lane_id() == 0) sh[threadIdx.x / kWave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < kWavesPerBlock; ++i) m = fmaxf(m, sh[i]);
    out[blockIdx.x] = m;
  }
}

__global__ void k_quantize(const float* __restrict__ w, int64_t E, int s, int64_t* __restrict__ wq) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i < E) wq[i] = static_cast<int64_t>(rint(ldexp(static_cast<double>(w[i]), s)));
}

// weighted degree per vertex (wave per vertex, exact) and the degree
__global__ __launch_bounds__(kThreads) void k_degrees(const int64_t* __restrict__ indptr, const int64_t* __restrict__ wq,
                                                      int64_t n, int64_t* __restrict__ k) {
  const int64_t v = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave;
  if (v >= n) return;
  int64_t s = 0;
  for (int64_t e = indptr[v] + lane_id(); e < indptr[v + 1]; e += kWave) s += wq[e];
  s = wave_sum_i64(s);
  if (lane_id() == 0) k[v] = s;
}

// block partial sums (int64, exact) and max degree
__global__ void k_sum_i64_max_deg(const int64_t* __restrict__ k, const int64_t* __restrict__ indptr, int64_t n,
                                  int64_t* __restrict__ out /*[0]=sum,[1]=maxdeg*/) {
  int64_t s = 0, md = 0;
#pragma unroll 8
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    s += k[i];
    md = max(md, indptr[i + 1] - indptr[i]);
  }
  s = wave_sum_i64(s);
#pragma unroll
  for (int m = 1; m < kWave; m <<= 1) md = max(md, static_cast<int64_t>(shfl_xor_u64(static_cast<uint64_t>(md), m)));
  __shared__ int64_t ss[kWavesPerBlock], sm[kWavesPerBlock];
  if (lane_id() == 0) { ss[threadIdx.x / kWave] = s; sm[threadIdx.x / kWave] = md; }
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t a = 0, b = 0;
    for (int i = 0; i < kWavesPerBlock; ++i) { a += ss[i]; b = max(b, sm[i]); }
    out[0] = a;
    out[1] = b;
  }
}

// ------------------------------------------------------------------------------------------
// grouped sums by label (sort-based, exact): out[c] = sum_{v: lab[v]=c} val[v]
// ------------------------------------------------------------------------------------------
struct GroupScratch {
  uint32_t *keys, *vals, *skeys, *svals;
  int64_t* gathered;  // n + 1
  int64_t* prefix;    // n + 1
  int64_t* starts;    // C + 1
  void* sort_ws;
  size_t sort_ws_bytes;
  void* scan_ws;
  size_t scan_ws_bytes;
};

__global__ void k_label_keys(const int32_t* __restrict__ lab, int64_t n, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i < n) {
    keys[i] = static_cast<uint32_t>(lab[i]);
    vals[i] = static_cast<uint32_t>(i);
  }
}
__global__ void k_gather_i64_perm(const int64_t* __restrict__ src, const uint32_t* __restrict__ perm, int64_t n,
                                  int64_t* __restrict__ dst) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i < n) dst[i] = src[perm[i]];
  if (i == n) dst[n] = 0;
}
__global__ void k_seg_diff(const int64_t* __restrict__ prefix, const int64_t* __restrict__ starts, int64_t C,
                           int64_t* __restrict__ out) {
  const int64_t c = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (c < C) out[c] = prefix[starts[c + 1]] - prefix[starts[c]];
}

// sort (label, vertex) once; then any number of value arrays can be grouped
static int group_sort(const int32_t* lab, int64_t n, int64_t C, GroupScratch& g, hipStream_t st) {
  k_label_keys<<<grid_for(n), kThreads, 0, st>>>(lab, n, g.keys, g.vals);
  MSC_LAUNCH_CHECK();
  MSC_TRY(prims::radix_sort_pairs(g.keys, g.vals, g.skeys, g.svals, n, prims::bits_for(std::max<int64_t>(C - 1, 1)),
                                  g.sort_ws, g.sort_ws_bytes, st));
  MSC_TRY(prims::segment_starts(g.skeys, n, C, g.starts, st));
  return kOk;
}
static int group_sum(const int64_t* val, int64_t n, int64_t C, GroupScratch& g, int64_t* out, hipStream_t st) {
  k_gather_i64_perm<<<grid_for(n + 1), kThreads, 0, st>>>(val, g.svals, n, g.gathered);
  MSC_LAUNCH_CHECK();
  MSC_TRY(prims::exclusive_scan<int64_t>(g.gathered, g.prefix, n + 1, g.scan_ws, g.scan_ws_bytes, st));
  k_seg_diff<<<grid_for(C), kThreads, 0, st>>>(g.prefix, g.starts, C, out);
  MSC_LAUNCH_CHECK();
  return kOk;
}

// ------------------------------------------------------------------------------------------
// best move, light vertices (degree <= 256): one wave per vertex
// ------------------------------------------------------------------------------------------
// Lane l holds neighbours l, l + 64, ... (NI of them) with their community, edge weight and
// that community's sigma_tot, all gathered up front.  Each round takes the lowest pending
// (item, lane), sums the weights of every neighbour in its community (exact int64 wave sum)
// and scores the move; rounds = distinct neighbour communities.
template <int NI>
__device__ __forceinline__ void move_one(int64_t v, int64_t e0, int64_t deg, int32_t curr, int64_t kv,
                                         const int32_t* __restrict__ indices, const int64_t* __restrict__ wq,
                                         const int32_t* __restrict__ comm, const int32_t* __restrict__ part,
                                         const int64_t* __restrict__ sigtot, const GainParams& gp_all,
                                         int32_t* __restrict__ target) {
  const GainParams gp = for_vertex(gp_all, v);
  const int l = lane_id();
  const int32_t pv = part ? part[v] : 0;
  bool ok[NI];
  int32_t c[NI];
  int64_t w[NI], sg[NI];
  uint64_t pend[NI];
#pragma unroll
  for (int i = 0; i < NI; ++i) {
    ok[i] = false;
    c[i] = -1;
    w[i] = 0;
    sg[i] = 0;
    const int64_t j = l + kWave * i;
    if (j < deg) {
      const int32_t u = indices[e0 + j];
      ok[i] = u != static_cast<int32_t>(v) && (!part || part[u] == pv);
      if (ok[i]) {
        c[i] = comm[u];
        w[i] = wq[e0 + j];
      }
    }
  }
#pragma unroll
  for (int i = 0; i < NI; ++i) {
    if (ok[i]) sg[i] = sigtot[c[i]];
    pend[i] = __ballot(ok[i]);
  }
  double best_g = -1e300;
  int32_t best_c = curr;
  double stay = 0.0;
  while (true) {
    int it = -1;
#pragma unroll
    for (int i = NI - 1; i >= 0; --i)
      if (pend[i]) it = i;
    if (it < 0) break;
    int32_t cc = 0;
    int64_t sig = 0;
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      if (i == it) {
        const int leader = __ffsll(static_cast<long long>(pend[i])) - 1;
        cc = __shfl(c[i], leader, kWave);
        sig = static_cast<int64_t>(shfl_u64(static_cast<uint64_t>(sg[i]), leader));
      }
    }
    int64_t part_w = 0;
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const bool m = ok[i] && c[i] == cc;
      part_w += m ? w[i] : 0;
      pend[i] &= ~__ballot(m);
    }
    const int64_t wc = wave_sum_i64(part_w);
    sig -= (cc == curr ? kv : 0);
    const double g = gain_of(wc, sig, kv, gp);
    if (cc == curr) stay = g;
    if (better(g, cc, best_g, best_c)) {
      best_g = g;
      best_c = cc;
    }
  }
  if (l == 0) target[v] = (best_g > stay + kTol) ? best_c : curr;
}

// Same decision as move_one, at a cost independent of the number of distinct communities:
// the wave sorts keys (community << IB | slot) of its up to 64*NI eligible neighbours
// (bitonic network, slots = item * 64 + lane), sums the weights of each run of equal
// community with an int64 prefix sum (staged in the wave's LDS), scores each community
// once at the end of its run and takes the wave argmax (ties to the smaller community).
// Requires community ids < 2^(32 - IB) so that a key fits 32 bits.
constexpr int64_t kSortedMaxN = int64_t(1) << 24;  // NI <= 4
// Listed vertices of degree 257..4096 take one 256-thread block each (k_move), in four
// classes: class i holds degrees (256 << i, 512 << i], sorted as N = 512 << i keys of
// (community << (9 + i) | slot), ITEMS = 2 << i per thread.  Class i needs community ids below
// 2^(23 - i) (strictly below, so that no key equals kNoKey), so it is used for graphs of fewer
// than hub_class_max_n(i) vertices.
constexpr int kHubClasses = 4;
constexpr int kHubMaxDeg = 512 << (kHubClasses - 1);  // 4096
__host__ __device__ constexpr int64_t hub_class_max_n(int i) { return int64_t(1) << (23 - i); }
constexpr int kMoveGroup = 8;  // consecutive vertices per wave on the light path of k_move

// Value of lane l ^ M for M = 2^j - 1 (the mirror image of l within its block of 2^j lanes):
// M = 1, 3: DPP quad_perm; 7: row_half_mirror; 15: row_mirror; 31, 63: row_mirror followed
// by the stride-16 (and stride-32) exchange.
template <int M>
__device__ __forceinline__ uint32_t lane_mirror(uint32_t v) {
  const int x = static_cast<int>(v);
  if constexpr (M == 1) {
    return lane_xor<1>(v);
  } else if constexpr (M == 3) {
    return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, x, 0x1B, 0xF, 0xF, true));
  } else if constexpr (M == 7) {
    return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, x, 0x141, 0xF, 0xF, true));
  } else if constexpr (M == 15) {
    return static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, x, 0x140, 0xF, 0xF, true));
  } else if constexpr (M == 31) {
    return lane_xor<16>(lane_mirror<15>(v));
  } else {
    static_assert(M == 63, "lane_mirror: M must be 2^j - 1 below 64");
    return lane_xor<32>(lane_xor<16>(lane_mirror<15>(v)));
  }
}

// Bitonic sort of the 64 * NI slots e = i * 64 + l of a wave (keys in registers), ascending,
// in the form whose comparators all point the same way: merge phase K first compares slot e
// with its mirror e ^ (K - 1) inside the block of K slots, then e with e ^ J for J = K / 4 ..
// 1; the lower slot of each pair keeps the minimum.  The choice between minimum and maximum
// then depends on the stride only (six lane masks), not on the stage.  The stage sequence
// is unrolled by template recursion so that every register index is a compile-time constant.
template <int NI, int J>
__device__ __forceinline__ void half_clean(uint32_t (&key)[NI], int l) {
  if constexpr (J >= kWave) {  // partner slot in item i ^ (J / 64) of the same lane
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const int ip = i ^ (J / kWave);
      if (ip > i) {
        const uint32_t a = key[i], b = key[ip];
        key[i] = a < b ? a : b;
        key[ip] = a < b ? b : a;
      }
    }
  } else {  // partner slot in lane l ^ J of the same item
    const bool lower = (l & J) == 0;
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const uint32_t p = lane_xor<J>(key[i]);
      const uint32_t lo = key[i] < p ? key[i] : p, hi = key[i] < p ? p : key[i];
      key[i] = lower ? lo : hi;
    }
  }
  if constexpr (J > 1) half_clean<NI, J / 2>(key, l);
}
template <int NI, int K>
__device__ __forceinline__ void bitonic_sort_slots(uint32_t (&key)[NI], int l) {
  if constexpr (K <= kWave) {  // mirror partner in lane l ^ (K - 1) of the same item
    const bool lower = (l & (K / 2)) == 0;
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const uint32_t p = lane_mirror<K - 1>(key[i]);
      const uint32_t lo = key[i] < p ? key[i] : p, hi = key[i] < p ? p : key[i];
      key[i] = lower ? lo : hi;
    }
  } else {  // mirror partner in item i ^ (K / 64 - 1), lane l ^ 63; item i < ip is the lower one
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const int ip = i ^ (K / kWave - 1);
      if (ip > i) {
        const uint32_t t = lane_mirror<63>(key[ip]);
        const uint32_t lo = key[i] < t ? key[i] : t, hi = key[i] < t ? t : key[i];
        key[i] = lo;
        key[ip] = lane_mirror<63>(hi);
      }
    }
  }
  if constexpr (K >= 4) half_clean<NI, K / 4>(key, l);
  if constexpr (K < NI * kWave) bitonic_sort_slots<NI, 2 * K>(key, l);
}

constexpr uint32_t kNoKey = 0xFFFFFFFFu;

__device__ __forceinline__ int64_t readlane_i64(int64_t x, int lane) {
  const uint64_t u = static_cast<uint64_t>(x);
  const uint32_t lo = __builtin_amdgcn_readlane(static_cast<int>(u), lane);
  const uint32_t hi = __builtin_amdgcn_readlane(static_cast<int>(u >> 32), lane);
  return static_cast<int64_t>((static_cast<uint64_t>(hi) << 32) | lo);
}

__device__ __forceinline__ int64_t sig_of(const int4& r) {
  return static_cast<int64_t>((static_cast<uint64_t>(static_cast<uint32_t>(r.w)) << 32) |
                              static_cast<uint32_t>(r.z));
}

// one butterfly step of the wave maximum of a double
template <int M>
__device__ __forceinline__ double max_step(double g) {
  const double o = __longlong_as_double(
      static_cast<long long>(lane_xor_u64<M>(static_cast<uint64_t>(__double_as_longlong(g)))));
  return o > g ? o : g;
}
template <int M>
__device__ __forceinline__ uint32_t min_step(uint32_t c) {
  const uint32_t o = lane_xor<M>(c);
  return o < c ? o : c;
}

// Wave argmax over (gain, community), ties to the smaller community: the largest gain, then
// the smallest community among the lanes that hold it (the lexicographic maximum of
// (gain, -community), the order of better()).  Communities are non-negative.
__device__ __forceinline__ void wave_argmax(double& g, int32_t& c) {
  double m = max_step<1>(g);
  m = max_step<2>(m);
  m = max_step<4>(m);
  m = max_step<8>(m);
  m = max_step<16>(m);
  m = max_step<32>(m);
  uint32_t k = g == m ? static_cast<uint32_t>(c) : 0x7FFFFFFFu;
  k = min_step<1>(k);
  k = min_step<2>(k);
  k = min_step<4>(k);
  k = min_step<8>(k);
  k = min_step<16>(k);
  k = min_step<32>(k);
  g = m;
  c = static_cast<int32_t>(k);
}

// The decision for vertex v from its eligible neighbour slots e = i * 64 + lane: key[i] =
// (community << IB | e) or kNoKey, their weights in lw[e] and sigma_tot in lsg[e].
template <int NI>
__device__ __forceinline__ void move_core(int64_t v, uint32_t (&key)[NI], int nvalid, int32_t curr, int64_t kv,
                                          const GainParams& gp_all, int32_t* __restrict__ target,
                                          const int64_t* __restrict__ lw, const int64_t* __restrict__ lsg,
                                          int64_t* __restrict__ lp) {
  const GainParams gp = for_vertex(gp_all, v);
  constexpr int IB = NI == 1 ? 6 : NI == 2 ? 7 : 8;
  static_assert(NI == 1 || NI == 2 || NI == 4, "NI");
  const int l = lane_id();
  bitonic_sort_slots<NI, 2>(key, l);  // ascending in slot order e = i * 64 + l

  // per slot: weight prefix and run heads (community = key >> IB)
  uint64_t hmask[NI];
  int64_t carry = 0;
  int32_t c_last = -1;  // community of slot (i * 64 - 1)
#pragma unroll
  for (int i = 0; i < NI; ++i) {
    const int e = i * kWave + l;
    const bool valid = e < nvalid;
    const uint32_t src = key[i] & ((1u << IB) - 1u);
    const int32_t ci = valid ? static_cast<int32_t>(key[i] >> IB) : -1;
    const int64_t w = valid ? lw[src] : 0;
    // inclusive prefix of w over the slots
    const int64_t x = wave_incl_sum_i64(w) + carry;
    carry = readlane_i64(x, kWave - 1);
    lp[e] = x;
    int32_t cprev = __shfl_up(ci, 1, kWave);
    if (l == 0) cprev = c_last;
    hmask[i] = __ballot(valid && (e == 0 || cprev != ci));
    c_last = __builtin_amdgcn_readlane(ci, kWave - 1);
  }

  double best_g = -1e300, stay = 0.0;
  int32_t best_c = 0x7FFFFFFF;
  bool has_stay = false;
  int64_t head_carry = 0;  // last run head before item i
#pragma unroll
  for (int i = 0; i < NI; ++i) {
    const int e = i * kWave + l;
    const bool valid = e < nvalid;
    const uint64_t hmi = hmask[i];                                   // run heads of this item
    const uint64_t hmn = i + 1 < NI ? hmask[i + 1 < NI ? i + 1 : i] : 0ull;  // and of the next one
    const bool next_head = (l < kWave - 1) ? ((hmi >> (l + 1)) & 1ull) != 0 : (hmn & 1ull) != 0;
    const bool end = valid && (e + 1 == nvalid || next_head);
    const uint64_t upto = (l == kWave - 1) ? ~0ull : ((2ull << l) - 1ull);
    const uint64_t hm = hmi & upto;
    const int64_t start = hm ? static_cast<int64_t>(i * kWave + 63 - __clzll(static_cast<long long>(hm))) : head_carry;
    if (hmi) head_carry = i * kWave + 63 - __clzll(static_cast<long long>(hmi));
    if (end) {
      const int32_t ci = static_cast<int32_t>(key[i] >> IB);
      const int64_t wc = lp[e] - (start > 0 ? lp[start - 1] : 0);
      const int64_t sig = lsg[key[i] & ((1u << IB) - 1u)] - (ci == curr ? kv : 0);
      const double g = gain_of(wc, sig, kv, gp);
      if (ci == curr) { stay = g; has_stay = true; }
      if (better(g, ci, best_g, best_c)) { best_g = g; best_c = ci; }
    }
  }
  wave_argmax(best_g, best_c);
  const uint64_t sm = __ballot(has_stay);
  if (sm) {
    const int src = __ffsll(static_cast<long long>(sm)) - 1;
    stay = __longlong_as_double(static_cast<long long>(readlane_i64(__double_as_longlong(stay), src)));
  }
  if (l == 0) target[v] = (best_g > stay + kTol) ? best_c : curr;
}

// one neighbour slot: (community, weight, sigma_tot) of neighbour j of v when eligible
template <int IB>
__device__ __forceinline__ void gather_slot(int64_t v, int64_t e0, int64_t deg, int64_t j, int slot, int32_t pv,
                                            const int32_t* __restrict__ indices, const int64_t* __restrict__ wq,
                                            const CommSig* __restrict__ cs, const int32_t* __restrict__ part,
                                            uint32_t& key, int64_t& w, int64_t& sg) {
  key = kNoKey;
  w = 0;
  sg = 0;
  if (j < deg) {
    const int32_t u = indices[e0 + j];
    const int64_t wj = wq[e0 + j];
    const int4 r = reinterpret_cast<const int4*>(cs)[u];  // (comm, part, sigtot lo, hi)
    if (u != static_cast<int32_t>(v) && (!part || r.y == pv)) {
      w = wj;
      sg = sig_of(r);
      key = (static_cast<uint32_t>(r.x) << IB) | static_cast<uint32_t>(slot);
    }
  }
}

template <int NI>
__device__ __forceinline__ void move_sorted(int64_t v, int64_t e0, int64_t deg, int32_t curr, int64_t kv,
                                            const int32_t* __restrict__ indices, const int64_t* __restrict__ wq,
                                            const CommSig* __restrict__ cs, const int32_t* __restrict__ part,
                                            const GainParams& gp, int32_t* __restrict__ target,
                                            int64_t* __restrict__ lw, int64_t* __restrict__ lsg,
                                            int64_t* __restrict__ lp) {
  constexpr int IB = NI == 1 ? 6 : NI == 2 ? 7 : 8;
  const int l = lane_id();
  const int32_t pv = part ? part[v] : 0;
  uint32_t key[NI];
  int nvalid = 0;
#pragma unroll
  for (int i = 0; i < NI; ++i) {
    int64_t w, sg;
    gather_slot<IB>(v, e0, deg, l + kWave * i, i * kWave + l, pv, indices, wq, cs, part, key[i], w, sg);
    lw[i * kWave + l] = w;
    lsg[i * kWave + l] = sg;
    nvalid += __popcll(__ballot(key[i] != kNoKey));
  }
  move_core<NI>(v, key, nvalid, curr, kv, gp, target, lw, lsg, lp);
}

__device__ __forceinline__ void move_vertex(int64_t v, const int64_t* __restrict__ indptr,
                                            const int32_t* __restrict__ indices, const int64_t* __restrict__ wq,
                                            const int64_t* __restrict__ kdeg, int64_t n,
                                            const int32_t* __restrict__ comm, const int32_t* __restrict__ part,
                                            const int64_t* __restrict__ sigtot, const CommSig* __restrict__ cs,
                                            const GainParams& gp, int32_t* __restrict__ target,
                                            int64_t (*lds)[kLightDeg]) {
  const int64_t e0 = indptr[v], deg = indptr[v + 1] - e0;
  const int32_t curr = comm[v];
  const int64_t kv = kdeg[v];
  if (n < kSortedMaxN) {
    int64_t* lw = lds[0];
    int64_t* lsg = lds[1];
    int64_t* lp = lds[2];
    if (deg <= kWave) move_sorted<1>(v, e0, deg, curr, kv, indices, wq, cs, part, gp, target, lw, lsg, lp);
    else if (deg <= 2 * kWave) move_sorted<2>(v, e0, deg, curr, kv, indices, wq, cs, part, gp, target, lw, lsg, lp);
    else move_sorted<4>(v, e0, deg, curr, kv, indices, wq, cs, part, gp, target, lw, lsg, lp);
  } else {
    if (deg <= kWave) move_one<1>(v, e0, deg, curr, kv, indices, wq, comm, part, sigtot, gp, target);
    else if (deg <= 2 * kWave) move_one<2>(v, e0, deg, curr, kv, indices, wq, comm, part, sigtot, gp, target);
    else move_one<4>(v, e0, deg, curr, kv, indices, wq, comm, part, sigtot, gp, target);
  }
}

// The wave's light vertices of NI slots per lane (listed in `mask`, lanes of the wave's
// kMoveGroup vertices), kMoveGroup / NI at a time: the neighbour ids and weights of the whole
// batch are loaded, then every (comm, part, sigma) record, and only then does the first
// decision start, so a wave keeps a batch's gathers in flight together.  Item j NI + s holds
// slot s 64 + l of the batch's j-th vertex.  The same decisions as move_sorted<NI>.
template <int NI>
__device__ __forceinline__ void move_batched(uint64_t mask, int64_t v0, int64_t e0v, int64_t degv, int32_t currv,
                                             int64_t kvv, int32_t pvv, const int32_t* __restrict__ indices,
                                             const int64_t* __restrict__ wq, const CommSig* __restrict__ cs,
                                             const int32_t* __restrict__ part, const GainParams& gp,
                                             int32_t* __restrict__ target, int64_t (*wl)[kLightDeg]) {
  constexpr int NB = kMoveGroup / NI;  // vertices per batch
  constexpr int IB = NI == 1 ? 6 : NI == 2 ? 7 : 8;
  const int l = lane_id();
  while (mask) {
    int vi[NB];  // wave-uniform lane of each vertex of the batch, -1 past the last
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      vi[j] = mask ? __ffsll(static_cast<long long>(mask)) - 1 : -1;
      mask &= mask ? mask - 1 : 0ull;
    }
    int32_t u[kMoveGroup];
    int64_t wr[kMoveGroup];
    int4 r[kMoveGroup];
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const int64_t e0 = vi[j] >= 0 ? readlane_i64(e0v, vi[j]) : 0;
      const int deg = vi[j] >= 0 ? static_cast<int>(readlane_i64(degv, vi[j])) : 0;
#pragma unroll
      for (int s = 0; s < NI; ++s) {
        const int it = j * NI + s, slot = s * kWave + l;
        u[it] = -1;
        wr[it] = 0;
        if (slot < deg) {
          u[it] = indices[e0 + slot];
          wr[it] = wq[e0 + slot];
        }
      }
    }
    // Unconditional gathers (lanes without a neighbour read record 0, which `ok` below never
    // uses): a conditional load lets the compiler move the first use of its value into the
    // branch, and the wait it then needs there serializes the gathers.
#pragma unroll
    for (int it = 0; it < kMoveGroup; ++it) r[it] = reinterpret_cast<const int4*>(cs)[u[it] >= 0 ? u[it] : 0];
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      if (vi[j] >= 0) {
        const int64_t v = v0 + vi[j];
        const int32_t pv = __builtin_amdgcn_readlane(pvv, vi[j]);
        uint32_t key[NI];
        int nvalid = 0;
#pragma unroll
        for (int s = 0; s < NI; ++s) {
          const int it = j * NI + s, slot = s * kWave + l;
          const bool ok = u[it] >= 0 && u[it] != static_cast<int32_t>(v) && (!part || r[it].y == pv);
          key[s] = ok ? ((static_cast<uint32_t>(r[it].x) << IB) | static_cast<uint32_t>(slot)) : kNoKey;
          wl[0][slot] = ok ? wr[it] : 0;
          wl[1][slot] = ok ? sig_of(r[it]) : 0;
          nvalid += __popcll(__ballot(ok));
        }
        move_core<NI>(v, key, nvalid, __builtin_amdgcn_readlane(currv, vi[j]), readlane_i64(kvv, vi[j]), gp,
                      target, wl[0], wl[1], wl[2]);
      }
    }
  }
}

// The light vertices of light block lb (degree <= 256): each wave owns kMoveGroup consecutive
// vertices; inactive ones keep their community (one coalesced store), active ones are moved
// one after another.  lds: 3 x kLightDeg int64 per wave.
__device__ __forceinline__ void move_light_block(
    int64_t lb, const int64_t* __restrict__ indptr, const int32_t* __restrict__ indices,
    const int64_t* __restrict__ wq, const int64_t* __restrict__ kdeg, int64_t n, const int32_t* __restrict__ comm,
    const int32_t* __restrict__ part, const int64_t* __restrict__ sigtot, const CommSig* __restrict__ cs,
    const uint8_t* __restrict__ active, const GainParams& gp, int32_t* __restrict__ target,
    int64_t (*lds)[3][kLightDeg]) {
  const int l = lane_id();
  const int64_t v0 = (lb * kWavesPerBlock + threadIdx.x / kWave) * kMoveGroup;
  bool todo = false;
  int64_t e0v = 0, degv = 0, kvv = 0;
  int32_t currv = 0, pvv = 0;
  if (l < kMoveGroup && v0 + l < n) {
    const int64_t vv = v0 + l;
    e0v = indptr[vv];
    degv = indptr[vv + 1] - e0v;
    currv = comm[vv];
    const bool light = degv <= kLightDeg;  // heavy path writes target[v]
    const bool act = active[vv];
    if (light && !act) target[vv] = currv;
    todo = light && act;
    kvv = kdeg[vv];
    pvv = part ? part[vv] : 0;
  }
  uint64_t pending = __ballot(todo);
  int64_t(*wl)[kLightDeg] = lds[threadIdx.x / kWave];
  if (n < kSortedMaxN) {
    // degree <= 64, 65..128 and 129..256: batches of 8, 4 and 2 vertices (one, two and four
    // slots per lane)
    const uint64_t small = __ballot(todo && degv <= kWave);
    const uint64_t mid2 = __ballot(todo && degv > kWave && degv <= 2 * kWave);
    const uint64_t mid4 = __ballot(todo && degv > 2 * kWave);
    move_batched<1>(small, v0, e0v, degv, currv, kvv, pvv, indices, wq, cs, part, gp, target, wl);
    move_batched<2>(mid2, v0, e0v, degv, currv, kvv, pvv, indices, wq, cs, part, gp, target, wl);
    move_batched<4>(mid4, v0, e0v, degv, currv, kvv, pvv, indices, wq, cs, part, gp, target, wl);
    pending = 0;
  }
  while (pending) {
    const int64_t v = v0 + __ffsll(static_cast<long long>(pending)) - 1;
    pending &= pending - 1;
    move_vertex(v, indptr, indices, wq, kdeg, n, comm, part, sigtot, cs, gp, target, lds[threadIdx.x / kWave]);
  }
}

// Degree 257..512, 513..1024, 1025..2048 and 2049..4096 (ITEMS = 2, 4, 8, 16): one 256-thread
// block per listed vertex, ITEMS neighbour slots per thread (N = 256 ITEMS keys).  The same decision as
// move_sorted, atomics-free: keys (community << IB | slot) sorted by a bitonic network whose
// stages run where the partner lives -- inside the thread (stride < ITEMS), across the wave by
// shuffles (stride < 64 ITEMS) or through LDS with barriers (the two or three widest strides)
// -- then the weights of each run of equal community summed in sorted order: inside a thread
// by a running sum that restarts at each run head, across threads by a block segmented scan
// (a run that starts in an earlier thread takes the scan's carry).  Each community is scored
// once at the end of its run and the block argmax taken (ties to the smaller community).
// Weights and sigma_tot are re-read by slot from global memory; the LDS holds only the keys.

// Exclusive segmented sum over the threads of the block.  Thread t gives (f, x): f = it holds
// a run head, x = the sum of its weights after its last head (all of them when f is false).
// Returns, for thread t, the sum of the weights from the last run head before its first
// position up to that position (exclusive): (f1, x1) . (f2, x2) = (f1 | f2, f2 ? x2 : x1 + x2).
template <int NW>
__device__ __forceinline__ int64_t block_excl_segsum_i64(bool f, int64_t x, int64_t* __restrict__ wsum,
                                                         int32_t* __restrict__ wflag) {
  const int l = lane_id(), wv = threadIdx.x / kWave;
  int fi = f ? 1 : 0;
  int64_t xi = x;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const int of = __shfl_up(fi, d, kWave);
    const int64_t ox = static_cast<int64_t>(shfl_up_u64(static_cast<uint64_t>(xi), d));
    if (l >= d) {
      if (!fi) xi += ox;
      fi |= of;
    }
  }
  if (l == kWave - 1) {
    wsum[wv] = xi;
    wflag[wv] = fi;
  }
  __syncthreads();
  int64_t cx = 0;  // carry of the earlier waves
  for (int w = 0; w < wv; ++w) cx = wflag[w] ? wsum[w] : cx + wsum[w];
  __syncthreads();  // wsum / wflag are free for the next scan
  int ef = __shfl_up(fi, 1, kWave);
  int64_t ex = static_cast<int64_t>(shfl_up_u64(static_cast<uint64_t>(xi), 1));
  if (l == 0) {
    ef = 0;
    ex = 0;
  }
  return ef ? ex : cx + ex;
}

// one compare-exchange of position e (value mine) with its partner p (value other) in the
// stage of block size k: ascending where (e & k) == 0
__device__ __forceinline__ uint32_t bitonic_pick(uint32_t mine, uint32_t other, int e, int p, int k) {
  const uint32_t lo = mine < other ? mine : other, hi = mine < other ? other : mine;
  return ((e < p) == ((e & k) == 0)) ? lo : hi;
}

// bitonic stage of stride J < ITEMS: the partner of item s is item s ^ J of the same thread
template <int ITEMS, int J>
__device__ __forceinline__ void stage_in_thread(uint32_t (&key)[ITEMS], int eb, int k) {
#pragma unroll
  for (int s = 0; s < ITEMS; ++s) {
    if ((s & J) == 0) {
      const bool asc = ((eb + s) & k) == 0;
      const uint32_t a = key[s], b = key[s | J];
      const uint32_t lo = a < b ? a : b, hi = a < b ? b : a;
      key[s] = asc ? lo : hi;
      key[s | J] = asc ? hi : lo;
    }
  }
}

template <int THREADS, int ITEMS>
struct BlockMoveSmem {
  static constexpr int NW = THREADS / kWave;
  uint32_t skey[THREADS * ITEMS];
  int64_t wsum[NW];
  int32_t wflag[NW];
  double wbest_g[NW];
  int32_t wbest_c[NW];
  double stay;
};

__host__ __device__ constexpr int ilog2(int x) { return x <= 1 ? 0 : 1 + ilog2(x / 2); }

template <int THREADS, int ITEMS>
__device__ __forceinline__ void block_move(int64_t v, const int64_t* __restrict__ indptr,
                                           const int32_t* __restrict__ indices, const int64_t* __restrict__ wq,
                                           const int64_t* __restrict__ kdeg, const int32_t* __restrict__ comm,
                                           const int32_t* __restrict__ part, const int64_t* __restrict__ sigtot,
                                           const uint8_t* __restrict__ active, const GainParams& gp_all,
                                           int32_t* __restrict__ target, BlockMoveSmem<THREADS, ITEMS>& sm) {
  static_assert(ITEMS == 2 || ITEMS == 4 || ITEMS == 8 || ITEMS == 16, "ITEMS");
  static_assert(THREADS == 256 || THREADS == 1024, "THREADS");
  constexpr int NW = THREADS / kWave;
  constexpr int N = THREADS * ITEMS;
  constexpr int IB = ilog2(N);
  constexpr int kWaveSpan = ITEMS * kWave;  // positions held by one wave
  constexpr uint32_t kSlot = (1u << IB) - 1u;
  const int t = threadIdx.x, l = lane_id(), wv = t / kWave;
  const GainParams gp = for_vertex(gp_all, v);
  const int32_t curr = comm[v];
  if (!active[v]) {  // uniform over the block
    if (t == 0) target[v] = curr;
    return;
  }
  const int64_t e0 = indptr[v], deg = indptr[v + 1] - e0;
  const int64_t kv = kdeg[v];
  const int32_t pv = part ? part[v] : 0;
  const int eb = ITEMS * t;  // this thread's first position
  int32_t u[ITEMS];
#pragma unroll
  for (int s = 0; s < ITEMS; ++s) u[s] = eb + s < deg ? indices[e0 + eb + s] : -1;
  uint32_t key[ITEMS];
#pragma unroll
  for (int s = 0; s < ITEMS; ++s) {
    key[s] = kNoKey;
    if (u[s] >= 0 && u[s] != static_cast<int32_t>(v) && (!part || part[u[s]] == pv))
      key[s] = (static_cast<uint32_t>(comm[u[s]]) << IB) | static_cast<uint32_t>(eb + s);
  }
  if (t == 0) sm.stay = 0.0;
  for (int k = 2; k <= N; k <<= 1)
    for (int j = k >> 1; j > 0; j >>= 1) {
      if (j >= kWaveSpan) {  // partner in another wave
        __syncthreads();     // earlier readers of skey are done
#pragma unroll
        for (int s = 0; s < ITEMS; ++s) sm.skey[eb + s] = key[s];
        __syncthreads();
#pragma unroll
        for (int s = 0; s < ITEMS; ++s)
          key[s] = bitonic_pick(key[s], sm.skey[(eb + s) ^ j], eb + s, (eb + s) ^ j, k);
      } else if (j >= ITEMS) {  // partner in lane l ^ (j / ITEMS) of this wave
#pragma unroll
        for (int s = 0; s < ITEMS; ++s)
          key[s] = bitonic_pick(key[s], __shfl_xor(key[s], j / ITEMS, kWave), eb + s, (eb + s) ^ j, k);
      } else if (j == 1) {
        stage_in_thread<ITEMS, 1>(key, eb, k);
      } else if constexpr (ITEMS > 2) {
        if (j == 2) stage_in_thread<ITEMS, 2>(key, eb, k);
        else if constexpr (ITEMS > 4) {
          if (j == 4) stage_in_thread<ITEMS, 4>(key, eb, k);
          else if constexpr (ITEMS > 8) stage_in_thread<ITEMS, 8>(key, eb, k);
        }
      }
    }
  __syncthreads();  // the last LDS stage's readers are done
#pragma unroll
  for (int s = 0; s < ITEMS; ++s) sm.skey[eb + s] = key[s];
  __syncthreads();
  const uint32_t before = t > 0 ? sm.skey[eb - 1] : kNoKey;
  const uint32_t after = eb + ITEMS < N ? sm.skey[eb + ITEMS] : kNoKey;
  double best_g = -1e300, stay = 0.0;
  int32_t best_c = 0x7FFFFFFF;
  bool has_stay = false;
  auto score = [&](int64_t wc, uint32_t kk) {
    const int32_t ci = static_cast<int32_t>(kk >> IB);
    const int64_t sig = sigtot[ci] - (ci == curr ? kv : 0);
    const double g = gain_of(wc, sig, kv, gp);
    if (ci == curr) {
      stay = g;
      has_stay = true;
    }
    if (better(g, ci, best_g, best_c)) {
      best_g = g;
      best_c = ci;
    }
  };
  // running sum from the last run head; a run end whose head lies in an earlier thread (at
  // most one per thread: the first run end) waits for the block scan's carry
  int64_t acc = 0, open_acc = 0;
  bool seen_head = false;
  int open_s = -1;
#pragma unroll
  for (int s = 0; s < ITEMS; ++s) {
    if (key[s] == kNoKey) continue;  // valid keys are a prefix of the sorted order
    const uint32_t prev = s > 0 ? key[s > 0 ? s - 1 : 0] : before;
    if (eb + s == 0 || prev == kNoKey || (prev >> IB) != (key[s] >> IB)) {
      acc = 0;
      seen_head = true;
    }
    acc += wq[e0 + (key[s] & kSlot)];
    const uint32_t next = s + 1 < ITEMS ? key[s + 1 < ITEMS ? s + 1 : s] : after;
    if (next == kNoKey || (next >> IB) != (key[s] >> IB)) {  // end of its run
      if (seen_head) score(acc, key[s]);
      else {
        open_acc = acc;
        open_s = s;
      }
    }
  }
  const int64_t carry = block_excl_segsum_i64<NW>(seen_head, acc, sm.wsum, sm.wflag);
  if (open_s >= 0) {
    uint32_t kk = key[0];
#pragma unroll
    for (int s = 1; s < ITEMS; ++s)
      if (s == open_s) kk = key[s];
    score(carry + open_acc, kk);
  }
#pragma unroll
  for (int m = 1; m < kWave; m <<= 1) {
    const double og = __shfl_xor(best_g, m, kWave);
    const int32_t oc = __shfl_xor(best_c, m, kWave);
    if (better(og, oc, best_g, best_c)) {
      best_g = og;
      best_c = oc;
    }
  }
  if (l == 0) {
    sm.wbest_g[wv] = best_g;
    sm.wbest_c[wv] = best_c;
  }
  if (has_stay) sm.stay = stay;  // one run holds the current community
  __syncthreads();
  if (t == 0) {
    double bg = sm.wbest_g[0];
    int32_t bc = sm.wbest_c[0];
    for (int w = 1; w < NW; ++w)
      if (better(sm.wbest_g[w], sm.wbest_c[w], bg, bc)) {
        bg = sm.wbest_g[w];
        bc = sm.wbest_c[w];
      }
    target[v] = (bg > sm.stay + kTol) ? bc : curr;
  }
}

constexpr size_t kLightSmem = sizeof(int64_t) * kWavesPerBlock * 3 * kLightDeg;
constexpr size_t kMoveSmem =
    kLightSmem > sizeof(BlockMoveSmem<kThreads, 16>) ? kLightSmem : sizeof(BlockMoveSmem<kThreads, 16>);

// The listed vertices of the four hub classes: one array, the largest class first.
struct HubLists {
  const int32_t* verts;
  int64_t count[kHubClasses];  // per class i; verts holds class 3's vertices, then 2, 1, 0
};

// Best move of every vertex of degree <= 4096 in ONE launch (256-thread blocks): the first
// blocks take the listed hub vertices, one block each, largest class first, so that they start
// while the light blocks are still being dispatched; the rest take the light vertices in
// groups of kMoveGroup per wave.  Each vertex's target[] is written by exactly one block;
// heavier vertices are left to the sort path.
__global__ __launch_bounds__(kThreads) void k_move(
    HubLists hubs, const int64_t* __restrict__ indptr, const int32_t* __restrict__ indices,
    const int64_t* __restrict__ wq, const int64_t* __restrict__ kdeg, int64_t n, const int32_t* __restrict__ comm,
    const int32_t* __restrict__ part, const int64_t* __restrict__ sigtot, const CommSig* __restrict__ cs,
    const uint8_t* __restrict__ active, GainParams gp, int32_t* __restrict__ target) {
  __shared__ __align__(16) unsigned char smem[kMoveSmem];
  int64_t b = blockIdx.x;
  const int32_t* verts = hubs.verts;
#define MSC_HUB_CLASS(I, ITEMS)                                                                  \
  if (b < hubs.count[I]) {                                                                       \
    block_move<kThreads, ITEMS>(verts[b], indptr, indices, wq, kdeg, comm, part, sigtot, active, gp, target, \
                                *reinterpret_cast<BlockMoveSmem<kThreads, ITEMS>*>(smem));             \
    return;                                                                                      \
  }                                                                                              \
  b -= hubs.count[I];                                                                            \
  verts += hubs.count[I];
  MSC_HUB_CLASS(3, 16)
  MSC_HUB_CLASS(2, 8)
  MSC_HUB_CLASS(1, 4)
  MSC_HUB_CLASS(0, 2)
#undef MSC_HUB_CLASS
  move_light_block(b, indptr, indices, wq, kdeg, n, comm, part, sigtot, cs, active, gp, target,
                   reinterpret_cast<int64_t(*)[3][kLightDeg]>(smem));
}

// Degree 4097..16384 (graphs below 2^18 vertices, so that (community << 14 | slot) fits a
// key): one 1024-thread block per listed vertex, 16 slots per thread, the same decision as the
// hub classes of k_move.  Launched after it on the same stream.
constexpr int kHugeThreads = 1024;
constexpr int kHugeDeg = kHugeThreads * 16;
constexpr int64_t kHugeMaxN = int64_t(1) << 18;
__global__ __launch_bounds__(kHugeThreads) void k_move_huge(
    const int32_t* __restrict__ verts, const int64_t* __restrict__ indptr, const int32_t* __restrict__ indices,
    const int64_t* __restrict__ wq, const int64_t* __restrict__ kdeg, const int32_t* __restrict__ comm,
    const int32_t* __restrict__ part, const int64_t* __restrict__ sigtot, const uint8_t* __restrict__ active,
    GainParams gp, int32_t* __restrict__ target) {
  __shared__ BlockMoveSmem<kHugeThreads, 16> sm;
  block_move<kHugeThreads, 16>(verts[blockIdx.x], indptr, indices, wq, kdeg, comm, part, sigtot, active, gp, target,
                               sm);
}

// ------------------------------------------------------------------------------------------
// best move, heavy vertices (degree above the one-wave paths): sort their eligible edges by
// (vertex, community)
// ------------------------------------------------------------------------------------------
struct Heavy {
  int64_t nh = 0, eh = 0;
  int32_t* verts = nullptr;   // heavy vertex ids, ascending (nh)
  int64_t* eoff = nullptr;    // edge offsets per heavy vertex (nh + 1)
  int32_t* eh_h = nullptr;    // heavy-edge -> local heavy index (eh)
  int64_t* eh_e = nullptr;    // heavy-edge -> CSR edge id (eh)
  // per-pass scratch (eh + 1)
  uint32_t *kc, *kv, *skc, *skv, *kh, *skh, *skv2;
  int64_t* wsorted;
  int64_t* prefix;
  int32_t* flag;
  int32_t* segid;
  int64_t* segstart;   // nseg + 1 (<= eh + 1)
  uint32_t* seg_h;     // per segment heavy index
  int64_t* hseg;       // nh + 1: first segment per heavy vertex
  int64_t* hstart;     // nh + 1: first sorted heavy edge per heavy vertex
};

// hub class of each vertex (kLightDeg < degree <= hub_max: class i for degrees in
// (256 << i, 512 << i]; hub_max < degree <= heavy_min: class 4, k_move_huge; else -1); heavy:
// degree above heavy_min
__global__ void k_heavy_flags(const int64_t* __restrict__ indptr, int64_t n, int64_t hub_max, int64_t heavy_min,
                              int32_t* __restrict__ flag, int64_t* __restrict__ deg_h, int32_t* __restrict__ hclass) {
  const int64_t v = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (v < n) {
    const int64_t d = indptr[v + 1] - indptr[v];
    flag[v] = d > heavy_min ? 1 : 0;
    deg_h[v] = d > heavy_min ? d : 0;
    int cl = -1;
    if (d > kLightDeg && d <= hub_max) cl = d <= 512 ? 0 : d <= 1024 ? 1 : d <= 2048 ? 2 : 3;
    else if (d > hub_max && d <= heavy_min) cl = 4;
    hclass[v] = cl;
  }
  if (v == n) {
    flag[n] = 0;
    deg_h[n] = 0;
    hclass[n] = -1;
  }
}
__global__ void k_class_flag(const int32_t* __restrict__ hclass, int64_t n, int cl, int32_t* __restrict__ flag) {
  const int64_t v = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (v <= n) flag[v] = (v < n && hclass[v] == cl) ? 1 : 0;
}
__global__ void k_heavy_compact(const int32_t* __restrict__ flag, const int32_t* __restrict__ pos, int64_t n,
                                int32_t* __restrict__ verts) {
  const int64_t v = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (v < n && flag[v]) verts[pos[v]] = static_cast<int32_t>(v);
}
// one wave per heavy vertex: write its edge ids and local index
__global__ void k_heavy_edges(const int64_t* __restrict__ indptr, const int32_t* __restrict__ verts,
                              const int64_t* __restrict__ degoff_all, int64_t nh, int64_t* __restrict__ eoff,
                              int32_t* __restrict__ eh_h, int64_t* __restrict__ eh_e, int64_t eh) {
  const int64_t h = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave;
  if (h > nh) return;
  if (h == nh) {
    if (lane_id() == 0) eoff[nh] = eh;
    return;
  }
  const int32_t v = verts[h];
  const int64_t base = degoff_all[v];
  if (lane_id() == 0) eoff[h] = base;
  const int64_t e0 = indptr[v], d = indptr[v + 1] - e0;
  for (int64_t j = lane_id(); j < d; j += kWave) {
    eh_h[base + j] = static_cast<int32_t>(h);
    eh_e[base + j] = e0 + j;
  }
}

__global__ void k_heavy_keys(const int32_t* __restrict__ eh_h, const int64_t* __restrict__ eh_e, int64_t eh,
                             const int32_t* __restrict__ verts, const int32_t* __restrict__ indices,
                             const int32_t* __restrict__ comm, const int32_t* __restrict__ part,
                             const uint8_t* __restrict__ active, int64_t n, uint32_t* __restrict__ kc,
                             uint32_t* __restrict__ kvv) {
  const int64_t p = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (p >= eh) return;
  const int32_t v = verts[eh_h[p]];
  const int32_t u = indices[eh_e[p]];
  const bool ok = active[v] && u != v && (!part || part[u] == part[v]);
  kc[p] = ok ? static_cast<uint32_t>(comm[u]) : static_cast<uint32_t>(n);
  kvv[p] = static_cast<uint32_t>(p);
}
__global__ void k_heavy_hkeys(const uint32_t* __restrict__ perm, const int32_t* __restrict__ eh_h, int64_t eh,
                              uint32_t* __restrict__ kh) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i < eh) kh[i] = static_cast<uint32_t>(eh_h[perm[i]]);
}
// after both sorts: perm2[i] = heavy position p in (h, comm) order
__global__ void k_heavy_seg_flags(const uint32_t* __restrict__ perm2, const uint32_t* __restrict__ skh, int64_t eh,
                                  const int64_t* __restrict__ eh_e, const int32_t* __restrict__ indices,
                                  const int32_t* __restrict__ comm, const uint32_t* __restrict__ kc_unsorted,
                                  const int64_t* __restrict__ wq, int64_t n, int32_t* __restrict__ flag,
                                  int64_t* __restrict__ wsorted) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i > eh) return;
  if (i == eh) {
    wsorted[eh] = 0;
    flag[eh] = 0;
    return;
  }
  const uint32_t p = perm2[i];
  const uint32_t c = kc_unsorted[p];
  const bool ok = c != static_cast<uint32_t>(n);
  wsorted[i] = ok ? wq[eh_e[p]] : 0;
  bool start = ok;
  if (ok && i > 0) {
    const uint32_t pp = perm2[i - 1];
    start = !(skh[i - 1] == skh[i] && kc_unsorted[pp] == c);
  }
  flag[i] = start ? 1 : 0;
}
__global__ void k_heavy_seg_starts(const int32_t* __restrict__ flag, const int32_t* __restrict__ pos, int64_t eh,
                                   const uint32_t* __restrict__ skh, int64_t* __restrict__ segstart,
                                   uint32_t* __restrict__ seg_h) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i < eh && flag[i]) {
    segstart[pos[i]] = i;
    seg_h[pos[i]] = skh[i];
  }
  if (i == eh) segstart[pos[eh]] = eh;
}
__global__ void k_heavy_hseg(const int64_t* __restrict__ hstart, const int32_t* __restrict__ segid, int64_t nh,
                             int64_t* __restrict__ hseg) {
  const int64_t h = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (h <= nh) hseg[h] = segid[hstart[h]];
}
// one wave per heavy vertex over its segments
__global__ __launch_bounds__(kThreads) void k_heavy_argmax(
    const int32_t* __restrict__ verts, int64_t nh, const int64_t* __restrict__ hseg,
    const int64_t* __restrict__ segstart, const int64_t* __restrict__ prefix, const uint32_t* __restrict__ perm2,
    const uint32_t* __restrict__ kc_unsorted, const int64_t* __restrict__ kdeg, const int32_t* __restrict__ comm,
    const int64_t* __restrict__ sigtot, GainParams gp_all, int32_t* __restrict__ target) {
  const int64_t h = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave;
  if (h >= nh) return;
  const int32_t v = verts[h];
  const GainParams gp = for_vertex(gp_all, v);
  const int32_t curr = comm[v];
  const int64_t kv = kdeg[v];
  double best_g = -1e300, stay = 0.0;
  int32_t best_c = curr;
  bool has_stay = false;
  for (int64_t s = hseg[h] + lane_id(); s < hseg[h + 1]; s += kWave) {
    const int64_t a = segstart[s], b = segstart[s + 1];
    const int32_t c = static_cast<int32_t>(kc_unsorted[perm2[a]]);
    const int64_t wc = prefix[b] - prefix[a];
    const int64_t sig = sigtot[c] - (c == curr ? kv : 0);
    const double g = gain_of(wc, sig, kv, gp);
    if (c == curr) { stay = g; has_stay = true; }
    if (better(g, c, best_g, best_c)) { best_g = g; best_c = c; }
  }
#pragma unroll
  for (int m = 1; m < kWave; m <<= 1) {
    const double og = __shfl_xor(best_g, m, kWave);
    const int32_t oc = __shfl_xor(best_c, m, kWave);
    if (better(og, oc, best_g, best_c)) { best_g = og; best_c = oc; }
    const double os = __shfl_xor(stay, m, kWave);
    const int ohs = __shfl_xor(static_cast<int>(has_stay), m, kWave);
    if (ohs && !has_stay) { stay = os; has_stay = true; }
  }
  if (lane_id() == 0) target[v] = (best_g > stay + kTol) ? best_c : curr;
}

// ------------------------------------------------------------------------------------------
// commit, frontier, relabel, modularity, contraction
// ------------------------------------------------------------------------------------------
// sigma_tot is kept current across passes by the committed vertices alone: each one
// contributes (previous community, -k_v) and (new community, +k_v).  Communities fall into
// kBuckets buckets by a hash of their id; k_commit lays each block's entries out grouped by
// bucket, and k_sigtot_apply gives every bucket one workgroup, which sums the bucket's entries
// per community in an LDS hash table and adds each total to sigtot.  A community is written
// by exactly one workgroup (plain read-modify-write, no global atomics) and integer sums do
// not depend on the order of the adds, so sigtot equals a recomputation from comm exactly.
constexpr int kCommitThreads = 1024;
constexpr int kBucketBits = 8;
constexpr int kBuckets = 1 << kBucketBits;
constexpr int kApplyThreads = 1024;
constexpr int kApplyChunk = 2048;          // entries per table fill
constexpr int kApplySlotBits = 12;         // 4096 slots: load <= 1/2, so every probe ends
constexpr int kApplySlots = 1 << kApplySlotBits;
__device__ __forceinline__ int comm_bucket(int32_t c) {
  return static_cast<int>((static_cast<uint32_t>(c) * 2654435761u) >> (32 - kBucketBits));
}
__device__ __forceinline__ uint32_t comm_slot(int32_t c) {
  return (static_cast<uint32_t>(c) * 2246822519u) >> (32 - kApplySlotBits);
}

// Commit the coin-approved moves (louvain.py:355-362) and start the next active set
// (louvain.py:364-372): commit | held = (wants & coin) | (wants & ~coin) = wants, so the
// vertex's own part is its wants bit; k_frontier_push adds N(commit) afterwards.  Block blk
// writes its sigma_tot entries to [blk * 2 * kCommitThreads, + 2 * committed) grouped by
// bucket, bucket j at offsets boff[blk * (kBuckets + 1) + j] (order inside a bucket is free).
__global__ __launch_bounds__(kCommitThreads) void k_commit(
    int32_t* __restrict__ comm, const int32_t* __restrict__ target, int64_t n, uint64_t pkey, uint64_t thr,
    uint8_t* __restrict__ commit, uint8_t* __restrict__ active, int32_t* __restrict__ wants_flag,
    const int64_t* __restrict__ kdeg, int32_t* __restrict__ ekey, int64_t* __restrict__ eval,
    int32_t* __restrict__ boff, const int32_t* __restrict__ vsid, const int64_t* __restrict__ vbase) {
  __shared__ int hist[kBuckets];
  __shared__ int wtot[kBuckets / kWave];
  const int64_t v = static_cast<int64_t>(blockIdx.x) * kCommitThreads + threadIdx.x;
  if (threadIdx.x < kBuckets) hist[threadIdx.x] = 0;
  __syncthreads();
  bool cm = false;
  int32_t c = 0, t = 0;
  int b0 = 0, b1 = 0, r0 = 0, r1 = 0;
  if (v < n) {
    t = target[v];
    c = comm[v];
    const bool wants = t != c;
    // a batched run draws each sample's coins by the vertex's index within its sample
    const int64_t vl = vsid ? v - vbase[vsid[v]] : v;
    const bool coin = static_cast<uint64_t>(coin_u32(pkey, static_cast<uint64_t>(vl))) < thr;
    cm = wants && coin;
    if (cm) {
      comm[v] = t;
      b0 = comm_bucket(c);
      b1 = comm_bucket(t);
      r0 = atomicAdd(&hist[b0], 1);  // LDS: a place inside the bucket
      r1 = atomicAdd(&hist[b1], 1);
    }
    commit[v] = cm;
    active[v] = wants;
    if (wants) *wants_flag = 1;  // every writer stores the same value
  }
  __syncthreads();
  // exclusive scan of the bucket counts (the first kBuckets threads, one bucket each)
  const int l = lane_id(), w = threadIdx.x / kWave;
  int x = 0, incl = 0;
  if (threadIdx.x < kBuckets) {
    x = hist[threadIdx.x];
    incl = x;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      const int y = __shfl_up(incl, d, kWave);
      if (l >= d) incl += y;
    }
    if (l == kWave - 1) wtot[w] = incl;
  }
  __syncthreads();
  int32_t* bo = boff + static_cast<int64_t>(blockIdx.x) * (kBuckets + 1);
  if (threadIdx.x < kBuckets) {
    int base = 0, total = 0;
#pragma unroll
    for (int i = 0; i < kBuckets / kWave; ++i) {
      base += (i < w) ? wtot[i] : 0;
      total += wtot[i];
    }
    const int ex = base + incl - x;
    hist[threadIdx.x] = ex;  // each thread rewrites only its own bucket
    bo[threadIdx.x] = ex;
    if (threadIdx.x == 0) bo[kBuckets] = total;
  }
  __syncthreads();
  if (cm) {
    const int64_t e0 = static_cast<int64_t>(blockIdx.x) * (2 * kCommitThreads);
    const int64_t kv = kdeg[v];
    ekey[e0 + hist[b0] + r0] = c;
    eval[e0 + hist[b0] + r0] = -kv;
    ekey[e0 + hist[b1] + r1] = t;
    eval[e0 + hist[b1] + r1] = kv;
  }
}

// One workgroup per bucket: thread i takes commit blocks i, i + kApplyThreads, ...; the
// bucket's entries are numbered by a block scan of the per-thread counts and summed
// kApplyChunk at a time (at most kApplyChunk distinct keys in kApplySlots slots).
__global__ __launch_bounds__(kApplyThreads) void k_sigtot_apply(const int32_t* __restrict__ ekey,
                                                                const int64_t* __restrict__ eval,
                                                                const int32_t* __restrict__ boff, int nblk,
                                                                int64_t* __restrict__ sigtot) {
  __shared__ int32_t hkey[kApplySlots];
  __shared__ unsigned long long hval[kApplySlots];
  __shared__ int wtot[kApplyThreads / kWave];
  const int bk = blockIdx.x, t = threadIdx.x, l = lane_id(), w = t / kWave;
  int mine = 0;
  for (int blk = t; blk < nblk; blk += kApplyThreads) {
    const int32_t* bo = boff + static_cast<int64_t>(blk) * (kBuckets + 1);
    mine += bo[bk + 1] - bo[bk];
  }
  int incl = mine;
#pragma unroll
  for (int d = 1; d < kWave; d <<= 1) {
    const int y = __shfl_up(incl, d, kWave);
    if (l >= d) incl += y;
  }
  if (l == kWave - 1) wtot[w] = incl;
  constexpr int32_t kEmpty = -1;  // community ids are >= 0
  for (int i = t; i < kApplySlots; i += kApplyThreads) {
    hkey[i] = kEmpty;
    hval[i] = 0;
  }
  __syncthreads();
  int first = incl - mine, total = 0;
#pragma unroll
  for (int i = 0; i < kApplyThreads / kWave; ++i) {
    first += (i < w) ? wtot[i] : 0;
    total += wtot[i];
  }
  for (int s0 = 0; s0 < total; s0 += kApplyChunk) {
    // this thread's entries numbered [first, first + mine); those in [s0, s0 + kApplyChunk)
    if (first < s0 + kApplyChunk && first + mine > s0) {
      int idx = first;
      for (int blk = t; blk < nblk; blk += kApplyThreads) {
        const int32_t* bo = boff + static_cast<int64_t>(blk) * (kBuckets + 1);
        const int lo = bo[bk], hi = bo[bk + 1];
        const int64_t e0 = static_cast<int64_t>(blk) * (2 * kCommitThreads);
        for (int j = lo; j < hi; ++j, ++idx) {
          if (idx < s0 || idx >= s0 + kApplyChunk) continue;
          const int32_t ck = ekey[e0 + j];
          const unsigned long long cv = static_cast<unsigned long long>(eval[e0 + j]);  // two's complement
          uint32_t h = comm_slot(ck);
          for (;;) {
            const int32_t prev = atomicCAS(&hkey[h], kEmpty, ck);
            if (prev == kEmpty || prev == ck) {
              atomicAdd(&hval[h], cv);
              break;
            }
            h = (h + 1) & (kApplySlots - 1);
          }
        }
      }
    }
    __syncthreads();
    for (int i = t; i < kApplySlots; i += kApplyThreads) {
      const int32_t ck = hkey[i];
      if (ck != kEmpty) {
        sigtot[ck] += static_cast<int64_t>(hval[i]);
        hkey[i] = kEmpty;
        hval[i] = 0;
      }
    }
    __syncthreads();
  }
}

__device__ __forceinline__ void refresh_one(int64_t v, const int32_t* __restrict__ comm,
                                            const int32_t* __restrict__ part, const int64_t* __restrict__ sigtot,
                                            CommSig* __restrict__ cs) {
  const int32_t c = comm[v];
  const int64_t s = sigtot[c];
  reinterpret_cast<int4*>(cs)[v] = make_int4(c, part ? part[v] : 0, static_cast<int32_t>(s),
                                             static_cast<int32_t>(static_cast<uint64_t>(s) >> 32));
}
// cs[v] = (comm[v], part[v], sigtot[comm[v]]) after sigma_tot has been rebuilt
__global__ void k_refresh_cs(int64_t n, const int32_t* __restrict__ comm, const int32_t* __restrict__ part,
                             const int64_t* __restrict__ sigtot, CommSig* __restrict__ cs) {
  const int64_t v = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (v < n) refresh_one(v, comm, part, sigtot, cs);
}

// next active set = commit | held | N(commit) (louvain.py:364-372); the graph stores both
// directions, so marking the neighbours of each committed vertex equals testing every
// vertex's neighbours.  Concurrent marks all store 1.
// one wave per kMoveGroup consecutive vertices; only committed ones touch their edges.
// refresh != 0 (sigma_tot already updated for this commit): also rewrite cs of the wave's
// vertices, as k_refresh_cs does.
__global__ __launch_bounds__(kThreads) void k_frontier_push(const int64_t* __restrict__ indptr,
                                                            const int32_t* __restrict__ indices, int64_t n,
                                                            const uint8_t* __restrict__ commit,
                                                            uint8_t* __restrict__ active, int refresh,
                                                            const int32_t* __restrict__ comm,
                                                            const int32_t* __restrict__ part,
                                                            const int64_t* __restrict__ sigtot,
                                                            CommSig* __restrict__ cs) {
  // one vertex per lane for the record rewrite and the commit scan; each committed vertex's
  // neighbours are then marked by the whole wave
  const int l = lane_id();
  const int64_t v0 = (static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave) * kWave;
  if (refresh && v0 + l < n) refresh_one(v0 + l, comm, part, sigtot, cs);
  uint64_t pending = __ballot(v0 + l < n && commit[v0 + l]);
  while (pending) {
    const int64_t v = v0 + __ffsll(static_cast<long long>(pending)) - 1;
    pending &= pending - 1;
    for (int64_t e = indptr[v] + l; e < indptr[v + 1]; e += kWave) active[indices[e]] = 1;
  }
}

__global__ void k_mark_used(const int32_t* __restrict__ comm, int64_t n, int32_t* __restrict__ used) {
  const int64_t v = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (v < n) used[comm[v]] = 1;  // every writer stores the same value
}
__global__ void k_apply_map(const int32_t* __restrict__ comm, const int32_t* __restrict__ map, int64_t n,
                            int32_t* __restrict__ dense) {
  const int64_t v = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (v < n) dense[v] = map[comm[v]];
}

// internal weight per vertex: sum of w over neighbours (self-loops included) in the same label
__global__ __launch_bounds__(kThreads) void k_internal(const int64_t* __restrict__ indptr, const int32_t* __restrict__ indices,
                                                       const int64_t* __restrict__ wq, const int32_t* __restrict__ lab,
                                                       int64_t n, int64_t* __restrict__ iw) {
  const int64_t v = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave;
  if (v >= n) return;
  const int32_t lv = lab[v];
  int64_t s = 0;
  for (int64_t e = indptr[v] + lane_id(); e < indptr[v + 1]; e += kWave)
    if (lab[indices[e]] == lv) s += wq[e];
  s = wave_sum_i64(s);
  if (lane_id() == 0) iw[v] = s;
}
__global__ void k_q_terms(const int64_t* __restrict__ tot, const int64_t* __restrict__ in, int64_t C, double inv_s,
                          double twom_real, double gamma, double* __restrict__ terms) {
  const int64_t c = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (c < C) {
    const double a = (static_cast<double>(tot[c]) * inv_s) / twom_real;
    terms[c] = (static_cast<double>(in[c]) * inv_s) / twom_real - gamma * (a * a);
  }
}

// contraction keys
__global__ void k_edge_src_dst_keys(const int64_t* __restrict__ indptr, const int32_t* __restrict__ indices,
                                    const int32_t* __restrict__ R, int64_t n, uint32_t* __restrict__ kdst,
                                    uint32_t* __restrict__ eid) {
  // one wave per source row
  const int64_t v = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave;
  if (v >= n) return;
  for (int64_t e = indptr[v] + lane_id(); e < indptr[v + 1]; e += kWave) {
    kdst[e] = static_cast<uint32_t>(R[indices[e]]);
    eid[e] = static_cast<uint32_t>(e);
  }
}
__global__ __launch_bounds__(kThreads) void k_edge_src_of(const int64_t* __restrict__ indptr, const int32_t* __restrict__ R,
                                                          int64_t n, int32_t* __restrict__ esrc) {
  const int64_t v = static_cast<int64_t>(blockIdx.x) * kWavesPerBlock + threadIdx.x / kWave;
  if (v >= n) return;
  const int32_t r = R[v];
  for (int64_t e = indptr[v] + lane_id(); e < indptr[v + 1]; e += kWave) esrc[e] = r;
}
__global__ void k_gather_src_keys(const uint32_t* __restrict__ perm, const int32_t* __restrict__ esrc, int64_t E,
                                  uint32_t* __restrict__ ksrc) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i < E) ksrc[i] = static_cast<uint32_t>(esrc[perm[i]]);
}
// sorted by (src, dst): flags of segment starts, weights in sorted order
__global__ void k_contract_flags(const uint32_t* __restrict__ perm2, const uint32_t* __restrict__ ssrc,
                                 const uint32_t* __restrict__ kdst_unsorted, const int64_t* __restrict__ wq, int64_t E,
                                 int32_t* __restrict__ flag, int64_t* __restrict__ wsorted) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i > E) return;
  if (i == E) {
    flag[E] = 0;
    wsorted[E] = 0;
    return;
  }
  const uint32_t e = perm2[i];
  wsorted[i] = wq[e];
  bool start = true;
  if (i > 0) start = !(ssrc[i - 1] == ssrc[i] && kdst_unsorted[perm2[i - 1]] == kdst_unsorted[e]);
  flag[i] = start ? 1 : 0;
}
__global__ void k_contract_emit(const int32_t* __restrict__ flag, const int32_t* __restrict__ pos, int64_t E,
                                const uint32_t* __restrict__ perm2, const uint32_t* __restrict__ ssrc,
                                const uint32_t* __restrict__ kdst_unsorted, const int64_t* __restrict__ prefix,
                                int32_t* __restrict__ out_idx, uint32_t* __restrict__ out_src,
                                int64_t* __restrict__ segstart) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i < E && flag[i]) {
    const int32_t s = pos[i];
    out_idx[s] = static_cast<int32_t>(kdst_unsorted[perm2[i]]);
    out_src[s] = ssrc[i];
    segstart[s] = i;
  }
  if (i == E) segstart[pos[E]] = E;
}
__global__ void k_contract_weights(const int64_t* __restrict__ segstart, const int64_t* __restrict__ prefix, int64_t m,
                                   int64_t* __restrict__ out_w) {
  const int64_t s = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (s < m) out_w[s] = prefix[segstart[s + 1]] - prefix[segstart[s]];
}
__global__ void k_scatter_pagg(const int32_t* __restrict__ R, const int32_t* __restrict__ P, int64_t n,
                               int32_t* __restrict__ Pagg) {
  const int64_t v = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (v < n) Pagg[R[v]] = P[v];  // refined communities lie inside one P community: same value
}

// ------------------------------------------------------------------------------------------
// batched runs: many independent samples as one block-diagonal graph
// ------------------------------------------------------------------------------------------
// sample of vertex v: the last s with vbase[s] <= v
__device__ __forceinline__ int32_t sample_of(const int64_t* __restrict__ vbase, int S, int64_t v) {
  int lo = 0, hi = S - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (vbase[mid] <= v) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}
__global__ void k_fill_sid(const int64_t* __restrict__ vbase, int S, int64_t n, int32_t* __restrict__ vsid) {
  const int64_t v = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (v < n) vsid[v] = sample_of(vbase, S, v);
}
__global__ void k_active_alive(uint8_t* __restrict__ active, int64_t n, const int32_t* __restrict__ vsid,
                               const uint8_t* __restrict__ alive) {
  const int64_t v = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (v < n) active[v] = alive[vsid[v]];
}
// out[s] = src[idx[s]] for s <= S
__global__ void k_gather_at(const int64_t* __restrict__ src64, const int32_t* __restrict__ src32,
                            const int64_t* __restrict__ idx, int S, int64_t* __restrict__ out) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s <= S) out[s] = src64 ? src64[idx[s]] : static_cast<int64_t>(src32[idx[s]]);
}
// per-sample maximum weight over the sample's edge range (block per sample; +inf for a
// negative or non-finite weight, as k_max_weight)
__global__ void k_sample_max_weight(const float* __restrict__ w, const int64_t* __restrict__ eoff,
                                    float* __restrict__ out) {
  float m = 0.0f;
  for (int64_t i = eoff[blockIdx.x] + threadIdx.x; i < eoff[blockIdx.x + 1]; i += blockDim.x) {
    const float x = w[i];
    m = fmaxf(m, (isfinite(x) && x >= 0.0f) ? x : __uint_as_float(0x7F800000u));
  }
  m = wave_max_f32(m);
  __shared__ float sh[kWavesPerBlock];
  if (lane_id() == 0) sh[threadIdx.x / kWave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < kWavesPerBlock; ++i) m = fmaxf(m, sh[i]);
    out[blockIdx.x] = m;
  }
}
__global__ void k_sample_quantize(const float* __restrict__ w, const int64_t* __restrict__ eoff,
                                  const int32_t* __restrict__ sexp, int64_t* __restrict__ wq) {
  const int se = sexp[blockIdx.x];
  for (int64_t i = eoff[blockIdx.x] + threadIdx.x; i < eoff[blockIdx.x + 1]; i += blockDim.x)
    wq[i] = static_cast<int64_t>(rint(ldexp(static_cast<double>(w[i]), se)));
}
// per-sample sum of k over the sample's vertex range (block per sample; exact)
__global__ void k_sample_sum(const int64_t* __restrict__ k, const int64_t* __restrict__ vbase,
                             int64_t* __restrict__ out) {
  int64_t t = 0;
  for (int64_t v = vbase[blockIdx.x] + threadIdx.x; v < vbase[blockIdx.x + 1]; v += blockDim.x) t += k[v];
  t = wave_sum_i64(t);
  __shared__ int64_t sh[kWavesPerBlock];
  if (lane_id() == 0) sh[threadIdx.x / kWave] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int i = 1; i < kWavesPerBlock; ++i) t += sh[i];
    out[blockIdx.x] = t;
  }
}
// out[o] = lab[orig2node[o]] - base[s] for the original vertices o of the samples s with
// mask[s] set (labels relative to the sample's own range)
__global__ void k_local_labels(const int32_t* __restrict__ lab, const int32_t* __restrict__ orig2node,
                               const int32_t* __restrict__ vsid0, const int64_t* __restrict__ base,
                               const uint8_t* __restrict__ mask, int64_t n0, int32_t* __restrict__ out) {
  const int64_t o = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (o >= n0) return;
  const int32_t sid = vsid0[o];
  if (mask == nullptr || mask[sid])
    out[o] = static_cast<int32_t>(static_cast<int64_t>(lab[orig2node ? orig2node[o] : o]) - base[sid]);
}
__global__ void k_copy_masked(const int32_t* __restrict__ src, const int32_t* __restrict__ vsid0,
                              const uint8_t* __restrict__ mask, int64_t n0, int32_t* __restrict__ dst) {
  const int64_t o = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (o < n0 && mask[vsid0[o]]) dst[o] = src[o];
}
// Pagg holds this level's move labels (sample s: [cb[s], cb[s+1])); the next level's vertices
// of sample s are [rb[s], rb[s+1]).  Community ids must lie in their own sample's vertex range,
// so they move to rb[s] + (label - cb[s]) (order within the sample kept).
__global__ void k_rebase_labels(const int32_t* __restrict__ src, const int32_t* __restrict__ vsid, int64_t n,
                                const int64_t* __restrict__ from, const int64_t* __restrict__ to,
                                int32_t* __restrict__ dst) {
  const int64_t v = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (v < n) {
    const int32_t sid = vsid[v];
    dst[v] = static_cast<int32_t>(static_cast<int64_t>(src[v]) - from[sid] + to[sid]);
  }
}
// out[o] = local[o] + vbase0[s]: per-sample labels into the level-0 id space
__global__ void k_global_labels(const int32_t* __restrict__ local, const int32_t* __restrict__ vsid0,
                                const int64_t* __restrict__ vbase0, int64_t n0, int32_t* __restrict__ out) {
  const int64_t o = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (o < n0) out[o] = static_cast<int32_t>(local[o] + vbase0[vsid0[o]]);
}
// modularity terms with each community's sample scale and total weight (k_q_terms per sample)
__global__ void k_q_terms_batched(const int64_t* __restrict__ tot, const int64_t* __restrict__ in, int64_t C,
                                  const int64_t* __restrict__ cbase, int S, const double* __restrict__ s_inv_s,
                                  const double* __restrict__ s_twom_real, double gamma, double* __restrict__ terms) {
  const int64_t c = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (c < C) {
    const int32_t sid = sample_of(cbase, S, c);
    const double inv_s = s_inv_s[sid], twom_real = s_twom_real[sid];
    const double a = (static_cast<double>(tot[c]) * inv_s) / twom_real;
    terms[c] = (static_cast<double>(in[c]) * inv_s) / twom_real - gamma * (a * a);
  }
}

// ------------------------------------------------------------------------------------------
// the driver
// ------------------------------------------------------------------------------------------
struct Ctx {
  hipStream_t st;
  int64_t n0, E0;
  double gamma;
  uint64_t thr;
  int s;           // fixed-point exponent
  double inv_s;
  // per-vertex (n0)
  int32_t *comm, *target, *P, *R, *Pagg, *orig2node, *final_lab, *prev_lab, *dense, *used, *map, *tmp_lab;
  uint8_t *active, *commit;
  int64_t *sigtot, *iw, *tot, *inw, *kbuf[2];
  double* terms;
  int32_t* pass_info;       // 2 * kMaxPasses: per pass, [0] = any vertex wanted to move
  int32_t *ekey, *boff;     // sigma_tot entries of the committed vertices, per commit block by bucket
  int64_t* eval;
  CommSig* cs;              // (comm, part, sigtot[comm]) per vertex for the light path of k_move
  int64_t* red;             // small reduction outputs
  float* fred;
  // graph slots
  int64_t* gindptr[2];
  int32_t* gindices[2];
  int64_t* gwq[2];
  GroupScratch grp;
  // contraction scratch (E0 + 1)
  uint32_t *ckd, *ceid, *cskd, *cseid, *cks, *csks, *cseid2;
  int32_t *cesrc, *cflag, *cpos;
  int64_t *cw, *cprefix, *cseg;
  uint32_t* csrc_out;
  void* sort_ws;
  size_t sort_ws_bytes;
  void* scan_ws;
  size_t scan_ws_bytes;
  Heavy hv;
  int32_t *hflag, *hpos;
  int64_t *hdeg, *hdegoff;
  int32_t *hub_class, *hub_flag, *hub_pos, *hub_verts;  // listed hub vertices (k_move), by class
  int64_t n_hub[kHubClasses] = {0, 0, 0, 0};
  int64_t n_huge = 0, huge_off = 0;  // k_move_huge's vertices: hub_verts + huge_off
  msc_leiden_info* info;
  // move-kernel timing (info->time_moves): an event pair per pass in a ring, read at the
  // host syncs of the pass loop
  static constexpr int kTimeRing = 16;  // pairs; > kSyncEvery
  bool time_moves = false;
  hipEvent_t tev[2 * kTimeRing] = {};
  int t_pending = 0, t_next = 0;
  double t_bytes_pending[kTimeRing] = {};
  // batched run (samples laid out block-diagonally); null for a single graph
  int32_t* vsid = nullptr;          // sample of each vertex of the current level graph
  int64_t* d_vbase = nullptr;       // sample vertex offsets of the current level graph (S + 1)
  double* s_inv_s = nullptr;        // per-sample 2^-s
  double* s_inv = nullptr;          // per-sample 1 / (twom * 2^-s)
  uint8_t* s_alive = nullptr;       // per sample: 0 once its run has ended (its vertices stay put)
};

static void carve_ctx(Carve& cv, Ctx& c, int64_t n, int64_t E) {
  const int64_t n1 = n + 1, E1 = E + 1;
  c.comm = cv.take<int32_t>(n1); c.target = cv.take<int32_t>(n1); c.P = cv.take<int32_t>(n1);
  c.R = cv.take<int32_t>(n1); c.Pagg = cv.take<int32_t>(n1); c.orig2node = cv.take<int32_t>(n1);
  c.final_lab = cv.take<int32_t>(n1); c.prev_lab = cv.take<int32_t>(n1); c.dense = cv.take<int32_t>(n1);
  c.used = cv.take<int32_t>(n1); c.map = cv.take<int32_t>(n1); c.tmp_lab = cv.take<int32_t>(n1);
  c.active = cv.take<uint8_t>(n1); c.commit = cv.take<uint8_t>(n1);
  c.sigtot = cv.take<int64_t>(n1); c.iw = cv.take<int64_t>(n1); c.tot = cv.take<int64_t>(n1);
  c.inw = cv.take<int64_t>(n1); c.kbuf[0] = cv.take<int64_t>(n1); c.kbuf[1] = cv.take<int64_t>(n1);
  c.terms = cv.take<double>(n1);
  c.pass_info = cv.take<int32_t>(2 * kMaxPasses);
  {
    const int64_t nb = ceil_div(n1, kCommitThreads);
    c.ekey = cv.take<int32_t>(2 * nb * kCommitThreads); c.eval = cv.take<int64_t>(2 * nb * kCommitThreads);
    c.boff = cv.take<int32_t>(nb * (kBuckets + 1));
  }
  c.cs = cv.take<CommSig>(n1);
  c.red = cv.take<int64_t>(8);
  c.fred = cv.take<float>(kMaxBlocks + 8);
  for (int i = 0; i < 2; ++i) {
    c.gindptr[i] = cv.take<int64_t>(n1);
    c.gindices[i] = cv.take<int32_t>(E1);
    c.gwq[i] = cv.take<int64_t>(E1);
  }
  GroupScratch& g = c.grp;
  g.keys = cv.take<uint32_t>(n1); g.vals = cv.take<uint32_t>(n1);
  g.skeys = cv.take<uint32_t>(n1); g.svals = cv.take<uint32_t>(n1);
  g.gathered = cv.take<int64_t>(n1); g.prefix = cv.take<int64_t>(n1); g.starts = cv.take<int64_t>(n1 + 1);
  c.ckd = cv.take<uint32_t>(E1); c.ceid = cv.take<uint32_t>(E1); c.cskd = cv.take<uint32_t>(E1);
  c.cseid = cv.take<uint32_t>(E1); c.cks = cv.take<uint32_t>(E1); c.csks = cv.take<uint32_t>(E1);
  c.cseid2 = cv.take<uint32_t>(E1);
  c.cesrc = cv.take<int32_t>(E1); c.cflag = cv.take<int32_t>(E1); c.cpos = cv.take<int32_t>(E1);
  c.cw = cv.take<int64_t>(E1); c.cprefix = cv.take<int64_t>(E1); c.cseg = cv.take<int64_t>(E1 + 1);
  c.csrc_out = cv.take<uint32_t>(E1);
  c.sort_ws_bytes = prims::radix_workspace(std::max(E1, n1));
  c.sort_ws = cv.take<char>(c.sort_ws_bytes);
  c.scan_ws_bytes = std::max(prims::scan_workspace<int64_t>(E1 + 1), prims::scan_workspace<int32_t>(E1 + 1));
  c.scan_ws = cv.take<char>(c.scan_ws_bytes);
  g.sort_ws = c.sort_ws; g.sort_ws_bytes = c.sort_ws_bytes;
  g.scan_ws = c.scan_ws; g.scan_ws_bytes = c.scan_ws_bytes;
  Heavy& h = c.hv;
  h.verts = cv.take<int32_t>(n1); h.eoff = cv.take<int64_t>(n1 + 1);
  h.eh_h = cv.take<int32_t>(E1); h.eh_e = cv.take<int64_t>(E1);
  h.kc = cv.take<uint32_t>(E1); h.kv = cv.take<uint32_t>(E1); h.skc = cv.take<uint32_t>(E1);
  h.skv = cv.take<uint32_t>(E1); h.kh = cv.take<uint32_t>(E1); h.skh = cv.take<uint32_t>(E1);
  h.skv2 = cv.take<uint32_t>(E1);
  h.wsorted = cv.take<int64_t>(E1); h.prefix = cv.take<int64_t>(E1);
  h.flag = cv.take<int32_t>(E1); h.segid = cv.take<int32_t>(E1);
  h.segstart = cv.take<int64_t>(E1 + 1); h.seg_h = cv.take<uint32_t>(E1); h.hseg = cv.take<int64_t>(n1 + 1);
  h.hstart = cv.take<int64_t>(n1 + 1);
  c.hflag = cv.take<int32_t>(n1); c.hpos = cv.take<int32_t>(n1);
  c.hdeg = cv.take<int64_t>(n1); c.hdegoff = cv.take<int64_t>(n1);
  c.hub_class = cv.take<int32_t>(n1); c.hub_flag = cv.take<int32_t>(n1); c.hub_pos = cv.take<int32_t>(n1);
  c.hub_verts = cv.take<int32_t>(n1);
}

template <typename T>
static int read_scalar(const T* dev, T* host, hipStream_t st) {
  MSC_HIP(hipMemcpyAsync(host, dev, sizeof(T), hipMemcpyDeviceToHost, st));
  MSC_HIP(hipStreamSynchronize(st));
  return kOk;
}

// graph derived quantities: k, twom, max degree, heavy structure
// destroys the timing events of a call on every return path
struct TimingGuard {
  Ctx& c;
  ~TimingGuard() {
    for (auto& e : c.tev)
      if (e) (void)hipEventDestroy(e);
  }
};

// adds the pending move-kernel timings to info (events complete: called after a host sync)
static int collect_move_times(Ctx& c) {
  for (int i = 0; i < c.t_pending; ++i) {
    const int slot = (c.t_next - c.t_pending + i + Ctx::kTimeRing) % Ctx::kTimeRing;
    float ms = 0.0f;
    MSC_HIP(hipEventElapsedTime(&ms, c.tev[2 * slot], c.tev[2 * slot + 1]));
    c.info->move_ms += ms;
    c.info->move_model_bytes += c.t_bytes_pending[slot];
    c.info->move_launches += 1;
  }
  c.t_pending = 0;
  return kOk;
}

static int graph_setup(Ctx& c, Graph& g) {
  hipStream_t st = c.st;
  k_degrees<<<grid_for(g.n * kWave), kThreads, 0, st>>>(g.indptr, g.wq, g.n, g.k);
  k_sum_i64_max_deg<<<1, kThreads, 0, st>>>(g.k, g.indptr, g.n, c.red);
  MSC_LAUNCH_CHECK();
  int64_t hostv[2];
  MSC_HIP(hipMemcpyAsync(hostv, c.red, 2 * sizeof(int64_t), hipMemcpyDeviceToHost, st));
  MSC_HIP(hipStreamSynchronize(st));
  g.twom = hostv[0];
  g.max_deg = hostv[1];
  // mid and heavy vertices
  Heavy& h = c.hv;
  h.nh = 0;
  h.eh = 0;
  for (int i = 0; i < kHubClasses; ++i) c.n_hub[i] = 0;
  c.n_huge = 0;
  if (g.max_deg > kLightDeg) {
    // the widest hub class whose community ids fit the sort key; above it k_move_huge (when
    // its keys fit too), then the sort path
    int64_t hub_max = kLightDeg;
    for (int i = 0; i < kHubClasses; ++i)
      if (g.n < hub_class_max_n(i)) hub_max = int64_t(512) << i;
    const int64_t heavy_min = (hub_max == kHubMaxDeg && g.n < kHugeMaxN) ? kHugeDeg : hub_max;
    k_heavy_flags<<<grid_for(g.n + 1), kThreads, 0, st>>>(g.indptr, g.n, hub_max, heavy_min, c.hflag, c.hdeg,
                                                          c.hub_class);
    MSC_LAUNCH_CHECK();
    int64_t off = 0;
    // hub_verts: class 3, 2, 1, 0 (k_move reads them in that order), then class 4 (k_move_huge)
    for (int o = 0; o <= kHubClasses; ++o) {
      const int i = o < kHubClasses ? kHubClasses - 1 - o : kHubClasses;
      if (i == kHubClasses ? heavy_min == hub_max || g.max_deg <= hub_max : int64_t(256) << i >= hub_max) continue;
      k_class_flag<<<grid_for(g.n + 1), kThreads, 0, st>>>(c.hub_class, g.n, i, c.hub_flag);
      MSC_LAUNCH_CHECK();
      MSC_TRY(prims::exclusive_scan<int32_t>(c.hub_flag, c.hub_pos, g.n + 1, c.scan_ws, c.scan_ws_bytes, st));
      int32_t cnt;
      MSC_TRY(read_scalar(c.hub_pos + g.n, &cnt, st));
      if (cnt > 0) {
        k_heavy_compact<<<grid_for(g.n), kThreads, 0, st>>>(c.hub_flag, c.hub_pos, g.n, c.hub_verts + off);
        MSC_LAUNCH_CHECK();
      }
      if (i == kHubClasses) {
        c.n_huge = cnt;
        c.huge_off = off;
      } else {
        c.n_hub[i] = cnt;
      }
      off += cnt;
    }
    MSC_TRY(prims::exclusive_scan<int32_t>(c.hflag, c.hpos, g.n + 1, c.scan_ws, c.scan_ws_bytes, st));
    MSC_TRY(prims::exclusive_scan<int64_t>(c.hdeg, c.hdegoff, g.n + 1, c.scan_ws, c.scan_ws_bytes, st));
    int32_t nh32;
    int64_t eh;
    MSC_TRY(read_scalar(c.hpos + g.n, &nh32, st));
    MSC_TRY(read_scalar(c.hdegoff + g.n, &eh, st));
    h.nh = nh32;
    h.eh = eh;
    k_heavy_compact<<<grid_for(g.n), kThreads, 0, st>>>(c.hflag, c.hpos, g.n, h.verts);
    k_heavy_edges<<<grid_for((h.nh + 1) * kWave), kThreads, 0, st>>>(g.indptr, h.verts, c.hdegoff, h.nh, h.eoff,
                                                                     h.eh_h, h.eh_e, h.eh);
    MSC_LAUNCH_CHECK();
  }
  return kOk;
}

static int sigma_tot(Ctx& c, const Graph& g, const int32_t* comm) {
  MSC_TRY(group_sort(comm, g.n, g.n, c.grp, c.st));
  MSC_TRY(group_sum(g.k, g.n, g.n, c.grp, c.sigtot, c.st));
  return kOk;
}

static int heavy_moves(Ctx& c, const Graph& g, const int32_t* comm, const int32_t* part, const GainParams& gp) {
  Heavy& h = c.hv;
  if (h.nh == 0) return kOk;
  hipStream_t st = c.st;
  const int64_t eh = h.eh;
  k_heavy_keys<<<grid_for(eh), kThreads, 0, st>>>(h.eh_h, h.eh_e, eh, h.verts, g.indices, comm, part, c.active,
                                                  g.n, h.kc, h.kv);
  MSC_LAUNCH_CHECK();
  MSC_TRY(prims::radix_sort_pairs(h.kc, h.kv, h.skc, h.skv, eh, prims::bits_for(g.n), c.sort_ws, c.sort_ws_bytes, st));
  k_heavy_hkeys<<<grid_for(eh), kThreads, 0, st>>>(h.skv, h.eh_h, eh, h.kh);
  MSC_LAUNCH_CHECK();
  MSC_TRY(prims::radix_sort_pairs(h.kh, h.skv, h.skh, h.skv2, eh, prims::bits_for(std::max<int64_t>(h.nh - 1, 1)),
                                  c.sort_ws, c.sort_ws_bytes, st));
  k_heavy_seg_flags<<<grid_for(eh + 1), kThreads, 0, st>>>(h.skv2, h.skh, eh, h.eh_e, g.indices, comm, h.kc, g.wq,
                                                           g.n, h.flag, h.wsorted);
  MSC_LAUNCH_CHECK();
  MSC_TRY(prims::exclusive_scan<int32_t>(h.flag, h.segid, eh + 1, c.scan_ws, c.scan_ws_bytes, st));
  MSC_TRY(prims::exclusive_scan<int64_t>(h.wsorted, h.prefix, eh + 1, c.scan_ws, c.scan_ws_bytes, st));
  k_heavy_seg_starts<<<grid_for(eh + 1), kThreads, 0, st>>>(h.flag, h.segid, eh, h.skh, h.segstart, h.seg_h);
  MSC_LAUNCH_CHECK();
  // first segment of each heavy vertex = segment id at its first sorted edge (no host sync)
  MSC_TRY(prims::segment_starts(h.skh, eh, h.nh, h.hstart, st));
  k_heavy_hseg<<<grid_for(h.nh + 1), kThreads, 0, st>>>(h.hstart, h.segid, h.nh, h.hseg);
  MSC_LAUNCH_CHECK();
  k_heavy_argmax<<<grid_for(h.nh * kWave), kThreads, 0, st>>>(h.verts, h.nh, h.hseg, h.segstart, h.prefix, h.skv2,
                                                              h.kc, g.k, comm, c.sigtot, gp, c.target);
  MSC_LAUNCH_CHECK();
  return kOk;
}

// synchronous local moving (part == nullptr) or refinement; comm updated in place
static int local_moving(Ctx& c, const Graph& g, int64_t seed, const int32_t* part, int* passes_out,
                        int* delta_passes_out) {
  hipStream_t st = c.st;
  GainParams gp;
  gp.inv_s = c.inv_s;
  gp.inv = 1.0 / (static_cast<double>(g.twom) * c.inv_s);
  gp.gamma = c.gamma;
  gp.vsid = c.vsid;
  gp.s_inv_s = c.s_inv_s;
  gp.s_inv = c.s_inv;
  const uint64_t key = stream_key(seed, part ? 1 : 0);
  if (c.vsid) k_active_alive<<<grid_for(g.n), kThreads, 0, st>>>(c.active, g.n, c.vsid, c.s_alive);
  else k_fill_u8<<<grid_for(g.n), kThreads, 0, st>>>(c.active, g.n, 1);
  k_fill_i32<<<grid_for(2 * kMaxPasses), kThreads, 0, st>>>(c.pass_info, 2 * kMaxPasses, 0);
  MSC_LAUNCH_CHECK();
  // sigma_tot is built from comm once (sort + segmented sums); every commit then keeps it
  // current (k_commit + k_sigtot_apply), and the frontier push rewrites each vertex's
  // (comm, part, sigma_tot) record after it
  const int nblk = static_cast<int>(ceil_div(g.n, kCommitThreads));
  MSC_TRY(sigma_tot(c, g, c.comm));
  k_refresh_cs<<<grid_for(g.n), kThreads, 0, st>>>(g.n, c.comm, part, c.sigtot, c.cs);
  MSC_LAUNCH_CHECK();
  int p = 0;
  for (; p < kMaxPasses; ++p) {
    // one launch moves every vertex up to the widest hub class: the listed hub vertices (one
    // block each) and the light ones; heavier vertices go through the sort path
    HubLists hubs;
    hubs.verts = c.hub_verts;
    int64_t nblocks = ceil_div(ceil_div(g.n, kMoveGroup) * kWave, kThreads);
    for (int i = 0; i < kHubClasses; ++i) {
      hubs.count[i] = c.n_hub[i];
      nblocks += c.n_hub[i];
    }
    const int tslot = c.t_next;
    if (c.time_moves) {
      if (c.t_pending == Ctx::kTimeRing) {  // ring full (no sync for a while): drain it
        MSC_HIP(hipStreamSynchronize(st));
        MSC_TRY(collect_move_times(c));
      }
      MSC_HIP(hipEventRecord(c.tev[2 * tslot], st));
    }
    k_move<<<static_cast<unsigned>(nblocks), kThreads, 0, st>>>(hubs, g.indptr, g.indices, g.wq, g.k, g.n, c.comm,
                                                               part, c.sigtot, c.cs, c.active, gp, c.target);
    MSC_LAUNCH_CHECK();
    if (c.time_moves) {
      MSC_HIP(hipEventRecord(c.tev[2 * tslot + 1], st));
      c.t_bytes_pending[tslot] = 8.0 * static_cast<double>(g.E) + 24.0 * static_cast<double>(g.n);
      c.t_next = (tslot + 1) % Ctx::kTimeRing;
      ++c.t_pending;
    }
    if (c.n_huge > 0) {
      k_move_huge<<<static_cast<unsigned>(c.n_huge), kHugeThreads, 0, st>>>(
          c.hub_verts + c.huge_off, g.indptr, g.indices, g.wq, g.k, c.comm, part, c.sigtot, c.active, gp, c.target);
      MSC_LAUNCH_CHECK();
    }
    MSC_TRY(heavy_moves(c, g, c.comm, part, gp));
    k_commit<<<static_cast<unsigned>(nblk), kCommitThreads, 0, st>>>(
        c.comm, c.target, g.n, pass_key(key, p), c.thr, c.commit, c.active, c.pass_info + 2 * p, g.k, c.ekey, c.eval,
        c.boff, c.vsid, c.d_vbase);
    k_sigtot_apply<<<kBuckets, kApplyThreads, 0, st>>>(c.ekey, c.eval, c.boff, nblk, c.sigtot);
    k_frontier_push<<<grid_for(g.n), kThreads, 0, st>>>(g.indptr, g.indices, g.n, c.commit, c.active, 1, c.comm, part,
                                                        c.sigtot, c.cs);
    MSC_LAUNCH_CHECK();
    if ((p % kSyncEvery == kSyncEvery - 1) || p == kMaxPasses - 1) {
      int32_t wants;  // any vertex wanted to move
      MSC_HIP(hipMemcpyAsync(&wants, c.pass_info + 2 * p, sizeof(wants), hipMemcpyDeviceToHost, st));
      MSC_HIP(hipStreamSynchronize(st));
      if (c.time_moves) MSC_TRY(collect_move_times(c));
      if (!wants) {
        ++p;
        break;
      }
    }
  }
  if (c.time_moves && c.t_pending > 0) {
    MSC_HIP(hipStreamSynchronize(st));
    MSC_TRY(collect_move_times(c));
  }
  *passes_out = std::min(p, kMaxPasses);
  // every pass after the first took sigma_tot from the previous commits alone
  if (delta_passes_out) *delta_passes_out = std::max(*passes_out - 1, 0);
  return kOk;
}

// dense relabel of lab (n) into out (may alias); returns number of labels
static int dense_relabel(Ctx& c, const int32_t* lab, int64_t n, int32_t* out, int64_t* C) {
  hipStream_t st = c.st;
  k_fill_i32<<<grid_for(n + 1), kThreads, 0, st>>>(c.used, n + 1, 0);
  k_mark_used<<<grid_for(n), kThreads, 0, st>>>(lab, n, c.used);
  MSC_LAUNCH_CHECK();
  MSC_TRY(prims::exclusive_scan<int32_t>(c.used, c.map, n + 1, c.scan_ws, c.scan_ws_bytes, st));
  k_apply_map<<<grid_for(n), kThreads, 0, st>>>(lab, c.map, n, out);
  MSC_LAUNCH_CHECK();
  int32_t nc;
  MSC_TRY(read_scalar(c.map + n, &nc, st));
  *C = nc;
  return kOk;
}

static int modularity(Ctx& c, const Graph& g, const int32_t* lab, int64_t C, double* q) {
  hipStream_t st = c.st;
  k_internal<<<grid_for(g.n * kWave), kThreads, 0, st>>>(g.indptr, g.indices, g.wq, lab, g.n, c.iw);
  MSC_LAUNCH_CHECK();
  MSC_TRY(group_sort(lab, g.n, C, c.grp, st));
  MSC_TRY(group_sum(g.k, g.n, C, c.grp, c.tot, st));
  MSC_TRY(group_sum(c.iw, g.n, C, c.grp, c.inw, st));
  const double twom_real = static_cast<double>(g.twom) * c.inv_s;
  k_q_terms<<<grid_for(C), kThreads, 0, st>>>(c.tot, c.inw, C, c.inv_s, twom_real, c.gamma, c.terms);
  MSC_LAUNCH_CHECK();
  std::vector<double> terms(static_cast<size_t>(C));
  MSC_HIP(hipMemcpyAsync(terms.data(), c.terms, C * sizeof(double), hipMemcpyDeviceToHost, st));
  MSC_HIP(hipStreamSynchronize(st));
  double s = 0.0;
  for (double t : terms) s += t;  // sequential, label order
  *q = s;
  return kOk;
}

// contract g on dense labels R (nR) into slot `dst`
static int contract(Ctx& c, const Graph& g, const int32_t* R, int64_t nR, int dst, Graph* out) {
  hipStream_t st = c.st;
  const int64_t E = g.E;
  k_edge_src_dst_keys<<<grid_for(g.n * kWave), kThreads, 0, st>>>(g.indptr, g.indices, R, g.n, c.ckd, c.ceid);
  k_edge_src_of<<<grid_for(g.n * kWave), kThreads, 0, st>>>(g.indptr, R, g.n, c.cesrc);
  MSC_LAUNCH_CHECK();
  const int bits = prims::bits_for(std::max<int64_t>(nR - 1, 1));
  MSC_TRY(prims::radix_sort_pairs(c.ckd, c.ceid, c.cskd, c.cseid, E, bits, c.sort_ws, c.sort_ws_bytes, st));
  k_gather_src_keys<<<grid_for(E), kThreads, 0, st>>>(c.cseid, c.cesrc, E, c.cks);
  MSC_LAUNCH_CHECK();
  MSC_TRY(prims::radix_sort_pairs(c.cks, c.cseid, c.csks, c.cseid2, E, bits, c.sort_ws, c.sort_ws_bytes, st));
  k_contract_flags<<<grid_for(E + 1), kThreads, 0, st>>>(c.cseid2, c.csks, c.ckd, g.wq, E, c.cflag, c.cw);
  MSC_LAUNCH_CHECK();
  MSC_TRY(prims::exclusive_scan<int32_t>(c.cflag, c.cpos, E + 1, c.scan_ws, c.scan_ws_bytes, st));
  MSC_TRY(prims::exclusive_scan<int64_t>(c.cw, c.cprefix, E + 1, c.scan_ws, c.scan_ws_bytes, st));
  int32_t m;
  MSC_TRY(read_scalar(c.cpos + E, &m, st));
  k_contract_emit<<<grid_for(E + 1), kThreads, 0, st>>>(c.cflag, c.cpos, E, c.cseid2, c.csks, c.ckd, c.cprefix,
                                                        c.gindices[dst], c.csrc_out, c.cseg);
  MSC_LAUNCH_CHECK();
  k_contract_weights<<<grid_for(m), kThreads, 0, st>>>(c.cseg, c.cprefix, m, c.gwq[dst]);
  MSC_LAUNCH_CHECK();
  MSC_TRY(prims::segment_starts(c.csrc_out, m, nR, c.gindptr[dst], st));
  out->n = nR;
  out->E = m;
  out->indptr = c.gindptr[dst];
  out->indices = c.gindices[dst];
  out->wq = c.gwq[dst];
  out->k = c.kbuf[dst];
  return graph_setup(c, *out);
}

// one multilevel pass (_leiden_pass); init (n0) may be null; result dense labels in c.final_lab
static int leiden_pass(Ctx& c, Graph& g0, int64_t seed, const int32_t* init) {
  hipStream_t st = c.st;
  const int64_t n0 = g0.n;
  Graph g = g0;
  int slot = 0;
  k_iota_i32<<<grid_for(n0), kThreads, 0, st>>>(c.orig2node, n0);
  if (init) k_copy_i32<<<grid_for(n0), kThreads, 0, st>>>(init, c.P, n0);
  else k_iota_i32<<<grid_for(n0), kThreads, 0, st>>>(c.P, n0);
  k_copy_i32<<<grid_for(n0), kThreads, 0, st>>>(c.P, c.final_lab, n0);  // P[orig2node] with identity map
  MSC_LAUNCH_CHECK();
  double q_prev = -1.0;
  for (int level = 0; level < kMaxLevels; ++level) {
    k_gather_i32<<<grid_for(n0), kThreads, 0, st>>>(c.P, c.orig2node, c.prev_lab, n0);
    k_copy_i32<<<grid_for(g.n), kThreads, 0, st>>>(c.P, c.comm, g.n);
    MSC_LAUNCH_CHECK();
    int mp = 0, rp = 0, md = 0, rd = 0;
    MSC_TRY(local_moving(c, g, seed + level, nullptr, &mp, &md));
    int64_t ncomm;
    MSC_TRY(dense_relabel(c, c.comm, g.n, c.P, &ncomm));
    double q;
    MSC_TRY(modularity(c, g, c.P, ncomm, &q));
    msc_leiden_level* rec = nullptr;
    if (c.info && c.info->n_levels < MSC_LEIDEN_MAX_RECORDS) {
      rec = &c.info->levels[c.info->n_levels++];
      rec->n = g.n; rec->n_edges = g.E; rec->move_passes = mp; rec->refine_passes = 0;
      rec->n_comms = ncomm; rec->n_refined = 0; rec->modularity = q;
      rec->move_delta_passes = md; rec->refine_delta_passes = 0;
    }
    if (q <= q_prev + kTol) {
      k_copy_i32<<<grid_for(n0), kThreads, 0, st>>>(c.prev_lab, c.final_lab, n0);
      MSC_LAUNCH_CHECK();
      break;
    }
    q_prev = q;
    k_gather_i32<<<grid_for(n0), kThreads, 0, st>>>(c.P, c.orig2node, c.final_lab, n0);
    MSC_LAUNCH_CHECK();
    if (ncomm == g.n) break;
    k_iota_i32<<<grid_for(g.n), kThreads, 0, st>>>(c.comm, g.n);
    MSC_LAUNCH_CHECK();
    MSC_TRY(local_moving(c, g, seed + level, c.P, &rp, &rd));
    int64_t nR;
    MSC_TRY(dense_relabel(c, c.comm, g.n, c.R, &nR));
    if (rec) { rec->refine_passes = rp; rec->n_refined = nR; rec->refine_delta_passes = rd; }
    if (nR == g.n) break;
    k_scatter_pagg<<<grid_for(g.n), kThreads, 0, st>>>(c.R, c.P, g.n, c.Pagg);
    k_gather_i32<<<grid_for(n0), kThreads, 0, st>>>(c.R, c.orig2node, c.tmp_lab, n0);
    k_copy_i32<<<grid_for(n0), kThreads, 0, st>>>(c.tmp_lab, c.orig2node, n0);
    MSC_LAUNCH_CHECK();
    Graph ng;
    MSC_TRY(contract(c, g, c.R, nR, slot, &ng));
    slot ^= 1;
    g = ng;
    k_copy_i32<<<grid_for(nR), kThreads, 0, st>>>(c.Pagg, c.P, nR);
    MSC_LAUNCH_CHECK();
  }
  int64_t nf;
  MSC_TRY(dense_relabel(c, c.final_lab, n0, c.final_lab, &nf));
  return kOk;
}

// multilevel Louvain (graph/louvain.py:397-437): every level moves from singletons with seed
// seed + 100 * level, relabels densely and stops, keeping the previous level's labels, when
// no vertex moved or Q did not rise by more than kTol; otherwise it aggregates on the moved
// partition itself (no refinement).  Result (dense) in c.final_lab.
static int louvain_pass(Ctx& c, Graph& g0, int64_t seed, int max_levels) {
  hipStream_t st = c.st;
  const int64_t n0 = g0.n;
  Graph g = g0;
  int slot = 0;
  k_iota_i32<<<grid_for(n0), kThreads, 0, st>>>(c.orig2node, n0);
  MSC_LAUNCH_CHECK();
  double q_prev = -1.0;
  for (int level = 0; level < max_levels; ++level) {
    k_iota_i32<<<grid_for(g.n), kThreads, 0, st>>>(c.comm, g.n);
    MSC_LAUNCH_CHECK();
    int mp = 0, md = 0;
    MSC_TRY(local_moving(c, g, seed + 100 * static_cast<int64_t>(level), nullptr, &mp, &md));
    int64_t ncomm;
    MSC_TRY(dense_relabel(c, c.comm, g.n, c.P, &ncomm));
    double q;
    MSC_TRY(modularity(c, g, c.P, ncomm, &q));
    if (c.info && c.info->n_levels < MSC_LEIDEN_MAX_RECORDS) {
      msc_leiden_level* rec = &c.info->levels[c.info->n_levels++];
      rec->n = g.n; rec->n_edges = g.E; rec->move_passes = mp; rec->refine_passes = 0;
      rec->n_comms = ncomm; rec->n_refined = ncomm; rec->modularity = q;
      rec->move_delta_passes = md; rec->refine_delta_passes = 0;
    }
    if (ncomm == g.n || q <= q_prev + kTol) break;
    k_gather_i32<<<grid_for(n0), kThreads, 0, st>>>(c.P, c.orig2node, c.tmp_lab, n0);
    k_copy_i32<<<grid_for(n0), kThreads, 0, st>>>(c.tmp_lab, c.orig2node, n0);
    MSC_LAUNCH_CHECK();
    Graph ng;
    MSC_TRY(contract(c, g, c.P, ncomm, slot, &ng));
    slot ^= 1;
    g = ng;
    q_prev = q;
  }
  // orig2node composes dense surjective maps, so it is dense already; the relabel keeps the
  // output contract independent of that argument
  int64_t nf;
  MSC_TRY(dense_relabel(c, c.orig2node, n0, c.final_lab, &nf));
  return kOk;
}

static int validate(int64_t n, int64_t nnz) {
  MSC_REQUIRE(n >= 1 && n < (int64_t(1) << 31) - 1, "leiden: n out of range");
  MSC_REQUIRE(nnz >= 0 && nnz < (int64_t(1) << 31) - 1, "leiden: nnz out of range (< 2^31)");
  return kOk;
}

static int ceil_log2(double x) {
  int e;
  const double m = std::frexp(x, &e);
  return m == 0.5 ? e - 1 : e;
}

// set-up shared by msc_leiden and msc_louvain: arguments, workspace, the fixed-point
// exponent and the quantised level-0 graph g0
static int begin_run(Ctx& c, Graph& g0, const int64_t* indptr, const int32_t* indices, const float* weights,
                     int64_t n, int64_t nnz, double resolution, double commit_prob, int32_t* out_labels,
                     msc_leiden_info* info, void* workspace, size_t workspace_bytes, void* stream) {
  MSC_TRY(validate(n, nnz));
  MSC_REQUIRE(indptr && out_labels && workspace && (nnz == 0 || (indices && weights)),
              "leiden: null pointer argument");
  MSC_REQUIRE(commit_prob > 0.0 && commit_prob <= 1.0, "leiden: commit_prob must be in (0, 1]");
  c.st = static_cast<hipStream_t>(stream);
  c.n0 = n;
  c.E0 = nnz;
  c.gamma = resolution;
  c.thr = static_cast<uint64_t>(std::min(std::floor(commit_prob * 4294967296.0), 4294967296.0));
  c.info = info;
  if (info) {
    info->n_levels = 0;
    info->move_launches = 0;
    info->move_ms = 0.0;
    info->move_model_bytes = 0.0;
    c.time_moves = info->time_moves != 0;
  }
  if (c.time_moves)
    for (auto& e : c.tev) MSC_HIP(hipEventCreate(&e));
  Carve cv(workspace, workspace_bytes);
  carve_ctx(cv, c, n, nnz);
  int64_t* wq0 = cv.take<int64_t>(nnz + 1);
  int64_t* k0 = cv.take<int64_t>(n + 1);
  if (!cv.ok()) {
    set_error("leiden: workspace too small (%zu < %zu)", workspace_bytes, cv.used);
    return kWorkspaceTooSmall;
  }
  hipStream_t st = c.st;
  // fixed-point exponent from max weight and edge count (both exact)
  float wmax = 0.0f;
  if (nnz > 0) {
    const int nb = static_cast<int>(std::min<int64_t>(kMaxBlocks, ceil_div(nnz, kThreads)));
    k_max_weight<<<nb, kThreads, 0, st>>>(weights, nnz, c.fred);
    k_max_weight<<<1, kThreads, 0, st>>>(c.fred, nb, c.fred + kMaxBlocks);
    MSC_LAUNCH_CHECK();
    MSC_TRY(read_scalar(c.fred + kMaxBlocks, &wmax, st));
  }
  MSC_REQUIRE(std::isfinite(wmax), "leiden: weights must be finite and >= 0");
  c.s = (nnz > 0 && wmax > 0.0f) ? 52 - ceil_log2(wmax) - ceil_log2(static_cast<double>(nnz)) : 0;
  c.inv_s = std::ldexp(1.0, -c.s);
  if (nnz > 0) {
    k_quantize<<<grid_for(nnz), kThreads, 0, st>>>(weights, nnz, c.s, wq0);
    MSC_LAUNCH_CHECK();
  }
  g0.n = n;
  g0.E = nnz;
  g0.indptr = const_cast<int64_t*>(indptr);
  g0.indices = const_cast<int32_t*>(indices);
  g0.wq = wq0;
  g0.k = k0;
  MSC_TRY(graph_setup(c, g0));
  MSC_REQUIRE(g0.twom > 0, "leiden: graph has no positive edge weight");
  return kOk;
}

// modularity of out_labels on g0, scale records, final synchronisation
static int end_run(Ctx& c, const Graph& g0, const int32_t* out_labels, double* out_modularity,
                   msc_leiden_info* info) {
  if (out_modularity) {
    int64_t C;
    MSC_TRY(dense_relabel(c, out_labels, g0.n, c.tmp_lab, &C));
    MSC_TRY(modularity(c, g0, c.tmp_lab, C, out_modularity));
  }
  if (info) {
    info->scale_exponent = c.s;
    info->twom_fixed = g0.twom;
  }
  MSC_HIP(hipStreamSynchronize(c.st));
  return kOk;
}

// ---- batched run ---------------------------------------------------------------------
// Many independent samples as one block-diagonal graph.  Every sample follows exactly the
// steps msc_leiden takes on it alone: its weights are quantised with its own exponent, its
// gains use its own total weight, its coins are drawn by its vertices' indices within the
// sample, and the pass loop runs in lockstep (a sample whose vertices stopped wanting to
// move makes no further change, which is what ending its own loop there would give).  The
// dense relabels keep each sample's communities contiguous and in their own order, so level
// graphs, labels and modularity are per sample what a single run produces.  A sample whose
// multilevel loop ends (no modularity gain, no merge, or no refinement) is frozen: its
// vertices stay inactive and its labels fixed while the others continue.
struct BatchRun {
  int S = 0;
  std::vector<int64_t> vb0, vb;       // sample vertex offsets: level 0 and the current level
  std::vector<double> twom_real;      // per sample
  std::vector<double> q_prev;
  std::vector<uint8_t> alive;
  int64_t* d_vb0 = nullptr;           // device copies
  int32_t* d_vsid0 = nullptr;         // sample of each original vertex
  int64_t* d_base = nullptr;          // per-sample label base of the current P numbering
  int64_t* d_tmp = nullptr;           // S + 1 gathered values
  int64_t* d_eoff = nullptr;
  double* d_twom_real = nullptr;
  uint8_t* d_mask = nullptr;          // per-sample mask for the label kernels
  float* d_wmax = nullptr;
  int32_t* d_sexp = nullptr;
  int64_t* d_twom = nullptr;
};

static void carve_batch(Carve& cv, Ctx& c, BatchRun& b, int64_t n, int S) {
  const size_t S1 = static_cast<size_t>(S) + 1;
  c.vsid = cv.take<int32_t>(n + 1);
  c.d_vbase = cv.take<int64_t>(S1);
  c.s_inv_s = cv.take<double>(S1);
  c.s_inv = cv.take<double>(S1);
  c.s_alive = cv.take<uint8_t>(S1);
  b.d_vb0 = cv.take<int64_t>(S1);
  b.d_vsid0 = cv.take<int32_t>(n + 1);
  b.d_base = cv.take<int64_t>(S1);
  b.d_tmp = cv.take<int64_t>(S1);
  b.d_eoff = cv.take<int64_t>(S1);
  b.d_twom_real = cv.take<double>(S1);
  b.d_mask = cv.take<uint8_t>(S1);
  b.d_wmax = cv.take<float>(S1);
  b.d_sexp = cv.take<int32_t>(S1);
  b.d_twom = cv.take<int64_t>(S1);
}

template <typename T>
static int upload(T* dev, const std::vector<T>& host, hipStream_t st) {
  MSC_HIP(hipMemcpyAsync(dev, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice, st));
  // host vectors may change right after: the copy must have consumed them
  MSC_HIP(hipStreamSynchronize(st));
  return kOk;
}

// per-sample first index of a dense relabel: map[vb[s]] for s <= S (map = c.map of the last relabel)
static int sample_bases(Ctx& c, BatchRun& b, std::vector<int64_t>& out) {
  hipStream_t st = c.st;
  k_gather_at<<<static_cast<unsigned>(ceil_div(b.S + 1, kThreads)), kThreads, 0, st>>>(nullptr, c.map, c.d_vbase, b.S,
                                                                                      b.d_tmp);
  MSC_LAUNCH_CHECK();
  out.resize(b.S + 1);
  MSC_HIP(hipMemcpyAsync(out.data(), b.d_tmp, (b.S + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, st));
  MSC_HIP(hipStreamSynchronize(st));
  return kOk;
}

static int set_mask(BatchRun& b, const std::vector<uint8_t>& m, hipStream_t st) { return upload(b.d_mask, m, st); }

// modularity of dense labels lab (C labels, sample s owning [cb[s], cb[s+1])) per sample
static int modularity_batched(Ctx& c, BatchRun& b, const Graph& g, const int32_t* lab, int64_t C,
                              const std::vector<int64_t>& cb, std::vector<double>& q) {
  hipStream_t st = c.st;
  k_internal<<<grid_for(g.n * kWave), kThreads, 0, st>>>(g.indptr, g.indices, g.wq, lab, g.n, c.iw);
  MSC_LAUNCH_CHECK();
  MSC_TRY(group_sort(lab, g.n, C, c.grp, st));
  MSC_TRY(group_sum(g.k, g.n, C, c.grp, c.tot, st));
  MSC_TRY(group_sum(c.iw, g.n, C, c.grp, c.inw, st));
  MSC_TRY(upload(b.d_tmp, cb, st));
  k_q_terms_batched<<<grid_for(C), kThreads, 0, st>>>(c.tot, c.inw, C, b.d_tmp, b.S, c.s_inv_s, b.d_twom_real,
                                                       c.gamma, c.terms);
  MSC_LAUNCH_CHECK();
  std::vector<double> terms(static_cast<size_t>(C));
  MSC_HIP(hipMemcpyAsync(terms.data(), c.terms, C * sizeof(double), hipMemcpyDeviceToHost, st));
  MSC_HIP(hipStreamSynchronize(st));
  q.assign(b.S, 0.0);
  for (int s = 0; s < b.S; ++s) {
    double acc = 0.0;
    for (int64_t t = cb[s]; t < cb[s + 1]; ++t) acc += terms[t];  // sequential, label order
    q[s] = acc;
  }
  return kOk;
}

static bool any_alive(const BatchRun& b) {
  for (uint8_t a : b.alive)
    if (a) return true;
  return false;
}

// one multilevel pass of every sample; init (global level-0 ids) may be null; the result,
// per-sample dense labels, in c.final_lab
static int leiden_pass_batched(Ctx& c, Graph& g0, int64_t seed, const int32_t* init, BatchRun& b) {
  hipStream_t st = c.st;
  const int64_t n0 = g0.n;
  const int S = b.S;
  Graph g = g0;
  int slot = 0;
  k_iota_i32<<<grid_for(n0), kThreads, 0, st>>>(c.orig2node, n0);
  if (init) k_copy_i32<<<grid_for(n0), kThreads, 0, st>>>(init, c.P, n0);
  else k_iota_i32<<<grid_for(n0), kThreads, 0, st>>>(c.P, n0);
  MSC_LAUNCH_CHECK();
  // the current level's offsets, sample ids and label base (level 0: the vertex offsets)
  b.vb = b.vb0;
  MSC_TRY(upload(c.d_vbase, b.vb, st));
  MSC_TRY(upload(b.d_base, b.vb0, st));
  k_copy_i32<<<grid_for(n0), kThreads, 0, st>>>(b.d_vsid0, c.vsid, n0);
  k_local_labels<<<grid_for(n0), kThreads, 0, st>>>(c.P, nullptr, b.d_vsid0, b.d_base, nullptr, n0, c.final_lab);
  MSC_LAUNCH_CHECK();
  b.q_prev.assign(S, -1.0);
  b.alive.assign(S + 1, 1);
  b.alive[S] = 0;
  MSC_TRY(upload(c.s_alive, b.alive, st));
  std::vector<int64_t> cb, rb;
  std::vector<double> q;
  std::vector<uint8_t> accept(S + 1, 0), reject(S + 1, 0);
  for (int level = 0; level < kMaxLevels; ++level) {
    // prev labels of the running samples (relative to their own P range)
    MSC_TRY(set_mask(b, b.alive, st));
    k_local_labels<<<grid_for(n0), kThreads, 0, st>>>(c.P, c.orig2node, b.d_vsid0, b.d_base, b.d_mask, n0,
                                                       c.prev_lab);
    k_copy_i32<<<grid_for(g.n), kThreads, 0, st>>>(c.P, c.comm, g.n);
    MSC_LAUNCH_CHECK();
    int mp = 0, md = 0;
    MSC_TRY(local_moving(c, g, seed + level, nullptr, &mp, &md));
    int64_t ncomm;
    MSC_TRY(dense_relabel(c, c.comm, g.n, c.P, &ncomm));
    MSC_TRY(sample_bases(c, b, cb));
    MSC_TRY(modularity_batched(c, b, g, c.P, ncomm, cb, q));
    std::fill(accept.begin(), accept.end(), 0);
    std::fill(reject.begin(), reject.end(), 0);
    for (int s = 0; s < S; ++s) {
      if (!b.alive[s]) continue;
      if (q[s] <= b.q_prev[s] + kTol) {
        reject[s] = 1;
        b.alive[s] = 0;
        continue;
      }
      b.q_prev[s] = q[s];
      accept[s] = 1;
      if (cb[s + 1] - cb[s] == b.vb[s + 1] - b.vb[s]) b.alive[s] = 0;
    }
    MSC_TRY(set_mask(b, reject, st));
    k_copy_masked<<<grid_for(n0), kThreads, 0, st>>>(c.prev_lab, b.d_vsid0, b.d_mask, n0, c.final_lab);
    MSC_LAUNCH_CHECK();
    MSC_TRY(set_mask(b, accept, st));
    MSC_TRY(upload(b.d_base, cb, st));
    k_local_labels<<<grid_for(n0), kThreads, 0, st>>>(c.P, c.orig2node, b.d_vsid0, b.d_base, b.d_mask, n0,
                                                       c.final_lab);
    MSC_LAUNCH_CHECK();
    if (!any_alive(b)) break;
    MSC_TRY(upload(c.s_alive, b.alive, st));
    k_iota_i32<<<grid_for(g.n), kThreads, 0, st>>>(c.comm, g.n);
    MSC_LAUNCH_CHECK();
    int rp = 0, rd = 0;
    MSC_TRY(local_moving(c, g, seed + level, c.P, &rp, &rd));
    int64_t nR;
    MSC_TRY(dense_relabel(c, c.comm, g.n, c.R, &nR));
    MSC_TRY(sample_bases(c, b, rb));
    for (int s = 0; s < S; ++s)
      if (b.alive[s] && rb[s + 1] - rb[s] == b.vb[s + 1] - b.vb[s]) b.alive[s] = 0;
    if (!any_alive(b)) break;
    MSC_TRY(upload(c.s_alive, b.alive, st));
    k_scatter_pagg<<<grid_for(g.n), kThreads, 0, st>>>(c.R, c.P, g.n, c.Pagg);
    k_gather_i32<<<grid_for(n0), kThreads, 0, st>>>(c.R, c.orig2node, c.tmp_lab, n0);
    k_copy_i32<<<grid_for(n0), kThreads, 0, st>>>(c.tmp_lab, c.orig2node, n0);
    MSC_LAUNCH_CHECK();
    Graph ng;
    MSC_TRY(contract(c, g, c.R, nR, slot, &ng));
    slot ^= 1;
    g = ng;
    // the next level's vertices are the refined communities (offsets rb); its starting
    // communities, this level's move labels, move into each sample's own vertex range
    b.vb = rb;
    MSC_TRY(upload(c.d_vbase, b.vb, st));
    k_fill_sid<<<grid_for(g.n), kThreads, 0, st>>>(c.d_vbase, S, g.n, c.vsid);
    MSC_LAUNCH_CHECK();
    MSC_TRY(upload(b.d_base, cb, st));
    k_rebase_labels<<<grid_for(nR), kThreads, 0, st>>>(c.Pagg, c.vsid, nR, b.d_base, c.d_vbase, c.P);
    MSC_LAUNCH_CHECK();
    MSC_TRY(upload(b.d_base, rb, st));
  }
  return kOk;
}

static int begin_batched(Ctx& c, BatchRun& b, Graph& g0, const int64_t* indptr, const int32_t* indices,
                         const float* weights, int64_t n, int64_t nnz, const int64_t* offsets, int S,
                         double resolution, double commit_prob, int32_t* out_labels, void* workspace,
                         size_t workspace_bytes, void* stream) {
  MSC_TRY(validate(n, nnz));
  MSC_REQUIRE(indptr && indices && weights && out_labels && workspace && offsets, "leiden_batched: null pointer");
  MSC_REQUIRE(S >= 1 && offsets[0] == 0 && offsets[S] == n, "leiden_batched: offsets must run from 0 to n");
  for (int s = 0; s < S; ++s)
    MSC_REQUIRE(offsets[s + 1] > offsets[s], "leiden_batched: sample %d is empty", s);
  MSC_REQUIRE(commit_prob > 0.0 && commit_prob <= 1.0, "leiden_batched: commit_prob must be in (0, 1]");
  c.st = static_cast<hipStream_t>(stream);
  hipStream_t st = c.st;
  c.n0 = n;
  c.E0 = nnz;
  c.gamma = resolution;
  c.thr = static_cast<uint64_t>(std::min(std::floor(commit_prob * 4294967296.0), 4294967296.0));
  c.info = nullptr;
  Carve cv(workspace, workspace_bytes);
  carve_ctx(cv, c, n, nnz);
  int64_t* wq0 = cv.take<int64_t>(nnz + 1);
  int64_t* k0 = cv.take<int64_t>(n + 1);
  carve_batch(cv, c, b, n, S);
  if (!cv.ok()) {
    set_error("leiden_batched: workspace too small (%zu < %zu)", workspace_bytes, cv.used);
    return kWorkspaceTooSmall;
  }
  b.S = S;
  b.vb0.assign(offsets, offsets + S + 1);
  MSC_TRY(upload(b.d_vb0, b.vb0, st));
  k_fill_sid<<<grid_for(n), kThreads, 0, st>>>(b.d_vb0, S, n, b.d_vsid0);
  k_gather_at<<<static_cast<unsigned>(ceil_div(S + 1, kThreads)), kThreads, 0, st>>>(indptr, nullptr, b.d_vb0, S,
                                                                                    b.d_eoff);
  k_sample_max_weight<<<S, kThreads, 0, st>>>(weights, b.d_eoff, b.d_wmax);
  MSC_LAUNCH_CHECK();
  std::vector<int64_t> eoff(S + 1);
  std::vector<float> wmax(S);
  MSC_HIP(hipMemcpyAsync(eoff.data(), b.d_eoff, (S + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, st));
  MSC_HIP(hipMemcpyAsync(wmax.data(), b.d_wmax, S * sizeof(float), hipMemcpyDeviceToHost, st));
  MSC_HIP(hipStreamSynchronize(st));
  MSC_REQUIRE(eoff[0] == 0 && eoff[S] == nnz, "leiden_batched: indptr does not match the offsets");
  std::vector<int32_t> sexp(S);
  std::vector<double> inv_s(S);
  for (int s = 0; s < S; ++s) {
    MSC_REQUIRE(std::isfinite(wmax[s]), "leiden_batched: weights must be finite and >= 0 (sample %d)", s);
    const int64_t Es = eoff[s + 1] - eoff[s];
    sexp[s] = (Es > 0 && wmax[s] > 0.0f) ? 52 - ceil_log2(wmax[s]) - ceil_log2(static_cast<double>(Es)) : 0;
    inv_s[s] = std::ldexp(1.0, -sexp[s]);
  }
  MSC_TRY(upload(b.d_sexp, sexp, st));
  k_sample_quantize<<<S, kThreads, 0, st>>>(weights, b.d_eoff, b.d_sexp, wq0);
  MSC_LAUNCH_CHECK();
  g0.n = n;
  g0.E = nnz;
  g0.indptr = const_cast<int64_t*>(indptr);
  g0.indices = const_cast<int32_t*>(indices);
  g0.wq = wq0;
  g0.k = k0;
  MSC_TRY(graph_setup(c, g0));
  k_sample_sum<<<S, kThreads, 0, st>>>(g0.k, b.d_vb0, b.d_twom);
  MSC_LAUNCH_CHECK();
  std::vector<int64_t> twom(S);
  MSC_HIP(hipMemcpyAsync(twom.data(), b.d_twom, S * sizeof(int64_t), hipMemcpyDeviceToHost, st));
  MSC_HIP(hipStreamSynchronize(st));
  std::vector<double> s_inv(S);
  b.twom_real.resize(S);
  for (int s = 0; s < S; ++s) {
    MSC_REQUIRE(twom[s] > 0, "leiden_batched: sample %d has no positive edge weight", s);
    s_inv[s] = 1.0 / (static_cast<double>(twom[s]) * inv_s[s]);
    b.twom_real[s] = static_cast<double>(twom[s]) * inv_s[s];
  }
  MSC_TRY(upload(c.s_inv_s, inv_s, st));
  MSC_TRY(upload(c.s_inv, s_inv, st));
  MSC_TRY(upload(b.d_twom_real, b.twom_real, st));
  c.s = 0;
  c.inv_s = 1.0;  // unused: every vertex takes its sample's scale
  return kOk;
}

}  // namespace leiden
}  // namespace msc

using namespace msc;
using namespace msc::leiden;

extern "C" int msc_leiden_workspace_size(int64_t n, int64_t nnz, size_t* bytes) {
  MSC_TRY(validate(n, nnz));
  MSC_REQUIRE(bytes, "leiden: bytes is null");
  Carve cv(nullptr, 0);
  Ctx c{};
  carve_ctx(cv, c, n, nnz);
  cv.take<int64_t>(nnz + 1);  // level-0 quantised weights
  cv.take<int64_t>(n + 1);    // level-0 degrees
  *bytes = cv.used + 256;
  return kOk;
}

extern "C" int msc_louvain_workspace_size(int64_t n, int64_t nnz, size_t* bytes) {
  return msc_leiden_workspace_size(n, nnz, bytes);
}

extern "C" int msc_leiden(const int64_t* indptr, const int32_t* indices, const float* weights, int64_t n,
                          int64_t nnz, double resolution, int64_t seed, int n_iterations, double commit_prob,
                          const int32_t* init_labels, int32_t* out_labels, double* out_modularity,
                          msc_leiden_info* info, void* workspace, size_t workspace_bytes, void* stream) {
  Ctx c{};
  TimingGuard timing_guard{c};
  Graph g0;
  MSC_TRY(begin_run(c, g0, indptr, indices, weights, n, nnz, resolution, commit_prob, out_labels, info, workspace,
                    workspace_bytes, stream));
  hipStream_t st = c.st;
  const int iters = std::max(1, n_iterations);
  for (int it = 0; it < iters; ++it) {
    const int32_t* init = nullptr;
    if (it > 0 || init_labels) {
      if (it == 0) {
        k_copy_i32<<<grid_for(n), kThreads, 0, st>>>(init_labels, out_labels, n);
        MSC_LAUNCH_CHECK();
      }
      init = out_labels;
    }
    // leiden_pass reads init before overwriting anything it aliases (init copied into P first)
    MSC_TRY(graph_setup(c, g0));
    MSC_TRY(leiden_pass(c, g0, seed + 17 * it, init));
    k_copy_i32<<<grid_for(n), kThreads, 0, st>>>(c.final_lab, out_labels, n);
    MSC_LAUNCH_CHECK();
  }
  return end_run(c, g0, out_labels, out_modularity, info);
}

extern "C" int msc_louvain(const int64_t* indptr, const int32_t* indices, const float* weights, int64_t n,
                           int64_t nnz, double resolution, int64_t seed, int max_levels, double commit_prob,
                           int32_t* out_labels, double* out_modularity, msc_leiden_info* info, void* workspace,
                           size_t workspace_bytes, void* stream) {
  MSC_REQUIRE(max_levels >= 1 && max_levels <= kMaxLevels, "louvain: max_levels must be in [1, %d]", kMaxLevels);
  Ctx c{};
  TimingGuard timing_guard{c};
  Graph g0;
  MSC_TRY(begin_run(c, g0, indptr, indices, weights, n, nnz, resolution, commit_prob, out_labels, info, workspace,
                    workspace_bytes, stream));
  MSC_TRY(louvain_pass(c, g0, seed, max_levels));
  k_copy_i32<<<grid_for(n), kThreads, 0, c.st>>>(c.final_lab, out_labels, n);
  MSC_LAUNCH_CHECK();
  return end_run(c, g0, out_labels, out_modularity, info);
}

extern "C" int msc_leiden_batched_workspace_size(int64_t n, int64_t nnz, int n_samples, size_t* bytes) {
  MSC_TRY(validate(n, nnz));
  MSC_REQUIRE(bytes && n_samples >= 1, "leiden_batched: bad arguments");
  Carve cv(nullptr, 0);
  Ctx c{};
  BatchRun b;
  carve_ctx(cv, c, n, nnz);
  cv.take<int64_t>(nnz + 1);
  cv.take<int64_t>(n + 1);
  carve_batch(cv, c, b, n, n_samples);
  *bytes = cv.used + 256;
  return kOk;
}

extern "C" int msc_leiden_batched(const int64_t* indptr, const int32_t* indices, const float* weights, int64_t n,
                                  int64_t nnz, const int64_t* offsets, int n_samples, double resolution, int64_t seed,
                                  int n_iterations, double commit_prob, int32_t* out_labels, double* out_modularity,
                                  void* workspace, size_t workspace_bytes, void* stream) {
  Ctx c{};
  BatchRun b;
  Graph g0;
  MSC_TRY(begin_batched(c, b, g0, indptr, indices, weights, n, nnz, offsets, n_samples, resolution, commit_prob,
                        out_labels, workspace, workspace_bytes, stream));
  hipStream_t st = c.st;
  const int iters = std::max(1, n_iterations);
  for (int it = 0; it < iters; ++it) {
    const int32_t* init = nullptr;
    if (it > 0) {  // the previous labels, in the level-0 id space
      k_global_labels<<<grid_for(n), kThreads, 0, st>>>(out_labels, b.d_vsid0, b.d_vb0, n, c.tmp_lab);
      MSC_LAUNCH_CHECK();
      init = c.tmp_lab;
    }
    MSC_TRY(graph_setup(c, g0));
    MSC_TRY(leiden_pass_batched(c, g0, seed + 17 * it, init, b));
    k_copy_i32<<<grid_for(n), kThreads, 0, st>>>(c.final_lab, out_labels, n);
    MSC_LAUNCH_CHECK();
  }
  if (out_modularity) {
    k_global_labels<<<grid_for(n), kThreads, 0, st>>>(out_labels, b.d_vsid0, b.d_vb0, n, c.tmp_lab);
    MSC_LAUNCH_CHECK();
    int64_t C;
    MSC_TRY(dense_relabel(c, c.tmp_lab, n, c.tmp_lab, &C));
    b.vb = b.vb0;
    MSC_TRY(upload(c.d_vbase, b.vb, st));
    std::vector<int64_t> cb;
    MSC_TRY(sample_bases(c, b, cb));
    std::vector<double> q;
    MSC_TRY(modularity_batched(c, b, g0, c.tmp_lab, C, cb, q));
    for (int s = 0; s < n_samples; ++s) out_modularity[s] = q[s];
  }
  MSC_HIP(hipStreamSynchronize(st));
  return kOk;
}

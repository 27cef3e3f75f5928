// Partition layout for the pruned exact kNN select (see part.h).
//
// Steps (one stream; the host reads a handful of counts in between):
//   1. centres   P0 = clamp(n / 256, 16, 512) rows at a fixed stride, then kRounds rounds of
//                D^2 sampling: row i is drawn when u_i < P0 D^2(i) / sum_j D^2(j), u_i a
//                counter-based hash of (round, i) and D(i) the distance to its nearest centre
//                so far (the oversampling rounds of k-means||, Bahmani et al., VLDB 2012).
//                A round's draws (at most 2 P0, lowest rows first) become centres and every
//                row moves to a new centre that is nearer.
//   2. radii     r(i) = |x_i - c(i)| in float32 direct differences; rows sorted by
//                (partition, r) with two stable radix sorts.
//   3. peeling   rows with r > 2 x their partition's median radius become singleton
//                partitions (when there are at most kPeelMax of them), so that a stray point
//                does not stretch a partition's ball over other clusters.
//   4. groups    connected components of the partitions whose balls (c_p, R_p) overlap, R_p
//                the largest radius kept in p, among the partitions with R_p <= 3 x the median
//                R over the rows (a small partition that straddles two clusters links nothing: as a
//                candidate it costs little, as a link it would merge the clusters' groups);
//                one pooled group for the groups under kMinGroup points (placed first: its
//                scattered rows need the most candidates), then the others, largest first (at most
//                kMaxGroups).
//   5. order     rows sorted by (group, partition, r); query blocks of <= 128 rows inside one
//                group.
//   6. bounds    UB(q) = (r(q) + rk(p(q))) (1 + eps), rk(p) the largest radius among the k
//                innermost members of q's partition p: those k points lie within UB(q) of q,
//                so q's k-th neighbour distance is at most UB(q).  (Rows of partitions with
//                fewer than k members take min over the other partitions p' of
//                |q - c_p'| + rk(p').)  A point x of partition p has |q - x| >= |q - c_p| - R_p,
//                so a block may skip p when |q - c_p| (1 - eps) - R_p (1 + eps) > UB(q) for
//                each of its queries; a ball test on the block's mean and radius screens most
//                partitions before the per-query test.  eps = 8 (d + 4) 2^-24 covers the float32
//                rounding of these distances and of the exact squared distance that orders the
//                neighbours, so a skipped point can neither beat nor tie the k-th neighbour.
//   7. ranges    a block's needed partitions as row ranges: consecutive partitions merged, then
//                the kMaxRanges - 1 widest gaps kept as cuts (scanning more rows is always safe).
//
// No result depends on these choices: centres, peeling, groups and the ball test only decide
// which candidates are scanned, and a candidate is skipped only on the bound above.
#include "part.h"

#include <algorithm>
#include <cmath>

#include "prims.h"

namespace msc {
namespace part {
namespace {

constexpr int kT = 256;
constexpr int kCompThreads = 1024;
constexpr int kPAll = kP0Max * (1 + 2 * kRounds) + kPeelMax;  // partitions, at most
constexpr int kPWords = (kPAll + 63) / 64;                    // 64-bit need words per block
constexpr int kSortCap = 8192;                                // >= kPAll, a power of two
static_assert(kSortCap >= kPAll, "component sort capacity");
constexpr int kMaxRuns = 1024;

__device__ __forceinline__ float inf_f() { return __builtin_inff(); }

// float32 direct-difference distance in four partial sums (dimensions t = u mod 4 in sum u;
// every use here carries the eps margin, which covers any summation order)
__device__ __forceinline__ float dist_f32(const float* __restrict__ a, const float* __restrict__ b, int d) {
  float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
  int t = 0;
  for (; t + 4 <= d; t += 4) {
    const float x0 = a[t] - b[t], x1 = a[t + 1] - b[t + 1], x2 = a[t + 2] - b[t + 2], x3 = a[t + 3] - b[t + 3];
    s0 = s0 + x0 * x0;
    s1 = s1 + x1 * x1;
    s2 = s2 + x2 * x2;
    s3 = s3 + x3 * x3;
  }
  for (; t < d; ++t) {
    const float x = a[t] - b[t];
    s0 = s0 + x * x;
  }
  return sqrtf((s0 + s1) + (s2 + s3));
}

// ---- 1. centres -----------------------------------------------------------------------------
__global__ void k_seed(const float* __restrict__ X, int64_t n, int d, int p0, float* __restrict__ C) {
  const int64_t e = blockIdx.x * static_cast<int64_t>(kT) + threadIdx.x;
  if (e >= static_cast<int64_t>(p0) * kDMax) return;
  const int64_t j = e / kDMax;
  const int t = static_cast<int>(e % kDMax);
  const int64_t row = ((2 * j + 1) * n) / (2 * static_cast<int64_t>(p0));
  C[e] = t < d ? X[row * d + t] : 0.0f;
}

__global__ void k_cnorm(const float* __restrict__ C, int c0, int c1, float* __restrict__ cn2) {
  const int c = c0 + blockIdx.x * kT + threadIdx.x;
  if (c >= c1) return;
  float s = 0.0f;
  for (int t = 0; t < kDMax; ++t) s = fmaf(C[static_cast<int64_t>(c) * kDMax + t], C[static_cast<int64_t>(c) * kDMax + t], s);
  cn2[c] = s;
}

// Nearest centre in [cb, ce) by |c|^2 - 2 x.c (+ |x|^2); a row keeps its previous centre unless
// a new one scores strictly lower.  Thread per row, the row in registers, 64 centres per LDS
// chunk read by broadcast, one centre at a time (two partial dot products).
template <int D4>
__global__ __launch_bounds__(kT) void k_assign(const float* __restrict__ X, int64_t n, int d,
                                               const float* __restrict__ C, const float* __restrict__ cn2,
                                               int cb, int ce, int first, int32_t* __restrict__ part,
                                               float* __restrict__ dmin2) {
  constexpr int CH = 64;
  __shared__ float4 cs[CH][D4];
  __shared__ float cns[CH];
  const int64_t i = blockIdx.x * static_cast<int64_t>(kT) + threadIdx.x;
  const bool valid = i < n;
  float4 xr[D4];
  float xn = 0.0f;
#pragma unroll
  for (int u = 0; u < D4; ++u) {
    float v[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int t = 4 * u + e;
      v[e] = (valid && t < d) ? X[i * d + t] : 0.0f;
      xn = fmaf(v[e], v[e], xn);
    }
    xr[u] = make_float4(v[0], v[1], v[2], v[3]);
  }
  float best = inf_f();
  int bp = 0;
  if (!first && valid) {
    best = dmin2[i];
    bp = part[i];
  }
  const float4* C4 = reinterpret_cast<const float4*>(C);
  for (int c0 = cb; c0 < ce; c0 += CH) {
    __syncthreads();
    for (int e = threadIdx.x; e < CH * D4; e += kT) {
      const int c = c0 + e / D4;
      cs[e / D4][e % D4] = c < ce ? C4[static_cast<int64_t>(c) * 16 + e % D4] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
    if (threadIdx.x < CH) cns[threadIdx.x] = (c0 + static_cast<int>(threadIdx.x) < ce) ? cn2[c0 + threadIdx.x] : inf_f();
    __syncthreads();
    const int cm = min(CH, ce - c0);
#pragma unroll 2
    for (int c = 0; c < cm; ++c) {
      float a0 = 0.f, a1 = 0.f;
#pragma unroll
      for (int u = 0; u < D4; ++u) {
        const float4 cv = cs[c][u];
        a0 = fmaf(xr[u].x, cv.x, a0);
        a1 = fmaf(xr[u].y, cv.y, a1);
        a0 = fmaf(xr[u].z, cv.z, a0);
        a1 = fmaf(xr[u].w, cv.w, a1);
      }
      const float sc = xn + cns[c] - 2.0f * (a0 + a1);
      if (sc < best) {
        best = sc;
        bp = c0 + c;
      }
    }
  }
  if (valid) {
    part[i] = bp;
    dmin2[i] = fmaxf(best, 0.0f);
  }
}

// The same choice on the 32x32x16 BF16 MFMA (d <= 62): scores |c|^2 - 2 x.c from one-term
// bf16 operands, |c|^2 carried as two bf16 terms against constant 1s in the query operand.  A
// block holds 128 rows (4 waves x 32, the B operand in registers) and streams the centres
// through LDS 32 at a time (the A operand).  Output layout of the MFMA: lane l holds row
// (l mod 32) against centres 8 (i div 4) + 4 (l div 32) + (i mod 4) of the tile, i < 16.  The
// rounding only changes which centre a row joins: every bound uses the exact float32 radius of the
// row to the centre it actually joined.
typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(16 * sizeof(float)))) float f32x16;
constexpr int kAsgStride = 72;  // bf16 per LDS centre row: 64 terms + 8 (an odd number of 16-B groups)

__global__ __launch_bounds__(kT) void k_assign_mfma(const float* __restrict__ X, int64_t n, int d,
                                                    const float* __restrict__ C, const float* __restrict__ cn2,
                                                    int cb, int ce, int first, int32_t* __restrict__ part,
                                                    float* __restrict__ dmin2) {
  __shared__ __attribute__((aligned(16))) __bf16 ct[32][kAsgStride];
  const int l = lane_id(), wv = threadIdx.x / kWave, h = l >> 5, lq = l & 31;
  const int64_t q = static_cast<int64_t>(blockIdx.x) * (4 * 32) + wv * 32 + lq;
  const bool valid = q < n;
  bf16x8 qb[4];
  float xn = 0.0f;
#pragma unroll
  for (int st = 0; st < 4; ++st) {
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int k = 16 * st + 8 * h + e;
      float v = 0.0f;
      if (k < d) {
        const float x = valid ? X[q * d + k] : 0.0f;
        xn = fmaf(x, x, xn);
        v = -2.0f * x;
      } else if (k == d || k == d + 1) {
        v = 1.0f;
      }
      qb[st][e] = static_cast<__bf16>(v);
    }
  }
  xn += __shfl_xor(xn, 32, kWave);  // the other half of the row's terms
  float best = inf_f();
  int bp = 0;
  if (!first && valid) {
    best = dmin2[q];
    bp = part[q];
  }
  for (int c0 = cb; c0 < ce; c0 += 32) {
    __syncthreads();
    {  // thread t converts 8 terms of centre t / 8: [c (d terms), hi(|c|^2), lo(|c|^2), 0 ...]
      const int m = threadIdx.x >> 3, j = threadIdx.x & 7, c = c0 + m;
      bf16x8 v;
      const float nrm = c < ce ? cn2[c] : 0.0f;
      const float hi = static_cast<float>(static_cast<__bf16>(nrm));
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int k = 8 * j + e;
        float x = 0.0f;
        if (c >= ce) x = (k == d) ? inf_f() : 0.0f;  // past the last centre: score +inf
        else if (k < d) x = C[static_cast<int64_t>(c) * kDMax + k];
        else if (k == d) x = hi;
        else if (k == d + 1) x = nrm - hi;
        v[e] = static_cast<__bf16>(x);
      }
      *reinterpret_cast<bf16x8*>(&ct[m][8 * j]) = v;
    }
    __syncthreads();
    f32x16 acc = f32x16{};
#pragma unroll
    for (int st = 0; st < 4; ++st) {
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(&ct[lq][16 * st + 8 * h]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, qb[st], acc, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const float sc = acc[i] + xn;  // D^2 = |c|^2 - 2 x.c + |x|^2
      if (sc < best) {  // centres ascend within a lane: ties keep the lower index
        best = sc;
        bp = c0 + 8 * (i >> 2) + 4 * h + (i & 3);
      }
    }
  }
  // the two lanes of a row: the smaller (D^2, centre)
  const float ob = __shfl_xor(best, 32, kWave);
  const int op = __shfl_xor(bp, 32, kWave);
  if (ob < best || (ob == best && op < bp)) {
    best = ob;
    bp = op;
  }
  if (valid && h == 0) {
    part[q] = bp;
    dmin2[q] = fmaxf(best, 0.0f);
  }
}

constexpr int kSumBlocks = 256;
// deterministic float64 sum of n floats: fixed grid, fixed per-thread stride, butterfly waves
__global__ __launch_bounds__(kT) void k_sum_partial(const float* __restrict__ v, int64_t n, double* __restrict__ partial) {
  double s = 0.0;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(kT) + threadIdx.x; i < n; i += static_cast<int64_t>(kSumBlocks) * kT)
    s += static_cast<double>(v[i]);
  s = wave_sum_f64(s);
  __shared__ double ws[kT / kWave];
  if (lane_id() == 0) ws[threadIdx.x / kWave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kT / kWave; ++w) t += ws[w];
    partial[blockIdx.x] = t;
  }
}
__global__ void k_sum_final(const double* __restrict__ partial, double* __restrict__ out) {
  if (threadIdx.x != 0) return;
  double t = 0.0;
  for (int b = 0; b < kSumBlocks; ++b) t += partial[b];
  *out = t;
}

__global__ void k_sample(const float* __restrict__ dmin2, int64_t n, const double* __restrict__ sum, double ell,
                         uint64_t key, int32_t* __restrict__ flags) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(kT) + threadIdx.x;
  if (i > n) return;
  int f = 0;
  if (i < n) {
    const double s = *sum;
    if (s > 0.0) {
      const double p = ell * static_cast<double>(dmin2[i]) / s;
      const double u = static_cast<double>(mix64(key ^ (static_cast<uint64_t>(i) * 0x9E3779B97F4A7C15ull)) >> 11) *
                       1.1102230246251565e-16;  // 2^-53
      f = u < p;
    }
  }
  flags[i] = f;
}

__global__ void k_take(const float* __restrict__ X, int64_t n, int d, const int32_t* __restrict__ flags,
                       const int32_t* __restrict__ pos, int cap, float* __restrict__ Cout) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(kT) + threadIdx.x;
  if (i >= n || !flags[i] || pos[i] >= cap) return;
  float* c = Cout + static_cast<int64_t>(pos[i]) * kDMax;
  for (int t = 0; t < kDMax; ++t) c[t] = t < d ? X[i * d + t] : 0.0f;
}

// ---- 2. radii and the (partition, r) order ----------------------------------------------------
__global__ void k_radius(const float* __restrict__ X, int64_t n, int d, const float* __restrict__ C,
                         const int32_t* __restrict__ part, float* __restrict__ r, uint32_t* __restrict__ rbits,
                         uint32_t* __restrict__ iota) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(kT) + threadIdx.x;
  if (i >= n) return;
  const float v = dist_f32(X + i * d, C + static_cast<int64_t>(part[i]) * kDMax, d);
  r[i] = v;
  rbits[i] = __float_as_uint(v);  // r >= 0: the bit pattern orders like the value
  iota[i] = static_cast<uint32_t>(i);
}

__global__ void k_keys_of(const int32_t* __restrict__ key, const uint32_t* __restrict__ idx, int64_t n,
                          uint32_t* __restrict__ out) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(kT) + threadIdx.x;
  if (i < n) out[i] = static_cast<uint32_t>(key[idx[i]]);
}

__global__ void k_gather_f(const float* __restrict__ v, const uint32_t* __restrict__ idx, int64_t n,
                           float* __restrict__ out) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(kT) + threadIdx.x;
  if (i < n) out[i] = v[idx[i]];
}

// per partition (rows [s, e) of the (partition, r) order): kept rows [s, cut), R = largest kept
// radius, rk = largest radius among the k innermost rows (inf with fewer than k rows)
__global__ void k_part_stats(const float* __restrict__ rs, const int64_t* __restrict__ pstart, int P, int k, int peel,
                             int64_t* __restrict__ cut, float* __restrict__ Rp, float* __restrict__ rk,
                             int32_t* __restrict__ size) {
  const int p = blockIdx.x * kT + threadIdx.x;
  if (p >= P) return;
  const int64_t s = pstart[p], e = pstart[p + 1];
  if (s == e) {
    cut[p] = s;
    Rp[p] = 0.0f;
    rk[p] = inf_f();
    size[p] = 0;
    return;
  }
  int64_t c = e;
  if (peel) {
    const float thr = 2.0f * rs[s + (e - s) / 2];
    int64_t lo = s, hi = e;
    while (lo < hi) {  // first row with rs > thr (rs ascending within the partition)
      const int64_t mid = (lo + hi) >> 1;
      if (rs[mid] > thr) hi = mid;
      else lo = mid + 1;
    }
    c = lo;  // > s: the first row is at most the median
  }
  cut[p] = c;
  Rp[p] = rs[c - 1];
  size[p] = static_cast<int32_t>(c - s);
  rk[p] = (c - s >= k) ? rs[s + k - 1] : inf_f();
}

__global__ void k_flag_out(const uint32_t* __restrict__ pk, const int64_t* __restrict__ cut, int64_t n,
                           int32_t* __restrict__ flags) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(kT) + threadIdx.x;
  if (i > n) return;
  flags[i] = (i < n && i >= cut[pk[i]]) ? 1 : 0;
}

// ---- 3. peeling: outlying rows become singleton partitions P + j ------------------------------
__global__ void k_peel(const float* __restrict__ X, int64_t n, int d, const uint32_t* __restrict__ ord,
                       const uint32_t* __restrict__ pk, const float* __restrict__ rs, const int32_t* __restrict__ flags,
                       const int32_t* __restrict__ pos, int P, int peel, int k, float* __restrict__ C,
                       float* __restrict__ Rp, float* __restrict__ rk, int32_t* __restrict__ size,
                       int32_t* __restrict__ part2, float* __restrict__ r2) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(kT) + threadIdx.x;
  if (i >= n) return;
  if (peel && flags[i]) {
    const int id = P + pos[i];
    part2[i] = id;
    r2[i] = 0.0f;
    float* c = C + static_cast<int64_t>(id) * kDMax;
    const int64_t row = ord[i];
    for (int t = 0; t < kDMax; ++t) c[t] = t < d ? X[row * d + t] : 0.0f;
    Rp[id] = 0.0f;
    rk[id] = k <= 1 ? 0.0f : inf_f();
    size[id] = 1;
  } else {
    part2[i] = static_cast<int32_t>(pk[i]);
    r2[i] = rs[i];
  }
}

// ---- 4. groups ---------------------------------------------------------------------------------
__device__ void block_bitonic_u64(uint64_t* a, int np2);

// rlim[0] = 3 x the median of R over the rows (each partition's R weighted by its size, so the
// singleton partitions, R = 0, barely move it); one workgroup
__global__ __launch_bounds__(kCompThreads) void k_radius_limit(const float* __restrict__ Rp,
                                                               const int32_t* __restrict__ size, int P,
                                                               float* __restrict__ rlim) {
  __shared__ uint64_t keys[kSortCap];
  int np2 = 2;
  while (np2 < P) np2 <<= 1;
  for (int i = threadIdx.x; i < np2; i += blockDim.x) {
    const bool ok = i < P && size[i] > 0;
    keys[i] = ok ? ((static_cast<uint64_t>(__float_as_uint(Rp[i])) << 32) | static_cast<uint32_t>(i)) : ~0ull;
  }
  __syncthreads();
  block_bitonic_u64(keys, np2);
  if (threadIdx.x == 0) {
    int64_t total = 0;
    for (int i = 0; i < P && keys[i] != ~0ull; ++i) total += size[static_cast<uint32_t>(keys[i])];
    int64_t acc = 0;
    float med = 0.0f;
    for (int i = 0; i < P && keys[i] != ~0ull; ++i) {
      acc += size[static_cast<uint32_t>(keys[i])];
      if (2 * acc >= total) {
        med = __uint_as_float(static_cast<uint32_t>(keys[i] >> 32));
        break;
      }
    }
    rlim[0] = 3.0f * med;
  }
}

// adj[p][w] bit b: partitions p and q = 32 w + b both non-empty, both with R <= rlim, with
// overlapping balls
__global__ void k_adjacency(const float* __restrict__ C, int P, int d, const float* __restrict__ Rp,
                            const int32_t* __restrict__ size, const float* __restrict__ rlim, int W,
                            uint32_t* __restrict__ adj) {
  const int64_t e = blockIdx.x * static_cast<int64_t>(kT) + threadIdx.x;
  if (e >= static_cast<int64_t>(P) * W) return;
  const int p = static_cast<int>(e / W), w = static_cast<int>(e % W);
  const float lim = rlim[0];
  uint32_t bits = 0;
  if (size[p] > 0 && Rp[p] <= lim) {
    const float* cp = C + static_cast<int64_t>(p) * kDMax;
    for (int b = 0; b < 32; ++b) {
      const int q = 32 * w + b;
      if (q < P && size[q] > 0 && Rp[q] <= lim) {
        const float D = dist_f32(cp, C + static_cast<int64_t>(q) * kDMax, d);
        if (D <= Rp[p] + Rp[q]) bits |= 1u << b;
      }
    }
  }
  adj[e] = bits;
}

__device__ void block_bitonic_u64(uint64_t* a, int np2) {
  for (int kk = 2; kk <= np2; kk <<= 1) {
    for (int j = kk >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < np2; i += blockDim.x) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const bool asc = (i & kk) == 0;
          const uint64_t x = a[i], y = a[ixj];
          if ((x > y) == asc) {
            a[i] = y;
            a[ixj] = x;
          }
        }
      }
      __syncthreads();
    }
  }
}

// One workgroup: component labels (the smallest partition id of each component, by min-label
// sweeps with pointer jumping until a sweep changes nothing), group sizes, the group order (rank
// 0: the pooled group; then the groups of >= kMinGroup points, largest first, then the id, at
// most kMaxGroups), and the partition order (group rank, id).
__global__ __launch_bounds__(kCompThreads) void k_components(const uint32_t* __restrict__ adj, int W, int P,
                                                           const int32_t* __restrict__ size,
                                                           int32_t* __restrict__ ppos, int32_t* __restrict__ porder,
                                                           uint32_t* __restrict__ pgr, int64_t* __restrict__ counts) {
  __shared__ int lab[kPAll];
  __shared__ int gsz[kPAll];
  __shared__ uint64_t keys[kSortCap];
  __shared__ int changed;
  __shared__ int nbig;
  for (int p = threadIdx.x; p < P; p += blockDim.x) lab[p] = p;
  __syncthreads();
  for (int it = 0; it <= P + 1; ++it) {
    if (threadIdx.x == 0) changed = 0;
    __syncthreads();
    for (int p = threadIdx.x; p < P; p += blockDim.x) {
      int m = lab[p];
      const uint32_t* row = adj + static_cast<int64_t>(p) * W;
      for (int w = 0; w < W; ++w) {
        uint32_t bits = row[w];
        while (bits) {
          const int b = __builtin_ctz(bits);
          bits &= bits - 1;
          m = min(m, lab[32 * w + b]);
        }
      }
      if (m < lab[p]) {
        lab[p] = m;
        changed = 1;
      }
    }
    __syncthreads();
    for (int rep = 0; rep < 4; ++rep) {
      for (int p = threadIdx.x; p < P; p += blockDim.x) lab[p] = lab[lab[p]];
      __syncthreads();
    }
    if (!changed) break;
    __syncthreads();
  }
  // group sizes by root
  for (int p = threadIdx.x; p < P; p += blockDim.x) gsz[p] = 0;
  __syncthreads();
  for (int p = threadIdx.x; p < P; p += blockDim.x)
    if (size[p] > 0) atomicAdd(&gsz[lab[p]], size[p]);
  __syncthreads();
  int np2 = 2;
  while (np2 < P) np2 <<= 1;
  for (int i = threadIdx.x; i < np2; i += blockDim.x) {
    uint64_t key = ~0ull;
    if (i < P && lab[i] == i && gsz[i] >= kMinGroup)
      key = (static_cast<uint64_t>(0x7FFFFFFF - gsz[i]) << 32) | static_cast<uint32_t>(i);
    keys[i] = key;
  }
  __syncthreads();
  block_bitonic_u64(keys, np2);
  if (threadIdx.x == 0) {
    int c = 0;
    while (c < P && c < kMaxGroups && keys[c] != ~0ull) ++c;
    nbig = c;
  }
  __syncthreads();
  const int G = nbig;
  // gsz[root] <- group rank: the pooled group first (rank 0: its queries are scattered rows
  // whose blocks scan the most candidates, so they start first), then the groups by size
  for (int p = threadIdx.x; p < P; p += blockDim.x) gsz[p] = 0;
  __syncthreads();
  for (int i = threadIdx.x; i < G; i += blockDim.x) gsz[static_cast<uint32_t>(keys[i])] = i + 1;
  __syncthreads();
  for (int i = threadIdx.x; i < np2; i += blockDim.x)
    keys[i] = i < P ? ((static_cast<uint64_t>(gsz[lab[i]]) << 32) | static_cast<uint32_t>(i)) : ~0ull;
  __syncthreads();
  block_bitonic_u64(keys, np2);
  for (int i = threadIdx.x; i < P; i += blockDim.x) {
    const int p = static_cast<int>(static_cast<uint32_t>(keys[i]));
    porder[i] = p;
    ppos[p] = i;
    pgr[i] = static_cast<uint32_t>(keys[i] >> 32);
  }
  if (threadIdx.x == 0) counts[0] = G;
}

// ---- 5. final order, groups' rows, query blocks ----------------------------------------------
__global__ void k_pos_keys(const int32_t* __restrict__ part2, const int32_t* __restrict__ ppos, int64_t n,
                           uint32_t* __restrict__ keys, uint32_t* __restrict__ iota) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(kT) + threadIdx.x;
  if (i >= n) return;
  keys[i] = static_cast<uint32_t>(ppos[part2[i]]);
  iota[i] = static_cast<uint32_t>(i);
}

__global__ void k_final_rows(const uint32_t* __restrict__ ord, const float* __restrict__ r2,
                             const uint32_t* __restrict__ fv, int64_t n, int32_t* __restrict__ perm,
                             float* __restrict__ rfin) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(kT) + threadIdx.x;
  if (i >= n) return;
  const uint32_t v = fv[i];
  perm[i] = static_cast<int32_t>(ord[v]);
  rfin[i] = r2[v];
}

__global__ void k_gather_rows(const float* __restrict__ X, const int32_t* __restrict__ perm, int64_t n, int d,
                              float* __restrict__ Xs) {
  const int64_t e = blockIdx.x * static_cast<int64_t>(kT) + threadIdx.x;
  if (e >= n * d) return;
  const int64_t i = e / d;
  Xs[e] = X[static_cast<int64_t>(perm[i]) * d + (e - i * d)];
}

// partition arrays in position order
__global__ void k_perm_parts(const int32_t* __restrict__ porder, int P, const float* __restrict__ C,
                             const float* __restrict__ Rp, const float* __restrict__ rk, const int32_t* __restrict__ size,
                             float* __restrict__ Cpos, float* __restrict__ Rpos, float* __restrict__ rkpos,
                             int32_t* __restrict__ szpos) {
  const int64_t e = blockIdx.x * static_cast<int64_t>(kT) + threadIdx.x;
  if (e >= static_cast<int64_t>(P) * kDMax) return;
  const int pos = static_cast<int>(e / kDMax), t = static_cast<int>(e % kDMax);
  const int p = porder[pos];
  Cpos[e] = C[static_cast<int64_t>(p) * kDMax + t];
  if (t == 0) {
    Rpos[pos] = Rp[p];
    rkpos[pos] = rk[p];
    szpos[pos] = size[p];
  }
}

// group g's rows start at fstart[gpstart[g]] (g = 0..G+1; the pooled group 0 may be empty)
__global__ void k_group_rows(const int64_t* __restrict__ fstart, const int64_t* __restrict__ gpstart, int ng,
                             int64_t* __restrict__ grow) {
  const int g = blockIdx.x * kT + threadIdx.x;
  if (g < ng) grow[g] = fstart[gpstart[g]];
}

// blocks per group and their offsets (one workgroup; ng <= kMaxGroups + 1 groups)
__global__ __launch_bounds__(kCompThreads) void k_block_offsets(const int64_t* __restrict__ grow, int ng,
                                                                int64_t* __restrict__ boff, int64_t* __restrict__ counts) {
  __shared__ int64_t s[kMaxGroups + 2];
  for (int g = threadIdx.x; g < ng; g += blockDim.x) s[g] = ceil_div(grow[g + 1] - grow[g], kBlockQ);
  __syncthreads();
  if (threadIdx.x == 0) {
    int64_t acc = 0;
    for (int g = 0; g < ng; ++g) {
      const int64_t v = s[g];
      boff[g] = acc;
      acc += v;
    }
    boff[ng] = acc;
    counts[1] = acc;
  }
}

__global__ void k_blocks_fill(const int64_t* __restrict__ grow, const int64_t* __restrict__ boff, int ng, int64_t nblk,
                              int32_t* __restrict__ qrows, int32_t* __restrict__ bfirst, int32_t* __restrict__ bcnt) {
  const int64_t s = blockIdx.x * static_cast<int64_t>(kT) + threadIdx.x;
  if (s >= nblk * kBlockQ) return;
  const int64_t b = s / kBlockQ;
  const int j = static_cast<int>(s % kBlockQ);
  int lo = 0, hi = ng;  // last g with boff[g] <= b
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (boff[mid] <= b) lo = mid;
    else hi = mid;
  }
  const int64_t first = grow[lo] + kBlockQ * (b - boff[lo]);
  const int64_t cnt = min(static_cast<int64_t>(kBlockQ), grow[lo + 1] - first);
  qrows[s] = j < cnt ? static_cast<int32_t>(first + j) : -1;
  if (j == 0) {
    bfirst[b] = static_cast<int32_t>(first);
    bcnt[b] = static_cast<int32_t>(cnt);
  }
}

// ---- 6. bounds -----------------------------------------------------------------------------------
__global__ void k_ub(const float* __restrict__ rfin, const uint32_t* __restrict__ fk, const float* __restrict__ rkpos,
                     int64_t n, double eps, float* __restrict__ ubq) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(kT) + threadIdx.x;
  if (i >= n) return;
  const float rk = rkpos[fk[i]];
  float ub = inf_f();
  if (!isinf(rk))  // rounded up past the float conversion
    ub = static_cast<float>((static_cast<double>(rfin[i]) + static_cast<double>(rk)) * (1.0 + eps) * (1.0 + 1e-6));
  ubq[i] = ub;
}

// rows whose own partition holds fewer than k points: the best bound over the partitions that
// do (one wave per row; rows with a finite bound return at once)
__global__ __launch_bounds__(kT) void k_ub_fix(const float* __restrict__ Xs, int64_t n, int d,
                                               const float* __restrict__ Cpos, const float* __restrict__ rkpos,
                                               int P, double eps, float* __restrict__ ubq) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(kT / kWave) + threadIdx.x / kWave;
  if (i >= n || !isinf(ubq[i])) return;
  const int l = lane_id();
  double m = __builtin_inf();
  for (int p = l; p < P; p += kWave) {
    const float rk = rkpos[p];
    if (isinf(rk)) continue;
    const double D = static_cast<double>(dist_f32(Xs + i * d, Cpos + static_cast<int64_t>(p) * kDMax, d));
    m = fmin(m, (D + static_cast<double>(rk)) * (1.0 + eps));
  }
  // m >= 0: the bit patterns order like the values
  uint64_t bits = static_cast<uint64_t>(__double_as_longlong(m));
  for (int o = 32; o >= 1; o >>= 1) bits = umin64(bits, shfl_xor_u64(bits, o));
  m = __longlong_as_double(static_cast<long long>(bits));
  if (l == 0) ubq[i] = isinf(m) ? inf_f() : static_cast<float>(m * (1.0 + 1e-6));
}

// ---- 7. needed partitions -> candidate ranges (workgroup per query block) ----------------------
// distance between two zero-padded rows of D4 float4 (four partial sums, like dist_f32)
template <int D4>
__device__ __forceinline__ float dist4(const float4* __restrict__ a, const float4 (&b)[D4]) {
  float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
#pragma unroll
  for (int u = 0; u < D4; ++u) {
    const float4 x = a[u];
    const float d0 = x.x - b[u].x, d1 = x.y - b[u].y, d2 = x.z - b[u].z, d3 = x.w - b[u].w;
    s0 = s0 + d0 * d0;
    s1 = s1 + d1 * d1;
    s2 = s2 + d2 * d2;
    s3 = s3 + d3 * d3;
  }
  return sqrtf((s0 + s1) + (s2 + s3));
}

template <int D4>
__global__ __launch_bounds__(kT) void k_need(const float* __restrict__ Xs, int d, const float* __restrict__ ubq,
                                             const int32_t* __restrict__ bfirst, const int32_t* __restrict__ bcnt,
                                             const float* __restrict__ Cpos, const float* __restrict__ Rpos,
                                             const int32_t* __restrict__ szpos, const int64_t* __restrict__ fstart,
                                             int P, double eps, int64_t* __restrict__ ranges,
                                             int64_t* __restrict__ counts, unsigned long long* __restrict__ pairs) {
  __shared__ float4 qs[kBlockQ][D4];
  __shared__ float ub[kBlockQ];
  __shared__ float4 mean[D4];
  __shared__ float red[kT / kWave];
  __shared__ uint64_t need[kPWords];
  __shared__ int rs_[kMaxRuns], re_[kMaxRuns];
  __shared__ int wcnt[kT];
  __shared__ int nruns;
  const int64_t b = blockIdx.x;
  const int64_t first = bfirst[b];
  const int cnt = bcnt[b];
  const int tid = threadIdx.x;
  const int l = lane_id(), wv = tid / kWave;
  constexpr int DW = 4 * D4;
  for (int e = tid; e < kBlockQ * DW; e += kT) {
    const int q = e / DW, t = e % DW;
    reinterpret_cast<float*>(qs[q])[t] = (q < cnt && t < d) ? Xs[(first + q) * d + t] : 0.0f;
  }
  for (int q = tid; q < kBlockQ; q += kT) ub[q] = q < cnt ? ubq[first + q] : -inf_f();
  __syncthreads();
  if (tid < DW) {
    float sm = 0.0f;
    for (int q = 0; q < cnt; ++q) sm += reinterpret_cast<const float*>(qs[q])[tid];
    reinterpret_cast<float*>(mean)[tid] = sm / static_cast<float>(cnt);
  }
  float u = tid < kBlockQ ? ub[tid] : -inf_f();
  u = wave_max_f32(u);
  if (l == 0) red[wv] = u;
  __syncthreads();
  const float UB = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
  float rho = 0.0f;
  if (tid < cnt) {
    float4 qv[D4];
#pragma unroll
    for (int v = 0; v < D4; ++v) qv[v] = qs[tid][v];
    rho = dist4<D4>(mean, qv);
  }
  rho = wave_max_f32(rho);
  __syncthreads();
  if (l == 0) red[wv] = rho;
  __syncthreads();
  const double rhob = static_cast<double>(fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]))) * (1.0 + eps);
  const double ubb = static_cast<double>(UB);
  const double lo_f = 1.0 - eps, hi_f = 1.0 + eps;
  // a partition is needed (1) unless the bound clears it; empty partitions count as needed (no
  // rows: they only join neighbouring runs)
  const int PW = (P + 63) / 64;
  const float4* C4 = reinterpret_cast<const float4*>(Cpos);
  for (int base = 0; base < PW * 64; base += kT) {
    const int pos = base + tid;
    bool nd = false;
    if (pos < P) {
      if (szpos[pos] == 0 || isinf(UB)) {
        nd = true;
      } else {
        float4 c[D4];
#pragma unroll
        for (int v = 0; v < D4; ++v) c[v] = C4[static_cast<int64_t>(pos) * 16 + v];
        const double R = static_cast<double>(Rpos[pos]) * hi_f;
        const double Dm = static_cast<double>(dist4<D4>(mean, c));
        if (Dm * lo_f - rhob - R <= ubb) {
          for (int q = 0; q < cnt; ++q) {
            const double Dq = static_cast<double>(dist4<D4>(qs[q], c));
            if (Dq * lo_f - R <= static_cast<double>(ub[q])) {
              nd = true;
              break;
            }
          }
        }
      }
    }
    const uint64_t m = __builtin_amdgcn_ballot_w64(nd);
    if (l == 0 && (base + kWave * wv) / 64 < PW) need[(base + kWave * wv) / 64] = m;
  }
  __syncthreads();
  // runs of needed positions: thread tid owns positions [tid * CH, (tid + 1) * CH)
  const int CH = (P + kT - 1) / kT;
  auto nbit = [&](int pos) { return pos >= 0 && pos < P && ((need[pos >> 6] >> (pos & 63)) & 1ull); };
  int mine = 0;
  for (int pos = tid * CH; pos < min(P, (tid + 1) * CH); ++pos)
    if (nbit(pos) && !nbit(pos - 1)) ++mine;
  wcnt[tid] = mine;
  __syncthreads();
  if (tid == 0) {
    int acc = 0;
    for (int t = 0; t < kT; ++t) {
      const int v = wcnt[t];
      wcnt[t] = acc;
      acc += v;
    }
    nruns = acc;
  }
  __syncthreads();
  const int R_ = nruns;
  if (R_ <= kMaxRuns) {
    int o = wcnt[tid];
    for (int pos = tid * CH; pos < min(P, (tid + 1) * CH); ++pos) {
      if (nbit(pos) && !nbit(pos - 1)) {
        int e = pos;
        while (nbit(e + 1)) ++e;
        rs_[o] = pos;
        re_[o] = e;
        ++o;
      }
    }
  }
  __syncthreads();
  if (tid == 0) {
    // row ranges of the runs, zero-row runs dropped; then keep the widest kMaxRanges - 1 gaps
    int nr = 0;
    int64_t ra[kMaxRanges], rb[kMaxRanges];
    if (R_ > kMaxRuns) {  // too many runs: one range over all of them (a superset)
      int lo = 0, hi = P - 1;
      while (lo < P && !nbit(lo)) ++lo;
      while (hi >= 0 && !nbit(hi)) --hi;
      if (lo <= hi && fstart[hi + 1] > fstart[lo]) {
        ra[0] = fstart[lo];
        rb[0] = fstart[hi + 1];
        nr = 1;
      }
    } else {
      // the kMaxRanges - 1 widest gaps between consecutive non-empty runs (ties: earliest)
      int64_t gw[kMaxRanges - 1];
      int gi[kMaxRanges - 1];
      for (int j = 0; j < kMaxRanges - 1; ++j) {
        gw[j] = 0;
        gi[j] = -1;
      }
      int64_t prev_end = -1;
      for (int i = 0; i < R_; ++i) {
        const int64_t s = fstart[rs_[i]], e = fstart[re_[i] + 1];
        if (e == s) continue;
        if (prev_end >= 0) {
          const int64_t gap = s - prev_end;
          // insert (gap, i) into the top list (strictly wider replaces)
          for (int j = 0; j < kMaxRanges - 1; ++j) {
            if (gap > gw[j]) {
              for (int m = kMaxRanges - 2; m > j; --m) {
                gw[m] = gw[m - 1];
                gi[m] = gi[m - 1];
              }
              gw[j] = gap;
              gi[j] = i;
              break;
            }
          }
        }
        prev_end = e;
      }
      // walk the runs again, cutting before the runs listed in gi
      int64_t cur_s = -1, cur_e = -1;
      for (int i = 0; i < R_; ++i) {
        const int64_t s = fstart[rs_[i]], e = fstart[re_[i] + 1];
        if (e == s) continue;
        bool cut_here = false;
        for (int j = 0; j < kMaxRanges - 1; ++j) cut_here |= (gi[j] == i);
        if (cur_s < 0) {
          cur_s = s;
          cur_e = e;
        } else if (cut_here) {
          ra[nr] = cur_s;
          rb[nr] = cur_e;
          ++nr;
          cur_s = s;
          cur_e = e;
        } else {
          cur_e = e;
        }
      }
      if (cur_s >= 0) {
        ra[nr] = cur_s;
        rb[nr] = cur_e;
        ++nr;
      }
    }
    int64_t rows = 0;
    for (int j = 0; j < kMaxRanges; ++j) {
      const int64_t s = j < nr ? ra[j] : 0, e = j < nr ? rb[j] : 0;
      ranges[(b * kMaxRanges + j) * 2] = s;
      ranges[(b * kMaxRanges + j) * 2 + 1] = e;
      rows += e - s;
    }
    atomicMax(reinterpret_cast<unsigned long long*>(&counts[2]), static_cast<unsigned long long>(nr));
    atomicAdd(pairs, static_cast<unsigned long long>(rows) * static_cast<unsigned long long>(cnt));
  }
}

__global__ void k_zero_i64(int64_t* __restrict__ p, int m) {
  const int i = threadIdx.x;
  if (i < m) p[i] = 0;
}

struct Bufs {
  float* C;
  float* cn2;
  int32_t* part;
  float* dmin2;
  double* partial;
  double* dsum;
  int32_t* flags;
  int32_t* pos;
  float* r;
  uint32_t* kA;
  uint32_t* vA;
  uint32_t* kB;
  uint32_t* vB;
  float* rs;
  int64_t* pstart;
  int64_t* cut;
  float* Rp;
  float* rk;
  int32_t* size;
  int32_t* part2;
  float* r2;
  uint32_t* adj;
  int32_t* ppos;
  int32_t* porder;
  uint32_t* pgr;
  int32_t* perm;
  float* rfin;
  float* Xs;
  int64_t* fstart;
  int64_t* gpstart;
  int64_t* grow;
  int64_t* boff;
  int32_t* qrows;
  int32_t* bfirst;
  int32_t* bcnt;
  float* Cpos;
  float* Rpos;
  float* rkpos;
  int32_t* szpos;
  float* ubq;
  int64_t* ranges;
  int64_t* counts;
  float* rlim;
  unsigned long long* pairs;
  void* scan_ws;
  size_t scan_bytes;
  void* sort_ws;
  size_t sort_bytes;
};

Bufs carve_bufs(Carve& cv, int64_t n, int d) {
  Bufs B{};
  const int pm = pmax_for(n);
  const int W = (pm + 31) / 32;
  const int64_t nb = nblk_max(n);
  B.C = cv.take<float>(static_cast<size_t>(pm) * kDMax);
  B.cn2 = cv.take<float>(pm);
  B.part = cv.take<int32_t>(n);
  B.dmin2 = cv.take<float>(n);
  B.partial = cv.take<double>(kSumBlocks);
  B.dsum = cv.take<double>(1);
  B.flags = cv.take<int32_t>(n + 1);
  B.pos = cv.take<int32_t>(n + 1);
  B.r = cv.take<float>(n);
  B.kA = cv.take<uint32_t>(n);
  B.vA = cv.take<uint32_t>(n);
  B.kB = cv.take<uint32_t>(n);
  B.vB = cv.take<uint32_t>(n);
  B.rs = cv.take<float>(n);
  B.pstart = cv.take<int64_t>(pm + 1);
  B.cut = cv.take<int64_t>(pm);
  B.Rp = cv.take<float>(pm);
  B.rk = cv.take<float>(pm);
  B.size = cv.take<int32_t>(pm);
  B.part2 = cv.take<int32_t>(n);
  B.r2 = cv.take<float>(n);
  B.adj = cv.take<uint32_t>(static_cast<size_t>(pm) * W);
  B.ppos = cv.take<int32_t>(pm);
  B.porder = cv.take<int32_t>(pm);
  B.pgr = cv.take<uint32_t>(pm);
  B.perm = cv.take<int32_t>(n);
  B.rfin = cv.take<float>(n);
  B.Xs = cv.take<float>(static_cast<size_t>(n) * d);
  B.fstart = cv.take<int64_t>(pm + 1);
  B.gpstart = cv.take<int64_t>(kMaxGroups + 3);
  B.grow = cv.take<int64_t>(kMaxGroups + 3);
  B.boff = cv.take<int64_t>(kMaxGroups + 3);
  B.qrows = cv.take<int32_t>(static_cast<size_t>(nb) * kBlockQ);
  B.bfirst = cv.take<int32_t>(nb);
  B.bcnt = cv.take<int32_t>(nb);
  B.Cpos = cv.take<float>(static_cast<size_t>(pm) * kDMax);
  B.Rpos = cv.take<float>(pm);
  B.rkpos = cv.take<float>(pm);
  B.szpos = cv.take<int32_t>(pm);
  B.ubq = cv.take<float>(n);
  B.ranges = cv.take<int64_t>(static_cast<size_t>(nb) * kMaxRanges * 2);
  B.counts = cv.take<int64_t>(8);
  B.rlim = cv.take<float>(1);
  B.pairs = cv.take<unsigned long long>(1);
  B.scan_bytes = prims::scan_workspace<int32_t>(n + 1);
  B.scan_ws = cv.take<char>(B.scan_bytes);
  B.sort_bytes = prims::radix_workspace(n);
  B.sort_ws = cv.take<char>(B.sort_bytes);
  return B;
}

unsigned grid_for(int64_t m) { return static_cast<unsigned>(ceil_div(std::max<int64_t>(m, 1), kT)); }

int launch_assign(const float* X, int64_t n, int d, const float* C, const float* cn2, int cb, int ce, int first,
                  int32_t* part, float* dmin2, hipStream_t st) {
  if (d + 2 <= 64) {
    k_assign_mfma<<<static_cast<unsigned>(ceil_div(n, 4 * 32)), kT, 0, st>>>(X, n, d, C, cn2, cb, ce, first, part,
                                                                             dmin2);
    MSC_LAUNCH_CHECK();
    return kOk;
  }
  const int D4 = (d + 3) / 4;
  const unsigned g = grid_for(n);
  switch (D4) {
#define MSC_ASSIGN(V) \
  case V: k_assign<V><<<g, kT, 0, st>>>(X, n, d, C, cn2, cb, ce, first, part, dmin2); break;
    MSC_ASSIGN(1) MSC_ASSIGN(2) MSC_ASSIGN(3) MSC_ASSIGN(4) MSC_ASSIGN(5) MSC_ASSIGN(6) MSC_ASSIGN(7) MSC_ASSIGN(8)
    MSC_ASSIGN(9) MSC_ASSIGN(10) MSC_ASSIGN(11) MSC_ASSIGN(12) MSC_ASSIGN(13) MSC_ASSIGN(14) MSC_ASSIGN(15)
    MSC_ASSIGN(16)
#undef MSC_ASSIGN
    default: set_error("partition: d=%d above %d", d, kDMax); return kInvalidArgument;
  }
  MSC_LAUNCH_CHECK();
  return kOk;
}

int read_count(const int64_t* dev, int64_t* host, hipStream_t st) {
  MSC_HIP(hipMemcpyAsync(host, dev, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  MSC_HIP(hipStreamSynchronize(st));
  return kOk;
}
int read_count32(const int32_t* dev, int64_t* host, hipStream_t st) {
  int32_t v = 0;
  MSC_HIP(hipMemcpyAsync(&v, dev, sizeof(int32_t), hipMemcpyDeviceToHost, st));
  MSC_HIP(hipStreamSynchronize(st));
  *host = v;
  return kOk;
}

}  // namespace

size_t workspace(int64_t n, int d) {
  Carve cv(nullptr, 0);
  carve_bufs(cv, n, d);
  return cv.used + 256;
}

int build(const float* X, int64_t n, int d, int k, void* ws, size_t ws_bytes, Layout* L, hipStream_t st) {
  MSC_REQUIRE(d >= 1 && d <= kDMax, "partition: d=%d outside [1, %d]", d, kDMax);
  MSC_REQUIRE(n >= 1 && n < (int64_t(1) << 31) - kBlockQ, "partition: n=%lld out of range", (long long)n);
  Carve cv(ws, ws_bytes);
  Bufs B = carve_bufs(cv, n, d);
  if (!cv.ok()) {
    set_error("partition: workspace too small (%zu < %zu)", ws_bytes, cv.used);
    return kWorkspaceTooSmall;
  }
  const int p0 = static_cast<int>(std::min<int64_t>(p0_for(n), n));
  const int pm = pmax_for(n);
  const double eps = 8.0 * (d + 4) * 5.9604644775390625e-08;

  // 1. centres
  k_seed<<<grid_for(static_cast<int64_t>(p0) * kDMax), kT, 0, st>>>(X, n, d, p0, B.C);
  k_cnorm<<<grid_for(p0), kT, 0, st>>>(B.C, 0, p0, B.cn2);
  MSC_LAUNCH_CHECK();
  MSC_TRY(launch_assign(X, n, d, B.C, B.cn2, 0, p0, 1, B.part, B.dmin2, st));
  int P = p0;
  for (int round = 1; round <= kRounds; ++round) {
    k_sum_partial<<<kSumBlocks, kT, 0, st>>>(B.dmin2, n, B.partial);
    k_sum_final<<<1, 64, 0, st>>>(B.partial, B.dsum);
    k_sample<<<grid_for(n + 1), kT, 0, st>>>(B.dmin2, n, B.dsum, static_cast<double>(p0),
                                             mix64(0x6B6E6E5061727431ull + static_cast<uint64_t>(round)), B.flags);
    MSC_LAUNCH_CHECK();
    MSC_TRY(prims::exclusive_scan<int32_t>(B.flags, B.pos, n + 1, B.scan_ws, B.scan_bytes, st));
    int64_t drawn = 0;
    MSC_TRY(read_count32(B.pos + n, &drawn, st));
    const int cap = std::min(2 * p0, pm - kPeelMax - P);
    const int take = static_cast<int>(std::min<int64_t>(drawn, cap));
    if (take <= 0) break;
    k_take<<<grid_for(n), kT, 0, st>>>(X, n, d, B.flags, B.pos, take, B.C + static_cast<int64_t>(P) * kDMax);
    k_cnorm<<<grid_for(take), kT, 0, st>>>(B.C, P, P + take, B.cn2);
    MSC_LAUNCH_CHECK();
    MSC_TRY(launch_assign(X, n, d, B.C, B.cn2, P, P + take, 0, B.part, B.dmin2, st));
    P += take;
  }

  // 2. radii, (partition, r) order, partition statistics
  k_radius<<<grid_for(n), kT, 0, st>>>(X, n, d, B.C, B.part, B.r, B.kA, B.vA);
  MSC_LAUNCH_CHECK();
  MSC_TRY(prims::radix_sort_pairs(B.kA, B.vA, B.kB, B.vB, n, 32, B.sort_ws, B.sort_bytes, st));
  k_keys_of<<<grid_for(n), kT, 0, st>>>(B.part, B.vB, n, B.kA);
  MSC_LAUNCH_CHECK();
  // (pk, ord) = (partition, row) in (partition, r) order: kB / vA
  MSC_TRY(prims::radix_sort_pairs(B.kA, B.vB, B.kB, B.vA, n, prims::bits_for(P), B.sort_ws, B.sort_bytes, st));
  const uint32_t* pk = B.kB;
  const uint32_t* ord = B.vA;
  k_gather_f<<<grid_for(n), kT, 0, st>>>(B.r, ord, n, B.rs);
  MSC_LAUNCH_CHECK();
  MSC_TRY(prims::segment_starts(pk, n, P, B.pstart, st));
  k_part_stats<<<grid_for(P), kT, 0, st>>>(B.rs, B.pstart, P, k, 1, B.cut, B.Rp, B.rk, B.size);
  k_flag_out<<<grid_for(n + 1), kT, 0, st>>>(pk, B.cut, n, B.flags);
  MSC_LAUNCH_CHECK();
  MSC_TRY(prims::exclusive_scan<int32_t>(B.flags, B.pos, n + 1, B.scan_ws, B.scan_bytes, st));
  int64_t n_out = 0;
  MSC_TRY(read_count32(B.pos + n, &n_out, st));

  // 3. peeling (or, with too many outlying rows, none: the full partitions are kept)
  const bool peel = n_out > 0 && n_out <= kPeelMax;
  if (!peel && n_out > 0) {
    k_part_stats<<<grid_for(P), kT, 0, st>>>(B.rs, B.pstart, P, k, 0, B.cut, B.Rp, B.rk, B.size);
    MSC_LAUNCH_CHECK();
  }
  k_peel<<<grid_for(n), kT, 0, st>>>(X, n, d, ord, pk, B.rs, B.flags, B.pos, P, peel ? 1 : 0, k, B.C, B.Rp, B.rk,
                                     B.size, B.part2, B.r2);
  MSC_LAUNCH_CHECK();
  const int P2 = P + (peel ? static_cast<int>(n_out) : 0);

  // 4. groups
  const int W = (P2 + 31) / 32;
  k_radius_limit<<<1, kCompThreads, 0, st>>>(B.Rp, B.size, P2, B.rlim);
  k_adjacency<<<grid_for(static_cast<int64_t>(P2) * W), kT, 0, st>>>(B.C, P2, d, B.Rp, B.size, B.rlim, W, B.adj);
  k_components<<<1, kCompThreads, 0, st>>>(B.adj, W, P2, B.size, B.ppos, B.porder, B.pgr, B.counts);
  MSC_LAUNCH_CHECK();
  int64_t G = 0;
  MSC_TRY(read_count(B.counts, &G, st));

  // 5. final order, groups' rows, query blocks
  k_pos_keys<<<grid_for(n), kT, 0, st>>>(B.part2, B.ppos, n, B.kA, B.vB);
  MSC_LAUNCH_CHECK();
  // pk is dead after k_peel and part after k_radius: they take the sorted (position, index) pairs
  uint32_t* fk = B.kB;
  uint32_t* fv = reinterpret_cast<uint32_t*>(B.part);  // part[] is dead after k_radius
  MSC_TRY(prims::radix_sort_pairs(B.kA, B.vB, fk, fv, n, prims::bits_for(P2), B.sort_ws, B.sort_bytes, st));
  k_final_rows<<<grid_for(n), kT, 0, st>>>(ord, B.r2, fv, n, B.perm, B.rfin);
  k_gather_rows<<<grid_for(n * d), kT, 0, st>>>(X, B.perm, n, d, B.Xs);
  MSC_LAUNCH_CHECK();
  MSC_TRY(prims::segment_starts(fk, n, P2, B.fstart, st));
  MSC_TRY(prims::segment_starts(B.pgr, P2, G + 1, B.gpstart, st));
  const int ng = static_cast<int>(G) + 1;  // the pooled group first (possibly empty)
  k_group_rows<<<grid_for(ng + 1), kT, 0, st>>>(B.fstart, B.gpstart, ng + 1, B.grow);
  k_block_offsets<<<1, kCompThreads, 0, st>>>(B.grow, ng, B.boff, B.counts);
  MSC_LAUNCH_CHECK();
  int64_t nblk = 0;
  MSC_TRY(read_count(B.counts + 1, &nblk, st));
  MSC_REQUIRE(nblk >= 1 && nblk <= nblk_max(n), "partition: internal block count %lld", (long long)nblk);
  k_blocks_fill<<<grid_for(nblk * kBlockQ), kT, 0, st>>>(B.grow, B.boff, ng, nblk, B.qrows, B.bfirst, B.bcnt);
  MSC_LAUNCH_CHECK();

  // 6. bounds
  k_perm_parts<<<grid_for(static_cast<int64_t>(P2) * kDMax), kT, 0, st>>>(B.porder, P2, B.C, B.Rp, B.rk, B.size,
                                                                           B.Cpos, B.Rpos, B.rkpos, B.szpos);
  k_ub<<<grid_for(n), kT, 0, st>>>(B.rfin, fk, B.rkpos, n, eps, B.ubq);
  k_ub_fix<<<static_cast<unsigned>(ceil_div(n, kT / kWave)), kT, 0, st>>>(B.Xs, n, d, B.Cpos, B.rkpos, P2, eps,
                                                                          B.ubq);
  MSC_LAUNCH_CHECK();

  // 7. ranges
  k_zero_i64<<<1, 64, 0, st>>>(B.counts + 2, 2);
  k_zero_i64<<<1, 64, 0, st>>>(reinterpret_cast<int64_t*>(B.pairs), 1);
  switch ((d + 3) / 4) {
#define MSC_NEED(V)                                                                                        \
  case V:                                                                                                  \
    k_need<V><<<static_cast<unsigned>(nblk), kT, 0, st>>>(B.Xs, d, B.ubq, B.bfirst, B.bcnt, B.Cpos, B.Rpos, \
                                                           B.szpos, B.fstart, P2, eps, B.ranges, B.counts,  \
                                                           B.pairs);                                        \
    break;
    MSC_NEED(1) MSC_NEED(2) MSC_NEED(3) MSC_NEED(4) MSC_NEED(5) MSC_NEED(6) MSC_NEED(7) MSC_NEED(8) MSC_NEED(9)
    MSC_NEED(10) MSC_NEED(11) MSC_NEED(12) MSC_NEED(13) MSC_NEED(14) MSC_NEED(15) MSC_NEED(16)
#undef MSC_NEED
  }
  MSC_LAUNCH_CHECK();
  int64_t tail[2] = {0, 0};
  unsigned long long pairs = 0;
  MSC_HIP(hipMemcpyAsync(tail, B.counts + 2, 2 * sizeof(int64_t), hipMemcpyDeviceToHost, st));
  MSC_HIP(hipMemcpyAsync(&pairs, B.pairs, sizeof(pairs), hipMemcpyDeviceToHost, st));
  MSC_HIP(hipStreamSynchronize(st));

  L->Xs = B.Xs;
  L->perm = B.perm;
  L->qrows = B.qrows;
  L->blk_range = B.ranges;
  L->nblk = nblk;
  L->n_slots = nblk * kBlockQ;
  L->S = static_cast<int>(std::max<int64_t>(1, tail[0]));
  L->n_parts = P2;
  L->n_groups = static_cast<int>(G);
  L->n_peeled = peel ? n_out : 0;
  L->pairs = static_cast<double>(pairs);
  return kOk;
}

}  // namespace part
}  // namespace msc

// Moran's I with a permutation null on gfx950, float64 reductions.
//
// Replaces spatial.spatial_autocorr's Moran branch (src/metalsinglecell/spatial.py:357-406):
// I_g = (n/S0) sum_i Xc_ig (W Xc)_ig / sum_i Xc_ig^2, and for each permutation pi the
// reference's permuted-source scatter (:401-405), which equals
// sim_pg = (n/S0) sum_r Xc[pi(r), g] Y[r, g] / denom_g with Y = W Xc.
//
// Layout: X dense cells x genes float32 row-major (as the reference densifies, gr.py:452);
// W CSR with float64 values; Y stored float32 (one read per permutation batch).  Column
// means, denominators, numerators and the permuted sums accumulate in float64 in a fixed
// order (block partials over row chunks, then a sequential sum over chunks), so results
// are deterministic.
//
// Permutations: either caller-supplied (P x n int32, e.g. numpy default_rng(seed)
// .permutation(n) in sequence, the reference stream) or generated on the device by a
// keyed 4-round Feistel bijection with cycle walking (no storage, no host transfer).
#include <cmath>

#include "common.h"
#include "msc_hip.h"
#include "prims.h"

namespace msc {
namespace moran {

constexpr int kThreads = 256;
constexpr int kPermBatch = 8;

// ---- column statistics --------------------------------------------------------------------
// partial[b][g] over rows [b*per, (b+1)*per); thread per gene (strided), rows sequential
__global__ void k_col_partial(const float* __restrict__ X, int64_t n, int64_t G, int64_t per,
                              const double* __restrict__ mean, double* __restrict__ partial) {
  const int64_t r0 = blockIdx.x * per, r1 = min(n, r0 + per);
  for (int64_t g = threadIdx.x + static_cast<int64_t>(blockIdx.y) * blockDim.x; g < G;
       g += static_cast<int64_t>(gridDim.y) * blockDim.x) {
    double s = 0.0;
    if (mean) {
      const double m = mean[g];
#pragma unroll 8
      for (int64_t r = r0; r < r1; ++r) {
        const double d = static_cast<double>(X[r * G + g]) - m;
        s += d * d;
      }
    } else {
#pragma unroll 8
      for (int64_t r = r0; r < r1; ++r) s += static_cast<double>(X[r * G + g]);
    }
    partial[blockIdx.x * G + g] = s;
  }
}
__global__ void k_col_final(const double* __restrict__ partial, int nb, int64_t G, double scale,
                            double* __restrict__ out) {
  const int64_t g = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (g >= G) return;
  double s = 0.0;
#pragma unroll 8
  for (int b = 0; b < nb; ++b) s += partial[static_cast<int64_t>(b) * G + g];
  out[g] = s * scale;
}

// ---- weights --------------------------------------------------------------------------------
// row-normalise (spatial.py:268-279): w / sum|row|, zero rows stay zero; also row sums
__global__ void k_row_normalize(const int64_t* __restrict__ indptr, const double* __restrict__ w_in, int64_t n,
                                int transform, double* __restrict__ w_out) {
  const int64_t r = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (r >= n) return;
  double s = 0.0;
  for (int64_t e = indptr[r]; e < indptr[r + 1]; ++e) s += fabs(w_in[e]);
  const double inv = (transform && s != 0.0) ? 1.0 / s : 1.0;
  for (int64_t e = indptr[r]; e < indptr[r + 1]; ++e) w_out[e] = (transform ? (s != 0.0 ? w_in[e] * inv : 0.0) : w_in[e]);
}
// Fixed-order float64 sum in two launches: block b sums its contiguous range (threads
// strided, then the block's per-thread sums in thread order), then one block sums the block
// results in block order.  The order depends only on m, so results are deterministic.
constexpr int kSumBlocks = 512;
__global__ void k_sum_f64(const double* __restrict__ v, int64_t m, int64_t per, double* __restrict__ out) {
  const int64_t b0 = blockIdx.x * per, b1 = min(m, b0 + per);
  double s = 0.0;
#pragma unroll 8
  for (int64_t i = b0 + threadIdx.x; i < b1; i += blockDim.x) s += v[i];
  __shared__ double sh[kThreads];
  sh[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < kThreads; ++i) t += sh[i];
    out[blockIdx.x] = t;
  }
}
// out = sum of v[0, m); scratch holds kSumBlocks doubles
static void sum_f64(const double* v, int64_t m, double* out, double* scratch, hipStream_t st) {
  const int64_t nb = std::max<int64_t>(1, std::min<int64_t>(kSumBlocks, ceil_div(m, 4 * kThreads)));
  const int64_t per = ceil_div(std::max<int64_t>(m, 1), nb);
  k_sum_f64<<<static_cast<unsigned>(nb), kThreads, 0, st>>>(v, m, per, scratch);
  k_sum_f64<<<1, kThreads, 0, st>>>(scratch, nb, nb, out);
}

// Y[r, :] = sum_e w_e (X[j_e, :] - mean), float64 accumulate, stored float32; block per row
__global__ __launch_bounds__(kThreads) void k_spmm(const float* __restrict__ X, int64_t n, int64_t G,
                                                   const int64_t* __restrict__ indptr, const int32_t* __restrict__ indices,
                                                   const double* __restrict__ w, const double* __restrict__ mean,
                                                   float* __restrict__ Y) {
  const int64_t r = blockIdx.x;
  const int64_t e0 = indptr[r], e1 = indptr[r + 1];
  for (int64_t g = threadIdx.x; g < G; g += kThreads) {
    const double m = mean[g];
    double s = 0.0;
    for (int64_t e = e0; e < e1; ++e) s += w[e] * (static_cast<double>(X[static_cast<int64_t>(indices[e]) * G + g]) - m);
    Y[r * G + g] = static_cast<float>(s);
  }
}

// ---- permutations -----------------------------------------------------------------------------
struct Feistel {
  uint64_t key;
  int half;       // bits per half; domain 2^(2*half) >= n
  uint64_t mask;
};

__device__ __forceinline__ uint64_t feistel_once(uint64_t x, uint64_t key, int half, uint64_t mask) {
  uint64_t L = x >> half, R = x & mask;
#pragma unroll
  for (int round = 0; round < 4; ++round) {
    const uint64_t F = mix64(R ^ (key + static_cast<uint64_t>(round) * 0x9E3779B97F4A7C15ull)) & mask;
    const uint64_t nL = R;
    R = L ^ F;
    L = nL;
  }
  return (L << half) | R;
}
__device__ __forceinline__ int64_t feistel_perm(int64_t r, int64_t n, const Feistel& f, uint64_t pkey) {
  uint64_t x = static_cast<uint64_t>(r);
  do {
    x = feistel_once(x, pkey, f.half, f.mask);
  } while (x >= static_cast<uint64_t>(n));
  return static_cast<int64_t>(x);
}

// ---- observed numerator and permuted numerators -----------------------------------------------
// grid: (row chunks, perm batches, gene tiles); partial[chunk][p][g].  A thread owns VEC
// consecutive genes (VEC = 4: 16-byte loads, needs G % 4 == 0 and 16-byte aligned rows); rows
// of the chunk in order, kRowStep at a time: the block first writes the source rows pi_p(r)
// of the step's rows and kPermBatch permutations to LDS (each computed once), then every
// thread gathers X[pi_p(r), g] and accumulates x * Y[r, g] in float64, per (p, g) in row
// order.
constexpr int kRowStep = 64;
__device__ __forceinline__ uint64_t perm_key(const Feistel& f, int p) {
  return mix64(f.key + static_cast<uint64_t>(p + 1) * 0xD1B54A32D192ED03ull);
}
template <int VEC>
__device__ __forceinline__ void load_vec(const float* __restrict__ p, float (&v)[VEC]) {
  if constexpr (VEC == 4) {
    const float4 q = *reinterpret_cast<const float4*>(p);
    v[0] = q.x; v[1] = q.y; v[2] = q.z; v[3] = q.w;
  } else {
    v[0] = *p;
  }
}
template <int VEC>
__global__ __launch_bounds__(kThreads) void k_perm_partial(
    const float* __restrict__ X, const float* __restrict__ Y, int64_t n, int64_t G, const double* __restrict__ mean,
    int64_t per, int P, const int32_t* __restrict__ perms, Feistel fst, int observed, double* __restrict__ partial) {
  __shared__ int32_t jtab[kPermBatch][kRowStep];
  const int64_t r0 = blockIdx.x * per, r1 = min(n, r0 + per);
  const int p0 = blockIdx.y * kPermBatch;
  const int64_t g0 = (static_cast<int64_t>(blockIdx.z) * blockDim.x + threadIdx.x) * VEC;
  const bool gok = g0 < G;
  double m[VEC];
#pragma unroll
  for (int v = 0; v < VEC; ++v) m[v] = gok ? mean[g0 + v] : 0.0;
  double acc[kPermBatch][VEC];
#pragma unroll
  for (int b = 0; b < kPermBatch; ++b)
#pragma unroll
    for (int v = 0; v < VEC; ++v) acc[b][v] = 0.0;
  for (int64_t rb = r0; rb < r1; rb += kRowStep) {
    const int nr = static_cast<int>(min(static_cast<int64_t>(kRowStep), r1 - rb));
    if (observed) {
      if (!gok) continue;
      for (int rr = 0; rr < nr; ++rr) {
        const int64_t r = rb + rr;
        float x[VEC], y[VEC];
        load_vec<VEC>(X + r * G + g0, x);
        load_vec<VEC>(Y + r * G + g0, y);
#pragma unroll
        for (int v = 0; v < VEC; ++v) acc[0][v] += (static_cast<double>(x[v]) - m[v]) * static_cast<double>(y[v]);
      }
      continue;
    }
    __syncthreads();  // the previous step's readers are done with jtab
    for (int q = threadIdx.x; q < kPermBatch * kRowStep; q += blockDim.x) {
      const int b = q / kRowStep, rr = q % kRowStep;
      const int p = p0 + b;
      int32_t j = 0;
      if (p < P && rr < nr) {
        const int64_t r = rb + rr;
        j = perms ? perms[static_cast<int64_t>(p) * n + r]
                  : static_cast<int32_t>(feistel_perm(r, n, fst, perm_key(fst, p)));
      }
      jtab[b][rr] = j;
    }
    __syncthreads();
    if (!gok) continue;
    for (int rr = 0; rr < nr; ++rr) {
      float y[VEC];
      load_vec<VEC>(Y + (rb + rr) * G + g0, y);
      // All kPermBatch gathers are issued before the first use, unconditionally: a permutation
      // past P reads row 0 (its jtab entry) into an accumulator that is never written.  A
      // branch or a use between the loads would make each gather wait for the previous one.
      float x[kPermBatch][VEC];
#pragma unroll
      for (int b = 0; b < kPermBatch; ++b) load_vec<VEC>(X + static_cast<int64_t>(jtab[b][rr]) * G + g0, x[b]);
#pragma unroll
      for (int b = 0; b < kPermBatch; ++b)
#pragma unroll
        for (int v = 0; v < VEC; ++v) acc[b][v] += (static_cast<double>(x[b][v]) - m[v]) * static_cast<double>(y[v]);
    }
  }
  if (!gok) return;
  if (observed) {
#pragma unroll
    for (int v = 0; v < VEC; ++v) partial[blockIdx.x * G + g0 + v] = acc[0][v];
    return;
  }
#pragma unroll
  for (int b = 0; b < kPermBatch; ++b) {
    const int p = p0 + b;
    if (p < P)
#pragma unroll
      for (int v = 0; v < VEC; ++v) partial[(static_cast<int64_t>(blockIdx.x) * P + p) * G + g0 + v] = acc[b][v];
  }
}

// out[p][g] = (n/S0) * sum_chunks partial / denom
__global__ void k_finish(const double* __restrict__ partial, int nchunks, int P, int64_t G, double n,
                         const double* __restrict__ S0, const double* __restrict__ denom, double* __restrict__ out) {
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t >= static_cast<int64_t>(P) * G) return;
  const int64_t p = t / G, g = t % G;
  double s = 0.0;
  for (int c = 0; c < nchunks; ++c) s += partial[(static_cast<int64_t>(c) * P + p) * G + g];
  out[t] = (n / *S0) * s / denom[g];
}

struct Plan {
  int64_t n, G, E;
  int P;
  int nchunks;
  int64_t per;
  int gtiles;
  int vec;       // genes per thread in the permutation kernel (4 when G % 4 == 0)
  int pthreads;  // its block size
};

static Plan make_plan(int64_t n, int64_t G, int64_t E, int P) {
  Plan p{};
  p.n = n;
  p.G = G;
  p.E = E;
  p.P = P;
  p.vec = (G % 4 == 0) ? 4 : 1;
  p.pthreads = p.vec == 4 ? kThreads / 2 : kThreads;
  p.gtiles = static_cast<int>(ceil_div(G, static_cast<int64_t>(p.pthreads) * p.vec));
  const int pb = std::max(1, static_cast<int>(ceil_div(std::max(P, 1), kPermBatch)));
  // about 2048 blocks for the permutation kernel, chunks of >= 256 rows
  int64_t chunks = ceil_div(2048, static_cast<int64_t>(pb) * p.gtiles);
  chunks = std::max<int64_t>(1, std::min<int64_t>(chunks, ceil_div(n, 256)));
  chunks = std::min<int64_t>(chunks, 1024);
  p.per = ceil_div(n, chunks);
  p.nchunks = static_cast<int>(ceil_div(n, p.per));
  return p;
}

struct Bufs {
  double *mean, *denom, *w, *S0, *partial, *cpartial, *sum_scratch;
  float* Y;
};

constexpr int kColBlocks = 1024;

static Bufs carve(Carve& cv, const Plan& p) {
  Bufs b{};
  b.mean = cv.take<double>(p.G);
  b.denom = cv.take<double>(p.G);
  b.w = cv.take<double>(std::max<int64_t>(p.E, 1));
  b.S0 = cv.take<double>(1);
  b.sum_scratch = cv.take<double>(kSumBlocks);
  b.cpartial = cv.take<double>(static_cast<size_t>(kColBlocks) * p.G);
  b.partial = cv.take<double>(static_cast<size_t>(p.nchunks) * std::max(p.P, 1) * p.G);
  b.Y = cv.take<float>(static_cast<size_t>(p.n) * p.G);
  return b;
}

static int validate(int64_t n, int64_t G, int64_t E, int P) {
  MSC_REQUIRE(n >= 2 && G >= 1 && E >= 0 && P >= 0, "moran: need n >= 2, G >= 1, nnz >= 0, n_perms >= 0");
  MSC_REQUIRE(n < (int64_t(1) << 31), "moran: n too large");
  return kOk;
}

}  // namespace moran
}  // namespace msc

using namespace msc;
using namespace msc::moran;

extern "C" int msc_moran_workspace_size(int64_t n, int64_t n_genes, int64_t nnz, int n_perms, size_t* bytes) {
  MSC_TRY(validate(n, n_genes, nnz, n_perms));
  MSC_REQUIRE(bytes, "moran: bytes is null");
  Plan p = make_plan(n, n_genes, nnz, n_perms);
  Carve cv(nullptr, 0);
  carve(cv, p);
  *bytes = cv.used + 256;
  return kOk;
}

extern "C" int msc_moran(const float* X, int64_t n, int64_t n_genes, const int64_t* w_indptr,
                         const int32_t* w_indices, const double* w_data, int64_t nnz, int transformation,
                         int n_perms, const int32_t* perms, int64_t seed, double* out_I, double* out_sims,
                         double* out_S0, void* workspace, size_t workspace_bytes, void* stream) {
  MSC_TRY(validate(n, n_genes, nnz, n_perms));
  MSC_REQUIRE(X && w_indptr && out_I && workspace && (nnz == 0 || (w_indices && w_data)),
              "moran: null pointer argument");
  MSC_REQUIRE(n_perms == 0 || out_sims, "moran: out_sims required when n_perms > 0");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int64_t G = n_genes;
  Plan p = make_plan(n, G, nnz, n_perms);
  Carve cv(workspace, workspace_bytes);
  Bufs b = carve(cv, p);
  if (!cv.ok()) {
    set_error("moran: workspace too small (%zu < %zu)", workspace_bytes, cv.used);
    return kWorkspaceTooSmall;
  }
  const int64_t per = ceil_div(n, kColBlocks);
  const dim3 cg(kColBlocks, static_cast<unsigned>(std::min<int64_t>(ceil_div(G, kThreads), 64)));
  k_col_partial<<<cg, kThreads, 0, st>>>(X, n, G, per, nullptr, b.cpartial);
  k_col_final<<<static_cast<unsigned>(ceil_div(G, kThreads)), kThreads, 0, st>>>(b.cpartial, kColBlocks, G,
                                                                                   1.0 / static_cast<double>(n), b.mean);
  k_col_partial<<<cg, kThreads, 0, st>>>(X, n, G, per, b.mean, b.cpartial);
  k_col_final<<<static_cast<unsigned>(ceil_div(G, kThreads)), kThreads, 0, st>>>(b.cpartial, kColBlocks, G, 1.0,
                                                                                   b.denom);
  MSC_LAUNCH_CHECK();
  k_row_normalize<<<static_cast<unsigned>(ceil_div(n, kThreads)), kThreads, 0, st>>>(w_indptr, w_data, n,
                                                                                       transformation, b.w);
  sum_f64(b.w, nnz, b.S0, b.sum_scratch, st);
  MSC_LAUNCH_CHECK();
  k_spmm<<<static_cast<unsigned>(n), kThreads, 0, st>>>(X, n, G, w_indptr, w_indices, b.w, b.mean, b.Y);
  MSC_LAUNCH_CHECK();
  Feistel fst{};
  {
    int bits = 2;
    while ((int64_t(1) << bits) < n) bits += 2;
    fst.half = bits / 2;
    fst.mask = (uint64_t(1) << fst.half) - 1;
    fst.key = stream_key(seed, 7);
  }
  // 16-byte gathers need 16-byte aligned rows of X (Y comes from the workspace, 256-aligned)
  const bool vec4 = p.vec == 4 && (reinterpret_cast<uintptr_t>(X) % 16 == 0);
  auto perm_launch = [&](dim3 grid, int P_, const int32_t* pr, int observed, double* part) {
    if (vec4) k_perm_partial<4><<<grid, p.pthreads, 0, st>>>(X, b.Y, n, G, b.mean, p.per, P_, pr, fst, observed, part);
    else {
      const dim3 g1(grid.x, grid.y, static_cast<unsigned>(ceil_div(G, kThreads)));
      k_perm_partial<1><<<g1, kThreads, 0, st>>>(X, b.Y, n, G, b.mean, p.per, P_, pr, fst, observed, part);
    }
  };
  // observed statistic
  {
    const dim3 grid(static_cast<unsigned>(p.nchunks), 1, static_cast<unsigned>(p.gtiles));
    perm_launch(grid, 1, nullptr, 1, b.partial);
    k_finish<<<static_cast<unsigned>(ceil_div(G, kThreads)), kThreads, 0, st>>>(
        b.partial, p.nchunks, 1, G, static_cast<double>(n), b.S0, b.denom, out_I);
    MSC_LAUNCH_CHECK();
  }
  if (n_perms > 0) {
    const dim3 grid(static_cast<unsigned>(p.nchunks), static_cast<unsigned>(ceil_div(n_perms, kPermBatch)),
                    static_cast<unsigned>(p.gtiles));
    perm_launch(grid, n_perms, perms, 0, b.partial);
    k_finish<<<static_cast<unsigned>(ceil_div(static_cast<int64_t>(n_perms) * G, kThreads)), kThreads, 0, st>>>(
        b.partial, p.nchunks, n_perms, G, static_cast<double>(n), b.S0, b.denom, out_sims);
    MSC_LAUNCH_CHECK();
  }
  if (out_S0) {
    MSC_HIP(hipMemcpyAsync(out_S0, b.S0, sizeof(double), hipMemcpyDeviceToHost, st));
    MSC_HIP(hipStreamSynchronize(st));
  }
  return kOk;
}

// ------------------------------------------------------------------------------------------
// CSR (cells x genes) -> dense row-major float32 on the device (gr.py:452 densifies on the host)
// ------------------------------------------------------------------------------------------
namespace msc {
namespace moran {
__global__ void k_fill_f32(float* __restrict__ p, int64_t m) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i < m) p[i] = 0.0f;
}
__global__ __launch_bounds__(kThreads) void k_scatter_rows(const int64_t* __restrict__ indptr,
                                                           const int32_t* __restrict__ indices,
                                                           const float* __restrict__ data, int64_t n,
                                                           int64_t G, float* __restrict__ out) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * (kThreads / kWave) + threadIdx.x / kWave;
  if (r >= n) return;
  for (int64_t e = indptr[r] + lane_id(); e < indptr[r + 1]; e += kWave) out[r * G + indices[e]] = data[e];
}
}  // namespace moran
}  // namespace msc

extern "C" int msc_csr_to_dense(const int64_t* indptr, const int32_t* indices, const float* data, int64_t n,
                                int64_t n_cols, float* out, void* stream) {
  MSC_REQUIRE(indptr && out && n >= 0 && n_cols >= 1, "csr_to_dense: bad arguments");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int64_t m = n * n_cols;
  if (m > 0) {
    k_fill_f32<<<static_cast<unsigned>(ceil_div(m, kThreads)), kThreads, 0, st>>>(out, m);
    k_scatter_rows<<<static_cast<unsigned>(ceil_div(n, kThreads / kWave)), kThreads, 0, st>>>(indptr, indices, data,
                                                                                             n, n_cols, out);
    MSC_LAUNCH_CHECK();
  }
  return kOk;
}

// ------------------------------------------------------------------------------------------
// pysal weight moments (spatial.py:282-290) on the device:
//   s0 = sum w,  s1 = 1/2 sum_ij (w_ij + w_ji)^2,  s2 = sum_i (row_i + col_i)^2
// Reverse entries are found by scanning the (short) row of the destination; column sums
// come from a per-destination sequential sum after a stable sort by destination.
// ------------------------------------------------------------------------------------------
namespace msc {
namespace moran {
// per directed edge: contribution to 2*s1 (pairs counted from both ends)
__global__ void k_s1_terms(const int64_t* __restrict__ indptr, const int32_t* __restrict__ indices,
                           const double* __restrict__ w, int64_t n, double* __restrict__ rowterm) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  double acc = 0.0;
  for (int64_t e = indptr[i]; e < indptr[i + 1]; ++e) {
    const int64_t j = indices[e];
    double rev = 0.0;
    bool has = false;
    for (int64_t f = indptr[j]; f < indptr[j + 1]; ++f)
      if (indices[f] == i) { rev += w[f]; has = true; }
    const double a = w[e];
    // the (i,j) entry of W + W^T is a + rev; when j->i is absent the (j,i) entry is a too
    acc += has ? (a + rev) * (a + rev) : 2.0 * a * a;
  }
  rowterm[i] = acc;
}
__global__ void k_edge_dst(const int64_t* __restrict__ indptr, const int32_t* __restrict__ indices, int64_t n,
                           uint32_t* __restrict__ key, uint32_t* __restrict__ val) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  for (int64_t e = indptr[i]; e < indptr[i + 1]; ++e) {
    key[e] = static_cast<uint32_t>(indices[e]);
    val[e] = static_cast<uint32_t>(e);
  }
}
__global__ void k_s2_terms(const int64_t* __restrict__ indptr, const double* __restrict__ w,
                           const int64_t* __restrict__ cstart, const uint32_t* __restrict__ sorted_e, int64_t n,
                           double* __restrict__ out) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  double r = 0.0, c = 0.0;
  for (int64_t e = indptr[i]; e < indptr[i + 1]; ++e) r += w[e];
  for (int64_t p = cstart[i]; p < cstart[i + 1]; ++p) c += w[sorted_e[p]];
  out[i] = (r + c) * (r + c);
}
}  // namespace moran
}  // namespace msc

extern "C" int msc_graph_moments_workspace_size(int64_t n, int64_t nnz, size_t* bytes) {
  MSC_REQUIRE(bytes && n >= 1 && nnz >= 0, "graph_moments: bad arguments");
  Carve cv(nullptr, 0);
  cv.take<double>(n + 1);
  cv.take<double>(4);
  cv.take<uint32_t>(nnz + 1); cv.take<uint32_t>(nnz + 1); cv.take<uint32_t>(nnz + 1); cv.take<uint32_t>(nnz + 1);
  cv.take<int64_t>(n + 2);
  cv.take<char>(msc::prims::radix_workspace(nnz + 1));
  cv.take<double>(kSumBlocks);
  *bytes = cv.used + 256;
  return kOk;
}

extern "C" int msc_graph_moments(const int64_t* indptr, const int32_t* indices, const double* w, int64_t n,
                                 int64_t nnz, double* out_s /*host: s0, s1, s2*/, void* workspace,
                                 size_t workspace_bytes, void* stream) {
  MSC_REQUIRE(indptr && out_s && workspace && n >= 1 && (nnz == 0 || (indices && w)),
              "graph_moments: bad arguments");
  hipStream_t st = static_cast<hipStream_t>(stream);
  Carve cv(workspace, workspace_bytes);
  double* terms = cv.take<double>(n + 1);
  double* red = cv.take<double>(4);
  uint32_t* key = cv.take<uint32_t>(nnz + 1);
  uint32_t* val = cv.take<uint32_t>(nnz + 1);
  uint32_t* skey = cv.take<uint32_t>(nnz + 1);
  uint32_t* sval = cv.take<uint32_t>(nnz + 1);
  int64_t* cstart = cv.take<int64_t>(n + 2);
  const size_t rsb = msc::prims::radix_workspace(nnz + 1);
  void* rws = cv.take<char>(rsb);
  double* scratch = cv.take<double>(kSumBlocks);
  if (!cv.ok()) {
    set_error("graph_moments: workspace too small");
    return kWorkspaceTooSmall;
  }
  const unsigned g = static_cast<unsigned>(ceil_div(n, kThreads));
  double hs[3];
  sum_f64(w, nnz, red, scratch, st);
  k_s1_terms<<<g, kThreads, 0, st>>>(indptr, indices, w, n, terms);
  sum_f64(terms, n, red + 1, scratch, st);
  MSC_LAUNCH_CHECK();
  k_edge_dst<<<g, kThreads, 0, st>>>(indptr, indices, n, key, val);
  MSC_LAUNCH_CHECK();
  MSC_TRY(msc::prims::radix_sort_pairs(key, val, skey, sval, nnz, msc::prims::bits_for(n), rws, rsb, st));
  MSC_TRY(msc::prims::segment_starts(skey, nnz, n, cstart, st));
  k_s2_terms<<<g, kThreads, 0, st>>>(indptr, w, cstart, sval, n, terms);
  sum_f64(terms, n, red + 2, scratch, st);
  MSC_LAUNCH_CHECK();
  MSC_HIP(hipMemcpyAsync(hs, red, 3 * sizeof(double), hipMemcpyDeviceToHost, st));
  MSC_HIP(hipStreamSynchronize(st));
  out_s[0] = hs[0];
  out_s[1] = 0.5 * hs[1];
  out_s[2] = hs[2];
  return kOk;
}

// ------------------------------------------------------------------------------------------
// Geary's C (spatial.py:388-393): C_g = (n - 1) / (2 S0) * sum_e w_e (x_src - x_dst)^2 / denom_g.
// Observed numerator summed directly over the edges in float64.  Permuted numerators (the
// reference scatters from pi[src], :401-405) are expanded per source row r,
//   N_pi = sum_r [ a_pi(r)^2 R_r - 2 a_pi(r) Y_r ] + Z,   a = x - mean, R_r = row sum of w,
//   Y = W a (float64), Z = sum_r sum_e w_e a_dst^2 (independent of pi),
// with Z taken from the observed numerator and the identity's expansion, all in float64.
// ------------------------------------------------------------------------------------------
namespace msc {
namespace moran {

// Y64[r, g] = sum_e w_e a_dst, D[r, g] = sum_e w_e (x_r - x_dst)^2, R[r] = sum_e w_e; block per row
__global__ __launch_bounds__(kThreads) void k_spmm_geary(const float* __restrict__ X, int64_t G,
                                                         const int64_t* __restrict__ indptr,
                                                         const int32_t* __restrict__ indices,
                                                         const double* __restrict__ w, const double* __restrict__ mean,
                                                         double* __restrict__ Y64, double* __restrict__ D,
                                                         double* __restrict__ R) {
  const int64_t r = blockIdx.x;
  const int64_t e0 = indptr[r], e1 = indptr[r + 1];
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int64_t e = e0; e < e1; ++e) s += w[e];
    R[r] = s;
  }
  for (int64_t g = threadIdx.x; g < G; g += kThreads) {
    const double m = mean[g];
    const double xr = static_cast<double>(X[r * G + g]);
    double y = 0.0, dd = 0.0;
    for (int64_t e = e0; e < e1; ++e) {
      const double xj = static_cast<double>(X[static_cast<int64_t>(indices[e]) * G + g]);
      y += w[e] * (xj - m);
      const double df = xr - xj;
      dd += w[e] * (df * df);
    }
    Y64[r * G + g] = y;
    D[r * G + g] = dd;
  }
}

// column sums of a float64 n x G matrix over row chunks (fixed order): partial[b][g]
__global__ void k_col_partial_f64(const double* __restrict__ M, int64_t n, int64_t G, int64_t per,
                                  double* __restrict__ partial) {
  const int64_t r0 = blockIdx.x * per, r1 = min(n, r0 + per);
  for (int64_t g = threadIdx.x + static_cast<int64_t>(blockIdx.y) * blockDim.x; g < G;
       g += static_cast<int64_t>(gridDim.y) * blockDim.x) {
    double s = 0.0;
#pragma unroll 8
    for (int64_t r = r0; r < r1; ++r) s += M[r * G + g];
    partial[blockIdx.x * G + g] = s;
  }
}

// grid (row chunks, perm batches, gene tiles): partA/partB[chunk][p][g] of sum a_pi(r)^2 R_r and
// sum a_pi(r) Y_r; `observed` runs the identity as permutation 0
template <int VEC>
__global__ __launch_bounds__(kThreads) void k_geary_partial(
    const float* __restrict__ X, const double* __restrict__ Y64, const double* __restrict__ R, int64_t n, int64_t G,
    const double* __restrict__ mean, int64_t per, int P, const int32_t* __restrict__ perms, Feistel fst, int observed,
    double* __restrict__ partA, double* __restrict__ partB) {
  __shared__ int32_t jtab[kPermBatch][kRowStep];
  const int64_t r0 = blockIdx.x * per, r1 = min(n, r0 + per);
  const int p0 = blockIdx.y * kPermBatch;
  const int64_t g0 = (static_cast<int64_t>(blockIdx.z) * blockDim.x + threadIdx.x) * VEC;
  const bool gok = g0 < G;
  double m[VEC];
#pragma unroll
  for (int v = 0; v < VEC; ++v) m[v] = gok ? mean[g0 + v] : 0.0;
  double accA[kPermBatch][VEC], accB[kPermBatch][VEC];
#pragma unroll
  for (int b = 0; b < kPermBatch; ++b)
#pragma unroll
    for (int v = 0; v < VEC; ++v) accA[b][v] = accB[b][v] = 0.0;
  for (int64_t rb = r0; rb < r1; rb += kRowStep) {
    const int nr = static_cast<int>(min(static_cast<int64_t>(kRowStep), r1 - rb));
    __syncthreads();  // the previous step's readers are done with jtab
    for (int q = threadIdx.x; q < kPermBatch * kRowStep; q += blockDim.x) {
      const int b = q / kRowStep, rr = q % kRowStep;
      const int p = p0 + b;
      int32_t j = 0;
      if (p < P && rr < nr) {
        const int64_t r = rb + rr;
        j = observed ? static_cast<int32_t>(r)
                     : (perms ? perms[static_cast<int64_t>(p) * n + r]
                              : static_cast<int32_t>(feistel_perm(r, n, fst, perm_key(fst, p))));
      }
      jtab[b][rr] = j;
    }
    __syncthreads();
    if (!gok) continue;
    for (int rr = 0; rr < nr; ++rr) {
      const int64_t r = rb + rr;
      const double Rr = R[r];
      double y[VEC];
#pragma unroll
      for (int v = 0; v < VEC; ++v) y[v] = Y64[r * G + g0 + v];
      // every gather issued before the first use, unconditionally (as in k_perm_partial)
      float x[kPermBatch][VEC];
#pragma unroll
      for (int b = 0; b < kPermBatch; ++b) load_vec<VEC>(X + static_cast<int64_t>(jtab[b][rr]) * G + g0, x[b]);
#pragma unroll
      for (int b = 0; b < kPermBatch; ++b) {
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
          const double a = static_cast<double>(x[b][v]) - m[v];
          accA[b][v] += (a * a) * Rr;
          accB[b][v] += a * y[v];
        }
      }
    }
  }
  if (!gok) return;
#pragma unroll
  for (int b = 0; b < kPermBatch; ++b) {
    const int p = p0 + b;
    if (p < P)
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        const int64_t o = (static_cast<int64_t>(blockIdx.x) * P + p) * G + g0 + v;
        partA[o] = accA[b][v];
        partB[o] = accB[b][v];
      }
  }
}

// observed: C_g from the direct numerator; Zc_g = Nobs - A_id + 2 B_id kept for the null
__global__ void k_geary_observed(const double* __restrict__ partA, const double* __restrict__ partB, int nchunks,
                                 int64_t G, double n, const double* __restrict__ S0, const double* __restrict__ denom,
                                 const double* __restrict__ Nobs, double* __restrict__ Zc, double* __restrict__ out) {
  const int64_t g = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (g >= G) return;
  double A = 0.0, B = 0.0;
  for (int c = 0; c < nchunks; ++c) {
    A += partA[static_cast<int64_t>(c) * G + g];
    B += partB[static_cast<int64_t>(c) * G + g];
  }
  Zc[g] = Nobs[g] - A + 2.0 * B;
  out[g] = ((n - 1.0) / (2.0 * *S0)) * Nobs[g] / denom[g];
}
__global__ void k_geary_finish(const double* __restrict__ partA, const double* __restrict__ partB, int nchunks, int P,
                               int64_t G, double n, const double* __restrict__ S0, const double* __restrict__ denom,
                               const double* __restrict__ Zc, double* __restrict__ out) {
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t >= static_cast<int64_t>(P) * G) return;
  const int64_t p = t / G, g = t % G;
  double A = 0.0, B = 0.0;
  for (int c = 0; c < nchunks; ++c) {
    A += partA[(static_cast<int64_t>(c) * P + p) * G + g];
    B += partB[(static_cast<int64_t>(c) * P + p) * G + g];
  }
  out[t] = ((n - 1.0) / (2.0 * *S0)) * ((A - 2.0 * B) + Zc[g]) / denom[g];
}

struct GearyBufs {
  Bufs b;  // mean, denom, w, S0, sums, column partials
  double *Y64, *D, *R, *Nobs, *Zc, *partA, *partB;
};
static GearyBufs carve_geary(Carve& cv, const Plan& p) {
  GearyBufs t{};
  Plan q = p;
  q.P = 0;      // the Moran partials / float32 Y are not used
  t.b = carve(cv, q);
  t.Y64 = cv.take<double>(static_cast<size_t>(p.n) * p.G);
  t.D = cv.take<double>(static_cast<size_t>(p.n) * p.G);
  t.R = cv.take<double>(p.n);
  t.Nobs = cv.take<double>(p.G);
  t.Zc = cv.take<double>(p.G);
  const size_t pp = static_cast<size_t>(p.nchunks) * std::max(p.P, 1) * p.G;
  t.partA = cv.take<double>(pp);
  t.partB = cv.take<double>(pp);
  return t;
}

}  // namespace moran
}  // namespace msc

extern "C" int msc_geary_workspace_size(int64_t n, int64_t n_genes, int64_t nnz, int n_perms, size_t* bytes) {
  MSC_TRY(validate(n, n_genes, nnz, n_perms));
  MSC_REQUIRE(bytes, "geary: bytes is null");
  Plan p = make_plan(n, n_genes, nnz, n_perms);
  Carve cv(nullptr, 0);
  carve_geary(cv, p);
  *bytes = cv.used + 256;
  return kOk;
}

extern "C" int msc_geary(const float* X, int64_t n, int64_t n_genes, const int64_t* w_indptr,
                         const int32_t* w_indices, const double* w_data, int64_t nnz, int transformation,
                         int n_perms, const int32_t* perms, int64_t seed, double* out_C, double* out_sims,
                         double* out_S0, void* workspace, size_t workspace_bytes, void* stream) {
  MSC_TRY(validate(n, n_genes, nnz, n_perms));
  MSC_REQUIRE(X && w_indptr && out_C && workspace && (nnz == 0 || (w_indices && w_data)),
              "geary: null pointer argument");
  MSC_REQUIRE(n_perms == 0 || out_sims, "geary: out_sims required when n_perms > 0");
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int64_t G = n_genes;
  Plan p = make_plan(n, G, nnz, n_perms);
  Carve cv(workspace, workspace_bytes);
  GearyBufs t = carve_geary(cv, p);
  if (!cv.ok()) {
    set_error("geary: workspace too small (%zu < %zu)", workspace_bytes, cv.used);
    return kWorkspaceTooSmall;
  }
  Bufs& b = t.b;
  const int64_t per = ceil_div(n, kColBlocks);
  const dim3 cg(kColBlocks, static_cast<unsigned>(std::min<int64_t>(ceil_div(G, kThreads), 64)));
  const unsigned gb = static_cast<unsigned>(ceil_div(G, kThreads));
  k_col_partial<<<cg, kThreads, 0, st>>>(X, n, G, per, nullptr, b.cpartial);
  k_col_final<<<gb, kThreads, 0, st>>>(b.cpartial, kColBlocks, G, 1.0 / static_cast<double>(n), b.mean);
  k_col_partial<<<cg, kThreads, 0, st>>>(X, n, G, per, b.mean, b.cpartial);
  k_col_final<<<gb, kThreads, 0, st>>>(b.cpartial, kColBlocks, G, 1.0, b.denom);
  MSC_LAUNCH_CHECK();
  k_row_normalize<<<static_cast<unsigned>(ceil_div(n, kThreads)), kThreads, 0, st>>>(w_indptr, w_data, n,
                                                                                       transformation, b.w);
  sum_f64(b.w, nnz, b.S0, b.sum_scratch, st);
  MSC_LAUNCH_CHECK();
  k_spmm_geary<<<static_cast<unsigned>(n), kThreads, 0, st>>>(X, G, w_indptr, w_indices, b.w, b.mean, t.Y64, t.D, t.R);
  k_col_partial_f64<<<cg, kThreads, 0, st>>>(t.D, n, G, per, b.cpartial);
  k_col_final<<<gb, kThreads, 0, st>>>(b.cpartial, kColBlocks, G, 1.0, t.Nobs);
  MSC_LAUNCH_CHECK();
  Feistel fst{};
  {
    int bits = 2;
    while ((int64_t(1) << bits) < n) bits += 2;
    fst.half = bits / 2;
    fst.mask = (uint64_t(1) << fst.half) - 1;
    fst.key = stream_key(seed, 7);
  }
  const bool vec4 = p.vec == 4 && (reinterpret_cast<uintptr_t>(X) % 16 == 0);
  auto launch = [&](dim3 grid, int P_, const int32_t* pr, int observed) {
    if (vec4)
      k_geary_partial<4><<<grid, p.pthreads, 0, st>>>(X, t.Y64, t.R, n, G, b.mean, p.per, P_, pr, fst, observed,
                                                      t.partA, t.partB);
    else {
      const dim3 g1(grid.x, grid.y, static_cast<unsigned>(ceil_div(G, kThreads)));
      k_geary_partial<1><<<g1, kThreads, 0, st>>>(X, t.Y64, t.R, n, G, b.mean, p.per, P_, pr, fst, observed, t.partA,
                                                  t.partB);
    }
  };
  launch(dim3(static_cast<unsigned>(p.nchunks), 1, static_cast<unsigned>(p.gtiles)), 1, nullptr, 1);
  k_geary_observed<<<gb, kThreads, 0, st>>>(t.partA, t.partB, p.nchunks, G, static_cast<double>(n), b.S0, b.denom,
                                            t.Nobs, t.Zc, out_C);
  MSC_LAUNCH_CHECK();
  if (n_perms > 0) {
    launch(dim3(static_cast<unsigned>(p.nchunks), static_cast<unsigned>(ceil_div(n_perms, kPermBatch)),
                static_cast<unsigned>(p.gtiles)),
           n_perms, perms, 0);
    k_geary_finish<<<static_cast<unsigned>(ceil_div(static_cast<int64_t>(n_perms) * G, kThreads)), kThreads, 0, st>>>(
        t.partA, t.partB, p.nchunks, n_perms, G, static_cast<double>(n), b.S0, b.denom, t.Zc, out_sims);
    MSC_LAUNCH_CHECK();
  }
  if (out_S0) {
    MSC_HIP(hipMemcpyAsync(out_S0, b.S0, sizeof(double), hipMemcpyDeviceToHost, st));
    MSC_HIP(hipStreamSynchronize(st));
  }
  return kOk;
}

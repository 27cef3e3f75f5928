// Exact 2-D / 3-D spatial kNN on a uniform grid (cell list), self excluded.
//
// Replaces neighbors._knn_grid with its grid_knn kernel (src/metalsinglecell/neighbors.py:
// 333-417, 143-196) as used by gr._knn_edges / spatial_neighbors_knn (gr.py:171-225).
// Differences by design: cells are built on the device (stable radix sort by cell id);
// queries run in cell order so a wave scans neighbouring cells together; instead of the
// reference's 3x3 scan plus brute-force fallback, each query scans Chebyshev shells of
// cells R = 0, 1, 2, ... until its k-th distance is provably below the shell radius.
//
// Order key (identical to oracle/grid.py): (float32 distance, index) with
//   dist = sqrt(fl(fl(dx*dx) + fl(dy*dy) [+ fl(dz*dz)])), dx = fl(x_j - x_i).
// Cell sizing follows the reference (neighbors.py:354-370): about max(2*want, 8) points per
// cell with want = k + 5 (gr.py:171,188), coarsened while the table exceeds 4n + 16 cells.
#include <cmath>

#include "common.h"
#include "msc_hip.h"
#include "prims.h"

namespace msc {
namespace grid {

constexpr int kThreads = 256;
constexpr uint64_t kEmpty = ~0ull;

struct Geo {
  float ox, oy, oz, cell;
  int gx, gy, gz, dim;
};

// per-block bounding box (grid-stride) into out + 8 * block: [min x3, max x3, any non-finite]
__global__ void k_bbox(const float* __restrict__ X, int64_t n, int dim, float* __restrict__ out) {
  float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  int bad = 0;
  out += 8 * blockIdx.x;
  for (int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * blockDim.x)
    for (int t = 0; t < dim; ++t) {
      const float v = X[i * dim + t];
      bad |= !isfinite(v);
      mn[t] = fminf(mn[t], v);
      mx[t] = fmaxf(mx[t], v);
    }
  bad = __any(bad);
  __shared__ int sbad[kThreads / kWave];
  if (lane_id() == 0) sbad[threadIdx.x / kWave] = bad;
  __shared__ float smn[3][kThreads / kWave], smx[3][kThreads / kWave];
  for (int t = 0; t < 3; ++t) {
    float a = mn[t], b = mx[t];
#pragma unroll
    for (int m = 1; m < kWave; m <<= 1) {
      a = fminf(a, __shfl_xor(a, m, kWave));
      b = fmaxf(b, __shfl_xor(b, m, kWave));
    }
    if (lane_id() == 0) {
      smn[t][threadIdx.x / kWave] = a;
      smx[t][threadIdx.x / kWave] = b;
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int any_bad = 0;
    for (int w = 0; w < kThreads / kWave; ++w) any_bad |= sbad[w];
    out[6] = any_bad ? 1.0f : 0.0f;
  }
  if (threadIdx.x == 0)
    for (int t = 0; t < 3; ++t) {
      float a = smn[t][0], b = smx[t][0];
      for (int w = 1; w < kThreads / kWave; ++w) {
        a = fminf(a, smn[t][w]);
        b = fmaxf(b, smx[t][w]);
      }
      out[t] = a;
      out[3 + t] = b;
    }
}
// combine nb per-block boxes (min / max / or are order independent)
__global__ void k_bbox_final(const float* __restrict__ part, int nb, float* __restrict__ out) {
  if (threadIdx.x != 0) return;
  float r[7] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY, 0.0f};
  for (int b = 0; b < nb; ++b) {
    const float* q = part + 8 * b;
    for (int t = 0; t < 3; ++t) {
      r[t] = fminf(r[t], q[t]);
      r[3 + t] = fmaxf(r[3 + t], q[3 + t]);
    }
    r[6] = fmaxf(r[6], q[6]);
  }
  for (int t = 0; t < 7; ++t) out[t] = r[t];
}
constexpr int kBBoxBlocks = 256;

__device__ __forceinline__ int cell_coord(float v, float o, float cell, int g) {
  int c = static_cast<int>(floorf((v - o) / cell));
  return c < 0 ? 0 : (c >= g ? g - 1 : c);
}

__global__ void k_cell_keys(const float* __restrict__ X, int64_t n, Geo geo, uint32_t* __restrict__ key,
                            uint32_t* __restrict__ val) {
  const int64_t i = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const int cx = cell_coord(X[i * geo.dim], geo.ox, geo.cell, geo.gx);
  const int cy = cell_coord(X[i * geo.dim + 1], geo.oy, geo.cell, geo.gy);
  const int cz = geo.dim == 3 ? cell_coord(X[i * geo.dim + 2], geo.oz, geo.cell, geo.gz) : 0;
  key[i] = static_cast<uint32_t>((static_cast<int64_t>(cz) * geo.gy + cy) * geo.gx + cx);
  val[i] = static_cast<uint32_t>(i);
}

template <int K>
__device__ __forceinline__ void insert_key(uint64_t (&lst)[K], uint64_t key) {
#pragma unroll
  for (int i = K - 1; i > 0; --i) {
    const uint64_t a = lst[i - 1], b = lst[i];
    lst[i] = (key < a) ? a : ((key < b) ? key : b);
  }
  lst[0] = key < lst[0] ? key : lst[0];
}

template <int K>
__device__ __forceinline__ uint64_t kth_of(const uint64_t (&lst)[K], int k) {
  uint64_t r = kEmpty;
#pragma unroll
  for (int i = 0; i < K; ++i)
    if (i == k - 1) r = lst[i];
  return r;
}

template <int K>
__global__ __launch_bounds__(kThreads) void k_grid_query(
    const float* __restrict__ X, int64_t n, int k, Geo geo, const uint32_t* __restrict__ order,
    const int64_t* __restrict__ cell_start, int32_t* __restrict__ out_idx, float* __restrict__ out_dist,
    int32_t* __restrict__ shells) {
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t >= n) return;
  const int dim = geo.dim;
  const int64_t i = order[t];
  const float qx = X[i * dim], qy = X[i * dim + 1], qz = dim == 3 ? X[i * dim + 2] : 0.0f;
  const int cx = cell_coord(qx, geo.ox, geo.cell, geo.gx);
  const int cy = cell_coord(qy, geo.oy, geo.cell, geo.gy);
  const int cz = dim == 3 ? cell_coord(qz, geo.oz, geo.cell, geo.gz) : 0;
  uint64_t lst[K];
#pragma unroll
  for (int j = 0; j < K; ++j) lst[j] = kEmpty;
  const int gmax = max(geo.gx, max(geo.gy, geo.gz));
  // absolute slack for cell assignment rounding at these coordinate magnitudes
  const float mag = fmaxf(fmaxf(fabsf(geo.ox), fabsf(geo.oy)), fabsf(geo.oz)) + geo.cell * gmax;
  const float slack = mag * 4.8e-7f;
  int R = 0;
  for (;; ++R) {
    const int zlo = dim == 3 ? -R : 0, zhi = dim == 3 ? R : 0;
    for (int dz = zlo; dz <= zhi; ++dz) {
      const int z = cz + dz;
      if (z < 0 || z >= geo.gz) continue;
      for (int dy = -R; dy <= R; ++dy) {
        const int y = cy + dy;
        if (y < 0 || y >= geo.gy) continue;
        const bool edge = (abs(dz) == R) || (abs(dy) == R);
        const int step = edge ? 1 : 2 * R;
        for (int dx = -R; dx <= R; dx += (step > 0 ? step : 1)) {
          const int x = cx + dx;
          if (x < 0 || x >= geo.gx) continue;
          const int64_t cell = (static_cast<int64_t>(z) * geo.gy + y) * geo.gx + x;
          for (int64_t p = cell_start[cell]; p < cell_start[cell + 1]; ++p) {
            const int64_t j = order[p];
            if (j == i) continue;
            const float ddx = __fsub_rn(X[j * dim], qx);
            const float ddy = __fsub_rn(X[j * dim + 1], qy);
            float d2 = __fadd_rn(__fmul_rn(ddx, ddx), __fmul_rn(ddy, ddy));
            if (dim == 3) {
              const float ddz = __fsub_rn(X[j * dim + 2], qz);
              d2 = __fadd_rn(d2, __fmul_rn(ddz, ddz));
            }
            const uint64_t key = (static_cast<uint64_t>(__float_as_uint(sqrtf(d2))) << 32) | static_cast<uint64_t>(j);
            if (key < lst[K - 1]) insert_key<K>(lst, key);
          }
          if (R == 0) break;
        }
      }
    }
    if (R >= gmax) break;
    const uint64_t kk = kth_of<K>(lst, k);
    if (R >= 1 && kk != kEmpty) {
      const float kd = __uint_as_float(static_cast<uint32_t>(kk >> 32));
      if (kd < static_cast<float>(R) * geo.cell * (1.0f - 1e-5f) - slack) break;
    }
  }
#pragma unroll
  for (int j = 0; j < K; ++j)
    if (j < k) {
      out_idx[i * k + j] = static_cast<int32_t>(lst[j] & 0xFFFFFFFFu);
      out_dist[i * k + j] = __uint_as_float(static_cast<uint32_t>(lst[j] >> 32));
    }
  if (shells) shells[i] = R;
}

// float64 edge lengths from float64 coordinates (gr.py:32-34)
__global__ void k_edge_len(const double* __restrict__ X, int64_t n, int dim, int k, const int32_t* __restrict__ idx,
                           double* __restrict__ out) {
  const int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (e >= n * k) return;
  const int64_t i = e / k, j = idx[e];
  double s = 0.0;
  for (int t = 0; t < dim; ++t) {
    const double d = X[i * dim + t] - X[j * dim + t];
    s += d * d;
  }
  out[e] = sqrt(s);
}

__global__ void k_max_i32(const int32_t* __restrict__ v, int64_t n, int32_t* __restrict__ out) {
  int32_t m = 0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) m = max(m, v[i]);
#pragma unroll
  for (int s = 1; s < kWave; s <<= 1) m = max(m, __shfl_xor(m, s, kWave));
  __shared__ int32_t sh[kThreads / kWave];
  if (lane_id() == 0) sh[threadIdx.x / kWave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kThreads / kWave; ++w) m = max(m, sh[w]);
    *out = max(m, sh[0]);
  }
}

static int64_t max_cells(int64_t n) { return 4 * n + 16; }

// cell geometry as the reference chooses it (neighbors.py:354-370)
static int choose_geometry(const float* mn, const float* mx, int64_t n, int dim, int want, Geo* geo) {
  double extent[3] = {0, 0, 0};
  bool nz[3] = {false, false, false};
  int eff = 0;
  double area = 1.0;
  for (int t = 0; t < dim; ++t) {
    extent[t] = static_cast<double>(mx[t]) - static_cast<double>(mn[t]);
    nz[t] = extent[t] > 0;
    if (nz[t]) {
      ++eff;
      area *= extent[t];
    }
  }
  geo->dim = dim;
  geo->ox = mn[0];
  geo->oy = mn[1];
  geo->oz = dim == 3 ? mn[2] : 0.0f;
  if (eff == 0) {  // all points coincide: one cell
    geo->cell = 1.0f;
    geo->gx = geo->gy = geo->gz = 1;
    return kOk;
  }
  const double density = static_cast<double>(n) / area;
  const double target = std::max(2 * want, 8);
  double cell = std::pow(target / density, 1.0 / eff);
  int64_t g[3] = {1, 1, 1};
  for (int it = 0; it < 64; ++it) {
    int64_t ncells = 1;
    for (int t = 0; t < dim; ++t) {
      g[t] = nz[t] ? static_cast<int64_t>(std::ceil(extent[t] / cell)) + 1 : 1;
      ncells *= g[t];
    }
    if (ncells <= max_cells(n)) break;
    cell *= std::pow(static_cast<double>(ncells) / (4.0 * n + 16.0), 1.0 / eff) * 1.0001;
  }
  geo->cell = static_cast<float>(cell);
  geo->gx = static_cast<int>(g[0]);
  geo->gy = static_cast<int>(g[1]);
  geo->gz = dim == 3 ? static_cast<int>(g[2]) : 1;
  return kOk;
}

struct Bufs {
  float* bbox;
  uint32_t *key, *val, *skey, *order;
  int64_t* cell_start;
  int32_t* shells;
  int32_t* red;
  void* sort_ws;
  size_t sort_ws_bytes;
};

static Bufs carve(Carve& cv, int64_t n) {
  Bufs b{};
  b.bbox = cv.take<float>(8 * (kBBoxBlocks + 1));
  b.key = cv.take<uint32_t>(n);
  b.val = cv.take<uint32_t>(n);
  b.skey = cv.take<uint32_t>(n);
  b.order = cv.take<uint32_t>(n);
  b.cell_start = cv.take<int64_t>(max_cells(n) + 2);
  b.shells = cv.take<int32_t>(n);
  b.red = cv.take<int32_t>(4);
  b.sort_ws_bytes = prims::radix_workspace(n);
  b.sort_ws = cv.take<char>(b.sort_ws_bytes);
  return b;
}

static int validate(int64_t n, int dim, int k) {
  MSC_REQUIRE(dim == 2 || dim == 3, "spatial_knn: dim must be 2 or 3");
  MSC_REQUIRE(k >= 1 && k <= 32, "spatial_knn: k must be in [1, 32]");
  MSC_REQUIRE(n > k && n < (int64_t(1) << 31) / 4, "spatial_knn: need k < n < 2^29");
  return kOk;
}

}  // namespace grid
}  // namespace msc

using namespace msc;
using namespace msc::grid;

extern "C" int msc_spatial_knn_workspace_size(int64_t n, int dim, int k, size_t* bytes) {
  MSC_TRY(validate(n, dim, k));
  MSC_REQUIRE(bytes, "spatial_knn: bytes is null");
  Carve cv(nullptr, 0);
  carve(cv, n);
  *bytes = cv.used + 256;
  return kOk;
}

extern "C" int msc_spatial_knn(const float* coords, int64_t n, int dim, int k, int32_t* out_idx, float* out_dist,
                               const double* coords64, double* out_len64, msc_spatial_info* info, void* workspace,
                               size_t workspace_bytes, void* stream) {
  MSC_TRY(validate(n, dim, k));
  MSC_REQUIRE(coords && out_idx && out_dist && workspace, "spatial_knn: null pointer argument");
  MSC_REQUIRE((coords64 == nullptr) == (out_len64 == nullptr), "spatial_knn: coords64 and out_len64 go together");
  hipStream_t st = static_cast<hipStream_t>(stream);
  Carve cv(workspace, workspace_bytes);
  Bufs b = carve(cv, n);
  if (!cv.ok()) {
    set_error("spatial_knn: workspace too small (%zu < %zu)", workspace_bytes, cv.used);
    return kWorkspaceTooSmall;
  }
  const int nbb = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(kBBoxBlocks, ceil_div(n, 4 * kThreads))));
  k_bbox<<<nbb, kThreads, 0, st>>>(coords, n, dim, b.bbox + 8);
  k_bbox_final<<<1, kWave, 0, st>>>(b.bbox + 8, nbb, b.bbox);
  MSC_LAUNCH_CHECK();
  float bb[8];
  MSC_HIP(hipMemcpyAsync(bb, b.bbox, 7 * sizeof(float), hipMemcpyDeviceToHost, st));
  MSC_HIP(hipStreamSynchronize(st));
  MSC_REQUIRE(bb[6] == 0.0f, "spatial_knn: coordinates must be finite");
  for (int t = 0; t < dim; ++t)
    MSC_REQUIRE(std::isfinite(bb[t]) && std::isfinite(bb[3 + t]), "spatial_knn: coordinates must be finite");
  Geo geo;
  MSC_TRY(choose_geometry(bb, bb + 3, n, dim, k + 5, &geo));
  const int64_t ncells = static_cast<int64_t>(geo.gx) * geo.gy * geo.gz;
  const unsigned g = static_cast<unsigned>(ceil_div(n, kThreads));
  k_cell_keys<<<g, kThreads, 0, st>>>(coords, n, geo, b.key, b.val);
  MSC_LAUNCH_CHECK();
  MSC_TRY(prims::radix_sort_pairs(b.key, b.val, b.skey, b.order, n, prims::bits_for(std::max<int64_t>(ncells - 1, 1)),
                                  b.sort_ws, b.sort_ws_bytes, st));
  MSC_TRY(prims::segment_starts(b.skey, n, ncells, b.cell_start, st));
  int32_t* shells = info ? b.shells : nullptr;
  if (k <= 8)
    k_grid_query<8><<<g, kThreads, 0, st>>>(coords, n, k, geo, b.order, b.cell_start, out_idx, out_dist, shells);
  else if (k <= 16)
    k_grid_query<16><<<g, kThreads, 0, st>>>(coords, n, k, geo, b.order, b.cell_start, out_idx, out_dist, shells);
  else
    k_grid_query<32><<<g, kThreads, 0, st>>>(coords, n, k, geo, b.order, b.cell_start, out_idx, out_dist, shells);
  MSC_LAUNCH_CHECK();
  if (coords64) {
    k_edge_len<<<static_cast<unsigned>(ceil_div(n * k, kThreads)), kThreads, 0, st>>>(coords64, n, dim, k, out_idx,
                                                                                      out_len64);
    MSC_LAUNCH_CHECK();
  }
  if (info) {
    k_max_i32<<<1, kThreads, 0, st>>>(b.shells, n, b.red);
    MSC_LAUNCH_CHECK();
    int32_t ms;
    MSC_HIP(hipMemcpyAsync(&ms, b.red, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    MSC_HIP(hipStreamSynchronize(st));
    info->max_shell = ms;
    info->grid_x = geo.gx;
    info->grid_y = geo.gy;
    info->grid_z = geo.gz;
    info->cell = geo.cell;
  }
  return kOk;
}

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
#include "wide.h"

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

// ---- fixed-radius neighbours ------------------------------------------------------------
// Pair (i, j), j != i, is kept when r2lo <= d2 <= r2hi with d2 the separately rounded float32
// sum of the kNN path.  One thread per query, queries in cell order; the cells within `ring`
// of the query's cell hold every kept pair (the cell side includes the rounding margin).
// COUNT writes counts[i]; otherwise columns go to out_col from indptr[i] on, in walk order.
template <bool COUNT>
__global__ __launch_bounds__(kThreads) void k_radius(const float* __restrict__ X, int64_t n, Geo geo, int ring,
                                                     float r2lo, float r2hi, const uint32_t* __restrict__ order,
                                                     const int64_t* __restrict__ cell_start,
                                                     int64_t* __restrict__ rows, uint32_t* __restrict__ out_col) {
  const int64_t t = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (t >= n) {
    if (COUNT && t == n) rows[n] = 0;  // the scan's last input
    return;
  }
  const int dim = geo.dim;
  const int64_t i = order[t];
  const float qx = X[i * dim], qy = X[i * dim + 1], qz = dim == 3 ? X[i * dim + 2] : 0.0f;
  const int cx = cell_coord(qx, geo.ox, geo.cell, geo.gx);
  const int cy = cell_coord(qy, geo.oy, geo.cell, geo.gy);
  const int cz = dim == 3 ? cell_coord(qz, geo.oz, geo.cell, geo.gz) : 0;
  int64_t w = COUNT ? 0 : rows[i];
  const int64_t wend = COUNT ? 0 : rows[i + 1];  // the fill never writes past its row
  int64_t cnt = 0;
  const int rz = dim == 3 ? ring : 0;
  for (int z = max(cz - rz, 0); z <= min(cz + rz, geo.gz - 1); ++z)
    for (int y = max(cy - ring, 0); y <= min(cy + ring, geo.gy - 1); ++y) {
      const int64_t row = (static_cast<int64_t>(z) * geo.gy + y) * geo.gx;
      // the cells x0..x1 of one grid row are contiguous in the cell list
      const int64_t p0 = cell_start[row + max(cx - ring, 0)];
      const int64_t p1 = cell_start[row + min(cx + ring, geo.gx - 1) + 1];
      for (int64_t p = p0; p < p1; ++p) {
        const int64_t j = order[p];
        if (j == i) continue;
        const float ddx = __fsub_rn(X[j * dim], qx);
        const float ddy = __fsub_rn(X[j * dim + 1], qy);
        float d2 = __fadd_rn(__fmul_rn(ddx, ddx), __fmul_rn(ddy, ddy));
        if (dim == 3) {
          const float ddz = __fsub_rn(X[j * dim + 2], qz);
          d2 = __fadd_rn(d2, __fmul_rn(ddz, ddz));
        }
        if (d2 < r2lo || d2 > r2hi) continue;
        if (COUNT) ++cnt;
        else if (w < wend) out_col[w++] = static_cast<uint32_t>(j);
      }
    }
  if (COUNT) rows[i] = cnt;
}

// row id of every CSR position (one wave per row)
__global__ __launch_bounds__(kThreads) void k_row_of(const int64_t* __restrict__ indptr, int64_t n,
                                                     uint32_t* __restrict__ row) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * (kThreads / kWave) + threadIdx.x / kWave;
  if (q >= n) return;
  for (int64_t p = indptr[q] + lane_id(); p < indptr[q + 1]; p += kWave) row[p] = static_cast<uint32_t>(q);
}
__global__ void k_iota_u32(uint32_t* __restrict__ v, int64_t m) {
  const int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (e < m) v[e] = static_cast<uint32_t>(e);
}
__global__ void k_gather_u32(const uint32_t* __restrict__ src, const uint32_t* __restrict__ idx, int64_t m,
                             uint32_t* __restrict__ dst) {
  const int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (e < m) dst[e] = src[idx[e]];
}
// final columns (sorted within rows) with the float32 distance and float64 length of each edge
__global__ void k_radius_emit(const float* __restrict__ X, const double* __restrict__ X64, int dim, int64_t m,
                              const uint32_t* __restrict__ row, const uint32_t* __restrict__ perm,
                              const uint32_t* __restrict__ col, int32_t* __restrict__ out_col,
                              float* __restrict__ out_dist, double* __restrict__ out_len64) {
  const int64_t e = blockIdx.x * static_cast<int64_t>(blockDim.x) + threadIdx.x;
  if (e >= m) return;
  const int64_t i = row[e], j = col[perm[e]];
  out_col[e] = static_cast<int32_t>(j);
  if (out_dist) {
    const float ddx = __fsub_rn(X[j * dim], X[i * dim]);
    const float ddy = __fsub_rn(X[j * dim + 1], X[i * dim + 1]);
    float d2 = __fadd_rn(__fmul_rn(ddx, ddx), __fmul_rn(ddy, ddy));
    if (dim == 3) {
      const float ddz = __fsub_rn(X[j * dim + 2], X[i * dim + 2]);
      d2 = __fadd_rn(d2, __fmul_rn(ddz, ddz));
    }
    out_dist[e] = sqrtf(d2);
  }
  if (out_len64) {
    double s = 0.0;
    for (int t = 0; t < dim; ++t) {
      const double d = X64[i * dim + t] - X64[j * dim + t];
      s += d * d;
    }
    out_len64[e] = sqrt(s);
  }
}

__global__ void k_max_i32(const int32_t* __restrict__ v, int64_t n, int32_t* __restrict__ out) {
  int32_t m = 0;
#pragma unroll 8
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

// radius geometry: cell side >= the reach of a query (r_hi, widened by the float32 rounding
// of the kept test and of the cell assignment at these coordinate magnitudes), coarsened while
// the table exceeds 4n + 16 cells; ring = cells to scan on each side.
static int radius_geometry(const float* mn, const float* mx, int64_t n, int dim, double r_hi, Geo* geo, int* ring) {
  double extent[3] = {0, 0, 0}, mag = 0.0;
  bool nz[3] = {false, false, false};
  int eff = 0;
  for (int t = 0; t < dim; ++t) {
    extent[t] = static_cast<double>(mx[t]) - static_cast<double>(mn[t]);
    nz[t] = extent[t] > 0;
    eff += nz[t];
    mag = std::max(mag, std::max(std::fabs(static_cast<double>(mn[t])), std::fabs(static_cast<double>(mx[t]))));
  }
  geo->dim = dim;
  geo->ox = mn[0];
  geo->oy = mn[1];
  geo->oz = dim == 3 ? mn[2] : 0.0f;
  *ring = 1;
  if (eff == 0) {
    geo->cell = 1.0f;
    geo->gx = geo->gy = geo->gz = 1;
    return kOk;
  }
  const double reach = r_hi * (1.0 + 1e-5) + 4.0 * 4.8e-7 * (mag + 1e-30);
  double cell = std::max(reach, 1e-30);
  int64_t g[3] = {1, 1, 1};
  for (int it = 0; it < 256; ++it) {
    int64_t ncells = 1;
    bool big = false;
    for (int t = 0; t < dim; ++t) {
      const double gt = nz[t] ? std::ceil(extent[t] / cell) + 1.0 : 1.0;
      big |= gt > 1e9;
      g[t] = big ? 1 : static_cast<int64_t>(gt);
      ncells *= g[t];
    }
    if (!big && ncells <= max_cells(n)) break;
    cell *= big ? 16.0 : std::pow(static_cast<double>(ncells) / (4.0 * n + 16.0), 1.0 / eff) * 1.0001;
  }
  MSC_REQUIRE(g[0] * g[1] * g[2] <= max_cells(n), "spatial_radius: no cell geometry fits the coordinate box");
  float cf = static_cast<float>(cell);
  if (static_cast<double>(cf) < cell) cf = std::nextafter(cf, INFINITY);  // keep the side >= the reach
  geo->cell = cf;
  geo->gx = static_cast<int>(g[0]);
  geo->gy = static_cast<int>(g[1]);
  geo->gz = dim == 3 ? static_cast<int>(g[2]) : 1;
  *ring = std::max(1, static_cast<int>(std::ceil(reach / static_cast<double>(geo->cell))));
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

// device bounding box -> host bb[0..2] min, bb[3..5] max; rejects non-finite coordinates
static int bounding_box(const float* coords, int64_t n, int dim, const Bufs& b, hipStream_t st, float* bb,
                        const char* who) {
  const int nbb = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(kBBoxBlocks, ceil_div(n, 4 * kThreads))));
  k_bbox<<<nbb, kThreads, 0, st>>>(coords, n, dim, b.bbox + 8);
  k_bbox_final<<<1, kWave, 0, st>>>(b.bbox + 8, nbb, b.bbox);
  MSC_LAUNCH_CHECK();
  MSC_HIP(hipMemcpyAsync(bb, b.bbox, 7 * sizeof(float), hipMemcpyDeviceToHost, st));
  MSC_HIP(hipStreamSynchronize(st));
  MSC_REQUIRE(bb[6] == 0.0f, "%s: coordinates must be finite", who);
  for (int t = 0; t < dim; ++t)
    MSC_REQUIRE(std::isfinite(bb[t]) && std::isfinite(bb[3 + t]), "%s: coordinates must be finite", who);
  return kOk;
}

// cell list: points sorted by cell id (stable: index order within a cell), cell starts
static int sort_cells(const float* coords, int64_t n, const Geo& geo, const Bufs& b, hipStream_t st) {
  const int64_t ncells = static_cast<int64_t>(geo.gx) * geo.gy * geo.gz;
  k_cell_keys<<<static_cast<unsigned>(ceil_div(n, kThreads)), kThreads, 0, st>>>(coords, n, geo, b.key, b.val);
  MSC_LAUNCH_CHECK();
  MSC_TRY(prims::radix_sort_pairs(b.key, b.val, b.skey, b.order, n, prims::bits_for(std::max<int64_t>(ncells - 1, 1)),
                                  b.sort_ws, b.sort_ws_bytes, st));
  return prims::segment_starts(b.skey, n, ncells, b.cell_start, st);
}

// The grid query holds at most 32 neighbours in registers and bins 2-D / 3-D cells; other
// shapes take the exact all-pairs path (the reference's brute-force fallback for k > 32 or
// d not in {2, 3}, neighbors.py:350-355).
static bool grid_shape(int dim, int k) { return (dim == 2 || dim == 3) && k <= 32; }

static int validate(int64_t n, int dim, int k) {
  MSC_REQUIRE(dim >= 1, "spatial_knn: dim must be >= 1");
  MSC_REQUIRE(k >= 1 && k < wide::kMaxK, "spatial_knn: k must be in [1, %d]", wide::kMaxK - 1);
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
  if (!grid_shape(dim, k)) {
    *bytes = 256;
    return kOk;
  }
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
  if (!grid_shape(dim, k)) {
    // k + 1 keys per row by (float32 distance, index), the point itself (key 0) dropped
    MSC_TRY(wide::knn(coords, n, dim, k + 1, true, 1, out_idx, out_dist, st));
    if (coords64) {
      k_edge_len<<<static_cast<unsigned>(ceil_div(n * k, kThreads)), kThreads, 0, st>>>(coords64, n, dim, k,
                                                                                        out_idx, out_len64);
      MSC_LAUNCH_CHECK();
    }
    if (info) {
      info->max_shell = -1;
      info->grid_x = info->grid_y = info->grid_z = 0;
      info->cell = 0.0f;
    }
    return kOk;
  }
  Carve cv(workspace, workspace_bytes);
  Bufs b = carve(cv, n);
  if (!cv.ok()) {
    set_error("spatial_knn: workspace too small (%zu < %zu)", workspace_bytes, cv.used);
    return kWorkspaceTooSmall;
  }
  float bb[8];
  MSC_TRY(bounding_box(coords, n, dim, b, st, bb, "spatial_knn"));
  Geo geo;
  MSC_TRY(choose_geometry(bb, bb + 3, n, dim, k + 5, &geo));
  MSC_TRY(sort_cells(coords, n, geo, b, st));
  const unsigned g = static_cast<unsigned>(ceil_div(n, kThreads));
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

// ---- fixed-radius neighbours (count, then fill) -------------------------------------------
namespace {
int radius_args(int64_t n, int dim, double r_min, double r_max) {
  MSC_REQUIRE(dim == 2 || dim == 3, "spatial_radius: dim must be 2 or 3");
  MSC_REQUIRE(n >= 1 && n < (int64_t(1) << 31) / 4, "spatial_radius: need 1 <= n < 2^29");
  MSC_REQUIRE(std::isfinite(r_min) && std::isfinite(r_max) && r_min >= 0.0 && r_max >= r_min,
              "spatial_radius: need 0 <= r_min <= r_max, finite");
  return kOk;
}

// radix scratch for the (row, column) sort of the filled CSR
struct RadiusSort {
  uint32_t *col, *row, *k1, *v0, *v1, *k2, *v2;
  void* ws;
  size_t ws_bytes;
};
RadiusSort carve_sort(Carve& cv, int64_t nnz) {
  RadiusSort r{};
  const size_t m = static_cast<size_t>(std::max<int64_t>(nnz, 1));
  r.col = cv.take<uint32_t>(m);
  r.row = cv.take<uint32_t>(m);
  r.k1 = cv.take<uint32_t>(m);
  r.v0 = cv.take<uint32_t>(m);
  r.v1 = cv.take<uint32_t>(m);
  r.k2 = cv.take<uint32_t>(m);
  r.v2 = cv.take<uint32_t>(m);
  r.ws_bytes = prims::radix_workspace(static_cast<int64_t>(m));
  r.ws = cv.take<char>(r.ws_bytes);
  return r;
}

// the grid of a radius query: bounding box, geometry, cell list
int radius_grid(const float* coords, int64_t n, int dim, double r_max, const Bufs& b, hipStream_t st, Geo* geo,
                int* ring) {
  float bb[8];
  MSC_TRY(bounding_box(coords, n, dim, b, st, bb, "spatial_radius"));
  MSC_TRY(radius_geometry(bb, bb + 3, n, dim, r_max, geo, ring));
  return sort_cells(coords, n, *geo, b, st);
}
}  // namespace

extern "C" int msc_spatial_radius_workspace_size(int64_t n, int dim, int64_t nnz, size_t* bytes) {
  MSC_TRY(radius_args(n, dim, 0.0, 0.0));
  MSC_REQUIRE(bytes && nnz >= 0 && nnz < (int64_t(1) << 31), "spatial_radius: bad nnz or null bytes");
  Carve cv(nullptr, 0);
  carve(cv, n);
  const size_t count_bytes = cv.used + prims::scan_workspace<int64_t>(n + 1) + 256;  // the count call's layout
  carve_sort(cv, nnz);
  *bytes = std::max(cv.used + 256, count_bytes);
  return kOk;
}

extern "C" int msc_spatial_radius_count(const float* coords, int64_t n, int dim, double r_min, double r_max,
                                        int64_t* out_indptr, int64_t* out_nnz, void* workspace,
                                        size_t workspace_bytes, void* stream) {
  MSC_TRY(radius_args(n, dim, r_min, r_max));
  MSC_REQUIRE(coords && out_indptr && out_nnz && workspace, "spatial_radius: null pointer argument");
  hipStream_t st = static_cast<hipStream_t>(stream);
  Carve cv(workspace, workspace_bytes);
  Bufs b = carve(cv, n);
  const size_t scan_bytes = prims::scan_workspace<int64_t>(n + 1);
  void* scan_ws = cv.take<char>(scan_bytes);
  if (!cv.ok()) {
    set_error("spatial_radius: workspace too small (%zu < %zu)", workspace_bytes, cv.used);
    return kWorkspaceTooSmall;
  }
  Geo geo;
  int ring;
  MSC_TRY(radius_grid(coords, n, dim, r_max, b, st, &geo, &ring));
  const float r2lo = static_cast<float>(r_min * r_min), r2hi = static_cast<float>(r_max * r_max);
  k_radius<true><<<static_cast<unsigned>(ceil_div(n + 1, kThreads)), kThreads, 0, st>>>(
      coords, n, geo, ring, r2lo, r2hi, b.order, b.cell_start, out_indptr, nullptr);
  MSC_LAUNCH_CHECK();
  MSC_TRY(prims::exclusive_scan<int64_t>(out_indptr, out_indptr, n + 1, scan_ws, scan_bytes, st));
  MSC_HIP(hipMemcpyAsync(out_nnz, out_indptr + n, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  MSC_HIP(hipStreamSynchronize(st));
  MSC_REQUIRE(*out_nnz < (int64_t(1) << 31), "spatial_radius: %lld pairs exceed 2^31 - 1", (long long)*out_nnz);
  return kOk;
}

extern "C" int msc_spatial_radius_fill(const float* coords, int64_t n, int dim, double r_min, double r_max,
                                       const int64_t* indptr, int64_t nnz, int32_t* out_indices, float* out_dist,
                                       const double* coords64, double* out_len64, void* workspace,
                                       size_t workspace_bytes, void* stream) {
  MSC_TRY(radius_args(n, dim, r_min, r_max));
  MSC_REQUIRE(coords && indptr && workspace && (nnz == 0 || out_indices), "spatial_radius: null pointer argument");
  MSC_REQUIRE((coords64 == nullptr) == (out_len64 == nullptr), "spatial_radius: coords64 and out_len64 go together");
  MSC_REQUIRE(nnz >= 0 && nnz < (int64_t(1) << 31), "spatial_radius: bad nnz");
  if (nnz == 0) return kOk;
  hipStream_t st = static_cast<hipStream_t>(stream);
  Carve cv(workspace, workspace_bytes);
  Bufs b = carve(cv, n);
  RadiusSort r = carve_sort(cv, nnz);
  if (!cv.ok()) {
    set_error("spatial_radius: workspace too small for %lld pairs (%zu < %zu)", (long long)nnz, workspace_bytes,
              cv.used);
    return kWorkspaceTooSmall;
  }
  Geo geo;
  int ring;
  MSC_TRY(radius_grid(coords, n, dim, r_max, b, st, &geo, &ring));
  const float r2lo = static_cast<float>(r_min * r_min), r2hi = static_cast<float>(r_max * r_max);
  k_radius<false><<<static_cast<unsigned>(ceil_div(n, kThreads)), kThreads, 0, st>>>(
      coords, n, geo, ring, r2lo, r2hi, b.order, b.cell_start, const_cast<int64_t*>(indptr), r.col);
  MSC_LAUNCH_CHECK();
  // (row, column) order: stable sort by column, then stable sort by row
  const unsigned ge = static_cast<unsigned>(ceil_div(nnz, kThreads));
  k_row_of<<<static_cast<unsigned>(ceil_div(n, kThreads / kWave)), kThreads, 0, st>>>(indptr, n, r.row);
  k_iota_u32<<<ge, kThreads, 0, st>>>(r.v0, nnz);
  MSC_LAUNCH_CHECK();
  const int bits = prims::bits_for(std::max<int64_t>(n - 1, 1));
  MSC_TRY(prims::radix_sort_pairs(r.col, r.v0, r.k1, r.v1, nnz, bits, r.ws, r.ws_bytes, st));
  k_gather_u32<<<ge, kThreads, 0, st>>>(r.row, r.v1, nnz, r.k1);
  MSC_LAUNCH_CHECK();
  MSC_TRY(prims::radix_sort_pairs(r.k1, r.v1, r.k2, r.v2, nnz, bits, r.ws, r.ws_bytes, st));
  k_radius_emit<<<ge, kThreads, 0, st>>>(coords, coords64, dim, nnz, r.k2, r.v2, r.col, out_indices, out_dist,
                                         out_len64);
  MSC_LAUNCH_CHECK();
  return kOk;
}

// Exact kNN for the shapes the MFMA select does not take: more than 64 neighbours, or more
// dimensions than its K padding holds, and the spatial kNN beyond k = 32 or outside 2-D/3-D.
#pragma once

#include "common.h"

namespace msc {
namespace wide {

constexpr int kMaxK = 512;  // neighbours per row (the LDS list holds a power of two >= k)

// Self-inclusive exact kNN of every row of X (n x d float32): out_idx / out_dist n x (k - skip),
// row q listing its k smallest keys from position `skip` on.  The key is (j != q, value, j)
// with value the float32 direct-difference squared distance (sqrt_key = false, the kNN order
// of neighbors.py:420-462) or its float32 square root (sqrt_key = true, the (distance, index)
// order of the spatial builders, gr.py:171-195); out_dist is the square root either way.
int knn(const float* X, int64_t n, int d, int k, bool sqrt_key, int skip, int32_t* out_idx,
        float* out_dist, hipStream_t st);

}  // namespace wide
}  // namespace msc

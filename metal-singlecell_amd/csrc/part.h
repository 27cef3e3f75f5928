// Partition-pruned candidate ranges for the exact kNN select (knn.hip).
//
// build() reorders the rows of X so that points of one partition (a ball around a centre) are
// contiguous and partitions whose balls overlap sit next to each other, cuts the reordered
// rows into query blocks of at most 128 rows that never straddle two groups, and gives every
// block at most kMaxRanges candidate row ranges.  A partition is left out of a block's ranges
// only when the triangle inequality proves that none of its points can be among the k nearest
// neighbours of any of the block's queries, so the select + certificate that scan those ranges
// return the same exact kNN as a scan of all n rows.
#pragma once

#include "common.h"

namespace msc {
namespace part {

constexpr int kBlockQ = 128;      // queries per select block
constexpr int kMaxRanges = 4;     // candidate ranges per block (the select's splits)
constexpr int kDMax = 64;         // dimensions the partition kernels take
constexpr int kP0Max = 512;       // stride-sampled centres
constexpr int kRounds = 4;        // D^2-sampling rounds
constexpr int kPeelMax = 2048;    // outlying points turned into singleton partitions, at most
constexpr int kMinGroup = 64;     // groups below this many points are pooled into one
constexpr int kMaxGroups = 1024;  // groups of their own (the rest are pooled)

inline int p0_for(int64_t n) {
  int64_t p = n / 256;
  return static_cast<int>(p < 16 ? 16 : (p > kP0Max ? kP0Max : p));
}
inline int pmax_for(int64_t n) { return p0_for(n) * (1 + 2 * kRounds) + kPeelMax; }
inline int64_t nblk_max(int64_t n) { return ceil_div(n, kBlockQ) + kMaxGroups + 1; }

// Device arrays point into the workspace given to build(); host fields are filled by build().
struct Layout {
  const float* Xs;            // n x d, the rows of X in partition order
  const int32_t* perm;        // partition-order row -> row of X
  const int32_t* qrows;       // n_slots: query slot -> partition-order row, -1 for padding
  const int64_t* blk_range;   // [kMaxRanges][nblk][2] candidate row ranges [begin, end)
  int64_t nblk;               // query blocks
  int64_t n_slots;            // nblk x 128
  int S;                      // ranges used (max over blocks), 1..kMaxRanges
  int n_parts;                // partitions (centres, singletons included)
  int n_groups;               // groups of their own (the pooled group not counted)
  int64_t n_peeled;           // singleton partitions
  double pairs;               // sum over blocks of (queries x candidate rows)
};

size_t workspace(int64_t n, int d);
int build(const float* X, int64_t n, int d, int k, void* ws, size_t ws_bytes, Layout* L, hipStream_t st);

}  // namespace part
}  // namespace msc

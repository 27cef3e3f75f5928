/* libmsc_hip — C ABI of the MI355X (gfx950) neighbor-graph path.
 *
 * Conventions
 *   - Every pointer argument named X, out_*, or *_dev is DEVICE memory owned by the caller
 *     (for example a torch tensor's data_ptr()).  Host pointers are marked `host`.
 *   - `stream` is a hipStream_t passed as void*; all work is enqueued on it, in order.
 *     Entry points that report host-side results (counts, info structs) synchronise it.
 *   - Scratch comes from a caller workspace sized by the matching *_workspace_size query;
 *     no entry point allocates device memory.
 *   - Return value: 0 on success, otherwise an MSC_E* code; msc_last_error() gives the
 *     calling thread's last message.
 *   - Results are deterministic: identical inputs give bit-identical outputs regardless of
 *     the stream, of other work on the device, or of the candidate-split count.
 */
#ifndef MSC_HIP_H_
#define MSC_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  MSC_OK = 0,
  MSC_EINVAL = 1,
  MSC_EHIP = 2,
  MSC_EWORKSPACE = 3,
  MSC_EUNSUPPORTED = 4,
  MSC_EINTERNAL = 5
};

enum { MSC_PRECISION_FP32 = 0, MSC_PRECISION_BF16_RERANK = 1 };

/* Library version (major*10000 + minor*100 + patch) and the thread's last error text. */
int msc_version(void);
const char* msc_last_error(void);

/* ---------------------------------------------------------------------------------------
 * (a) Exact Euclidean kNN.
 * Replaces neighbors._knn (src/metalsinglecell/neighbors.py:698-724) and its engines
 * _knn_gpu (:82-132, fp16 Gram + topk_rows kernel :25-79) and _knn_nndescent (:677-695).
 * X: n x d row-major float32.  Output: self-inclusive neighbours, column 0 = the point itself,
 * ascending by the exact float32 direct-difference squared distance, ties to the lower index.
 * out_idx n x k int32, out_dist n x k float32 (= sqrt of that squared distance).
 * precision: MSC_PRECISION_FP32 ranks candidates with FP32 MFMA; MSC_PRECISION_BF16_RERANK
 * ranks them with BF16 MFMA (FP32 accumulation).  Both re-rank 64 admitted candidates
 * exactly and certify each row; uncertified rows are recomputed by exact brute force, so
 * the two precisions return identical results.  Limits: 1 <= k <= 64, k <= n < 2^31,
 * d <= 127 (FP32) / 125 (BF16).
 * ------------------------------------------------------------------------------------- */
typedef struct {
  int timing;           /* in: nonzero -> record per-stage times (adds one stream sync) */
  int n_splits;         /* out: candidate splits used by the select kernel */
  int list_len;         /* out: admitted candidates per query */
  int64_t n_fallback;   /* out: rows recomputed by exact brute force */
  double prep_ms, select_ms, finalize_ms, fallback_ms; /* out (timing only) */
} msc_knn_info;

int msc_knn_workspace_size(int64_t n, int d, int k, int precision, size_t* bytes /*host*/);
int msc_knn_exact(const float* X, int64_t n, int d, int k, int precision, int32_t* out_idx,
                  float* out_dist, void* workspace, size_t workspace_bytes, void* stream,
                  msc_knn_info* info /*host, may be NULL*/);

/* obsp['distances'] as CSR from a self-inclusive kNN (neighbors.py:755-760): k-1 entries per
 * row (self column dropped), columns ascending, explicit zeros kept.
 * out_indptr n+1 int64, out_indices n*(k-1) int32, out_data n*(k-1) float32. */
int msc_knn_distances_csr(const int32_t* idx, const float* dist, int64_t n, int k,
                          int64_t* out_indptr, int32_t* out_indices, float* out_data,
                          void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MSC_HIP_H_ */

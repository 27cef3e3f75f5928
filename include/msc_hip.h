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
 * exactly and certify each row; rows that fail their certificate are selected again over
 * all candidates by a wide pass of the same precision (16 queries per wave, 128 admitted
 * candidates per query), and rows no pass certifies are recomputed by exact brute force, so
 * the two precisions return identical results.  Shapes beyond the MFMA select (k > 64, or
 * d > 127 in FP32 / d > 125 in BF16) compute every candidate's exact distance instead (the
 * reference's argpartition fallback for K > 32, neighbors.py:66-68); the result is the same
 * exact kNN and the workspace query returns a token size.  Limits: 1 <= k <= 512,
 * k <= n < 2^31, d >= 1.
 * Pruned candidate ranges (d <= 64, k <= 64, n >= 256): the rows are grouped into partitions
 * (balls around sampled centres) and every block of 128 queries scans only the partitions the
 * triangle inequality cannot rule out for any of its queries (a partition's ball against an upper
 * bound on each query's k-th neighbour distance, with float32 rounding margins).  The result is
 * the same exact kNN; only the work changes (about the sum of squared cluster fractions of n^2 on
 * well-separated clusters, all of n^2 on data without cluster structure).  info->prune: AUTO uses
 * them from 65,536 rows on, OFF never, ON whenever the shape allows (otherwise MSC_EINVAL).
 * info == NULL means AUTO.
 * ------------------------------------------------------------------------------------- */
enum { MSC_PRUNE_AUTO = 0, MSC_PRUNE_OFF = 1, MSC_PRUNE_ON = 2 };

typedef struct {
  int timing;           /* in: nonzero -> record per-stage times (adds one stream sync) */
  int n_splits;         /* out: candidate splits used by the select kernel */
  int list_len;         /* out: admitted candidates per query */
  int64_t n_second_pass;/* out: rows re-selected by the wide pass (certificate failed) */
  int64_t n_fallback;   /* out: rows recomputed by exact brute force */
  double prep_ms, select_ms, finalize_ms, second_pass_ms, fallback_ms; /* out (timing only) */
  int prune;            /* in (msc_knn_exact): MSC_PRUNE_AUTO / _OFF / _ON, see below */
  int n_partitions;     /* out: partitions of the pruned layout (0: every candidate scanned) */
  int n_groups;         /* out: groups of partitions with query blocks of their own */
  int n_ranges;         /* out: candidate row ranges per query block, at most */
  double scanned;       /* out: fraction of the n^2 (query, candidate) pairs the select scans */
  double partition_ms;  /* out (timing only): building the pruned layout */
} msc_knn_info;

int msc_knn_workspace_size(int64_t n, int d, int k, int precision, size_t* bytes /*host*/);
int msc_knn_exact(const float* X, int64_t n, int d, int k, int precision, int32_t* out_idx,
                  float* out_dist, void* workspace, size_t workspace_bytes, void* stream,
                  msc_knn_info* info /*host, may be NULL*/);

/* Batched exact kNN over many independent samples in one select launch (each query scans
 * only its own sample).  X: total x d rows, samples contiguous; offsets (host): n_samples+1
 * row offsets.  out_idx / out_dist: total x k, column ids GLOBAL (offset of the sample
 * included), so the result is the block-diagonal kNN of the concatenation.  Every sample
 * needs >= k cells; d <= 125.  The tile shape has two knobs, neither of which changes the
 * results: splits (1..16, 0 = 1): each sample's candidates are scanned in that many ranges
 * by separate blocks; list_len (0 = 16, 16 or 8; 8 needs k <= 32): keys each lane keeps per
 * query (64 or 32 admitted candidates per query; a shorter list costs less upkeep on short
 * candidate streams and may send more rows to the exact fallback).
 * Replaces per-sample calls of neighbors._knn (src/metalsinglecell/neighbors.py:698-724),
 * as bbknn's per-batch loop does (neighbors.py:491-559).
 *   _prepare uploads the layout tables into the workspace (synchronises);
 *   _launch enqueues the kernels only: no allocation, no host synchronisation, so it may be
 *   captured in a HIP graph; it needs a prior _prepare with the same offsets, d, k,
 *   precision, splits, list_len and workspace;
 *   msc_knn_batched = _prepare + _launch (+ the fallback count in info). */
int msc_knn_batched_workspace_size(const int64_t* offsets /*host*/, int n_samples, int d, int k,
                                   int precision, int splits, int list_len, size_t* bytes /*host*/);
int msc_knn_batched_prepare(const int64_t* offsets /*host*/, int n_samples, int d, int k, int precision,
                            int splits, int list_len, void* workspace, size_t workspace_bytes, void* stream);
int msc_knn_batched_launch(const float* X, const int64_t* offsets /*host*/, int n_samples, int d, int k,
                           int precision, int splits, int list_len, int32_t* out_idx, float* out_dist,
                           void* workspace, size_t workspace_bytes, void* stream);
int msc_knn_batched(const float* X, const int64_t* offsets /*host*/, int n_samples, int d, int k,
                    int precision, int splits, int list_len, int32_t* out_idx, float* out_dist, void* workspace,
                    size_t workspace_bytes, void* stream, msc_knn_info* info /*host, may be NULL*/);

/* Batch-balanced kNN: for every cell, its min(k, n_t) exact nearest cells within each batch
 * t, the per-batch lists concatenated in batch order and sorted by (float32 distance,
 * position in the concatenation).  Replaces the per-batch loop of neighbors.bbknn
 * (src/metalsinglecell/neighbors.py:491-541) as called by pp.bbknn (pp.py:809-833).
 * X: n x d float32 (device); batch (host): n codes in [0, n_batches), every code present,
 * n_batches <= 1024, k <= 64.  Width K = sum_t min(k, n_t) <= 1024 (returned in *out_width,
 * host; rows wider than 64 are combined by a sort in LDS).
 * out_idx / out_dist: n x K.  Within a batch the order is the kNN key (j != i, d2, j), so a
 * cell's own batch lists the cell itself first at distance 0.  d <= 125 (BF16) / 127.
 * Workspace grows as n_batches x n x 512 B (one candidate list per cell and batch). */
int msc_bbknn_workspace_size(const int32_t* batch /*host*/, int64_t n, int n_batches, int d, int k,
                             int precision, size_t* bytes /*host*/);
int msc_bbknn(const float* X, int64_t n, int d, const int32_t* batch /*host*/, int n_batches, int k,
              int precision, int32_t* out_idx, float* out_dist, int* out_width /*host*/,
              void* workspace, size_t workspace_bytes, void* stream,
              msc_knn_info* info /*host, may be NULL; n_fallback counts (cell, batch) lists*/);

/* obsp['distances'] of bbknn (neighbors.py:554-557): the K combined neighbours per row
 * without the row itself and without zero distances, columns ascending.  out_indptr n+1
 * int64 (nnz = out_indptr[n], on the device); out_indices / out_data capacity n*K.
 * 1 <= K <= 1024. */
int msc_bbknn_distances_workspace_size(int64_t n, size_t* bytes /*host*/);
int msc_bbknn_distances_csr(const int32_t* idx, const float* dist, int64_t n, int K,
                            int64_t* out_indptr, int32_t* out_indices, float* out_data,
                            void* workspace, size_t workspace_bytes, void* stream);

/* obsp['distances'] as CSR from a self-inclusive kNN (neighbors.py:755-760): k-1 entries per
 * row (self column dropped), columns ascending, explicit zeros kept.  1 <= k <= 513.
 * out_indptr n+1 int64, out_indices n*(k-1) int32, out_data n*(k-1) float32. */
int msc_knn_distances_csr(const int32_t* idx, const float* dist, int64_t n, int k,
                          int64_t* out_indptr, int32_t* out_indices, float* out_data,
                          void* stream);

/* ---------------------------------------------------------------------------------------
 * (c) UMAP fuzzy-simplicial-set connectivities.
 * Replaces the call of umap-learn's fuzzy_simplicial_set in neighbors.neighbors
 * (src/metalsinglecell/neighbors.py:742-753) with set_op_mix_ratio=1, local_connectivity=1.
 * knn_idx/knn_dist: n x k self-inclusive kNN (rows ascending, distinct indices per row).
 * Output: symmetric canonical CSR (columns ascending, no zeros, no duplicates) in
 * out_indptr (n+1 int64), out_indices / out_data (capacity entries; 2*n*(k-1) always
 * suffices), per-row sigma and rho (n float32 each), and *out_nnz (host).  1 <= k <= 1024,
 * n*k < 2^31.  Synchronises. */
int msc_fuzzy_workspace_size(int64_t n, int k, size_t* bytes /*host*/);
int msc_fuzzy_connectivities(const int32_t* knn_idx, const float* knn_dist, int64_t n, int k,
                             int64_t* out_indptr, int32_t* out_indices, float* out_data,
                             int64_t capacity, float* out_sigma, float* out_rho,
                             int64_t* out_nnz /*host*/, void* workspace, size_t workspace_bytes,
                             void* stream);

/* The same for a batch of samples whose kNN rows are concatenated (global column ids, as
 * msc_knn_batched writes them): the output is the block-diagonal CSR whose block s equals
 * msc_fuzzy_connectivities of sample s alone (the mean kNN distance used by the sigma floor
 * is taken per sample).  group_offsets (DEVICE, n_groups+1 row offsets).  capacity must be
 * >= 2*n*(k-1); nnz = out_indptr[n] stays on the device.  No allocation and no host
 * synchronisation: may be captured in a HIP graph.  Replaces per-sample calls of the
 * fuzzy_simplicial_set step (neighbors.py:742-753). */
int msc_fuzzy_batched_workspace_size(int64_t n, int k, int n_groups, size_t* bytes /*host*/);
int msc_fuzzy_connectivities_batched(const int32_t* knn_idx, const float* knn_dist, int64_t n, int k,
                                     const int64_t* group_offsets, int n_groups, int64_t* out_indptr,
                                     int32_t* out_indices, float* out_data, int64_t capacity,
                                     float* out_sigma, float* out_rho, void* workspace,
                                     size_t workspace_bytes, void* stream);

/* msc_fuzzy_connectivities_batched without a workspace argument.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Scratch is a static 64 MiB device
 * array per GPU, held until the process ends, so sizes whose workspace would exceed it
 * (msc_fuzzy_batched_workspace_size > 64 MiB, about n*k > 2.5M) are rejected with
 * MSC_EINVAL.  Calls that overlap on one GPU share that array and race: serialise them (one
 * stream, or events).  Every other argument, the result and the graph-capture property are
 * those of msc_fuzzy_connectivities_batched; the call allocates nothing and does not
 * synchronise. */
int msc_fuzzy_connectivities_batched_noalloc(const int32_t* knn_idx, const float* knn_dist, int64_t n, int k,
                                             const int64_t* group_offsets, int n_groups, int64_t* out_indptr,
                                             int32_t* out_indices, float* out_data, int64_t capacity,
                                             float* out_sigma, float* out_rho, void* stream);

/* Connectivity trimming of bbknn: neighbors._trim_graph (src/metalsinglecell/neighbors.py:
 * 465-488) on a symmetric canonical CSR.  Row i's threshold is its trim-th largest value
 * when it holds more than trim entries, else 0; entry (i, j) is kept when nonzero and >= the
 * thresholds of rows i and j.  Output canonical CSR (capacity = input nnz); *out_nnz (host,
 * may be NULL; when given the call synchronises).  trim >= 1. */
int msc_trim_graph_workspace_size(int64_t n, size_t* bytes /*host*/);
int msc_trim_graph(const int64_t* indptr, const int32_t* indices, const float* data, int64_t n,
                   int64_t trim, int64_t* out_indptr, int32_t* out_indices, float* out_data,
                   int64_t* out_nnz /*host*/, void* workspace, size_t workspace_bytes, void* stream);

/* ---------------------------------------------------------------------------------------
 * (d) Synchronous parallel Leiden (modularity, resolution gamma).
 * Replaces cluster.leiden(backend="gpu") -> graph.leiden.leiden
 * (src/metalsinglecell/cluster.py:36-54, graph/leiden.py:352-455) with its move / refine
 * kernels (graph/louvain.py:88-131, graph/leiden.py:100-144), pass loop
 * (louvain.py:282-394, leiden.py:251-349), contraction (louvain.py:440-453) and modularity
 * (graph/primitives.py:94-111).
 * Input: symmetric canonical CSR (indptr n+1 int64, indices nnz int32 columns ascending,
 * weights nnz float32 >= 0), both directions stored, self-loops allowed.
 * init_labels: n int32 labels to start the first pass from, or NULL (singletons).
 * out_labels: n int32 dense labels 0..C-1.  *out_modularity (host, may be NULL): Q of the
 * result.  Weights are quantised once to int64 fixed point (exponent reported in info), so
 * every reduction is an exact integer sum; labels are deterministic.  Synchronises. */
#define MSC_LEIDEN_MAX_RECORDS 64
typedef struct {
  int64_t n, n_edges;
  int32_t move_passes, refine_passes;
  int64_t n_comms, n_refined;
  double modularity;
  /* passes of each phase whose sigma_tot came from the previous passes' committed vertices
   * alone (exact int64 updates); the first pass of a phase builds it over all vertices */
  int32_t move_delta_passes, refine_delta_passes;
} msc_leiden_level;

typedef struct {
  int n_levels;                 /* records written (all iterations, in order) */
  int scale_exponent;           /* s: weights quantised as rint(w * 2^s) */
  int64_t twom_fixed;           /* total directed weight in fixed point */
  msc_leiden_level levels[MSC_LEIDEN_MAX_RECORDS];
  /* in: nonzero to time every launch of the move kernel (one per pass) with HIP events on
   * the call's stream; off, the three outputs below are zero.  Timing adds two event records
   * per pass and creates (and destroys) 16 events in the call. */
  int32_t time_moves;
  int32_t move_launches;        /* out: move-kernel launches timed */
  double move_ms;               /* out: their summed durations */
  double move_model_bytes;      /* out: sum over them of 8 E_l + 24 n_l (the level's edges and
                                 * vertices, every edge counted once per pass) */
} msc_leiden_info;

int msc_leiden_workspace_size(int64_t n, int64_t nnz, size_t* bytes /*host*/);
int msc_leiden(const int64_t* indptr, const int32_t* indices, const float* weights, int64_t n,
               int64_t nnz, double resolution, int64_t seed, int n_iterations,
               double commit_prob, const int32_t* init_labels, int32_t* out_labels,
               double* out_modularity /*host*/, msc_leiden_info* info /*host, may be NULL*/,
               void* workspace, size_t workspace_bytes, void* stream);

/* Leiden of many independent samples in one run: the samples are the diagonal blocks of one
 * CSR (sample s = rows offsets[s]..offsets[s+1]-1, host offsets, n_samples+1 of them, column
 * ids global, no edge between samples).  Sample s's labels (out_labels at its rows, dense
 * 0..C_s-1) and modularity (out_modularity[s], host) equal msc_leiden on that sample alone
 * with the same arguments: each sample is quantised with its own exponent, scored with its
 * own total weight, and draws its coins by its vertices' indices within the sample; the
 * passes run in lockstep and a sample whose multilevel loop ends is frozen while the rest
 * continue.  Replaces the per-sample loop of calls of cluster.leiden (cluster.py:36-54) that
 * a batched caller would make.  Synchronises. */
int msc_leiden_batched_workspace_size(int64_t n, int64_t nnz, int n_samples, size_t* bytes /*host*/);
int msc_leiden_batched(const int64_t* indptr, const int32_t* indices, const float* weights, int64_t n,
                       int64_t nnz, const int64_t* offsets /*host*/, int n_samples, double resolution,
                       int64_t seed, int n_iterations, double commit_prob, int32_t* out_labels,
                       double* out_modularity /*host, n_samples*/, void* workspace, size_t workspace_bytes,
                       void* stream);

/* Synchronous parallel multilevel Louvain (the same move kernels and contraction, no
 * refinement).  Replaces tl.louvain(backend="gpu") -> graph.louvain.louvain
 * (src/metalsinglecell/tl.py:76-91, graph/louvain.py:397-453) with variant="sync".
 * Level l moves from singletons with seed + 100*l; the run stops, keeping the previous
 * level's labels, when no vertex moved or Q rose by at most 1e-9; otherwise the graph is
 * aggregated on the moved partition.  1 <= max_levels <= 20.  Input, output, fixed-point
 * arithmetic, info records (one per level; n_refined = n_comms, refine_passes = 0) and the
 * workspace size (= msc_leiden_workspace_size) as for msc_leiden.  Synchronises. */
int msc_louvain_workspace_size(int64_t n, int64_t nnz, size_t* bytes /*host*/);
int msc_louvain(const int64_t* indptr, const int32_t* indices, const float* weights, int64_t n,
                int64_t nnz, double resolution, int64_t seed, int max_levels, double commit_prob,
                int32_t* out_labels, double* out_modularity /*host*/,
                msc_leiden_info* info /*host, may be NULL*/, void* workspace,
                size_t workspace_bytes, void* stream);

/* ---------------------------------------------------------------------------------------
 * (b) Exact spatial kNN on a uniform grid, self excluded (2-D or 3-D coordinates).
 * Replaces neighbors._knn_grid and its grid_knn kernel (src/metalsinglecell/neighbors.py:
 * 333-417, 143-196) as used by gr._knn_edges / gr.spatial_neighbors_knn (gr.py:171-225).
 * coords: n x dim float32.  out_idx / out_dist: n x k, ordered by (float32 distance, index).
 * coords64 / out_len64 (optional, together): n x dim float64 coordinates -> n x k float64
 * edge lengths (obsp['spatial_distances'] values, gr.py:32-34).  1 <= k <= 511, k < n,
 * dim >= 1.  k > 32 or dim outside {2, 3} compares every pair exactly (the reference's
 * brute-force fallback, neighbors.py:350-355), with the same order and distances. */
typedef struct {
  int32_t max_shell;     /* out: largest Chebyshev cell shell any query needed */
  int32_t grid_x, grid_y, grid_z;
  float cell;
} msc_spatial_info;

int msc_spatial_knn_workspace_size(int64_t n, int dim, int k, size_t* bytes /*host*/);
int msc_spatial_knn(const float* coords, int64_t n, int dim, int k, int32_t* out_idx,
                    float* out_dist, const double* coords64, double* out_len64,
                    msc_spatial_info* info /*host, may be NULL*/, void* workspace,
                    size_t workspace_bytes, void* stream);

/* Fixed-radius spatial neighbours, self excluded, as a canonical CSR (columns ascending).
 * Replaces neighbors._radius_grid and its grid_radius kernel (src/metalsinglecell/
 * neighbors.py:255-330, 199-252) as used by gr.spatial_neighbors_radius (gr.py:228-252).
 * Pair (i, j) is kept when r_min^2 <= d2 <= r_max^2 with d2 the float32 squared distance of
 * msc_spatial_knn and the bounds rounded to float32 from float64 squares.  Two calls:
 * _count writes out_indptr (n+1 int64) and *out_nnz (host; synchronises); _fill (same
 * coords, radii and indptr) writes out_indices (nnz int32), optionally out_dist (nnz
 * float32, sqrt(d2)) and, with coords64 (n x dim float64), out_len64 (nnz float64 edge
 * lengths).  Workspace: msc_spatial_radius_workspace_size(n, dim, nnz) — nnz = 0 suffices
 * for _count, the fill needs the pair count. */
int msc_spatial_radius_workspace_size(int64_t n, int dim, int64_t nnz, size_t* bytes /*host*/);
int msc_spatial_radius_count(const float* coords, int64_t n, int dim, double r_min, double r_max,
                             int64_t* out_indptr, int64_t* out_nnz /*host*/, void* workspace,
                             size_t workspace_bytes, void* stream);
int msc_spatial_radius_fill(const float* coords, int64_t n, int dim, double r_min, double r_max,
                            const int64_t* indptr, int64_t nnz, int32_t* out_indices, float* out_dist,
                            const double* coords64, double* out_len64, void* workspace,
                            size_t workspace_bytes, void* stream);

/* ---------------------------------------------------------------------------------------
 * (e) Moran's I with a permutation null (float64 reductions).
 * Replaces the Moran branch of spatial.spatial_autocorr (src/metalsinglecell/spatial.py:
 * 357-406) including row normalisation (spatial.py:268-279) when transformation != 0.
 * X: n x n_genes float32 row-major (cells x genes).  W: CSR (w_indptr n+1 int64, w_indices
 * nnz int32, w_data nnz float64).  n_perms permutations come from `perms` (n_perms x n
 * int32 device array, e.g. numpy default_rng(seed).permutation(n) drawn in sequence) or,
 * when perms is NULL, from a device Feistel bijection keyed by `seed`.
 * out_I: n_genes float64; out_sims: n_perms x n_genes float64; *out_S0 (host, may be NULL):
 * the weight sum after normalisation. */
int msc_moran_workspace_size(int64_t n, int64_t n_genes, int64_t nnz, int n_perms,
                             size_t* bytes /*host*/);
int msc_moran(const float* X, int64_t n, int64_t n_genes, const int64_t* w_indptr,
              const int32_t* w_indices, const double* w_data, int64_t nnz, int transformation,
              int n_perms, const int32_t* perms, int64_t seed, double* out_I, double* out_sims,
              double* out_S0 /*host*/, void* workspace, size_t workspace_bytes, void* stream);

/* Geary's C with the same permutation null: the Geary branch of spatial.spatial_autocorr
 * (src/metalsinglecell/spatial.py:388-393, 396-406) as called by gr.spatial_autocorr
 * (gr.py:455-503, mode="geary").  C_g = (n-1)/(2 S0) sum_e w_e (x_src - x_dst)^2 / denom_g,
 * the observed numerator summed directly over the edges in float64; each permuted numerator
 * (sources pi[src]) by its exact per-row expansion in float64.  Arguments as msc_moran;
 * out_C: n_genes float64. */
int msc_geary_workspace_size(int64_t n, int64_t n_genes, int64_t nnz, int n_perms,
                             size_t* bytes /*host*/);
int msc_geary(const float* X, int64_t n, int64_t n_genes, const int64_t* w_indptr,
              const int32_t* w_indices, const double* w_data, int64_t nnz, int transformation,
              int n_perms, const int32_t* perms, int64_t seed, double* out_C, double* out_sims,
              double* out_S0 /*host*/, void* workspace, size_t workspace_bytes, void* stream);

/* Cluster co-occurrence pair histogram: the fused kernel of spatial.co_occurrence
 * (src/metalsinglecell/spatial.py:123-171) as called by gr.co_occurrence (gr.py:514-530).
 * out_hist[c_i][c_j][r] (device int64, n_codes^2 x n_bins) counts the ordered pairs (i, j),
 * i != j, whose squared distance d2 has r = the first bin with thr2[r] >= d2; pairs beyond
 * thr2[n_bins-1] are not counted.  d2 is the float32 squared distance of msc_spatial_knn.
 * coords: n x dim float32 (dim 2 or 3); codes (host): n values in [0, n_codes); thr2 (host):
 * n_bins ascending float32.  n < 2^24; the LDS histogram needs n_codes x n_bins <= ~15000. */
int msc_co_occurrence_workspace_size(const int32_t* codes /*host*/, int64_t n, int dim, int n_codes,
                                     int n_bins, size_t* bytes /*host*/);
int msc_co_occurrence(const float* coords, int64_t n, int dim, const int32_t* codes /*host*/,
                      int n_codes, const float* thr2 /*host*/, int n_bins, int64_t* out_hist,
                      void* workspace, size_t workspace_bytes, void* stream);

/* Dense n x n_cols float32 row-major copy of a CSR matrix (duplicates: last write wins; a
 * canonical CSR has none).  Moran densifies the selected genes (gr.py:452). */
int msc_csr_to_dense(const int64_t* indptr, const int32_t* indices, const float* data, int64_t n,
                     int64_t n_cols, float* out, void* stream);

/* pysal weight moments of W (spatial.py:282-290): out_s[0] = s0 = sum w,
 * out_s[1] = s1 = 1/2 sum (W + W^T)^2, out_s[2] = s2 = sum_i (row_i + col_i)^2 (host). */
int msc_graph_moments_workspace_size(int64_t n, int64_t nnz, size_t* bytes /*host*/);
int msc_graph_moments(const int64_t* indptr, const int32_t* indices, const double* w, int64_t n,
                      int64_t nnz, double* out_s /*host, 3*/, void* workspace,
                      size_t workspace_bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MSC_HIP_H_ */

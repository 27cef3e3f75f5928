/* Exact Euclidean kNN oracle in C — TEST INFRASTRUCTURE ONLY.
 *
 * Same definition as oracle/knn.py (restating neighbors.py:420-462 and :82-132 of the
 * reference): squared distance = float32 sum over t in order of (x_t - y_t)^2, each
 * operation separately rounded (built with -ffp-contract=off); key (j != i, d2, j); self
 * first.  One bounded max-heap of k keys per query; OpenMP over queries.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { uint32_t flag; float d2; int64_t j; } okey_t;

static int key_lt(const okey_t *a, const okey_t *b) {
  if (a->flag != b->flag) return a->flag < b->flag;
  if (a->d2 != b->d2) return a->d2 < b->d2;
  return a->j < b->j;
}

static void sift_down(okey_t *h, int k, int i) {  /* max-heap on key order */
  for (;;) {
    int l = 2 * i + 1, r = l + 1, m = i;
    if (l < k && key_lt(&h[m], &h[l])) m = l;
    if (r < k && key_lt(&h[m], &h[r])) m = r;
    if (m == i) return;
    okey_t t = h[i]; h[i] = h[m]; h[m] = t; i = m;
  }
}

static int cmp_key(const void *a, const void *b) {
  const okey_t *x = (const okey_t *)a, *y = (const okey_t *)b;
  if (key_lt(x, y)) return -1;
  if (key_lt(y, x)) return 1;
  return 0;
}

int oracle_knn(const float *X, int64_t n, int d, int k, const int64_t *rows, int64_t m,
               int64_t *out_idx, float *out_d2, int nthreads) {
  if (k < 1 || k > n || d < 1) return 1;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  int err = 0;
#pragma omp parallel
  {
    okey_t *heap = (okey_t *)malloc(sizeof(okey_t) * (size_t)k);
    if (!heap) {
#pragma omp atomic write
      err = 2;
    }
#pragma omp for schedule(dynamic, 16)
    for (int64_t r = 0; r < m; ++r) {
      if (!heap) continue;
      const int64_t i = rows[r];
      const float *q = X + (size_t)i * d;
      int cnt = 0;
      for (int64_t j = 0; j < n; ++j) {
        const float *y = X + (size_t)j * d;
        float acc = 0.0f;
        for (int t = 0; t < d; ++t) {
          float diff = q[t] - y[t];
          float sq = diff * diff;
          acc = acc + sq;
        }
        okey_t kk = {(uint32_t)(j != i), acc, j};
        if (cnt < k) {
          heap[cnt++] = kk;
          if (cnt == k)
            for (int s = k / 2 - 1; s >= 0; --s) sift_down(heap, k, s);
        } else if (key_lt(&kk, &heap[0])) {
          heap[0] = kk;
          sift_down(heap, k, 0);
        }
      }
      qsort(heap, (size_t)k, sizeof(okey_t), cmp_key);
      for (int s = 0; s < k; ++s) {
        out_idx[r * k + s] = heap[s].j;
        out_d2[r * k + s] = heap[s].d2;
      }
    }
    free(heap);
  }
  return err;
}

/* Exact 2-D/3-D spatial kNN (self excluded), key (sqrt(d2), j), as oracle/grid.py. */
int oracle_spatial_knn(const float *X, int64_t n, int d, int k, int64_t *out_idx,
                       float *out_dist, int nthreads) {
  if (k < 1 || k >= n || (d != 2 && d != 3)) return 1;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  int err = 0;
#pragma omp parallel
  {
    okey_t *heap = (okey_t *)malloc(sizeof(okey_t) * (size_t)k);
    if (!heap) {
#pragma omp atomic write
      err = 2;
    }
#pragma omp for schedule(dynamic, 64)
    for (int64_t i = 0; i < n; ++i) {
      if (!heap) continue;
      const float *q = X + (size_t)i * d;
      int cnt = 0;
      for (int64_t j = 0; j < n; ++j) {
        if (j == i) continue;
        const float *y = X + (size_t)j * d;
        float dx = y[0] - q[0], dy = y[1] - q[1];
        float acc = dx * dx;
        acc = acc + dy * dy;
        if (d == 3) { float dz = y[2] - q[2]; acc = acc + dz * dz; }
        okey_t kk = {0u, sqrtf(acc), j};
        if (cnt < k) {
          heap[cnt++] = kk;
          if (cnt == k)
            for (int s = k / 2 - 1; s >= 0; --s) sift_down(heap, k, s);
        } else if (key_lt(&kk, &heap[0])) {
          heap[0] = kk;
          sift_down(heap, k, 0);
        }
      }
      qsort(heap, (size_t)k, sizeof(okey_t), cmp_key);
      for (int s = 0; s < k; ++s) {
        out_idx[i * k + s] = heap[s].j;
        out_dist[i * k + s] = heap[s].d2;
      }
    }
    free(heap);
  }
  return err;
}

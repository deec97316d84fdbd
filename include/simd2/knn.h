/*
 * simd2 -- k-nearest neighbours by pairwise squared L2 (the minus_square
 * semiring): C ABI of libsimd2.so.
 */
#ifndef SIMD2_KNN_H
#define SIMD2_KNN_H

#include <stdbool.h>
#include <stddef.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { SIMD2_KNN_SQRT = 1 /* return distances, not squared distances */ };

/* replaces modified_insertion_sort_linear + compute_sqrt_linear,
 * apps/pld/cuASR/knncuda.cu:170-237.  dist is the reference's Q x R layout
 * (dist[q + r * query_nb], the column-major D of cuasr_plusmiussquare_srsgemm)
 * and is not modified; knn_dist / knn_index receive, per query q and rank j,
 * the k smallest distances in ascending order with ties in index order at
 * [j * query_nb + q] (square-rooted when flags has SIMD2_KNN_SQRT, as
 * compute_sqrt_linear does).  1 <= k <= 32, k <= ref_nb. */
int simd2_knn_select(const float *dist, int query_nb, int ref_nb, int k, float *knn_dist,
                     int *knn_index, int flags, hipStream_t stream);

/* replaces knn_mxu, apps/pld/cuASR/knncuda.cu:399-549: host arrays ref
 * (ref_nb x dim, point-major) and query (dim x query_nb, dimension-major) as
 * that function takes them; computes the Q x R distances with
 * cuasr_plusmiussquare_srsgemm, selects the k nearest and returns their
 * distances (square-rooted) and indices in host arrays laid out [j*query_nb+q].
 * Allocates and frees its device buffers as the reference does; *kernel_time
 * accumulates the seconds spent in the distance kernel.  Returns false on error. */
bool knn_mxu(const float *ref, int ref_nb, const float *query, int query_nb, int dim, int k,
             float *knn_dist, int *knn_index, double *kernel_time);

/* Precision form (no counterpart in the reference): device arrays query
 * (query_nb x dim, leading dimension ldq) and ref (ref_nb x dim, ldr), FP32.
 * Candidates from ||q||^2 + ||r||^2 - 2 q.r with the dot product on BF16 MFMA
 * (FP32 accumulate, the Q x R matrix is never stored), the best 32 per query
 * re-ranked by the exact FP32 sum of (q - r)^2, FP64 for near ties (relative
 * gap <= 1e-5 at or above rank k).  Outputs row-major query_nb x k (distances
 * ascending, ties in index order; square-rooted with SIMD2_KNN_SQRT).
 * stats (optional, 2 device ints, accumulated): queries that needed FP64, and
 * candidates recomputed in FP64.  dim <= 160, 1 <= k <= 31, k <= ref_nb.
 * workspace must hold simd2_knn_bf16_workspace_size(query_nb, ref_nb, dim) bytes. */
size_t simd2_knn_bf16_workspace_size(int query_nb, int ref_nb, int dim);
int simd2_knn_bf16(const float *query, int ldq, const float *ref, int ldr, int query_nb,
                   int ref_nb, int dim, int k, float *knn_dist, int *knn_index, int flags,
                   void *workspace, size_t workspace_bytes, int *stats, hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* SIMD2_KNN_H */

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
 * compute_sqrt_linear does).  Any 1 <= k <= ref_nb.  NaN distances are
 * outside the contract.  workspace (device memory, kept by the caller) must
 * hold simd2_knn_select_workspace_size(query_nb, ref_nb, k) bytes (at most
 * about 256 MiB: queries are processed in batches); calls sharing a
 * workspace must be ordered. */
size_t simd2_knn_select_workspace_size(int query_nb, int ref_nb, int k);
int simd2_knn_select(const float *dist, int query_nb, int ref_nb, int k, float *knn_dist,
                     int *knn_index, int flags, void *workspace, size_t workspace_bytes,
                     hipStream_t stream);

/* replaces knn_mxu, apps/pld/cuASR/knncuda.cu:399-549: host arrays ref
 * (ref_nb x dim, point-major) and query (dim x query_nb, dimension-major) as
 * that function takes them; computes the Q x R distances with
 * cuasr_plusmiussquare_srsgemm, selects the k nearest and returns their
 * distances (square-rooted) and indices in host arrays laid out [j*query_nb+q];
 * any 1 <= k <= ref_nb.  Allocates and frees its device buffers as the
 * reference does; *kernel_time
 * accumulates the seconds spent in the distance kernel.  Returns false on error. */
bool knn_mxu(const float *ref, int ref_nb, const float *query, int query_nb, int dim, int k,
             float *knn_dist, int *knn_index, double *kernel_time);

/* Precision form (no counterpart in the reference): device arrays query
 * (query_nb x dim, leading dimension ldq) and ref (ref_nb x dim, ldr), FP32,
 * any dim.  Candidates from ||q - mu||^2 + ||r - mu||^2 - 2 (q - mu).(r - mu)
 * (mu = the references' mean) with the dot product on BF16 MFMA (hi/lo
 * split, FP32 accumulate; the Q x R matrix is never stored), kept wherever
 * that approximation, widened by a bound on its error, can reach the k
 * nearest; every candidate is then ranked by its exact FP32 squared distance,
 * and candidates whose FP32 distances are too close to order (near ties
 * reaching rank k) by their FP64 distance.  A query with more than 1024
 * candidates (thousands of equidistant points) is answered from the exact
 * FP32 distances to every reference.  Outputs row-major query_nb x k
 * (distances ascending, ties in index order; square-rooted with
 * SIMD2_KNN_SQRT).  stats (optional, 3 device ints, accumulated): queries
 * that needed FP64, candidates recomputed in FP64, queries answered by the
 * exact scan.  1 <= k <= 256, k <= ref_nb.  workspace must hold
 * simd2_knn_bf16_workspace_size(query_nb, ref_nb, dim) bytes. */
size_t simd2_knn_bf16_workspace_size(int query_nb, int ref_nb, int dim);
int simd2_knn_bf16(const float *query, int ldq, const float *ref, int ldr, int query_nb,
                   int ref_nb, int dim, int k, float *knn_dist, int *knn_index, int flags,
                   void *workspace, size_t workspace_bytes, int *stats, hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* SIMD2_KNN_H */

// C ABI of the k-nearest-neighbour entry points (declared in include/simd2/knn.h).
#include <hip/hip_runtime.h>

#include <chrono>

#include "../../include/simd2/knn.h"
#include "../../include/simd2/srgemm.h"
#include "internal.hpp"

using namespace simd2;

namespace {
int status_of(hipError_t e) {
  if (e == hipSuccess) return SIMD2_SUCCESS;
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return SIMD2_ERROR_NO_DEVICE;
  return SIMD2_ERROR_INTERNAL;
}
}  // namespace

extern "C" {

size_t simd2_knn_select_workspace_size(int query_nb, int ref_nb, int k) {
  return knn_select_workspace(query_nb, ref_nb, k);
}

int simd2_knn_select(const float *dist, int query_nb, int ref_nb, int k, float *knn_dist,
                     int *knn_index, int flags, void *workspace, size_t workspace_bytes,
                     hipStream_t stream) {
  if (query_nb < 0 || ref_nb < 0 || k < 1 || k > ref_nb) return SIMD2_ERROR_INVALID_PROBLEM;
  if (query_nb == 0) return SIMD2_SUCCESS;
  if (!dist || !knn_dist || !knn_index) return SIMD2_ERROR_NULL_POINTER;
  if (!workspace || workspace_bytes < knn_select_workspace(query_nb, ref_nb, k))
    return SIMD2_ERROR_WORKSPACE;
  if (!workspace_aligned(workspace)) return SIMD2_ERROR_MISALIGNED_OPERAND;
  return status_of(launch_knn_select(dist, query_nb, ref_nb, k, knn_dist, knn_index,
                                     (flags & SIMD2_KNN_SQRT) ? 1 : 0, workspace, stream));
}

size_t simd2_knn_bf16_workspace_size(int query_nb, int ref_nb, int dim) {
  if (query_nb <= 0 || ref_nb <= 0 || dim <= 0) return 0;
  return knn_bf16_workspace(query_nb, ref_nb, dim);
}

int simd2_knn_bf16(const float *query, int ldq, const float *ref, int ldr, int query_nb,
                   int ref_nb, int dim, int k, float *knn_dist, int *knn_index, int flags,
                   void *workspace, size_t workspace_bytes, int *stats, hipStream_t stream) {
  if (query_nb < 0 || ref_nb < 0 || dim < 1 || k < 1 || k > knn_bf16_kmax() || k > ref_nb)
    return SIMD2_ERROR_INVALID_PROBLEM;
  if (query_nb == 0) return SIMD2_SUCCESS;
  if (ldq < dim || ldr < dim) return SIMD2_ERROR_INVALID_PROBLEM;
  if (!query || !ref || !knn_dist || !knn_index) return SIMD2_ERROR_NULL_POINTER;
  if (!workspace || workspace_bytes < knn_bf16_workspace(query_nb, ref_nb, dim))
    return SIMD2_ERROR_WORKSPACE;
  if (!workspace_aligned(workspace)) return SIMD2_ERROR_MISALIGNED_OPERAND;
  return status_of(launch_knn_bf16(query, ldq, ref, ldr, query_nb, ref_nb, dim, k, knn_dist,
                                   knn_index, (flags & SIMD2_KNN_SQRT) ? 1 : 0, workspace, stats,
                                   stream));
}

bool knn_mxu(const float *ref, int ref_nb, const float *query, int query_nb, int dim, int k,
             float *knn_dist, int *knn_index, double *kernel_time) {
  if (!ref || !query || !knn_dist || !knn_index || ref_nb <= 0 || query_nb <= 0 || dim <= 0 ||
      k < 1 || k > ref_nb)
    return false;
  float *ref_d = nullptr, *query_d = nullptr, *dist_d = nullptr, *kd = nullptr;
  int *ki = nullptr;
  void *ws = nullptr;
  const size_t ws_bytes = knn_select_workspace(query_nb, ref_nb, k);
  bool ok = hipMalloc(&ws, ws_bytes) == hipSuccess &&
            hipMalloc((void **)&ref_d, (size_t)ref_nb * dim * sizeof(float)) == hipSuccess &&
            hipMalloc((void **)&query_d, (size_t)query_nb * dim * sizeof(float)) == hipSuccess &&
            hipMalloc((void **)&dist_d, (size_t)query_nb * ref_nb * sizeof(float)) == hipSuccess &&
            hipMalloc((void **)&kd, (size_t)query_nb * k * sizeof(float)) == hipSuccess &&
            hipMalloc((void **)&ki, (size_t)query_nb * k * sizeof(int)) == hipSuccess;
  ok = ok && hipMemcpy(ref_d, ref, (size_t)ref_nb * dim * sizeof(float),
                       hipMemcpyHostToDevice) == hipSuccess;
  ok = ok && hipMemcpy(query_d, query, (size_t)query_nb * dim * sizeof(float),
                       hipMemcpyHostToDevice) == hipSuccess;
  if (ok) {
    const auto t0 = std::chrono::steady_clock::now();
    // knncuda.cu:479-484: D(q, r) = sum_d (query(q, d) - ref(d, r))^2, column-major.
    ok = cuasr_plusmiussquare_srsgemm(query_nb, ref_nb, dim, query_d, query_nb, ref_d, dim,
                                      dist_d, query_nb, dist_d, false, nullptr) == SIMD2_SUCCESS;
    ok = ok && hipDeviceSynchronize() == hipSuccess;
    const auto t1 = std::chrono::steady_clock::now();
    if (kernel_time) *kernel_time += std::chrono::duration<double>(t1 - t0).count();
  }
  ok = ok && simd2_knn_select(dist_d, query_nb, ref_nb, k, kd, ki, SIMD2_KNN_SQRT, ws, ws_bytes,
                              nullptr) == SIMD2_SUCCESS;
  ok = ok && hipMemcpy(knn_dist, kd, (size_t)query_nb * k * sizeof(float),
                       hipMemcpyDeviceToHost) == hipSuccess;
  ok = ok && hipMemcpy(knn_index, ki, (size_t)query_nb * k * sizeof(int),
                       hipMemcpyDeviceToHost) == hipSuccess;
  (void)hipFree(ref_d);
  (void)hipFree(query_d);
  (void)hipFree(dist_d);
  (void)hipFree(kd);
  (void)hipFree(ki);
  (void)hipFree(ws);
  return ok;
}

}  // extern "C"

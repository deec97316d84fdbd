// C ABI of the learned schedule: features, vertex order, the MLP's forward /
// gradient / Adam entry points, and the learned converge loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <cstdint>
#include <numeric>
#include <vector>

#include "../../include/simd2/converge.h"
#include "../../include/simd2/learned.h"
#include "../../include/simd2/schedule.h"
#include "../../include/simd2/srgemm.h"
#include "internal.hpp"

using namespace simd2;

namespace {

constexpr int MAX_N = 32768;

int hip_status(hipError_t e) {
  if (e == hipSuccess) return SIMD2_SUCCESS;
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return SIMD2_ERROR_NO_DEVICE;
  return SIMD2_ERROR_INTERNAL;
}

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

bool op_ok(int op) {
  return op == SIMD2_MIN_PLUS || op == SIMD2_MAX_PLUS || op == SIMD2_MAX_MIN ||
         op == SIMD2_MIN_MAX;
}

// Features workspace: graph_features_workspace(n) plus n ints of levels.
size_t features_bytes(int n) { return align256(graph_features_workspace(n)); }

// Levels of the pseudo-peripheral breadth-first search -> vertex order
// (unreached vertices last), computed with the features workspace.
int order_into(int n, const float *X, int ldx, int *perm_dev, void *ws, bool after_features,
               hipStream_t s, std::vector<int> *perm_host) {
  char *w = static_cast<char *>(ws);
  int *level = reinterpret_cast<int *>(w + features_bytes(n));
  hipError_t e = launch_graph_levels(n, X, ldx, level, ws, after_features, s);
  std::vector<int> lv(n);
  if (e == hipSuccess)
    e = hipMemcpyAsync(lv.data(), level, n * sizeof(int), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return hip_status(e);
  std::vector<int> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  auto key = [&](int v) { return lv[v] < 0 ? MAX_N + 1 : lv[v]; };
  std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return key(a) < key(b); });
  e = hipMemcpy(perm_dev, perm.data(), n * sizeof(int), hipMemcpyHostToDevice);
  if (perm_host) *perm_host = perm;
  return hip_status(e);
}

constexpr int CFG_OF_CLASS[3] = {CFG_S64, CFG_M128x64, CFG_L128};  // by tile size
// The shipped model's training graphs have n in [256, 4096].  Above that its
// tile class is an extrapolation (at n = 8192 it chose 64x64 tiles for grids,
// 4% slower than the size rule's 128x128), so larger graphs take the size
// rule's tile and only the squaring count from the network.
constexpr int TILE_TRAINED_MAX_N = 4096;

}  // namespace

extern "C" {

size_t simd2_graph_features_workspace(int n) {
  if (n <= 0 || n > MAX_N) return 0;
  return features_bytes(n) + align256((size_t)n * sizeof(int));
}

int simd2_graph_features(int op, int n, const float *X, int ldx, int do_epilogue, float *features,
                         void *workspace, size_t workspace_bytes, hipStream_t stream) {
  if (!op_ok(op)) return SIMD2_ERROR_NOT_SUPPORTED;
  if (n <= 0 || n > MAX_N || ldx < n) return SIMD2_ERROR_INVALID_PROBLEM;
  if (!X || !features || !workspace) return SIMD2_ERROR_NULL_POINTER;
  if (workspace_bytes < simd2_graph_features_workspace(n)) return SIMD2_ERROR_WORKSPACE;
  if (!workspace_aligned(workspace)) return SIMD2_ERROR_MISALIGNED_OPERAND;
  return hip_status(launch_graph_features(op, n, X, ldx, do_epilogue, features, workspace, stream));
}

int simd2_graph_order(int op, int n, const float *X, int ldx, int *perm, void *workspace,
                      size_t workspace_bytes, hipStream_t stream) {
  if (!op_ok(op)) return SIMD2_ERROR_NOT_SUPPORTED;
  if (n <= 0 || n > MAX_N || ldx < n) return SIMD2_ERROR_INVALID_PROBLEM;
  if (!X || !perm || !workspace) return SIMD2_ERROR_NULL_POINTER;
  if (workspace_bytes < simd2_graph_features_workspace(n)) return SIMD2_ERROR_WORKSPACE;
  if (!workspace_aligned(workspace)) return SIMD2_ERROR_MISALIGNED_OPERAND;
  return order_into(n, X, ldx, perm, workspace, false, stream, nullptr);
}

int simd2_mlp_forward(const float *params, const float *x, int B, float *y, hipStream_t stream) {
  if (B < 0) return SIMD2_ERROR_INVALID_PROBLEM;
  if (B == 0) return SIMD2_SUCCESS;
  if (!params || !x || !y) return SIMD2_ERROR_NULL_POINTER;
  return hip_status(launch_mlp_forward(params, x, B, y, nullptr, stream));
}

size_t simd2_mlp_workspace(int B) {
  if (B <= 0) return 0;
  return align256(mlp_workspace(B)) + align256((size_t)SIMD2_MLP_PARAMS * sizeof(float));
}

int simd2_mlp_gradient(const float *params, const float *x, const float *t, const float *w, int B,
                       float *grad, float *loss, void *workspace, size_t workspace_bytes,
                       hipStream_t stream) {
  if (B <= 0) return SIMD2_ERROR_INVALID_PROBLEM;
  if (!params || !x || !t || !w || !workspace) return SIMD2_ERROR_NULL_POINTER;
  if (workspace_bytes < simd2_mlp_workspace(B)) return SIMD2_ERROR_WORKSPACE;
  if (!workspace_aligned(workspace)) return SIMD2_ERROR_MISALIGNED_OPERAND;
  return hip_status(launch_mlp_gradient(params, x, t, w, B, grad, loss, workspace, stream));
}

int simd2_mlp_train_step(float *params, float *m, float *v, int step, float lr, const float *x,
                         const float *t, const float *w, int B, float *loss, void *workspace,
                         size_t workspace_bytes, hipStream_t stream) {
  if (B <= 0 || step < 1) return SIMD2_ERROR_INVALID_PROBLEM;
  if (!params || !m || !v || !x || !t || !w || !workspace) return SIMD2_ERROR_NULL_POINTER;
  if (workspace_bytes < simd2_mlp_workspace(B)) return SIMD2_ERROR_WORKSPACE;
  if (!workspace_aligned(workspace)) return SIMD2_ERROR_MISALIGNED_OPERAND;
  float *grad = reinterpret_cast<float *>(static_cast<char *>(workspace) +
                                          align256(mlp_workspace(B)));
  hipError_t e = launch_mlp_gradient(params, x, t, w, B, grad, loss, workspace, stream);
  const float c1 = 1.0f / (1.0f - std::pow(0.9f, (float)step));
  const float c2 = 1.0f / (1.0f - std::pow(0.999f, (float)step));
  if (e == hipSuccess) e = launch_mlp_adam(params, m, v, grad, lr, c1, c2, stream);
  return hip_status(e);
}

size_t simd2_converge_learned_workspace(int n, int max_itr) {
  if (n <= 0 || n > MAX_N || max_itr < 0) return 0;
  size_t sched = 0;
  for (int c = 0; c < 3; ++c)
    sched = std::max(sched, simd2_converge_scheduled_workspace(n, c, max_itr));
  return simd2_graph_features_workspace(n) + align256(SIMD2_FEATURES * sizeof(float)) +
         align256(SIMD2_MLP_OUT * sizeof(float)) + align256((size_t)n * sizeof(int)) + sched;
}

int simd2_converge_learned(int op, int n, float *X, int ldx, int do_epilogue, int check_mode,
                           int max_itr, const float *model, void *workspace,
                           size_t workspace_bytes, hipStream_t stream, int *num_itr,
                           simd2_schedule *used, simd2_sched_report *report) {
  if (num_itr) *num_itr = 0;
  if (!op_ok(op)) return SIMD2_ERROR_NOT_SUPPORTED;
  if (check_mode < 0 || check_mode > 2) return SIMD2_ERROR_NOT_SUPPORTED;
  if (n < 0 || max_itr < 0) return SIMD2_ERROR_INVALID_PROBLEM;
  if (n == 0) return SIMD2_SUCCESS;
  if (n > MAX_N || ldx < n) return SIMD2_ERROR_INVALID_PROBLEM;
  if (!X || !model || !workspace) return SIMD2_ERROR_NULL_POINTER;
  if (workspace_bytes < simd2_converge_learned_workspace(n, max_itr))
    return SIMD2_ERROR_WORKSPACE;
  if (!workspace_aligned(workspace)) return SIMD2_ERROR_MISALIGNED_OPERAND;
  char *ws = static_cast<char *>(workspace);
  const size_t fbytes = simd2_graph_features_workspace(n);
  float *feat = reinterpret_cast<float *>(ws + fbytes);
  float *pred = reinterpret_cast<float *>(ws + fbytes + align256(SIMD2_FEATURES * 4));
  int *perm = reinterpret_cast<int *>(ws + fbytes + align256(SIMD2_FEATURES * 4) +
                                      align256(SIMD2_MLP_OUT * 4));
  char *sched_ws = reinterpret_cast<char *>(perm) + align256((size_t)n * sizeof(int));
  const size_t sched_bytes = workspace_bytes - (size_t)(sched_ws - ws);

  // features and the network's prediction in one launch sequence
  hipError_t e = launch_graph_features(op, n, X, ldx, do_epilogue, feat, ws, stream, model, pred);
  // features and predictions in one copy: pred follows feat in the workspace
  constexpr size_t pred_at = (SIMD2_FEATURES * 4 + 255) / 256 * 256;
  float hostbuf[pred_at / 4 + SIMD2_MLP_OUT];
  if (e == hipSuccess)
    e = hipMemcpyAsync(hostbuf, feat, pred_at + SIMD2_MLP_OUT * sizeof(float),
                       hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) return hip_status(e);
  float host[SIMD2_MLP_OUT + SIMD2_FEATURES];
  std::memcpy(host, hostbuf + pred_at / 4, SIMD2_MLP_OUT * sizeof(float));
  std::memcpy(host + SIMD2_MLP_OUT, hostbuf, SIMD2_FEATURES * sizeof(float));

  // Schedule from the prediction.  Enqueueing a squaring too many costs a
  // launch (it returns at once); one too few costs a flag read per extra
  // squaring, hence the rounding up.
  simd2_schedule sc{-1, 0, -1, -1};
  const float itr_pred = std::isfinite(host[0]) ? host[0] : 1.0f;
  sc.predicted_itr = std::max(1, std::min(max_itr, (int)std::ceil(itr_pred + 0.25f)));
  const float cls = std::isfinite(host[1]) ? host[1] : 2.0f;
  sc.tile_config = n > TILE_TRAINED_MAX_N
                       ? valu_pick_cfg(op, n, n, n)
                       : CFG_OF_CLASS[std::max(0, std::min(2, (int)std::lround(cls)))];
  // Vertex order: breadth-first levels when the given numbering has no
  // locality (mean |i - j| over edges above n / 20), one search reaches most
  // of the graph, and the graph is "long" (mean search depth above log2 n): a
  // small-world graph (Erdos-Renyi) has no order with locality, and a graph
  // already numbered with locality keeps it.
  const float *f = host + SIMD2_MLP_OUT;
  const double depth = std::exp2(16.0 * f[17]) - 1.0;
  const bool reorder = f[14] > 0.05f && f[19] > 0.5f && depth > std::log2((double)n);
  int st = SIMD2_SUCCESS;
  if (reorder) st = order_into(n, X, ldx, perm, ws, true, stream, nullptr);
  if (st != SIMD2_SUCCESS) return st;
  if (used) *used = sc;
  st = simd2_converge_scheduled(op, n, X, ldx, do_epilogue, check_mode, max_itr, &sc,
                                reorder ? perm : nullptr, sched_ws, sched_bytes, stream,
                                num_itr, report);
  if (report) report->host_syncs += reorder ? 2 : 1;
  return st;
}

}  // extern "C"

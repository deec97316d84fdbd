// C ABI of libsimd2.so: argument checks, the column-major -> row-major
// mapping, the reference-surface drop-ins and the repeated-squaring drivers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <utility>

#include "../../include/simd2/converge.h"
#include "../../include/simd2/srgemm.h"
#include "../../include/simd2/tensor_emu.h"
#include "internal.hpp"

using namespace simd2;

namespace {

constexpr int OK_NOTHING_TO_DO = -1;  // internal: valid call with an empty output

int hip_status(hipError_t e) {
  if (e == hipSuccess) return SIMD2_SUCCESS;
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return SIMD2_ERROR_NO_DEVICE;
  return SIMD2_ERROR_INTERNAL;
}

// Byte range [lo, hi) of a column-major rows x cols matrix with leading dim ld.
std::pair<uintptr_t, uintptr_t> span(const void *p, int rows, int cols, int ld, size_t es) {
  const uintptr_t lo = reinterpret_cast<uintptr_t>(p);
  if (rows <= 0 || cols <= 0) return {lo, lo};
  return {lo, lo + ((uint64_t)(cols - 1) * ld + rows) * es};
}

bool overlap(std::pair<uintptr_t, uintptr_t> a, std::pair<uintptr_t, uintptr_t> b) {
  return a.first < b.second && b.first < a.second;
}

// Checks of a column-major call: A MxK (lda), B KxN (ldb), C, D MxN.
int check_cm(int M, int N, int K, const void *A, int lda, const void *B, int ldb,
             const void *C, int ldc, const void *D, int ldd, bool epi, size_t es) {
  if (M < 0 || N < 0 || K < 0) return SIMD2_ERROR_INVALID_PROBLEM;
  if (M == 0 || N == 0) return OK_NOTHING_TO_DO;
  if (ldd < std::max(1, M)) return SIMD2_ERROR_INVALID_PROBLEM;
  if (epi && ldc < std::max(1, M)) return SIMD2_ERROR_INVALID_PROBLEM;
  if (K > 0 && (lda < std::max(1, M) || ldb < std::max(1, K))) return SIMD2_ERROR_INVALID_PROBLEM;
  if (!D || (K > 0 && (!A || !B)) || (epi && !C)) return SIMD2_ERROR_NULL_POINTER;
  if (K > 0) {
    const auto d = span(D, M, N, ldd, es);
    if (overlap(d, span(A, M, K, lda, es)) || overlap(d, span(B, K, N, ldb, es)))
      return SIMD2_ERROR_ALIASING;
  }
  return SIMD2_SUCCESS;
}

// Column-major D(MxN) = A(MxK) (x) B(KxN) is the row-major N x M problem with
// X = B (rows N, k contiguous, ldb) and Y = A (K rows, M contiguous, lda).
int srgemm_f32_cm(int op, int M, int N, int K, const float *A, int lda, const float *B,
                  int ldb, const float *C, int ldc, float *D, int ldd, bool epi,
                  hipStream_t s, int cfg = CFG_AUTO) {
  if (op < 0 || op > 8 || op == SIMD2_OR_AND) return SIMD2_ERROR_NOT_SUPPORTED;
  const int st = check_cm(M, N, K, A, lda, B, ldb, C, ldc, D, ldd, epi, sizeof(float));
  if (st == OK_NOTHING_TO_DO) return SIMD2_SUCCESS;
  if (st != SIMD2_SUCCESS) return st;
  hipError_t e;
  if (op == SIMD2_PLUS_MUL) {
    e = launch_plusmul_mfma(N, M, K, B, ldb, A, lda, D, ldd, epi ? C : nullptr, ldc, s);
  } else {
    Epilogue ep;
    if (epi) {
      ep.C = C;
      ep.ldc = ldc;
    }
    e = launch_valu_f32(op, N, M, K, B, ldb, A, lda, D, ldd, ep, s, cfg);
  }
  return hip_status(e);
}

int srgemm_i32_cm(int M, int N, int K, const int *A, int lda, const int *B, int ldb,
                  const int *C, int ldc, int *D, int ldd, bool epi, hipStream_t s) {
  const int st = check_cm(M, N, K, A, lda, B, ldb, C, ldc, D, ldd, epi, sizeof(int));
  if (st == OK_NOTHING_TO_DO) return SIMD2_SUCCESS;
  if (st != SIMD2_SUCCESS) return st;
  return hip_status(launch_orand_i32(N, M, K, B, ldb, A, lda, D, ldd, epi ? C : nullptr,
                                     ldc, nullptr, 0, nullptr, s));
}

}  // namespace

extern "C" {

const char *simd2_status_string(int status) {
  switch (status) {
  case SIMD2_SUCCESS: return "success";
  case SIMD2_ERROR_MISALIGNED_OPERAND: return "misaligned operand";
  case SIMD2_ERROR_INVALID_DATA_TYPE: return "invalid data type";
  case SIMD2_ERROR_INVALID_LAYOUT: return "invalid layout";
  case SIMD2_ERROR_INVALID_PROBLEM: return "invalid problem (negative size or leading dimension too small)";
  case SIMD2_ERROR_NOT_SUPPORTED: return "operation not supported";
  case SIMD2_ERROR_INTERNAL: return "internal error (kernel launch failed)";
  case SIMD2_ERROR_NULL_POINTER: return "null pointer";
  case SIMD2_ERROR_ALIASING: return "output overlaps an input";
  case SIMD2_ERROR_NO_DEVICE: return "no GPU device";
  case SIMD2_ERROR_NOT_CONVERGED: return "iteration cap reached before convergence";
  case SIMD2_ERROR_WORKSPACE: return "workspace missing or too small";
  default: return "unknown status";
  }
}

#define SIMD2_DROPIN_F32(name, opcode)                                                   \
  int name(int M, int N, int K, float const *A, int lda, float const *B, int ldb,       \
           float *C, int ldc, float *D, bool do_epilogue_min, hipStream_t stream) {     \
    return srgemm_f32_cm(opcode, M, N, K, A, lda, B, ldb, C, ldc, D, ldc,              \
                         do_epilogue_min, stream);                                      \
  }

SIMD2_DROPIN_F32(cuasr_minplus_srsgemm, SIMD2_MIN_PLUS)
SIMD2_DROPIN_F32(cuasr_plusmul_srsgemm, SIMD2_PLUS_MUL)
SIMD2_DROPIN_F32(cuasr_maxplus_srsgemm, SIMD2_MAX_PLUS)
SIMD2_DROPIN_F32(cuasr_maxmin_srsgemm, SIMD2_MAX_MIN)
SIMD2_DROPIN_F32(cuasr_maxmul_srsgemm, SIMD2_MAX_MUL)
SIMD2_DROPIN_F32(cuasr_minmul_srsgemm, SIMD2_MIN_MUL)
SIMD2_DROPIN_F32(cuasr_minmax_srsgemm, SIMD2_MIN_MAX)
SIMD2_DROPIN_F32(cuasr_plusmiussquare_srsgemm, SIMD2_MINUS_SQUARE)
#undef SIMD2_DROPIN_F32

int cuasr_orand_srsgemm(int M, int N, int K, int const *A, int lda, int const *B, int ldb,
                        int *C, int ldc, int *D, bool do_epilogue_min, hipStream_t stream) {
  return srgemm_i32_cm(M, N, K, A, lda, B, ldb, C, ldc, D, ldc, do_epilogue_min, stream);
}

int simd2_srgemm(int op, int M, int N, int K, const void *A, int lda, const void *B,
                 int ldb, const void *C, int ldc, void *D, int ldd, int do_epilogue,
                 hipStream_t stream) {
  if (op == SIMD2_OR_AND)
    return srgemm_i32_cm(M, N, K, static_cast<const int *>(A), lda,
                         static_cast<const int *>(B), ldb, static_cast<const int *>(C), ldc,
                         static_cast<int *>(D), ldd, do_epilogue != 0, stream);
  return srgemm_f32_cm(op, M, N, K, static_cast<const float *>(A), lda,
                       static_cast<const float *>(B), ldb, static_cast<const float *>(C),
                       ldc, static_cast<float *>(D), ldd, do_epilogue != 0, stream);
}

int simd2_srgemm_ex(int op, int M, int N, int K, const void *A, int lda, const void *B,
                    int ldb, const void *C, int ldc, void *D, int ldd, int do_epilogue,
                    int tile_config, hipStream_t stream) {
  if (tile_config < -1 || tile_config >= CFG_COUNT) return SIMD2_ERROR_NOT_SUPPORTED;
  if (op == SIMD2_OR_AND || op == SIMD2_PLUS_MUL)
    return simd2_srgemm(op, M, N, K, A, lda, B, ldb, C, ldc, D, ldd, do_epilogue, stream);
  return srgemm_f32_cm(op, M, N, K, static_cast<const float *>(A), lda,
                       static_cast<const float *>(B), ldb, static_cast<const float *>(C),
                       ldc, static_cast<float *>(D), ldd, do_epilogue != 0, stream,
                       tile_config);
}

int simd2_tile_config_count(void) { return CFG_COUNT; }

int simd2_pick_tile_config(int op, int M, int N, int K) { return valu_pick_cfg(op, N, M, K); }

int simd2_autotune(int op, int M, int N, int K, const void *A, int lda, const void *B, int ldb,
                   const void *C, int ldc, void *D, int ldd, int do_epilogue, hipStream_t stream,
                   int *best, float *ms) {
  if (best) *best = -1;
  if (op < 0 || op > 8 || op == SIMD2_OR_AND || op == SIMD2_PLUS_MUL) return SIMD2_ERROR_NOT_SUPPORTED;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return SIMD2_ERROR_INTERNAL;
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return SIMD2_ERROR_INTERNAL;
  }
  int st = SIMD2_SUCCESS, arg = -1;
  float best_ms = 1e30f;
  for (int cfg = 0; cfg < CFG_COUNT && st == SIMD2_SUCCESS; ++cfg) {
    st = simd2_srgemm_ex(op, M, N, K, A, lda, B, ldb, C, ldc, D, ldd, do_epilogue, cfg, stream);
    if (st != SIMD2_SUCCESS) break;
    float t = 0.f;
    if (hipEventRecord(e0, stream) != hipSuccess) { st = SIMD2_ERROR_INTERNAL; break; }
    for (int r = 0; r < 3 && st == SIMD2_SUCCESS; ++r)
      st = simd2_srgemm_ex(op, M, N, K, A, lda, B, ldb, C, ldc, D, ldd, do_epilogue, cfg, stream);
    if (st != SIMD2_SUCCESS) break;
    if (hipEventRecord(e1, stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
        hipEventElapsedTime(&t, e0, e1) != hipSuccess) {
      st = SIMD2_ERROR_INTERNAL;
      break;
    }
    t /= 3.f;
    if (ms) ms[cfg] = t;
    if (t < best_ms) {
      best_ms = t;
      arg = cfg;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st != SIMD2_SUCCESS) return st;
  valu_record_cfg(op, N, M, K, arg);
  if (best) *best = arg;
  return SIMD2_SUCCESS;
}

int simd2_orand_packed(int M, int N, int K, const uint32_t *A, int lda_words,
                       const uint32_t *B, int ldb_words, const uint32_t *C, int ldc_words,
                       uint32_t *D, int ldd_words, int do_epilogue, hipStream_t stream) {
  if (M < 0 || N < 0 || K < 0) return SIMD2_ERROR_INVALID_PROBLEM;
  if (M == 0 || N == 0) return SIMD2_SUCCESS;
  const int nw = (N + 31) / 32, kw = (K + 31) / 32;
  if (ldd_words < nw || (do_epilogue && ldc_words < nw) ||
      (K > 0 && (lda_words < kw || ldb_words < nw)))
    return SIMD2_ERROR_INVALID_PROBLEM;
  if (!D || (K > 0 && (!A || !B)) || (do_epilogue && !C)) return SIMD2_ERROR_NULL_POINTER;
  if (K > 0) {
    // row-major spans: rows x words with leading dimension ld
    const auto d = span(D, nw, M, ldd_words, 4);
    if (overlap(d, span(A, kw, M, lda_words, 4)) || overlap(d, span(B, nw, K, ldb_words, 4)))
      return SIMD2_ERROR_ALIASING;
  }
  return hip_status(launch_orand_packed(M, N, K, A, lda_words, B, ldb_words,
                                        do_epilogue ? C : nullptr, ldc_words, D, ldd_words,
                                        nullptr, 0, nullptr, stream));
}

size_t simd2_minus_square_workspace(int M, int N, int K) {
  if (M <= 0 || N <= 0 || K <= 0) return 0;
  return minus_square_exact_workspace(N, M, K);  // row-major N x M problem
}

int simd2_minus_square_exact(int M, int N, int K, const float *A, int lda, const float *B,
                             int ldb, const float *C, int ldc, float *D, int ldd,
                             int do_epilogue, void *workspace, size_t workspace_bytes,
                             hipStream_t stream) {
  const bool epi = do_epilogue != 0;
  const int st = check_cm(M, N, K, A, lda, B, ldb, C, ldc, D, ldd, epi, sizeof(float));
  if (st == OK_NOTHING_TO_DO) return SIMD2_SUCCESS;
  if (st != SIMD2_SUCCESS) return st;
  const size_t need = simd2_minus_square_workspace(M, N, K);
  if (need > 0 && (!workspace || workspace_bytes < need)) return SIMD2_ERROR_WORKSPACE;
  if (need > 0) {  // the workspace is written: it may not overlap an operand
    const uintptr_t w0 = reinterpret_cast<uintptr_t>(workspace);
    const std::pair<uintptr_t, uintptr_t> w{w0, w0 + need};
    if (overlap(w, span(A, M, K, lda, 4)) || overlap(w, span(B, K, N, ldb, 4)) ||
        overlap(w, span(D, M, N, ldd, 4)) || (epi && overlap(w, span(C, M, N, ldc, 4))))
      return SIMD2_ERROR_ALIASING;
    if (!workspace_aligned(workspace)) return SIMD2_ERROR_MISALIGNED_OPERAND;
  }
  return hip_status(launch_minus_square_exact(N, M, K, B, ldb, A, lda, D, ldd,
                                              epi ? C : nullptr, ldc, workspace,
                                              workspace_bytes, stream));
}

int simd2_minus_square_exact_noalloc(int M, int N, int K, const float *A, int lda,
                                     const float *B, int ldb, const float *C, int ldc, float *D,
                                     int ldd, int do_epilogue, hipStream_t stream) {
  static_assert(SIMD2_MSQ_NOALLOC_MAX_BYTES == MSQ_NOALLOC_BYTES, "scratch size");
  // The size first: a problem past the scratch is refused before anything else.
  if (simd2_minus_square_workspace(M, N, K) > SIMD2_MSQ_NOALLOC_MAX_BYTES)
    return SIMD2_ERROR_WORKSPACE;
  const bool epi = do_epilogue != 0;
  const int st = check_cm(M, N, K, A, lda, B, ldb, C, ldc, D, ldd, epi, sizeof(float));
  if (st == OK_NOTHING_TO_DO) return SIMD2_SUCCESS;
  if (st != SIMD2_SUCCESS) return st;
  return hip_status(minus_square_exact_noalloc(N, M, K, B, ldb, A, lda, D, ldd,
                                               epi ? C : nullptr, ldc, stream));
}

int simd2_minus_square_tiles(const void *workspace, int M, int N, int K, hipStream_t stream,
                             int *mfma_tiles, int *valu_tiles) {
  if (!mfma_tiles || !valu_tiles) return SIMD2_ERROR_NULL_POINTER;
  if (simd2_minus_square_workspace(M, N, K) > 0 && !workspace) return SIMD2_ERROR_WORKSPACE;
  return hip_status(
      minus_square_exact_tile_counts(workspace, N, M, K, stream, mfma_tiles, valu_tiles));
}

int simd2_device_info(int *cu_count, int *clock_khz, char *name, int name_len) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_status(e);
  hipDeviceProp_t p;
  e = hipGetDeviceProperties(&p, dev);
  if (e != hipSuccess) return hip_status(e);
  if (cu_count) *cu_count = p.multiProcessorCount;
  if (clock_khz) *clock_khz = p.clockRate;
  if (name && name_len > 0) {
    std::strncpy(name, p.gcnArchName, name_len - 1);
    name[name_len - 1] = 0;
  }
  return SIMD2_SUCCESS;
}

// ---- FP16 emulation --------------------------------------------------------

int simd2_tensor_srgemm(const simd2_half *a, const simd2_half *b, const float *c, float *d,
                        int m, int n, int k, int do_epilogue, int opcode, hipStream_t stream) {
  if (m < 0 || n < 0 || k < 0 || m % 16 || n % 16 || k % 16) return -1;
  if (opcode < 0 || opcode > 8) return SIMD2_ERROR_NOT_SUPPORTED;
  if (m == 0 || n == 0) return SIMD2_SUCCESS;
  if (!d || (k > 0 && (!a || !b)) || (do_epilogue && !c)) return SIMD2_ERROR_NULL_POINTER;
  return hip_status(launch_tensor_emu(opcode, a, b, c, d, m, n, k, do_epilogue, stream));
}

int tensor_srgemm(simd2_half *a, simd2_half *b, float *c, float *d, int m, int n, int k,
                  bool do_epilogue, int opcode) {
  return simd2_tensor_srgemm(a, b, c, d, m, n, k, do_epilogue, opcode, nullptr);
}

void f2h_device(simd2_half *fp16_d, float *fp32_d, int num_elements) {
  if (!fp16_d || !fp32_d || num_elements <= 0) return;
  if (launch_f2h(fp16_d, fp32_d, num_elements, nullptr) == hipSuccess)
    (void)hipDeviceSynchronize();
}

void h2f_device(float *fp32_d, simd2_half *fp16_d, int num_elements) {
  if (!fp16_d || !fp32_d || num_elements <= 0) return;
  if (launch_h2f(fp32_d, fp16_d, num_elements, nullptr) == hipSuccess)
    (void)hipDeviceSynchronize();
}

// ---- convergence --------------------------------------------------------

static bool comp_update_common(void *prev, void *cur, int *check, int *check_h, int M, int N,
                               int mode, bool is_int) {
  const int64_t n = (int64_t)M * N;
  if (check_h) check_h[0] = 0;
  if (!prev || !cur || !check || n <= 0) return false;
  hipError_t e = launch_fill_i32(check, 0, 1, nullptr);
  if (e == hipSuccess)
    e = is_int ? launch_comp_update_i32(static_cast<int *>(prev), static_cast<const int *>(cur),
                                        check, n, nullptr)
               : launch_comp_update_f32(static_cast<float *>(prev), static_cast<float *>(cur),
                                        check, n, mode, nullptr);
  int h = 0;
  if (e == hipSuccess) e = hipMemcpy(&h, check, sizeof(int), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return false;  // the bool result cannot carry an error
  if (check_h) check_h[0] = h;
  return h == 1;
}

bool comp_update(float *prev, float *cur, int *check, int *check_h, int M, int N) {
  return comp_update_common(prev, cur, check, check_h, M, N, CHECK_THRESHOLD, false);
}

bool comp_update2(float *prev, float *cur, int *check, int *check_h, int M, int N) {
  return comp_update_common(prev, cur, check, check_h, M, N, CHECK_APLP, false);
}

bool comp_update_int(int *prev, int *cur, int *check, int *check_h, int M, int N) {
  return comp_update_common(prev, cur, check, check_h, M, N, 0, true);
}

static int read_flag(int *flag_d, hipStream_t s, int *out) {
  int h = 0;
  hipError_t e = hipMemcpyAsync(&h, flag_d, sizeof(int), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  *out = h;
  return hip_status(e);
}

int simd2_converge_f32(int op, int n, float *X, int ldx, float *work, int ldw, int *flag_d,
                       int do_epilogue, int check_mode, int max_itr, int flags,
                       hipStream_t stream, int *num_itr) {
  if (num_itr) *num_itr = 0;
  if (op < 0 || op > 8 || op == SIMD2_OR_AND || op == SIMD2_PLUS_MUL ||
      op == SIMD2_MINUS_SQUARE)
    return SIMD2_ERROR_NOT_SUPPORTED;
  if (check_mode < 0 || check_mode > 2) return SIMD2_ERROR_NOT_SUPPORTED;
  if (n < 0 || max_itr < 0) return SIMD2_ERROR_INVALID_PROBLEM;
  if (n == 0) return SIMD2_SUCCESS;
  if (ldx < n || ldw < n) return SIMD2_ERROR_INVALID_PROBLEM;
  if (!X || !work || !flag_d) return SIMD2_ERROR_NULL_POINTER;
  if (overlap(span(X, n, n, ldx, 4), span(work, n, n, ldw, 4))) return SIMD2_ERROR_ALIASING;

  float *cur = X, *nxt = work;
  int64_t ldcur = ldx, ldnxt = ldw;
  int itr = 0, run = 1;
  hipError_t e = hipSuccess;
  while (run && itr < max_itr) {
    ++itr;
    e = launch_fill_i32(flag_d, 0, 1, stream);
    if (e != hipSuccess) break;
    Epilogue ep;
    if (do_epilogue) {
      ep.C = cur;
      ep.ldc = ldcur;
    }
    if (flags == SIMD2_CONVERGE_FUSED) {
      ep.P = cur;
      ep.ldp = ldcur;
      ep.flag = flag_d;
      ep.mode = check_mode;
      e = launch_valu_f32(op, n, n, n, cur, ldcur, cur, ldcur, nxt, ldnxt, ep, stream);
      if (e != hipSuccess) break;
      std::swap(cur, nxt);
      std::swap(ldcur, ldnxt);
    } else {
      e = launch_valu_f32(op, n, n, n, cur, ldcur, cur, ldcur, nxt, ldnxt, ep, stream);
      if (e != hipSuccess) break;
      if (ldcur == n && ldnxt == n) {
        e = launch_comp_update_f32(cur, nxt, flag_d, (int64_t)n * n, check_mode, stream);
      } else {
        // strided matrices: compare row by row through the same kernel
        for (int r = 0; r < n && e == hipSuccess; ++r)
          e = launch_comp_update_f32(cur + (int64_t)r * ldcur, nxt + (int64_t)r * ldnxt, flag_d,
                                     n, check_mode, stream);
      }
      if (e != hipSuccess) break;
    }
    const int st = read_flag(flag_d, stream, &run);
    if (st != SIMD2_SUCCESS) {
      if (num_itr) *num_itr = itr;
      return st;
    }
  }
  if (e == hipSuccess && cur != X)
    e = launch_copy_2d(X, (int64_t)ldx * 4, cur, ldcur * 4, n, n * 4, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (num_itr) *num_itr = itr;
  if (e != hipSuccess) return hip_status(e);
  return run ? SIMD2_ERROR_NOT_CONVERGED : SIMD2_SUCCESS;
}

int simd2_converge_i32(int n, int *X, int ldx, int *work, int ldw, int *flag_d, int max_itr,
                       int flags, hipStream_t stream, int *num_itr) {
  if (num_itr) *num_itr = 0;
  if (n < 0 || max_itr < 0) return SIMD2_ERROR_INVALID_PROBLEM;
  if (n == 0) return SIMD2_SUCCESS;
  if (ldx < n || ldw < n) return SIMD2_ERROR_INVALID_PROBLEM;
  if (!X || !work || !flag_d) return SIMD2_ERROR_NULL_POINTER;
  if (overlap(span(X, n, n, ldx, 4), span(work, n, n, ldw, 4))) return SIMD2_ERROR_ALIASING;
  int *cur = X, *nxt = work;
  int64_t ldcur = ldx, ldnxt = ldw;
  int itr = 0, run = 1;
  hipError_t e = hipSuccess;
  // Fused loop, n >= 64: every squaring on bit-packed rows -- the same 0/1
  // matrices bit for bit, 32x less to move.  The two packed copies live in
  // work's rows (row r's first 2 ceil(n/32) <= n ints; nothing past column n
  // is touched).  The first squaring's change test is the int compare's:
  // its 0/1 result differs from X wherever X holds a value other than 0 or 1,
  // which the packing reports into the same flag.  The result is unpacked
  // into X.
  if (flags == SIMD2_CONVERGE_FUSED && n >= 64 && max_itr > 0) {
    const int nw = (n + 31) / 32;
    uint32_t *pc = reinterpret_cast<uint32_t *>(work), *pn = pc + nw;
    e = launch_fill_i32(flag_d, 0, 1, stream);
    if (e == hipSuccess) e = launch_pack_i32(n, n, X, ldx, pc, ldw, flag_d, stream);
    while (e == hipSuccess && run && itr < max_itr) {
      ++itr;
      if (itr > 1) e = launch_fill_i32(flag_d, 0, 1, stream);
      if (e == hipSuccess)
        e = launch_orand_packed(n, n, n, pc, ldw, pc, ldw, pc, ldw, pn, ldw, pc, ldw, flag_d,
                                stream);
      if (e != hipSuccess) break;
      std::swap(pc, pn);
      const int st = read_flag(flag_d, stream, &run);
      if (st != SIMD2_SUCCESS) {
        if (num_itr) *num_itr = itr;
        return st;
      }
    }
    if (e == hipSuccess) e = launch_unpack_i32(n, n, pc, ldw, X, ldx, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (num_itr) *num_itr = itr;
    if (e != hipSuccess) return hip_status(e);
    return run ? SIMD2_ERROR_NOT_CONVERGED : SIMD2_SUCCESS;
  }
  while (run && itr < max_itr) {
    ++itr;
    e = launch_fill_i32(flag_d, 0, 1, stream);
    if (e != hipSuccess) break;
    if (flags == SIMD2_CONVERGE_FUSED) {
      e = launch_orand_i32(n, n, n, cur, ldcur, cur, ldcur, nxt, ldnxt, cur, ldcur, cur, ldcur,
                           flag_d, stream);
      if (e != hipSuccess) break;
      std::swap(cur, nxt);
      std::swap(ldcur, ldnxt);
    } else {
      e = launch_orand_i32(n, n, n, cur, ldcur, cur, ldcur, nxt, ldnxt, cur, ldcur, nullptr, 0,
                           nullptr, stream);
      if (e != hipSuccess) break;
      for (int r = 0; r < n && e == hipSuccess; ++r) {
        if (ldcur == n && ldnxt == n) {
          e = launch_comp_update_i32(cur, nxt, flag_d, (int64_t)n * n, stream);
          break;
        }
        e = launch_comp_update_i32(cur + (int64_t)r * ldcur, nxt + (int64_t)r * ldnxt, flag_d, n,
                                   stream);
      }
      if (e != hipSuccess) break;
    }
    const int st = read_flag(flag_d, stream, &run);
    if (st != SIMD2_SUCCESS) {
      if (num_itr) *num_itr = itr;
      return st;
    }
  }
  if (e == hipSuccess && cur != X)
    e = launch_copy_2d(X, (int64_t)ldx * 4, cur, ldcur * 4, n, n * 4, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (num_itr) *num_itr = itr;
  if (e != hipSuccess) return hip_status(e);
  return run ? SIMD2_ERROR_NOT_CONVERGED : SIMD2_SUCCESS;
}

int simd2_converge_orand_packed(int n, uint32_t *X, int ldx_words, uint32_t *work, int ldw_words,
                                int *flag_d, int max_itr, hipStream_t stream, int *num_itr) {
  if (num_itr) *num_itr = 0;
  if (n < 0 || max_itr < 0) return SIMD2_ERROR_INVALID_PROBLEM;
  if (n == 0) return SIMD2_SUCCESS;
  const int nw = (n + 31) / 32;
  if (ldx_words < nw || ldw_words < nw) return SIMD2_ERROR_INVALID_PROBLEM;
  if (!X || !work || !flag_d) return SIMD2_ERROR_NULL_POINTER;
  if (overlap(span(X, nw, n, ldx_words, 4), span(work, nw, n, ldw_words, 4)))
    return SIMD2_ERROR_ALIASING;
  uint32_t *cur = X, *nxt = work;
  int64_t ldcur = ldx_words, ldnxt = ldw_words;
  int itr = 0, run = 1;
  hipError_t e = hipSuccess;
  while (run && itr < max_itr) {
    ++itr;
    e = launch_fill_i32(flag_d, 0, 1, stream);
    if (e != hipSuccess) break;
    e = launch_orand_packed(n, n, n, cur, ldcur, cur, ldcur, cur, ldcur, nxt, ldnxt, cur, ldcur,
                            flag_d, stream);
    if (e != hipSuccess) break;
    std::swap(cur, nxt);
    std::swap(ldcur, ldnxt);
    const int st = read_flag(flag_d, stream, &run);
    if (st != SIMD2_SUCCESS) {
      if (num_itr) *num_itr = itr;
      return st;
    }
  }
  if (e == hipSuccess && cur != X)
    e = launch_copy_2d(X, (int64_t)ldx_words * 4, cur, ldcur * 4, n, nw * 4, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (num_itr) *num_itr = itr;
  if (e != hipSuccess) return hip_status(e);
  return run ? SIMD2_ERROR_NOT_CONVERGED : SIMD2_SUCCESS;
}

}  // extern "C"

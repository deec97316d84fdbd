/*
 * simd2 -- semiring GEMM on AMD Instinct MI355X (gfx950).
 *
 * C ABI of libsimd2.so.  Every entry point takes plain device pointers and
 * sizes, launches asynchronously on `stream` (0 = the null stream) unless
 * stated otherwise, allocates nothing, and returns a status code.
 *
 * Semantics shared by all srgemm entry points (the reference's CPU oracle,
 * kernel/host_srgemm.h:16-49, with D starting from the additive identity):
 *
 *     acc(i,j) = (+)_k  A(i,k) (x) B(k,j)          acc starts at the (+)-identity
 *     D(i,j)   = do_epilogue ? acc(i,j) (+) C(i,j) : acc(i,j)
 *
 * The (+)-identity is +inf for min, -inf for max, 0 for plus and or.
 * min/max ops and or_and are exact (each term is one IEEE rounding of (x),
 * and (+) selects), so results are bit-identical to the CPU oracle for inputs
 * without NaN and without -0.0.  plus_mul and minus_square are floating
 * point sums whose order differs from the oracle's.
 *
 * Aliasing: A, B and C may be the same buffer; C may equal D; D must not
 * overlap A or B (SIMD2_ERROR_ALIASING).
 */
#ifndef SIMD2_SRGEMM_H
#define SIMD2_SRGEMM_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
#define SIMD2_DEFAULT_STREAM = nullptr
extern "C" {
#else
#define SIMD2_DEFAULT_STREAM
#endif

/* Operator codes in the order of the reference's `enum op`
 * (kernel/tensor_srgemm.cuh:18-26). */
typedef enum simd2_op {
  SIMD2_MIN_PLUS = 0,
  SIMD2_PLUS_MUL = 1,
  SIMD2_MAX_PLUS = 2,
  SIMD2_MAX_MIN = 3,
  SIMD2_MAX_MUL = 4,
  SIMD2_MIN_MUL = 5,
  SIMD2_MIN_MAX = 6,
  SIMD2_OR_AND = 7,
  SIMD2_MINUS_SQUARE = 8
} simd2_op;

/* Status codes.  0..10 keep the numbering of the cutlass::Status values the
 * reference's entry points cast to int (kernel/srgemm.cuh:82-83); 11.. are
 * specific to this library. */
enum {
  SIMD2_SUCCESS = 0,
  SIMD2_ERROR_MISALIGNED_OPERAND = 1,
  SIMD2_ERROR_INVALID_DATA_TYPE = 2,
  SIMD2_ERROR_INVALID_LAYOUT = 3,
  SIMD2_ERROR_INVALID_PROBLEM = 4, /* negative size, leading dim too small */
  SIMD2_ERROR_NOT_SUPPORTED = 5,   /* unknown op, op not valid for the call */
  SIMD2_ERROR_INTERNAL = 7,        /* kernel launch failed */
  SIMD2_ERROR_NULL_POINTER = 11,
  SIMD2_ERROR_ALIASING = 12,
  SIMD2_ERROR_NO_DEVICE = 13,
  SIMD2_ERROR_NOT_CONVERGED = 14,
  SIMD2_ERROR_WORKSPACE = 15       /* workspace missing or too small */
};
/* Every workspace argument of this library is device memory whose base is
 * 256-byte aligned (hipMalloc's and PyTorch's allocations are); a misaligned
 * one is refused with SIMD2_ERROR_MISALIGNED_OPERAND. */

/* Human-readable name of a status code (static storage). */
const char *simd2_status_string(int status);

/* ------------------------------------------------------------------------
 * Drop-in entry points of the reference's operator surface
 * (kernel/srgemm.cuh).  Column-major A (MxK, lda), B (KxN, ldb), C and D
 * (MxN, both with leading dimension ldc, as the reference passes {D, ldc}).
 * ---------------------------------------------------------------------- */

/* replaces cuasr_minplus_srsgemm, kernel/srgemm.cuh:28-84 */
int cuasr_minplus_srsgemm(int M, int N, int K, float const *A, int lda,
                          float const *B, int ldb, float *C, int ldc, float *D,
                          bool do_epilogue_min, hipStream_t stream SIMD2_DEFAULT_STREAM);
/* replaces cuasr_plusmul_srsgemm, kernel/srgemm.cuh:89-141 (runs on MFMA) */
int cuasr_plusmul_srsgemm(int M, int N, int K, float const *A, int lda,
                          float const *B, int ldb, float *C, int ldc, float *D,
                          bool do_epilogue_min, hipStream_t stream SIMD2_DEFAULT_STREAM);
/* replaces cuasr_maxplus_srsgemm, kernel/srgemm.cuh:147-208 */
int cuasr_maxplus_srsgemm(int M, int N, int K, float const *A, int lda,
                          float const *B, int ldb, float *C, int ldc, float *D,
                          bool do_epilogue_min, hipStream_t stream SIMD2_DEFAULT_STREAM);
/* replaces cuasr_maxmin_srsgemm, kernel/srgemm.cuh:213-265 */
int cuasr_maxmin_srsgemm(int M, int N, int K, float const *A, int lda,
                         float const *B, int ldb, float *C, int ldc, float *D,
                         bool do_epilogue_min, hipStream_t stream SIMD2_DEFAULT_STREAM);
/* replaces cuasr_maxmul_srsgemm, kernel/srgemm.cuh:269-321 */
int cuasr_maxmul_srsgemm(int M, int N, int K, float const *A, int lda,
                         float const *B, int ldb, float *C, int ldc, float *D,
                         bool do_epilogue_min, hipStream_t stream SIMD2_DEFAULT_STREAM);
/* replaces cuasr_minmul_srsgemm, kernel/srgemm.cuh:326-378 */
int cuasr_minmul_srsgemm(int M, int N, int K, float const *A, int lda,
                         float const *B, int ldb, float *C, int ldc, float *D,
                         bool do_epilogue_min, hipStream_t stream SIMD2_DEFAULT_STREAM);
/* replaces cuasr_minmax_srsgemm, kernel/srgemm.cuh:384-436 */
int cuasr_minmax_srsgemm(int M, int N, int K, float const *A, int lda,
                         float const *B, int ldb, float *C, int ldc, float *D,
                         bool do_epilogue_min, hipStream_t stream SIMD2_DEFAULT_STREAM);
/* replaces cuasr_orand_srsgemm, kernel/srgemm.cuh:442-494.  Elements are
 * ints read as booleans (nonzero = true); D holds 0/1. */
int cuasr_orand_srsgemm(int M, int N, int K, int const *A, int lda,
                        int const *B, int ldb, int *C, int ldc, int *D,
                        bool do_epilogue_min, hipStream_t stream SIMD2_DEFAULT_STREAM);
/* replaces cuasr_plusmiussquare_srsgemm (the reference's spelling),
 * kernel/srgemm.cuh:552-626: D = [C +] sum_k (A(i,k) - B(k,j))^2, the direct
 * FP32 kernel.  Like every drop-in it allocates nothing; the exact int8 form
 * with a caller-owned workspace is simd2_minus_square_exact below. */
int cuasr_plusmiussquare_srsgemm(int M, int N, int K, float const *A, int lda,
                                 float const *B, int ldb, float *C, int ldc,
                                 float *D, bool do_epilogue_min,
                                 hipStream_t stream SIMD2_DEFAULT_STREAM);

/* ------------------------------------------------------------------------
 * Generic entry points.
 * ---------------------------------------------------------------------- */

/* One entry point for all nine ops with the reference's opcode order
 * (kernel/tensor_srgemm.cuh:18-26) and the column-major convention of
 * kernel/srgemm.cuh, with a separate leading dimension for D.  Elements are
 * float, or int for SIMD2_OR_AND. */
int simd2_srgemm(int op, int M, int N, int K, const void *A, int lda,
                 const void *B, int ldb, const void *C, int ldc, void *D,
                 int ldd, int do_epilogue, hipStream_t stream);

/* simd2_srgemm with an explicit tile configuration for the VALU semirings
 * (no counterpart in the reference, whose tile shape is fixed at compile
 * time, kernel/srgemm.cuh:574-583): -1 = automatic, 0 = 128x128 tiles (8x8
 * per thread), 1 = 64x64 (4x4), 2 = 64x128 (4x8 per thread; 64 along M and
 * 128 along N of the column-major call).  Used by the autotuner and benchmarks; results are
 * identical for every configuration. */
int simd2_srgemm_ex(int op, int M, int N, int K, const void *A, int lda, const void *B,
                    int ldb, const void *C, int ldc, void *D, int ldd, int do_epilogue,
                    int tile_config, hipStream_t stream);

/* Number of tile configurations, and the one chosen automatically for op on
 * a column-major M x N x K problem. */
int simd2_tile_config_count(void);
int simd2_pick_tile_config(int op, int M, int N, int K);

/* Autotuner: times every tile configuration on this exact call (each run
 * writes the same D, so D holds the correct result afterwards), records the
 * fastest for (op, M, N, K) -- later automatic calls of that shape use it --
 * and returns it in *best; ms[cfg] receives each configuration's mean time
 * when ms is not null.  Synchronises `stream`.  Not for or_and / plus_mul. */
int simd2_autotune(int op, int M, int N, int K, const void *A, int lda, const void *B, int ldb,
                   const void *C, int ldc, void *D, int ldd, int do_epilogue,
                   hipStream_t stream, int *best, float *ms);

/* Bit-packed or_and (no counterpart in the reference, whose or_and stores one
 * boolean per int32).  Row-major bit matrices: bit (k & 31) of word
 * A[i*lda_words + k/32] is A(i,k); likewise B (K x N bits) and C, D (M x N
 * bits).  D(i,:) = [C(i,:) |] OR_{k : A(i,k)} B(k,:).  Bits of D beyond
 * column N-1 in the last word of a row are written as zero. */
int simd2_orand_packed(int M, int N, int K, const uint32_t *A, int lda_words,
                       const uint32_t *B, int ldb_words, const uint32_t *C,
                       int ldc_words, uint32_t *D, int ldd_words,
                       int do_epilogue, hipStream_t stream);

/* minus_square on the int8 matrix cores, exactly (no counterpart in the
 * reference; column-major like simd2_srgemm).  Each 64x64 tile of D whose
 * rows of A and columns of B hold only integers in [-128, 127] and whose
 * sums are at most 2^24 is computed as |a|^2 + |b|^2 - 2 a.b in integer
 * arithmetic: for such a tile the FP32 sum k by k of host_srgemm is that
 * same integer, so D is bit-identical to cuasr_plusmiussquare_srsgemm's.
 * Other tiles run the direct FP32 kernel.  `workspace` (device memory, kept
 * by the caller) must hold simd2_minus_square_workspace(M, N, K) bytes, about
 * (M + N) * K; that size is 0 when K > 65536, and then every tile runs the
 * direct kernel.  Calls sharing a workspace must be ordered (one stream).
 * A workspace overlapping A, B, C or D is refused (SIMD2_ERROR_ALIASING). */
size_t simd2_minus_square_workspace(int M, int N, int K);
int simd2_minus_square_exact(int M, int N, int K, const float *A, int lda, const float *B,
                             int ldb, const float *C, int ldc, float *D, int ldd,
                             int do_epilogue, void *workspace, size_t workspace_bytes,
                             hipStream_t stream);

/* simd2_minus_square_exact without a workspace argument and without any
 * allocation (no counterpart in the reference): for callers that cannot
 * allocate device memory -- code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc.  The workspace is a scratch array
 * the library holds on each device it runs on (SIMD2_MSQ_NOALLOC_MAX_BYTES,
 * kept until the process ends); a problem whose
 * simd2_minus_square_workspace(M, N, K) exceeds it is refused with
 * SIMD2_ERROR_WORKSPACE (4096 x 4096 x 4096 needs about 32 MiB and fits).
 * Results equal simd2_minus_square_exact's bit for bit.  Every call on one
 * device shares that device's scratch: calls that can overlap on one device
 * (different streams, different host threads) race and must be ordered by
 * the caller. */
#define SIMD2_MSQ_NOALLOC_MAX_BYTES ((size_t)40 << 20)
int simd2_minus_square_exact_noalloc(int M, int N, int K, const float *A, int lda,
                                     const float *B, int ldb, const float *C, int ldc, float *D,
                                     int ldd, int do_epilogue, hipStream_t stream);

/* 64x64 tiles of the last simd2_minus_square_exact call on this workspace
 * and problem that ran on the int8 matrix cores and on the direct kernel
 * (for tests and reports).  Synchronises `stream`. */
int simd2_minus_square_tiles(const void *workspace, int M, int N, int K, hipStream_t stream,
                             int *mfma_tiles, int *valu_tiles);

/* Library / device information for reporting (clock, CU count). */
int simd2_device_info(int *cu_count, int *clock_khz, char *name, int name_len);

#ifdef __cplusplus
}
#endif
#endif /* SIMD2_SRGEMM_H */

/*
 * CPU oracle for the semiring-GEMM path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's host-side semiring GEMM
 * (kernel/host_srgemm.h:16-85) and of its convergence check / repeated
 * squaring driver (kernel/converge.cuh:6-85, apps/apsp/cuASR/apsp_cuASR.cu:37-54).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this library, and only as the checker or as the timed CPU baseline.  The
 * product path (simd2_amd, libsimd2.so) never calls it.
 *
 * Operator codes follow the reference's `enum op` (kernel/tensor_srgemm.cuh:18-26).
 */
#ifndef SIMD2_ORACLE_H
#define SIMD2_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum oracle_op {
  ORACLE_MIN_PLUS = 0,
  ORACLE_PLUS_MUL = 1,
  ORACLE_MAX_PLUS = 2,
  ORACLE_MAX_MIN = 3,
  ORACLE_MAX_MUL = 4,
  ORACLE_MIN_MUL = 5,
  ORACLE_MIN_MAX = 6,
  ORACLE_OR_AND = 7,
  ORACLE_MINUS_SQUARE = 8
};

/* host_srgemm (kernel/host_srgemm.h:16-49) for float, row-major, i-j-k order,
 * accumulating into D's prior content (the caller pre-fills D), then the
 * optional epilogue D = addop(D, C).  `op` selects the (addop, mulop) pair as
 * tests/cuAsr_kernel_test.cu pairs them.  Returns 0, or -1 for a bad op. */
int oracle_srgemm_f32(int op, int M, int N, int K, const float *A,
                      const float *B, const float *C, float *D,
                      int do_epilogue);

/* Same loops, rows split across OpenMP threads.  Every D element sees the
 * same operations in the same order, so the result is bit-identical. */
int oracle_srgemm_f32_omp(int op, int M, int N, int K, const float *A,
                          const float *B, const float *C, float *D,
                          int do_epilogue, int nthreads);

/* Row-sampled variant for timing: computes rows [r0, r1) only. */
int oracle_srgemm_f32_rows(int op, int M, int N, int K, const float *A,
                           const float *B, const float *C, float *D,
                           int do_epilogue, int r0, int r1, int nthreads);

/* host_srgemm<int> with bin_or / bin_and (kernel/host_srgemm.h:76-85). */
int oracle_srgemm_i32(int op, int M, int N, int K, const int32_t *A,
                      const int32_t *B, const int32_t *C, int32_t *D,
                      int do_epilogue);

/* The additive identity each op's D is pre-filled with (+inf for min,
 * -inf for max, 0 for plus and or). */
float oracle_add_identity_f32(int op);

/* comp_update (kernel/converge.cuh:6-30): changed if |prev-cur| > 1e-4;
 * prev <- cur.  Returns 1 if changed. */
int oracle_comp_update_f32(float *prev, const float *cur, int64_t n);
/* comp_update2 (kernel/converge.cuh:32-59): cur in (-1,1) -> -FLT_MAX (cur is
 * modified), changed if prev != cur; prev <- cur. */
int oracle_comp_update2_f32(float *prev, float *cur, int64_t n);
/* comp_update(int) (kernel/converge.cuh:61-85). */
int oracle_comp_update_i32(int32_t *prev, const int32_t *cur, int64_t n);
/* Exact mode: changed if any bit pattern differs; prev <- cur. */
int oracle_comp_update_exact_f32(float *prev, const float *cur, int64_t n);

/* Repeated squaring driver (apps/apsp/cuASR/apsp_cuASR.cu:37-54):
 *   while (run && itr < max_itr) { itr++; X' = X (+) (X (x) X) [epilogue per
 *   do_epilogue]; run = check(X, X') }
 * check_mode: 0 = comp_update (1e-4), 1 = comp_update2 (APLP), 2 = exact.
 * X (n*n, row-major) is updated in place; returns the iteration count, or -1. */
int oracle_converge_f32(int op, int n, float *X, int do_epilogue,
                        int check_mode, int max_itr, int nthreads);
int oracle_converge_i32(int n, int32_t *X, int max_itr);

/* PLD post-processing (apps/pld/cuASR/knncuda.cu:170-237), see knn.c. */
void oracle_knn_insertion_sort(float *dist, int32_t *index, int width, int height, int k);
void oracle_knn_sqrt(float *dist, int width, int k);

#ifdef __cplusplus
}
#endif
#endif

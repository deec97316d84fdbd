/*
 * simd2 -- convergence check and repeated-squaring driver (C ABI of libsimd2.so).
 *
 * Drop-ins for kernel/converge.cuh plus a driver that runs the apps'
 * repeated-squaring loop (apps/<app>/cuASR/<app>_cuASR.cu) on the GPU.
 */
#ifndef SIMD2_CONVERGE_H
#define SIMD2_CONVERGE_H

#include <stdbool.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Change tests (the `check_mode` argument below). */
enum {
  SIMD2_CHECK_EXACT = 0,     /* changed if any bit pattern differs */
  SIMD2_CHECK_THRESHOLD = 1, /* comp_update: changed if |prev-cur| > 1e-4 */
  SIMD2_CHECK_APLP = 2       /* comp_update2: cur in (-1,1) -> -FLT_MAX, then != */
};

/* replaces comp_update(float*, float*, int*, int*, int, int),
 * kernel/converge.cuh:16-30.  Synchronous on the null stream: sets *check to
 * 0, compares prev and cur (M*N elements), copies cur into prev, copies the
 * flag to *check_h and returns whether any element changed by more than 1e-4. */
bool comp_update(float *prev, float *cur, int *check, int *check_h, int M, int N);

/* replaces comp_update2, kernel/converge.cuh:45-59 (APLP: maps cur in (-1,1)
 * to -FLT_MAX in place, then exact compare). */
bool comp_update2(float *prev, float *cur, int *check, int *check_h, int M, int N);

/* replaces comp_update(int*, int*, int*, int*, int, int),
 * kernel/converge.cuh:71-85 (C has no overloads; srgemm.hpp restores the
 * overloaded name for C++ callers). */
bool comp_update_int(int *prev, int *cur, int *check, int *check_h, int M, int N);

/* Driver variants (the `flags` argument below). */
enum {
  /* Reference loop structure: full squaring into `work`, then a separate
   * comp_update pass that copies work -> X and raises the flag, then a
   * device-to-host flag read (apps/apsp/cuASR/apsp_cuASR.cu:42-54). */
  SIMD2_CONVERGE_REFERENCE = 0,
  /* Change test fused into the squaring's epilogue, X and work swap roles
   * each iteration; one 4-byte flag read per iteration. */
  SIMD2_CONVERGE_FUSED = 1
};

/* Repeated squaring X <- [X (+)] X (x) X until no element changes or
 * max_itr squarings ran (the reference caps at 500).  X is n x n row-major
 * (leading dimension ldx), `work` an n x n scratch matrix (ldw), `flag_d` one
 * device int.  do_epilogue: 1 for APSP/APLP/MCP/MaxRP/MinRP, 0 for MST.
 * On return X holds the result and *num_itr the number of squarings,
 * counting the final one that changed nothing, as the reference counts it.
 * Returns SIMD2_ERROR_NOT_CONVERGED if the cap was hit while X still
 * changed.  Blocks the calling thread (one flag read per squaring). */
int simd2_converge_f32(int op, int n, float *X, int ldx, float *work, int ldw,
                       int *flag_d, int do_epilogue, int check_mode,
                       int max_itr, int flags, hipStream_t stream,
                       int *num_itr);

/* Transitive closure on int32 booleans (apps/gtc/cuASR/gtc_cuASR.cu:40-51),
 * or_and with epilogue, exact compare. */
int simd2_converge_i32(int n, int *X, int ldx, int *work, int ldw,
                       int *flag_d, int max_itr, int flags,
                       hipStream_t stream, int *num_itr);

/* Transitive closure on bit-packed rows (ld in 32-bit words), fused only. */
int simd2_converge_orand_packed(int n, uint32_t *X, int ldx_words,
                                uint32_t *work, int ldw_words, int *flag_d,
                                int max_itr, hipStream_t stream,
                                int *num_itr);

#ifdef __cplusplus
}
#endif
#endif /* SIMD2_CONVERGE_H */

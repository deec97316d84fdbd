/*
 * CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates kernel/host_srgemm.h:16-85 and kernel/converge.cuh:6-85 in C.
 * Build with -ffp-contract=off so that plus_mul and minus_square keep one
 * rounding per operator call, as the reference's function-pointer calls do.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef float (*f32_binop)(float, float);
typedef int32_t (*i32_binop)(int32_t, int32_t);

/* Operators, kernel/host_srgemm.h:52-85.  minimum/maximum restate std::min /
 * std::max exactly: std::min(a,b) returns a unless b < a; std::max(a,b)
 * returns a unless a < b. */
static float op_plus(float l, float r) { return l + r; }
static float op_multiplies(float l, float r) { return l * r; }
static float op_minimum(float l, float r) { return (r < l) ? r : l; }
static float op_maximum(float l, float r) { return (l < r) ? r : l; }
static float op_minussquare(float l, float r) { return (l - r) * (l - r); }
static int32_t op_bin_and(int32_t l, int32_t r) { return l && r; }
static int32_t op_bin_or(int32_t l, int32_t r) { return l || r; }

/* (addop, mulop) per op, as tests/cuAsr_kernel_test.cu:72,... pairs them. */
static int ops_for(int op, f32_binop *add, f32_binop *mul) {
  switch (op) {
  case ORACLE_MIN_PLUS: *add = op_minimum; *mul = op_plus; return 0;
  case ORACLE_PLUS_MUL: *add = op_plus; *mul = op_multiplies; return 0;
  case ORACLE_MAX_PLUS: *add = op_maximum; *mul = op_plus; return 0;
  case ORACLE_MAX_MIN: *add = op_maximum; *mul = op_minimum; return 0;
  case ORACLE_MAX_MUL: *add = op_maximum; *mul = op_multiplies; return 0;
  case ORACLE_MIN_MUL: *add = op_minimum; *mul = op_multiplies; return 0;
  case ORACLE_MIN_MAX: *add = op_minimum; *mul = op_maximum; return 0;
  case ORACLE_MINUS_SQUARE: *add = op_plus; *mul = op_minussquare; return 0;
  default: return -1;
  }
}

float oracle_add_identity_f32(int op) {
  switch (op) {
  case ORACLE_MIN_PLUS: case ORACLE_MIN_MUL: case ORACLE_MIN_MAX:
    return INFINITY;
  case ORACLE_MAX_PLUS: case ORACLE_MAX_MIN: case ORACLE_MAX_MUL:
    return -INFINITY;
  default:
    return 0.0f;
  }
}

/* The i-j-k body of host_srgemm (kernel/host_srgemm.h:34-40) for rows
 * [r0, r1), followed by the epilogue rows (:42-48). */
static void srgemm_rows_f32(int N, int K, const float *A, const float *B,
                            const float *C, float *D, int do_epilogue,
                            f32_binop addop, f32_binop mulop, int r0, int r1) {
  for (int i = r0; i < r1; i++)
    for (int j = 0; j < N; j++)
      for (int k = 0; k < K; k++)
        D[(int64_t)i * N + j] = addop(D[(int64_t)i * N + j],
                                      mulop(A[(int64_t)i * K + k],
                                            B[(int64_t)k * N + j]));
  if (do_epilogue)
    for (int i = r0; i < r1; i++)
      for (int j = 0; j < N; j++)
        D[(int64_t)i * N + j] = addop(D[(int64_t)i * N + j], C[(int64_t)i * N + j]);
}

int oracle_srgemm_f32(int op, int M, int N, int K, const float *A,
                      const float *B, const float *C, float *D,
                      int do_epilogue) {
  f32_binop add, mul;
  if (ops_for(op, &add, &mul)) return -1;
  if (M < 0 || N < 0 || K < 0) return -1;
  srgemm_rows_f32(N, K, A, B, C, D, do_epilogue, add, mul, 0, M);
  return 0;
}

int oracle_srgemm_f32_rows(int op, int M, int N, int K, const float *A,
                           const float *B, const float *C, float *D,
                           int do_epilogue, int r0, int r1, int nthreads) {
  f32_binop add, mul;
  if (ops_for(op, &add, &mul)) return -1;
  if (M < 0 || N < 0 || K < 0 || r0 < 0 || r1 > M || r0 > r1) return -1;
#ifdef _OPENMP
  if (nthreads <= 0) nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
#endif
  for (int i = r0; i < r1; i++)
    srgemm_rows_f32(N, K, A, B, C, D, do_epilogue, add, mul, i, i + 1);
  (void)nthreads;
  return 0;
}

int oracle_srgemm_f32_omp(int op, int M, int N, int K, const float *A,
                          const float *B, const float *C, float *D,
                          int do_epilogue, int nthreads) {
  return oracle_srgemm_f32_rows(op, M, N, K, A, B, C, D, do_epilogue, 0, M,
                                nthreads);
}

int oracle_srgemm_i32(int op, int M, int N, int K, const int32_t *A,
                      const int32_t *B, const int32_t *C, int32_t *D,
                      int do_epilogue) {
  if (op != ORACLE_OR_AND) return -1;
  if (M < 0 || N < 0 || K < 0) return -1;
  i32_binop addop = op_bin_or, mulop = op_bin_and;
  for (int i = 0; i < M; i++)
    for (int j = 0; j < N; j++)
      for (int k = 0; k < K; k++)
        D[(int64_t)i * N + j] = addop(D[(int64_t)i * N + j],
                                      mulop(A[(int64_t)i * K + k],
                                            B[(int64_t)k * N + j]));
  if (do_epilogue)
    for (int i = 0; i < M; i++)
      for (int j = 0; j < N; j++)
        D[(int64_t)i * N + j] = addop(D[(int64_t)i * N + j], C[(int64_t)i * N + j]);
  return 0;
}

/* kernel/converge.cuh:6-14.  In device code abs(float) is the float overload,
 * i.e. fabsf; a NaN difference (inf - inf) compares false. */
int oracle_comp_update_f32(float *prev, const float *cur, int64_t n) {
  int changed = 0;
  for (int64_t i = 0; i < n; i++) {
    if (fabsf(prev[i] - cur[i]) > 0.0001) changed = 1;
    prev[i] = cur[i];
  }
  return changed;
}

/* kernel/converge.cuh:32-43. */
int oracle_comp_update2_f32(float *prev, float *cur, int64_t n) {
  int changed = 0;
  for (int64_t i = 0; i < n; i++) {
    if (cur[i] > -1 && cur[i] < 1) cur[i] = -FLT_MAX;
    if (prev[i] != cur[i]) changed = 1;
    prev[i] = cur[i];
  }
  return changed;
}

/* kernel/converge.cuh:61-69. */
int oracle_comp_update_i32(int32_t *prev, const int32_t *cur, int64_t n) {
  int changed = 0;
  for (int64_t i = 0; i < n; i++) {
    if (prev[i] != cur[i]) changed = 1;
    prev[i] = cur[i];
  }
  return changed;
}

int oracle_comp_update_exact_f32(float *prev, const float *cur, int64_t n) {
  int changed = memcmp(prev, cur, (size_t)n * sizeof(float)) != 0;
  memcpy(prev, cur, (size_t)n * sizeof(float));
  return changed;
}

/* apps/apsp/cuASR/apsp_cuASR.cu:37-54 with the operator and check of the
 * sibling apps (APLP uses comp_update2, see apps/aplp/cuASR/aplp_cuASR.cu:55).
 * The GPU library writes D without reading it, so the oracle pre-fills the
 * new matrix with the additive identity before host_srgemm accumulates. */
int oracle_converge_f32(int op, int n, float *X, int do_epilogue,
                        int check_mode, int max_itr, int nthreads) {
  if (n < 0 || op == ORACLE_OR_AND) return -1;
  int64_t nn = (int64_t)n * n;
  float *Xd = (float *)malloc((size_t)(nn ? nn : 1) * sizeof(float));
  if (!Xd) return -1;
  float id = oracle_add_identity_f32(op);
  int run = 1, itr = 0;
  while (run && itr < max_itr) {
    itr += 1;
    for (int64_t i = 0; i < nn; i++) Xd[i] = id;
    oracle_srgemm_f32_omp(op, n, n, n, X, X, X, Xd, do_epilogue, nthreads);
    if (check_mode == 1) run = oracle_comp_update2_f32(X, Xd, nn);
    else if (check_mode == 2) run = oracle_comp_update_exact_f32(X, Xd, nn);
    else run = oracle_comp_update_f32(X, Xd, nn);
  }
  free(Xd);
  return itr;
}

/* apps/gtc/cuASR/gtc_cuASR.cu:40-51 (or_and on int, comp_update(int)). */
int oracle_converge_i32(int n, int32_t *X, int max_itr) {
  if (n < 0) return -1;
  int64_t nn = (int64_t)n * n;
  int32_t *Xd = (int32_t *)malloc((size_t)(nn ? nn : 1) * sizeof(int32_t));
  if (!Xd) return -1;
  int run = 1, itr = 0;
  while (run && itr < max_itr) {
    itr += 1;
    memset(Xd, 0, (size_t)nn * sizeof(int32_t));
    oracle_srgemm_i32(ORACLE_OR_AND, n, n, n, X, X, X, Xd, 1);
    run = oracle_comp_update_i32(X, Xd, nn);
  }
  free(Xd);
  return itr;
}

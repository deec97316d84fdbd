// simd2 -- C++ surface over the C ABI of libsimd2.so.
//
// * srgemm<AddOp, MulOp, T>(M, N, K, A, lda, B, ldb, C, ldc, D, do_epilogue, stream):
//   the templated semiring GEMM shape of the reference's CPU oracle
//   (kernel/host_srgemm.h:16-30, operators :52-85) and of cuASR's
//   Srgemm<AdditionOp, MultiplicationOp, ...> (kernel/srgemm.cuh:54-65), with the
//   column-major convention of kernel/srgemm.cuh.  The (AddOp, MulOp) pair is
//   resolved at compile time to one of the nine supported semirings.
// * comp_update(int*, ...): the int overload of kernel/converge.cuh:71-85, which
//   the C header names comp_update_int.
// * half-precision overloads of tensor_srgemm / f2h_device / h2f_device taking
//   __half pointers, as kernel/tensor_srgemm.cuh and kernel/precision.cuh do.
#pragma once

#include <hip/hip_fp16.h>

#include <type_traits>

#include "converge.h"
#include "srgemm.h"
#include "tensor_emu.h"

namespace simd2 {

// Operator tags (kernel/host_srgemm.h:52-85 names).
template <typename T> struct plus {};
template <typename T> struct multiplies {};
template <typename T> struct minimum {};
template <typename T> struct maximum {};
template <typename T> struct minussquare {};
template <typename T> struct binary_and {};
template <typename T> struct binary_or {};

template <template <typename> class AddOp, template <typename> class MulOp, typename T>
struct semiring_code {
  static constexpr int value = -1;
};
#define SIMD2_SEMIRING(ADD, MUL, TY, CODE)              \
  template <> struct semiring_code<ADD, MUL, TY> {      \
    static constexpr int value = CODE;                  \
  };
SIMD2_SEMIRING(minimum, plus, float, SIMD2_MIN_PLUS)
SIMD2_SEMIRING(plus, multiplies, float, SIMD2_PLUS_MUL)
SIMD2_SEMIRING(maximum, plus, float, SIMD2_MAX_PLUS)
SIMD2_SEMIRING(maximum, minimum, float, SIMD2_MAX_MIN)
SIMD2_SEMIRING(maximum, multiplies, float, SIMD2_MAX_MUL)
SIMD2_SEMIRING(minimum, multiplies, float, SIMD2_MIN_MUL)
SIMD2_SEMIRING(minimum, maximum, float, SIMD2_MIN_MAX)
SIMD2_SEMIRING(binary_or, binary_and, int, SIMD2_OR_AND)
SIMD2_SEMIRING(plus, minussquare, float, SIMD2_MINUS_SQUARE)
#undef SIMD2_SEMIRING

// D = [C (+)] A (x) B, column-major; returns the status of simd2_srgemm.
template <template <typename> class AddOp, template <typename> class MulOp, typename T>
int srgemm(int M, int N, int K, const T *A, int lda, const T *B, int ldb, const T *C, int ldc,
           T *D, bool do_epilogue, hipStream_t stream = nullptr) {
  constexpr int code = semiring_code<AddOp, MulOp, T>::value;
  static_assert(code >= 0, "unsupported (AddOp, MulOp, T) semiring");
  return simd2_srgemm(code, M, N, K, A, lda, B, ldb, C, ldc, D, ldc, do_epilogue ? 1 : 0,
                      stream);
}

}  // namespace simd2

// kernel/converge.cuh:71-85 overload.
inline bool comp_update(int *prev, int *cur, int *check, int *check_h, int M, int N) {
  return comp_update_int(prev, cur, check, check_h, M, N);
}

// kernel/tensor_srgemm.cuh:197-218 / kernel/precision.cuh:21-33 with __half.
inline int tensor_srgemm(__half *a, __half *b, float *c, float *d, int m, int n, int k,
                         bool do_epilogue, simd2_op opcode) {
  return tensor_srgemm(reinterpret_cast<simd2_half *>(a), reinterpret_cast<simd2_half *>(b), c,
                       d, m, n, k, do_epilogue, static_cast<int>(opcode));
}
inline void f2h_device(__half *fp16_d, float *fp32_d, int num_elements) {
  f2h_device(reinterpret_cast<simd2_half *>(fp16_d), fp32_d, num_elements);
}
inline void h2f_device(float *fp32_d, __half *fp16_d, int num_elements) {
  h2f_device(fp32_d, reinterpret_cast<simd2_half *>(fp16_d), num_elements);
}

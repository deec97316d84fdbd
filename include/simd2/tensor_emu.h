/*
 * simd2 -- FP16-input semiring GEMM ("tensor core emulation") and FP32<->FP16
 * conversions: C ABI of libsimd2.so, drop-ins for kernel/tensor_srgemm.cuh
 * and kernel/precision.cuh.
 */
#ifndef SIMD2_TENSOR_EMU_H
#define SIMD2_TENSOR_EMU_H

#include <stdbool.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* IEEE binary16 storage (ABI-identical to __half). */
typedef uint16_t simd2_half;

/* replaces tensor_srgemm, kernel/tensor_srgemm.cuh:197-218.  Row-major
 * D (m x n, float) = [C (+)] A (m x k, half) (x) B (k x n, half) with the (x)
 * rounded to half and accumulated in float in k order; accumulators start at
 * +inf (min-based), 0 (max-based -- the reference's __float2half(FLT_MIN)) or
 * 0 (plus, or, minus_square).  opcode: the reference's enum op order.
 * Returns -1 when m, n or k is not a multiple of 16 (as the reference does),
 * SIMD2_ERROR_NOT_SUPPORTED for an unknown opcode, 0 on success.  Launches on
 * the null stream without waiting, like the reference. */
int tensor_srgemm(simd2_half *a, simd2_half *b, float *c, float *d, int m, int n, int k,
                  bool do_epilogue, int opcode);

/* tensor_srgemm on a given stream (no counterpart in the reference). */
int simd2_tensor_srgemm(const simd2_half *a, const simd2_half *b, const float *c, float *d,
                        int m, int n, int k, int do_epilogue, int opcode, hipStream_t stream);

/* replaces f2h_device, kernel/precision.cuh:21-26 (synchronous). */
void f2h_device(simd2_half *fp16_d, float *fp32_d, int num_elements);

/* replaces h2f_device, kernel/precision.cuh:28-33 (synchronous). */
void h2f_device(float *fp32_d, simd2_half *fp16_d, int num_elements);

#ifdef __cplusplus
}
#endif
#endif /* SIMD2_TENSOR_EMU_H */

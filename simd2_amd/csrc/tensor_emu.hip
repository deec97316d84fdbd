// FP16-input / FP32-accumulate semiring GEMM: the reference's "tensor core
// emulation" semantics (kernel/tensor_srgemm.cuh:37-218) and its FP32<->FP16
// conversions (kernel/precision.cuh:5-33).
//
// Per (i, j) and k in order 0..K-1: the (x) is evaluated in half precision
// (one FP16 rounding), converted to float and folded into a float
// accumulator with fminf / fmaxf / +.  Accumulators start where the
// reference starts them: +inf for min-based ops (__float2half(FLT_MAX)
// overflows to +inf), 0 for max-based ops (__float2half(FLT_MIN) underflows
// to 0 -- so max-based results are clamped below at 0), 0 for the rest.  The
// epilogue folds C the same way; or_and treats values > 0 as true.
//
// One workgroup computes a 16x16 tile (256 threads, one output each); A and B
// come through LDS in 16-deep half tiles.  K order is kept, so results are
// bit-identical to a sequential evaluation.
#include <hip/hip_fp16.h>

#include "internal.hpp"

namespace simd2 {

namespace {

constexpr int TE = 16;  // the reference's 16x16 tile; m, n, k are multiples of 16

__device__ __forceinline__ float emu_step(int op, float acc, __half a, __half b) {
  switch (op) {
  case 0: return fminf(acc, __half2float(__hadd(a, b)));
  case 1: return acc + __half2float(__hmul(a, b));
  case 2: return fmaxf(acc, __half2float(__hadd(a, b)));
  case 3: return fmaxf(acc, __half2float(__hmin(a, b)));
  case 4: return fmaxf(acc, __half2float(__hmul(a, b)));
  case 5: return fminf(acc, __half2float(__hmul(a, b)));
  case 6: return fminf(acc, __half2float(__hmax(a, b)));
  case 7: {
    const bool ab = __half2float(a) > 0.0f, bb = __half2float(b) > 0.0f;
    return (float)((acc > 0.0f) || (ab && bb));
  }
  default: {
    const __half d = __hsub(a, b);
    return acc + __half2float(__hmul(d, d));
  }
  }
}

__device__ __forceinline__ float emu_init(int op) {
  switch (op) {
  case 0: case 5: case 6: return __builtin_inff();
  default: return 0.0f;
  }
}

__device__ __forceinline__ float emu_epilogue(int op, float acc, float c) {
  switch (op) {
  case 0: case 5: case 6: return fminf(acc, c);
  case 2: case 3: case 4: return fmaxf(acc, c);
  case 7: return (float)((int)acc || (int)c);
  default: return acc + c;
  }
}

template <int OP>
__global__ __launch_bounds__(256) void tensor_emu_kernel(const __half *__restrict__ a,
                                                         const __half *__restrict__ b,
                                                         const float *c, float *d, int m,
                                                         int n, int k, int do_epilogue) {
  __shared__ __half As[TE][TE + 2];
  __shared__ __half Bs[TE][TE + 2];
  const int tx = tid_x() % TE;   // column within the tile
  const int ty = tid_x() / TE;   // row within the tile
  const int row0 = bid_y() * TE, col0 = bid_x() * TE;
  float acc = emu_init(OP);
  for (int k0 = 0; k0 < k; k0 += TE) {
    As[ty][tx] = a[(int64_t)(row0 + ty) * k + k0 + tx];
    Bs[ty][tx] = b[(int64_t)(k0 + ty) * n + col0 + tx];
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < TE; ++kk) acc = emu_step(OP, acc, As[ty][kk], Bs[kk][tx]);
    __syncthreads();
  }
  const int64_t idx = (int64_t)(row0 + ty) * n + col0 + tx;
  d[idx] = do_epilogue ? emu_epilogue(OP, acc, c[idx]) : acc;
}

__global__ void f2h_kernel(const float *in, __half *out, int64_t n) {
  for (int64_t i = (int64_t)bid_x() * blockDim.x + tid_x(); i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = __float2half(in[i]);
}

__global__ void h2f_kernel(const __half *in, float *out, int64_t n) {
  for (int64_t i = (int64_t)bid_x() * blockDim.x + tid_x(); i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = __half2float(in[i]);
}

}  // namespace

hipError_t launch_tensor_emu(int op, const void *a, const void *b, const float *c, float *d,
                             int m, int n, int k, int do_epilogue, hipStream_t s) {
  // m, n, k are multiples of 16 (checked by the caller).
  dim3 grid(n / TE, m / TE), block(256);
  const __half *A = static_cast<const __half *>(a);
  const __half *B = static_cast<const __half *>(b);
  switch (op) {
  case 0: tensor_emu_kernel<0><<<grid, block, 0, s>>>(A, B, c, d, m, n, k, do_epilogue); break;
  case 1: tensor_emu_kernel<1><<<grid, block, 0, s>>>(A, B, c, d, m, n, k, do_epilogue); break;
  case 2: tensor_emu_kernel<2><<<grid, block, 0, s>>>(A, B, c, d, m, n, k, do_epilogue); break;
  case 3: tensor_emu_kernel<3><<<grid, block, 0, s>>>(A, B, c, d, m, n, k, do_epilogue); break;
  case 4: tensor_emu_kernel<4><<<grid, block, 0, s>>>(A, B, c, d, m, n, k, do_epilogue); break;
  case 5: tensor_emu_kernel<5><<<grid, block, 0, s>>>(A, B, c, d, m, n, k, do_epilogue); break;
  case 6: tensor_emu_kernel<6><<<grid, block, 0, s>>>(A, B, c, d, m, n, k, do_epilogue); break;
  case 7: tensor_emu_kernel<7><<<grid, block, 0, s>>>(A, B, c, d, m, n, k, do_epilogue); break;
  default: tensor_emu_kernel<8><<<grid, block, 0, s>>>(A, B, c, d, m, n, k, do_epilogue); break;
  }
  return hipGetLastError();
}

hipError_t launch_f2h(void *out, const float *in, int64_t n, hipStream_t s) {
  int64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  f2h_kernel<<<(int)blocks, 256, 0, s>>>(in, static_cast<__half *>(out), n);
  return hipGetLastError();
}

hipError_t launch_h2f(float *out, const void *in, int64_t n, hipStream_t s) {
  int64_t blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (blocks < 1) blocks = 1;
  h2f_kernel<<<(int)blocks, 256, 0, s>>>(static_cast<const __half *>(in), out, n);
  return hipGetLastError();
}

}  // namespace simd2

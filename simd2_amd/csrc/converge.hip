// Element-wise kernels of the convergence check (kernel/converge.cuh) and
// small helpers.  The change flag is raised with one workgroup-reduced atomic
// per workgroup instead of the reference's racing plain stores.
#include <float.h>

#include "internal.hpp"

namespace simd2 {

namespace {

constexpr int EW_THREADS = 256;

// kernel/converge.cuh:6-14 (mode 1) and :32-43 (mode 2); mode 0 compares bits.
__device__ __forceinline__ bool update_one(float &c, float &p, int mode, bool &mapped) {
  mapped = false;
  if (mode == CHECK_APLP) {
    if (c > -1.0f && c < 1.0f) {
      c = -FLT_MAX;
      mapped = true;
    }
    const bool ch = p != c;
    p = c;
    return ch;
  }
  const bool bits = __float_as_uint(p) != __float_as_uint(c);
  const bool ch = (mode == CHECK_THRESHOLD) ? (bits && fabsf(p - c) > 0.0001f) : bits;
  p = c;
  return ch;
}

// VEC: prev and cur 16-byte aligned; the first n / 4 * 4 elements move as
// float4, the rest one at a time.  One flag atomic per workgroup at most.
template <bool VEC>
__global__ __launch_bounds__(EW_THREADS) void comp_update_f32_kernel(
    float *prev, float *cur, int *flag, int64_t n, int mode) {
  bool changed = false;
  const int64_t stride = (int64_t)gridDim.x * EW_THREADS;
  const int64_t t = (int64_t)bid_x() * EW_THREADS + tid_x();
  const int64_t n4 = VEC ? n / 4 : 0;
  for (int64_t i = t; i < n4; i += stride) {
    float4 c = reinterpret_cast<const float4 *>(cur)[i];
    float4 p = reinterpret_cast<const float4 *>(prev)[i];
    bool m0, m1, m2, m3;
    changed |= update_one(c.x, p.x, mode, m0) | update_one(c.y, p.y, mode, m1) |
               update_one(c.z, p.z, mode, m2) | update_one(c.w, p.w, mode, m3);
    if (m0 | m1 | m2 | m3) reinterpret_cast<float4 *>(cur)[i] = c;
    reinterpret_cast<float4 *>(prev)[i] = p;
  }
  for (int64_t i = 4 * n4 + t; i < n; i += stride) {
    float c = cur[i], p = prev[i];
    bool m;
    changed |= update_one(c, p, mode, m);
    if (m) cur[i] = c;
    prev[i] = p;
  }
  if (__syncthreads_or(changed) && tid_x() == 0) atomicOr(flag, 1);
}

// kernel/converge.cuh:61-69.
__global__ __launch_bounds__(EW_THREADS) void comp_update_i32_kernel(
    int *prev, const int *cur, int *flag, int64_t n) {
  bool changed = false;
  for (int64_t i = (int64_t)bid_x() * EW_THREADS + tid_x(); i < n;
       i += (int64_t)gridDim.x * EW_THREADS) {
    const int c = cur[i];
    changed |= (prev[i] != c);
    prev[i] = c;
  }
  if (__syncthreads_or(changed) && tid_x() == 0) atomicOr(flag, 1);
}

__global__ __launch_bounds__(EW_THREADS) void fill_i32_kernel(int *p, int v, int64_t n) {
  for (int64_t i = (int64_t)bid_x() * EW_THREADS + tid_x(); i < n;
       i += (int64_t)gridDim.x * EW_THREADS)
    p[i] = v;
}

__global__ __launch_bounds__(EW_THREADS) void copy_2d_kernel(
    unsigned char *dst, int64_t ldd, const unsigned char *src, int64_t lds,
    int rows, int row_bytes) {
  // One row per block-iteration, 4-byte granules (row_bytes % 4 == 0).
  const int words = row_bytes / 4;
  for (int r = bid_x(); r < rows; r += gridDim.x) {
    const uint32_t *s = reinterpret_cast<const uint32_t *>(src + (int64_t)r * lds);
    uint32_t *d = reinterpret_cast<uint32_t *>(dst + (int64_t)r * ldd);
    for (int w = tid_x(); w < words; w += EW_THREADS) d[w] = s[w];
  }
}

int ew_blocks(int64_t n) {
  int64_t b = (n + EW_THREADS - 1) / EW_THREADS;
  if (b > 4096) b = 4096;
  if (b < 1) b = 1;
  return (int)b;
}

}  // namespace

hipError_t launch_comp_update_f32(float *prev, float *cur, int *flag,
                                  int64_t n, int mode, hipStream_t s) {
  auto aligned = [](const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
  if (aligned(prev) && aligned(cur))
    comp_update_f32_kernel<true><<<ew_blocks((n + 3) / 4), EW_THREADS, 0, s>>>(prev, cur, flag, n,
                                                                             mode);
  else
    comp_update_f32_kernel<false><<<ew_blocks(n), EW_THREADS, 0, s>>>(prev, cur, flag, n, mode);
  return hipGetLastError();
}

hipError_t launch_comp_update_i32(int *prev, const int *cur, int *flag,
                                  int64_t n, hipStream_t s) {
  comp_update_i32_kernel<<<ew_blocks(n), EW_THREADS, 0, s>>>(prev, cur, flag, n);
  return hipGetLastError();
}

hipError_t launch_fill_i32(int *p, int v, int64_t n, hipStream_t s) {
  fill_i32_kernel<<<ew_blocks(n), EW_THREADS, 0, s>>>(p, v, n);
  return hipGetLastError();
}

hipError_t launch_copy_2d(void *dst, int64_t ldd, const void *src, int64_t lds,
                          int rows, int row_bytes, hipStream_t s) {
  if (rows <= 0 || row_bytes <= 0) return hipSuccess;
  copy_2d_kernel<<<rows < 4096 ? rows : 4096, EW_THREADS, 0, s>>>(
      static_cast<unsigned char *>(dst), ldd, static_cast<const unsigned char *>(src),
      lds, rows, row_bytes);
  return hipGetLastError();
}

}  // namespace simd2

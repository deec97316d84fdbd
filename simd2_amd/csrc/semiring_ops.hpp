// Semiring operator traits for the gfx950 VALU kernels.
//
// Each trait gives the multiplicative op `mul`, the additive op `add`, a
// three-way `add3(acc, x, y)` that lowers to one v_min3_f32 / v_max3_f32 for
// the min/max semirings, the additive identity, and the values that pad a
// tile past the end of K (chosen so that mul(pad_x, pad_y) is absorbed by
// `add`).
//
// Bit-exactness: the oracle (kernel/host_srgemm.h:52-85) applies std::min /
// std::max, which return the first argument on ties; v_min/v_max differ from
// that only for NaN and for -0.0 vs +0.0, both outside the input contract.
// These kernels are compiled with IEEE mode off (-mno-amdgpu-ieee) so that
// min/max of loaded values needs no canonicalising instruction.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace simd2 {

#define SIMD2_INF __builtin_inff()

// Three-input min / max as single instructions.  Written as (non-volatile)
// inline asm because with NaNs excluded the compiler treats fminf/fmaxf as
// associative and rebalances the k-chains of a tile into trees, which needs
// every partial term live at once (register spills) and turns each v_min3
// into two v_min.  The asm keeps one v_min3/v_max3 per two k-steps.
__device__ __forceinline__ float vmin3(float a, float b, float c) {
  float r;
  asm("v_min3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}
__device__ __forceinline__ float vmax3(float a, float b, float c) {
  float r;
  asm("v_max3_f32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}

// The plus-based ops use the same asm form: as plain C++ the register
// allocator rotates their accumulators across the k-loop back-edge, adding a
// copy per accumulator and spilling the 8x8 tile.
__device__ __forceinline__ float vadd(float a, float b) {
  float r;
  asm("v_add_f32 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
// c + a * b with one rounding, as the two-address v_fmac_f32: the
// accumulator needs no second register (89 instead of 103 VGPRs in the
// 128x64 minus_square tile; same rate as v_fma_f32).
__device__ __forceinline__ float vfma(float a, float b, float c) {
  asm("v_fmac_f32 %0, %1, %2" : "+v"(c) : "v"(a), "v"(b));
  return c;
}

struct MinPlus {
  static constexpr int code = 0;
  static constexpr float identity = SIMD2_INF;
  static constexpr float pad_x = SIMD2_INF, pad_y = SIMD2_INF;
  __device__ __forceinline__ static float mul(float a, float b) { return a + b; }
  __device__ __forceinline__ static float add(float a, float b) { return fminf(a, b); }
  __device__ __forceinline__ static float add3(float acc, float x, float y) {
    return vmin3(acc, x, y);
  }
};

struct PlusMul {  // VALU form; the default plus_mul path is the MFMA kernel
  static constexpr int code = 1;
  static constexpr float identity = 0.0f;
  static constexpr float pad_x = 0.0f, pad_y = 0.0f;
  // plain C product (an asm v_mul_f32 feeding the asm adds costs an s_nop
  // per pair); the asm adds keep it from being contracted into an fma
  __device__ __forceinline__ static float mul(float a, float b) { return a * b; }
  __device__ __forceinline__ static float add(float a, float b) { return a + b; }
  // Sequential in k with one rounding per op, as the oracle accumulates.
  __device__ __forceinline__ static float add3(float acc, float x, float y) {
    return vadd(vadd(acc, x), y);
  }
};

struct MaxPlus {
  static constexpr int code = 2;
  static constexpr float identity = -SIMD2_INF;
  static constexpr float pad_x = -SIMD2_INF, pad_y = -SIMD2_INF;
  __device__ __forceinline__ static float mul(float a, float b) { return a + b; }
  __device__ __forceinline__ static float add(float a, float b) { return fmaxf(a, b); }
  __device__ __forceinline__ static float add3(float acc, float x, float y) {
    return vmax3(acc, x, y);
  }
};

struct MaxMin {
  static constexpr int code = 3;
  static constexpr float identity = -SIMD2_INF;
  static constexpr float pad_x = -SIMD2_INF, pad_y = -SIMD2_INF;
  __device__ __forceinline__ static float mul(float a, float b) { return fminf(a, b); }
  __device__ __forceinline__ static float add(float a, float b) { return fmaxf(a, b); }
  __device__ __forceinline__ static float add3(float acc, float x, float y) {
    return vmax3(acc, x, y);
  }
};

struct MaxMul {
  static constexpr int code = 4;
  static constexpr float identity = -SIMD2_INF;
  // -inf * +inf = -inf, absorbed by max.
  static constexpr float pad_x = -SIMD2_INF, pad_y = SIMD2_INF;
  __device__ __forceinline__ static float mul(float a, float b) { return a * b; }
  __device__ __forceinline__ static float add(float a, float b) { return fmaxf(a, b); }
  __device__ __forceinline__ static float add3(float acc, float x, float y) {
    return vmax3(acc, x, y);
  }
};

struct MinMul {
  static constexpr int code = 5;
  static constexpr float identity = SIMD2_INF;
  static constexpr float pad_x = SIMD2_INF, pad_y = SIMD2_INF;
  __device__ __forceinline__ static float mul(float a, float b) { return a * b; }
  __device__ __forceinline__ static float add(float a, float b) { return fminf(a, b); }
  __device__ __forceinline__ static float add3(float acc, float x, float y) {
    return vmin3(acc, x, y);
  }
};

struct MinMax {
  static constexpr int code = 6;
  static constexpr float identity = SIMD2_INF;
  static constexpr float pad_x = SIMD2_INF, pad_y = SIMD2_INF;
  __device__ __forceinline__ static float mul(float a, float b) { return fmaxf(a, b); }
  __device__ __forceinline__ static float add(float a, float b) { return fminf(a, b); }
  __device__ __forceinline__ static float add3(float acc, float x, float y) {
    return vmin3(acc, x, y);
  }
};

struct MinusSquare {  // direct form: sum_k (a - b)^2, the square fused into the sum
  static constexpr int code = 8;
  static constexpr float identity = 0.0f;
  static constexpr float pad_x = 0.0f, pad_y = 0.0f;
  __device__ __forceinline__ static float mul(float a, float b) {
    float d = a - b;
    return d * d;
  }
  __device__ __forceinline__ static float add(float a, float b) { return a + b; }
  __device__ __forceinline__ static float add3(float acc, float x, float y) {
    return (acc + x) + y;
  }
  // Fused form used by the tiled kernel: acc + (a-b)^2 as one v_fma_f32.  The
  // difference is plain C: an asm v_sub_f32 feeding the asm v_fma_f32 makes
  // the compiler put an s_nop between every such pair.
  static constexpr bool fused = true;
  __device__ __forceinline__ static float fma_step(float acc, float a, float b) {
    const float d = a - b;
    return vfma(d, d, acc);
  }
};

template <class Op> struct is_fused { static constexpr bool value = false; };
template <> struct is_fused<MinusSquare> { static constexpr bool value = true; };

}  // namespace simd2

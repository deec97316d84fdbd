// plus_mul on the gfx950 matrix cores: v_mfma_f32_16x16x4_f32, FP32 in, FP32
// accumulate (the (+,*) semiring is the only one MFMA computes natively).
//
// A 256-thread workgroup computes a 128x128 tile; its 4 waves own 64x64
// quadrants, each a 4x4 grid of 16x16 MFMA blocks (64 accumulator registers
// per lane).  K advances in 16-deep slabs staged in LDS (X transposed to
// k-major), double-buffered with one barrier per slab; the slab is consumed
// in 4 MFMA k-steps of 4.  Operand layout of the 16x16x4 f32 instruction:
// lane l supplies A[l%16][l/16] and B[l/16][l%16]; accumulator element v of
// lane l is D[4*(l/16)+v][l%16].  LDS rows are padded by 16 floats so the
// four k-rows a wave reads at once land in distinct banks.
//
// Every shape runs here: VEC loads operands as float4 (leading dimensions
// and bases 16-byte aligned, K and N multiples of 4), otherwise one float at
// a time; rows, columns and k past the matrix are read as zeros and never
// stored (a zero contributes an exact +0 to every sum it enters).
#include "internal.hpp"

namespace simd2 {

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int MB = 128, NB = 128, KB = 16, LDP = 16;

template <bool VEC>
__global__ __launch_bounds__(256) void plusmul_mfma_kernel(
    int R, int Cn, int K, const float *__restrict__ X, int64_t ldx,
    const float *__restrict__ Y, int64_t ldy, float *O, int64_t ldo,
    const float *C, int64_t ldc, int tiles_c) {
  __shared__ __attribute__((aligned(16))) float Xs[2][KB][MB + LDP];
  __shared__ __attribute__((aligned(16))) float Ys[2][KB][NB + LDP];
  const int tid = tid_x();
  const int lane = tid % 64;
  const int wave = tid / 64;
  const int wr = wave / 2, wc = wave % 2;
  const int row0 = (bid_x() / tiles_c) * MB;
  const int col0 = (bid_x() % tiles_c) * NB;

  f32x4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  // Slab loads: X 128 rows x 16 k (2 x 4 floats per thread), Y 16 x 128 (2 x 4).
  const int xr = tid / 4, xk = (tid % 4) * 4;
  const int yk = tid / 32, yc = (tid % 32) * 4;
  float4 x0, x1, y0, y1;
  // 4 consecutive floats of row r of a matrix from column c; zeros past (rows, cols)
  auto ld4 = [&](const float *M, int64_t ld, int r, int c, int rows, int cols) {
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (r < rows) {
      const float *p = M + (int64_t)r * ld + c;
      if (VEC) {
        if (c < cols) v = *reinterpret_cast<const float4 *>(p);
      } else {
        if (c < cols) v.x = p[0];
        if (c + 1 < cols) v.y = p[1];
        if (c + 2 < cols) v.z = p[2];
        if (c + 3 < cols) v.w = p[3];
      }
    }
    return v;
  };
  auto fetch = [&](int k0) {
    x0 = ld4(X, ldx, row0 + xr, k0 + xk, R, K);
    x1 = ld4(X, ldx, row0 + xr + 64, k0 + xk, R, K);
    y0 = ld4(Y, ldy, k0 + yk, col0 + yc, K, Cn);
    y1 = ld4(Y, ldy, k0 + yk + 8, col0 + yc, K, Cn);
  };
  auto stage = [&](int b) {
    Xs[b][xk + 0][xr] = x0.x; Xs[b][xk + 1][xr] = x0.y;
    Xs[b][xk + 2][xr] = x0.z; Xs[b][xk + 3][xr] = x0.w;
    Xs[b][xk + 0][xr + 64] = x1.x; Xs[b][xk + 1][xr + 64] = x1.y;
    Xs[b][xk + 2][xr + 64] = x1.z; Xs[b][xk + 3][xr + 64] = x1.w;
    *reinterpret_cast<float4 *>(&Ys[b][yk][yc]) = y0;
    *reinterpret_cast<float4 *>(&Ys[b][yk + 8][yc]) = y1;
  };
  const int nkt = (K + KB - 1) / KB;
  if (nkt > 0) {
    fetch(0);
    stage(0);
  }
  __syncthreads();
  const int lr = lane % 16, lk = lane / 16;
  for (int kt = 0; kt < nkt; ++kt) {
    const int buf = kt & 1;
    const bool more = kt + 1 < nkt;
    if (more) fetch((kt + 1) * KB);
#pragma unroll
    for (int ks = 0; ks < KB; ks += 4) {
      float a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = Xs[buf][ks + lk][wr * 64 + i * 16 + lr];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Ys[buf][ks + lk][wc * 64 + j * 16 + lr];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    if (more) stage(buf ^ 1);
    __syncthreads();
  }

#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = row0 + wr * 64 + i * 16 + 4 * lk + v;
        const int c = col0 + wc * 64 + j * 16 + lr;
        if (r >= R || c >= Cn) continue;
        float o = acc[i][j][v];
        if (C) o = o + C[(int64_t)r * ldc + c];
        O[(int64_t)r * ldo + c] = o;
      }
}

}  // namespace

hipError_t launch_plusmul_mfma(int R, int Cn, int K, const float *X,
                               int64_t ldx, const float *Y, int64_t ldy,
                               float *O, int64_t ldo, const float *C,
                               int64_t ldc, hipStream_t s) {
  auto aligned = [](const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
  const bool vec = K % 4 == 0 && Cn % 4 == 0 && ldx % 4 == 0 && ldy % 4 == 0 && aligned(X) &&
                   aligned(Y);
  const int tiles_c = (Cn + NB - 1) / NB;
  const dim3 grid(((R + MB - 1) / MB) * tiles_c);
  if (vec)
    plusmul_mfma_kernel<true><<<grid, dim3(256), 0, s>>>(R, Cn, K, X, ldx, Y, ldy, O, ldo, C, ldc,
                                                         tiles_c);
  else
    plusmul_mfma_kernel<false><<<grid, dim3(256), 0, s>>>(R, Cn, K, X, ldx, Y, ldy, O, ldo, C, ldc,
                                                          tiles_c);
  return hipGetLastError();
}

}  // namespace simd2

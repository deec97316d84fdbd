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
#include "internal.hpp"

namespace simd2 {

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int MB = 128, NB = 128, KB = 16, LDP = 16;

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

  // Slab loads: X 128 rows x 16 k (2 float4 per thread), Y 16 x 128 (2 each).
  const int xr = tid / 4, xk = (tid % 4) * 4;
  const int yk = tid / 32, yc = (tid % 32) * 4;
  const float *xp0 = X + (int64_t)(row0 + xr) * ldx + xk;
  const float *xp1 = xp0 + 64 * ldx;
  const float *yp0 = Y + (int64_t)yk * ldy + col0 + yc;
  const float *yp1 = yp0 + 8 * ldy;
  float4 x0, x1, y0, y1;
  auto stage = [&](int b) {
    Xs[b][xk + 0][xr] = x0.x; Xs[b][xk + 1][xr] = x0.y;
    Xs[b][xk + 2][xr] = x0.z; Xs[b][xk + 3][xr] = x0.w;
    Xs[b][xk + 0][xr + 64] = x1.x; Xs[b][xk + 1][xr + 64] = x1.y;
    Xs[b][xk + 2][xr + 64] = x1.z; Xs[b][xk + 3][xr + 64] = x1.w;
    *reinterpret_cast<float4 *>(&Ys[b][yk][yc]) = y0;
    *reinterpret_cast<float4 *>(&Ys[b][yk + 8][yc]) = y1;
  };
  const int nkt = K / KB;
  if (nkt > 0) {
    x0 = *reinterpret_cast<const float4 *>(xp0);
    x1 = *reinterpret_cast<const float4 *>(xp1);
    y0 = *reinterpret_cast<const float4 *>(yp0);
    y1 = *reinterpret_cast<const float4 *>(yp1);
    stage(0);
  }
  __syncthreads();
  const int lr = lane % 16, lk = lane / 16;
  for (int kt = 0; kt < nkt; ++kt) {
    const int buf = kt & 1;
    const bool more = kt + 1 < nkt;
    if (more) {
      xp0 += KB; xp1 += KB; yp0 += KB * ldy; yp1 += KB * ldy;
      x0 = *reinterpret_cast<const float4 *>(xp0);
      x1 = *reinterpret_cast<const float4 *>(xp1);
      y0 = *reinterpret_cast<const float4 *>(yp0);
      y1 = *reinterpret_cast<const float4 *>(yp1);
    }
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
  const bool ok = R % MB == 0 && Cn % NB == 0 && K % KB == 0 && ldx % 4 == 0 &&
                  ldy % 4 == 0 && aligned(X) && aligned(Y);
  if (!ok) {  // shapes off the 128x128x16 grid use the VALU kernel
    Epilogue e;
    e.C = C;
    e.ldc = ldc;
    return launch_valu_f32(1, R, Cn, K, X, ldx, Y, ldy, O, ldo, e, s);
  }
  const int tiles_c = Cn / NB;
  plusmul_mfma_kernel<<<dim3((R / MB) * tiles_c), dim3(256), 0, s>>>(
      R, Cn, K, X, ldx, Y, ldy, O, ldo, C, ldc, tiles_c);
  return hipGetLastError();
}

}  // namespace simd2

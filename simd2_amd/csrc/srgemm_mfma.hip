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

constexpr int KB = 16, LDP = 16;

// TILE x TILE output tile (128, or 64 when the 128 tiles would leave CUs
// idle); each wave owns a (TILE/2)^2 quadrant of NB x NB 16x16 blocks.  An
// element's sum runs over the same k slabs in the same order for either
// tile size, so the results are identical.
template <int TILE, bool VEC, bool VEC_OUT>
__global__ __launch_bounds__(256) void plusmul_mfma_kernel(
    int R, int Cn, int K, const float *__restrict__ X, int64_t ldx,
    const float *__restrict__ Y, int64_t ldy, float *O, int64_t ldo,
    const float *C, int64_t ldc, int tiles_c) {
  constexpr int QB = TILE / 2;       // quadrant edge
  constexpr int NB = QB / 16;        // 16x16 blocks per quadrant edge
  constexpr int XPT = TILE / 64;     // float4 loads per thread per slab, X and Y each
  constexpr int YROWS = 1024 / TILE; // k rows of Y one pass of 256 threads covers
  __shared__ __attribute__((aligned(16))) float Xs[2][KB][TILE + LDP];
  __shared__ __attribute__((aligned(16))) float Ys[2][KB][TILE + LDP];
  const int tid = tid_x();
  const int lane = tid % 64;
  const int wave = tid / 64;
  const int wr = wave / 2, wc = wave % 2;
  const int row0 = (bid_x() / tiles_c) * TILE;
  const int col0 = (bid_x() % tiles_c) * TILE;

  f32x4 acc[NB][NB];
#pragma unroll
  for (int i = 0; i < NB; ++i)
#pragma unroll
    for (int j = 0; j < NB; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

  // Slab loads: X TILE rows x 16 k, Y 16 x TILE, XPT float4 of each per thread.
  const int xr = tid / 4, xk = (tid % 4) * 4;
  const int yk = tid / (TILE / 4), yc = (tid % (TILE / 4)) * 4;
  float4 xv[XPT], yv[XPT];
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
#pragma unroll
    for (int q = 0; q < XPT; ++q) {
      xv[q] = ld4(X, ldx, row0 + xr + 64 * q, k0 + xk, R, K);
      yv[q] = ld4(Y, ldy, k0 + yk + YROWS * q, col0 + yc, K, Cn);
    }
  };
  auto stage = [&](int b) {
#pragma unroll
    for (int q = 0; q < XPT; ++q) {
      Xs[b][xk + 0][xr + 64 * q] = xv[q].x;
      Xs[b][xk + 1][xr + 64 * q] = xv[q].y;
      Xs[b][xk + 2][xr + 64 * q] = xv[q].z;
      Xs[b][xk + 3][xr + 64 * q] = xv[q].w;
      *reinterpret_cast<float4 *>(&Ys[b][yk + YROWS * q][yc]) = yv[q];
    }
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
      float a[NB], b[NB];
#pragma unroll
      for (int i = 0; i < NB; ++i) a[i] = Xs[buf][ks + lk][wr * QB + i * 16 + lr];
#pragma unroll
      for (int j = 0; j < NB; ++j) b[j] = Ys[buf][ks + lk][wc * QB + j * 16 + lr];
#pragma unroll
      for (int i = 0; i < NB; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    if (more) stage(buf ^ 1);
    __syncthreads();
  }

  // Stores through LDS, 16 rows of the wave's quadrant at a time: written in
  // the accumulator layout (one column per lane, 4 rows per 16 lanes) to the
  // wave's own padded region and read back row-major, so C is read and O
  // written 16 bytes per lane along rows.  The loop's last barrier has
  // retired every read of Xs; only this wave touches its region.
  constexpr int TLD = QB + 4;
  static_assert(4 * 16 * TLD <= 2 * KB * (TILE + LDP), "store staging fits in Xs");
  float *T = &Xs[0][0][0] + wave * 16 * TLD;
  const int qr0 = row0 + wr * QB, qc0 = col0 + wc * QB;
#pragma unroll
  for (int i = 0; i < NB; ++i) {
#pragma unroll
    for (int j = 0; j < NB; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) T[(4 * lk + v) * TLD + j * 16 + lr] = acc[i][j][v];
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int h = 0; h < QB / 16; ++h) {
      const int idx = lane + 64 * h;
      const int rr = idx / (QB / 4), cc = (idx % (QB / 4)) * 4;
      const int r = qr0 + i * 16 + rr, c = qc0 + cc;
      f32x4 o = *reinterpret_cast<const f32x4 *>(&T[rr * TLD + cc]);
      if (r >= R) continue;
      if (VEC_OUT && c + 3 < Cn) {
        if (C) o += *reinterpret_cast<const f32x4 *>(C + (int64_t)r * ldc + c);
        *reinterpret_cast<f32x4 *>(O + (int64_t)r * ldo + c) = o;
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (c + q < Cn) O[(int64_t)r * ldo + c + q] = C ? o[q] + C[(int64_t)r * ldc + c + q] : o[q];
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
}

template <int TILE>
void launch_tile(bool vec, bool vec_out, int R, int Cn, int K, const float *X, int64_t ldx,
                 const float *Y, int64_t ldy, float *O, int64_t ldo, const float *C, int64_t ldc,
                 hipStream_t s) {
  const int tiles_c = (Cn + TILE - 1) / TILE;
  const dim3 grid(((R + TILE - 1) / TILE) * tiles_c), block(256);
  if (vec && vec_out)
    plusmul_mfma_kernel<TILE, true, true><<<grid, block, 0, s>>>(R, Cn, K, X, ldx, Y, ldy, O, ldo,
                                                                 C, ldc, tiles_c);
  else if (vec)
    plusmul_mfma_kernel<TILE, true, false><<<grid, block, 0, s>>>(R, Cn, K, X, ldx, Y, ldy, O,
                                                                  ldo, C, ldc, tiles_c);
  else if (vec_out)
    plusmul_mfma_kernel<TILE, false, true><<<grid, block, 0, s>>>(R, Cn, K, X, ldx, Y, ldy, O,
                                                                  ldo, C, ldc, tiles_c);
  else
    plusmul_mfma_kernel<TILE, false, false><<<grid, block, 0, s>>>(R, Cn, K, X, ldx, Y, ldy, O,
                                                                   ldo, C, ldc, tiles_c);
}

}  // namespace

hipError_t launch_plusmul_mfma(int R, int Cn, int K, const float *X,
                               int64_t ldx, const float *Y, int64_t ldy,
                               float *O, int64_t ldo, const float *C,
                               int64_t ldc, hipStream_t s) {
  auto aligned = [](const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
  const bool vec = K % 4 == 0 && Cn % 4 == 0 && ldx % 4 == 0 && ldy % 4 == 0 && aligned(X) &&
                   aligned(Y);
  // 16-byte stores (and C loads) when every row of O and C is 16-byte aligned
  const bool vec_out = ldo % 4 == 0 && aligned(O) && (!C || (ldc % 4 == 0 && aligned(C)));
  // 64x64 tiles when the 128x128 tiles would not give every CU one
  const int64_t tiles128 = (int64_t)((R + 127) / 128) * ((Cn + 127) / 128);
  if (tiles128 < device_cu_count())
    launch_tile<64>(vec, vec_out, R, Cn, K, X, ldx, Y, ldy, O, ldo, C, ldc, s);
  else
    launch_tile<128>(vec, vec_out, R, Cn, K, X, ldx, Y, ldy, O, ldo, C, ldc, s);
  return hipGetLastError();
}

}  // namespace simd2

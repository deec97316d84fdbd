// minus_square on the int8 matrix cores, exact, for operands that hold small
// integers.
//
// D = [C +] sum_k (X[r][k] - Y[k][c])^2 is computed as
//   ||x_r||^2 + ||y_c||^2 - 2 x_r . y_c
// with the dot products on v_mfma_i32_16x16x64_i8 (int8 in, int32
// accumulate) and the norms in int32.  When every X and Y entry is an integer
// in [-128, 127] every term is exact, so this is the exact integer S.  The
// sequential FP32 sum of the oracle (kernel/host_srgemm.h:34-48: acc = 0,
// acc += (a-b)^2 for k ascending) reaches S through partial sums that are
// integers no larger than S (the terms are >= 0); when S <= 2^24 each of them
// is a float, so the FP32 sum is S bit for bit and D = fl(S [+ C]) equals the
// oracle's.  A 64x64 output tile is written here only when all of its rows
// and columns passed the integer check and all of its sums are <= 2^24;
// otherwise its gate entry asks the VALU kernel to compute it the direct way.
//
// The caller supplies the workspace (minus_square_exact_workspace bytes);
// nothing here allocates.  Steps (one stream, no host synchronisation):
//   1. quantize: X rows and Y columns to int8 in fragment order (below),
//      zero-padded to 128 rows / columns and to a multiple of 64 in k, with
//      each row's / column's exact sum of squares (bit 31 set: some entry was
//      not an integer in [-128, 127]);
//   2. MFMA tiles 128x128 (4 waves x 64x64 quadrants of 4x4 16x16 blocks),
//      64-deep int8 slabs double-buffered in LDS, epilogue S -> float [+ C];
//   3. the VALU kernel on the 64x64 grid computes the tiles whose gate is 1
//      (srgemm_valu.hip; every other workgroup returns at once).
#include <stdint.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "internal.hpp"

// Scratch of the allocation-free entry point (minus_square_exact_noalloc):
// the same workspace layout, held by the library on every device it runs on
// until the process ends.  Host code never uses this name as an address; it
// takes the device address from hipGetSymbolAddress, per device.
static __device__ __align__(256) unsigned char scratch[simd2::MSQ_NOALLOC_BYTES];

namespace simd2 {

namespace {

typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int QT = 128;        // MFMA output tile
constexpr int QK = 64;         // k per slab = one MFMA k-step
constexpr int QSLAB = QT * QK; // bytes of one 128-row panel slab (8 fragment blocks)
constexpr int QR = 32;         // rows (or columns) per quantize workgroup
constexpr int QS = 256;        // k per quantize step
constexpr int QSLD = QS + 16;
constexpr int EXACT_LIMIT = 1 << 24;
constexpr int NUM_XCD = 8;

// Fragment order of the int8 operands.  A 16 x 64 block (16 rows or columns
// by 64 k) is 1 KB stored as the 64 lanes of v_mfma_i32_16x16x64_i8 read it:
// lane l = (row % 16) + 16 * (k % 64 / 16) holds bytes k % 16 = 0..15.  The
// eight blocks of a 128-row panel for one 64-k slab are consecutive (8 KB),
// and the slabs of a panel follow each other.  One slab of a workgroup tile
// is then one contiguous 8 KB run per operand, and global loads, LDS stores
// and fragment reads are all 16 bytes per lane at consecutive addresses.
__device__ __forceinline__ int64_t frag_offset(int row, int k, int Kp) {
  return ((int64_t)(row / QT) * (Kp / QK) + k / QK) * QSLAB + (row % QT / 16) * 1024 +
         ((row % 16) + 16 * (k % QK / 16)) * 16 + k % 16;
}

__device__ __forceinline__ bool small_int(float v) {
  return v == rintf(v) && v >= -128.0f && v <= 127.0f;
}

// Grid: Rq/32 workgroups for X rows, then Cq/32 for Y columns.  Rows in
// [n, Rq) are written as zeros; norms only for rows < n.  Every load of a
// step is issued before any is used (addresses clamped into the matrix,
// values outside it replaced by 0), so a step waits for memory once.
__global__ __launch_bounds__(256) void msq_quantize_kernel(
    int R, int Cn, int K, int Kp, const float *__restrict__ X, int64_t ldx,
    const float *__restrict__ Y, int64_t ldy, int8_t *__restrict__ Xq,
    int8_t *__restrict__ Yq, int *__restrict__ nx, int *__restrict__ ny, int bx, int kchunk) {
  __shared__ __attribute__((aligned(16))) int8_t T[QR][QSLD];
  __shared__ int rowbad[QR];
  const int tid = tid_x();
  const bool isx = bid_x() < bx;
  const int i0 = (isx ? bid_x() : bid_x() - bx) * QR;
  const int n = isx ? R : Cn;
  int8_t *Q = isx ? Xq : Yq;
  if (tid < QR) rowbad[tid] = 0;
  // write phase: row wi, bytes [wb, wb + 32) of the step
  const int wi = tid / 8, wb = (tid % 8) * 32;
  int norm = 0;
  uint32_t badmask = 0;  // X: bit j = row rg + 4j; Y: bit 0 = column cc
  constexpr int NV = QR * QS / 256;
  auto issue = [&](int k0, float (&v)[NV]) {
    if (isx) {
      // X[i][k], k contiguous: a wave reads 64 consecutive k of one row
      const int kk = tid % 64, rg = tid / 64;
#pragma unroll
      for (int j = 0; j < QR / 4; ++j) {
        const int i = min(i0 + rg + 4 * j, n - 1);
#pragma unroll
        for (int s = 0; s < QS / 64; ++s)
          v[j * (QS / 64) + s] = X[(int64_t)i * ldx + min(k0 + 64 * s + kk, K - 1)];
      }
    } else {
      // Y[k][c], c contiguous: half a wave reads 32 consecutive columns
      const int cc = tid % QR, kr = tid / QR;
      const int c = min(i0 + cc, n - 1);
#pragma unroll
      for (int j = 0; j < NV; ++j) v[j] = Y[(int64_t)min(k0 + kr + 8 * j, K - 1) * ldy + c];
    }
  };
  auto consume = [&](int k0, const float (&v)[NV]) {
    if (isx) {
      const int kk = tid % 64, rg = tid / 64;
#pragma unroll
      for (int j = 0; j < QR / 4; ++j) {
        const int ri = rg + 4 * j;
#pragma unroll
        for (int s = 0; s < QS / 64; ++s) {
          const int k = k0 + 64 * s + kk;
          const float x = (i0 + ri < n && k < K) ? v[j * (QS / 64) + s] : 0.0f;
          const bool ok = small_int(x);
          badmask |= ok ? 0u : 1u << j;
          T[ri][64 * s + kk] = ok ? (int8_t)(int)x : (int8_t)0;
        }
      }
    } else {
      const int cc = tid % QR, kr = tid / QR;
#pragma unroll
      for (int j = 0; j < NV; ++j) {
        const int k = k0 + kr + 8 * j;
        const float y = (i0 + cc < n && k < K) ? v[j] : 0.0f;
        const bool ok = small_int(y);
        badmask |= ok ? 0u : 1u;
        T[cc][kr + 8 * j] = ok ? (int8_t)(int)y : (int8_t)0;
      }
    }
    __syncthreads();
    const uint4 w0 = *reinterpret_cast<const uint4 *>(&T[wi][wb]);
    const uint4 w1 = *reinterpret_cast<const uint4 *>(&T[wi][wb + 16]);
    const uint32_t words[8] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
#pragma unroll
    for (int q = 0; q < 8; ++q) {
#pragma unroll
      for (int b = 0; b < 4; ++b) {
        const int e = (int)(int8_t)(words[q] >> (8 * b));
        norm += e * e;
      }
    }
    // Kp is a multiple of 64 (not of the 256 of a step): a 32-byte piece is
    // either wholly inside the row or wholly past it.
    if (k0 + wb < Kp) {
      *reinterpret_cast<uint4 *>(Q + frag_offset(i0 + wi, k0 + wb, Kp)) = w0;
      *reinterpret_cast<uint4 *>(Q + frag_offset(i0 + wi, k0 + wb + 16, Kp)) = w1;
    }
    __syncthreads();
  };
  // Two register sets: the loads of the next step are in flight while this
  // step is converted and stored.
  // This workgroup's share of k: [kbeg, kend), a whole number of step pairs
  // except at the end of the row.
  const int kbeg = bid_y() * kchunk, kend = min(Kp, kbeg + kchunk);
  float va[NV], vb[NV];
  issue(kbeg, va);
  for (int k0 = kbeg; k0 < kend; k0 += 2 * QS) {
    if (k0 + QS < kend) issue(k0 + QS, vb);
    consume(k0, va);
    if (k0 + QS < kend) {
      if (k0 + 2 * QS < kend) issue(k0 + 2 * QS, va);
      consume(k0 + QS, vb);
    }
  }
  if (badmask) {
    if (isx) {
#pragma unroll
      for (int j = 0; j < QR / 4; ++j)
        if (badmask & (1u << j)) rowbad[tid / 64 + 4 * j] = 1;
    } else {
      rowbad[tid % QR] = 1;
    }
  }
  __syncthreads();
  norm += __shfl_xor(norm, 1);
  norm += __shfl_xor(norm, 2);
  norm += __shfl_xor(norm, 4);
  // Partial norms of the k-chunks add up exactly (integers below 2^30, so
  // bit 31, the "not small integers" mark, is never reached by a carry).
  const int i = i0 + wi;
  if (tid % 8 == 0 && i < n) {
    int *dst = (isx ? nx : ny) + i;
    if (norm) atomicAdd(dst, norm);
    if (rowbad[wi]) atomicOr(dst, (int)0x80000000u);
  }
}

__device__ __forceinline__ void map_tile(int b, int tiles_r, int tiles_c, int &tr, int &tc) {
  const int nt = tiles_r * tiles_c;
  int t = b;
  if (nt % NUM_XCD == 0) t = (b % NUM_XCD) * (nt / NUM_XCD) + b / NUM_XCD;
  constexpr int GROUP = 4;  // 4 tile-rows per group: 5% faster than 8 at 4096^3
  const int width = GROUP * tiles_c;
  const int group = t / width;
  const int first = group * GROUP;
  const int gsz = min(tiles_r - first, GROUP);
  const int in = t - group * width;
  tr = first + in % gsz;
  tc = in / gsz;
}

// Fragment layout of v_mfma_i32_16x16x64_i8: lane l supplies 16 bytes of
// row l % 16 of the A block and 16 bytes of column l % 16 of the B block,
// both from k group l / 16; accumulator element v of lane l is block
// element (4 * (l / 16) + v, l % 16).  A and B take their k bytes in the same
// order, and integer sums are exact, so the order of k within a group does
// not change the result.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void msq_mfma_kernel(
    int R, int Cn, int Kp, const int8_t *__restrict__ Xq, const int8_t *__restrict__ Yq,
    const int *__restrict__ nx, const int *__restrict__ ny, float *O, int64_t ldo,
    const float *C, int64_t ldc, int *__restrict__ gate, int gate_ld, int tiles_r,
    int tiles_c, bool vec_out) {
  __shared__ __attribute__((aligned(16))) int8_t Xs[2][QSLAB];
  __shared__ __attribute__((aligned(16))) int8_t Ys[2][QSLAB];
  // Quadrant flags (rows 0-63, rows 64-127, columns 0-63, columns 64-127) in
  // the second X buffer, which nothing writes before the first slab's barrier.
  int *qbad = reinterpret_cast<int *>(&Xs[1][0]);
  const int tid = tid_x();
  const int lane = tid % 64;
  const int wave = tid / 64;
  const int wr = wave / 2, wc = wave % 2;
  int tr, tc;
  map_tile(bid_x(), tiles_r, tiles_c, tr, tc);
  const int row0 = tr * QT, col0 = tc * QT;

  // Quadrants with a row or column that failed the integer check go to the
  // VALU kernel; when all four do, skip the products.
  if (tid < 4) qbad[tid] = 0;
  __syncthreads();
  {
    const int t = tid % QT;
    if (tid < QT) {
      const int r = row0 + t;
      if (r < R && nx[r] < 0) qbad[t / 64] = 1;
    } else {
      const int c = col0 + t;
      if (c < Cn && ny[c] < 0) qbad[2 + t / 64] = 1;
    }
  }
  __syncthreads();
  int *g = gate + (int64_t)(tr * 2 + wr) * gate_ld + tc * 2 + wc;
  const bool mybad = qbad[wr] | qbad[2 + wc];
  if ((qbad[0] | qbad[2]) & (qbad[0] | qbad[3]) & (qbad[1] | qbad[2]) & (qbad[1] | qbad[3])) {
    if (lane == 0) *g = 1;
    return;
  }

  // slab ks of the tile's row panel: Xq + (tr * Kp / 64 + ks) * 8 KB
  const int8_t *xp = Xq + (int64_t)tr * (Kp / QK) * QSLAB + tid * 16;
  const int8_t *yp = Yq + (int64_t)tc * (Kp / QK) * QSLAB + tid * 16;
  uint4 x0, x1, y0, y1;
  auto load = [&]() {
    x0 = *reinterpret_cast<const uint4 *>(xp);
    x1 = *reinterpret_cast<const uint4 *>(xp + QSLAB / 2);
    y0 = *reinterpret_cast<const uint4 *>(yp);
    y1 = *reinterpret_cast<const uint4 *>(yp + QSLAB / 2);
  };
  auto stage = [&](int b) {
    *reinterpret_cast<uint4 *>(&Xs[b][tid * 16]) = x0;
    *reinterpret_cast<uint4 *>(&Xs[b][QSLAB / 2 + tid * 16]) = x1;
    *reinterpret_cast<uint4 *>(&Ys[b][tid * 16]) = y0;
    *reinterpret_cast<uint4 *>(&Ys[b][QSLAB / 2 + tid * 16]) = y1;
  };

  i32x4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = i32x4{0, 0, 0, 0};

  const int nks = Kp / QK;
  load();
  stage(0);
  __syncthreads();
  const int fr = lane % 16;
  for (int ks = 0; ks < nks; ++ks) {
    const int buf = ks & 1;
    const bool more = ks + 1 < nks;
    if (more) {
      xp += QSLAB;
      yp += QSLAB;
      load();
    }
    i32x4 a[4], b[4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
      a[i] = *reinterpret_cast<const i32x4 *>(&Xs[buf][(wr * 4 + i) * 1024 + lane * 16]);
#pragma unroll
    for (int j = 0; j < 4; ++j)
      b[j] = *reinterpret_cast<const i32x4 *>(&Ys[buf][(wc * 4 + j) * 1024 + lane * 16]);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[i], b[j], acc[i][j], 0, 0, 0);
    if (more) stage(buf ^ 1);
    __syncthreads();
  }

  // Epilogue of this wave's 64x64 quadrant: first whether every sum is
  // exact in FP32, then the stores.
  const int qr0 = row0 + wr * 64, qc0 = col0 + wc * 64;
  int nyv[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int c = qc0 + j * 16 + fr;
    nyv[j] = c < Cn ? ny[c] : 0;
  }
  // For valid rows and columns the sum s = sum_k (x - y)^2 is below
  // 65536 * 255^2 < 2^32, so it is exact in 32-bit unsigned arithmetic.
  bool bad = mybad;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = qr0 + i * 16 + (lane / 16) * 4 + v;
      const uint32_t nxv = r < R ? (uint32_t)nx[r] : 0u;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int c = qc0 + j * 16 + fr;
        const uint32_t s = nxv + (uint32_t)nyv[j] - 2u * (uint32_t)acc[i][j][v];
        if (r < R && c < Cn && s > (uint32_t)EXACT_LIMIT) bad = true;
      }
    }
  if (__ballot(bad) != 0) {
    if (lane == 0) *g = 1;
    return;
  }
  // Stores through LDS, one 16-row block of the quadrant at a time: the wave
  // writes the block's exact sums (as floats) to its own LDS region in the
  // accumulator layout and reads them back row-major, so the C loads and the
  // stores are 16 bytes per lane along rows.  The main loop's last barrier
  // has retired every read of Xs / Ys; only this wave touches its region, and
  // a wave's LDS operations complete in order, so no workgroup barrier is
  // needed (a wave may have left above).
  constexpr int TLD = 68;  // padded row: the 4 row groups of a store fall in distinct banks
  float *T = reinterpret_cast<float *>(wave < 2 ? &Xs[0][0] : &Ys[0][0]) + (wave & 1) * 16 * TLD;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int rr = (lane / 16) * 4 + v;
      const int r = qr0 + i * 16 + rr;
      const uint32_t nxv = r < R ? (uint32_t)nx[r] : 0u;
#pragma unroll
      for (int j = 0; j < 4; ++j)  // the gate loop's exact unsigned sum, <= 2^24 where stored
        T[rr * TLD + j * 16 + fr] = (float)(nxv + (uint32_t)nyv[j] - 2u * (uint32_t)acc[i][j][v]);
    }
    __builtin_amdgcn_wave_barrier();
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      const int idx = lane + 64 * h;
      const int rr = idx / 16, cc = (idx % 16) * 4;
      const int r = qr0 + i * 16 + rr, c = qc0 + cc;
      f32x4 d = *reinterpret_cast<const f32x4 *>(&T[rr * TLD + cc]);
      if (r >= R) continue;
      if (vec_out && c + 3 < Cn) {
        if (C) d += *reinterpret_cast<const f32x4 *>(C + (int64_t)r * ldc + c);
        *reinterpret_cast<f32x4 *>(O + (int64_t)r * ldo + c) = d;
      } else {
#pragma unroll
        for (int q = 0; q < 4; ++q)
          if (c + q < Cn) O[(int64_t)r * ldo + c + q] = C ? d[q] + C[(int64_t)r * ldc + c + q] : d[q];
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
  if (lane == 0) *g = 0;
}

// Workspace layout (caller's memory, 256-byte aligned pieces): Xq (Rq x Kp
// int8), Yq (Cq x Kp int8), nx (R ints), ny (Cn ints), gate ((Rq/64) x
// (Cq/64) ints).
struct WsLayout {
  int Kp, Rq, Cq, gate_ld, gate_n;
  size_t bx, by, bnx, bny, bgate;
  size_t bytes() const { return bx + by + bnx + bny + bgate; }
};

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

WsLayout ws_layout(int R, int Cn, int K) {
  WsLayout w;
  w.Kp = (K + QK - 1) / QK * QK;
  w.Rq = (R + QT - 1) / QT * QT;
  w.Cq = (Cn + QT - 1) / QT * QT;
  w.gate_ld = w.Cq / 64;
  w.gate_n = (w.Rq / 64) * w.gate_ld;
  w.bx = align256((size_t)w.Rq * w.Kp);
  w.by = align256((size_t)w.Cq * w.Kp);
  w.bnx = align256((size_t)R * 4);
  w.bny = align256((size_t)Cn * 4);
  w.bgate = align256((size_t)w.gate_n * 4);
  return w;
}

// K <= 65536 keeps every norm below 2^30 and every int32 dot product in range.
bool exact_eligible(int R, int Cn, int K) { return R > 0 && Cn > 0 && K > 0 && K <= 65536; }

}  // namespace

size_t minus_square_exact_workspace(int R, int Cn, int K) {
  return exact_eligible(R, Cn, K) ? ws_layout(R, Cn, K).bytes() : 0;
}

hipError_t launch_minus_square_exact(int R, int Cn, int K, const float *X, int64_t ldx,
                                     const float *Y, int64_t ldy, float *O, int64_t ldo,
                                     const float *C, int64_t ldc, void *ws, size_t ws_bytes,
                                     hipStream_t s) {
  Epilogue ep;
  ep.C = C;
  ep.ldc = ldc;
  if (!exact_eligible(R, Cn, K))
    return launch_valu_f32(8, R, Cn, K, X, ldx, Y, ldy, O, ldo, ep, s);
  const WsLayout w = ws_layout(R, Cn, K);
  if (!ws || ws_bytes < w.bytes()) return hipErrorInvalidValue;
  char *base = static_cast<char *>(ws);
  int8_t *Xq = reinterpret_cast<int8_t *>(base);
  int8_t *Yq = reinterpret_cast<int8_t *>(base + w.bx);
  int *nx = reinterpret_cast<int *>(base + w.bx + w.by);
  int *ny = reinterpret_cast<int *>(base + w.bx + w.by + w.bnx);
  int *gate = reinterpret_cast<int *>(base + w.bx + w.by + w.bnx + w.bny);

  // Quantize: rows / columns in groups of 32, K split over up to about 1024
  // workgroups in chunks of whole step pairs (512 k); the norms accumulate
  // into the zeroed nx / ny.
  const int qbx = w.Rq / QR, qby = w.Cq / QR;
  const int pairs = (w.Kp + 2 * QS - 1) / (2 * QS);
  const int want = std::max(1, std::min(pairs, (1024 + qbx + qby - 1) / (qbx + qby)));
  const int kchunk = ((pairs + want - 1) / want) * 2 * QS;
  const int ksplit = (w.Kp + kchunk - 1) / kchunk;
  hipError_t e = launch_fill_i32(nx, 0, R, s);
  if (e == hipSuccess) e = launch_fill_i32(ny, 0, Cn, s);
  if (e != hipSuccess) return e;
  msq_quantize_kernel<<<dim3(qbx + qby, ksplit), 256, 0, s>>>(R, Cn, K, w.Kp, X, ldx, Y, ldy, Xq,
                                                             Yq, nx, ny, qbx, kchunk);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int tiles_r = w.Rq / QT, tiles_c = w.Cq / QT;
  auto aligned = [](const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
  const bool vec_out = ldo % 4 == 0 && aligned(O) && (!C || (ldc % 4 == 0 && aligned(C)));
  msq_mfma_kernel<<<tiles_r * tiles_c, 256, 0, s>>>(R, Cn, w.Kp, Xq, Yq, nx, ny, O, ldo, C, ldc,
                                                    gate, w.gate_ld, tiles_r, tiles_c, vec_out);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  ep.gate = gate;
  ep.gate_ld = w.gate_ld;
  return launch_valu_f32(8, R, Cn, K, X, ldx, Y, ldy, O, ldo, ep, s, CFG_S64);
}

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex scratch_mutex;
void *scratch_ptr[MAX_DEVICES];  // device address of `scratch`, per device
}  // namespace

hipError_t minus_square_exact_noalloc(int R, int Cn, int K, const float *X, int64_t ldx,
                                      const float *Y, int64_t ldy, float *O, int64_t ldo,
                                      const float *C, int64_t ldc, hipStream_t s) {
  const size_t need = minus_square_exact_workspace(R, Cn, K);
  if (need > MSQ_NOALLOC_BYTES) return hipErrorInvalidValue;  // the caller checks first
  void *ws = nullptr;
  if (need > 0) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (!scratch_ptr[dev]) {
      // First call on this device (it is current): its copy of the array.
      void *p = nullptr;
      e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess) return e;
      if (!p) return hipErrorInvalidSymbol;
      scratch_ptr[dev] = p;
    }
    ws = scratch_ptr[dev];
  }
  // Every byte the kernels read is written earlier in the same call
  // (quantized operands with their padding, norms, one gate word per tile),
  // so the previous call's contents need no clearing.
  return launch_minus_square_exact(R, Cn, K, X, ldx, Y, ldy, O, ldo, C, ldc, ws, need, s);
}

hipError_t minus_square_exact_tile_counts(const void *ws, int R, int Cn, int K, hipStream_t s,
                                          int *mfma_tiles, int *valu_tiles) {
  *mfma_tiles = 0;
  *valu_tiles = 0;
  if (R <= 0 || Cn <= 0) return hipSuccess;
  const int tr = (R + 63) / 64, tc = (Cn + 63) / 64;
  if (!exact_eligible(R, Cn, K)) {
    *valu_tiles = tr * tc;
    return hipSuccess;
  }
  const WsLayout w = ws_layout(R, Cn, K);
  const char *gate = static_cast<const char *>(ws) + w.bx + w.by + w.bnx + w.bny;
  std::vector<int> g((size_t)tr * w.gate_ld);
  hipError_t e = hipStreamSynchronize(s);
  if (e == hipSuccess) e = hipMemcpy(g.data(), gate, g.size() * sizeof(int), hipMemcpyDeviceToHost);
  if (e != hipSuccess) return e;
  for (int r = 0; r < tr; ++r)
    for (int c = 0; c < tc; ++c) ++*(g[(size_t)r * w.gate_ld + c] ? valu_tiles : mfma_tiles);
  return hipSuccess;
}

}  // namespace simd2

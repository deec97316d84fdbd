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
//      zero-padded to 256 rows / 128 columns and to a multiple of 64 in k, with
//      each row's / column's exact sum of squares (bit 31 set: some entry was
//      not an integer in [-128, 127]);
//   2. MFMA tiles 256x128 (8 waves x 64x64 quadrants of 4x4 16x16 blocks),
//      64-deep int8 slabs copied straight into LDS, three in flight,
//      epilogue S -> float [+ C];
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

constexpr int QT = 128;        // rows (columns) of a fragment panel
constexpr int TR = 256;        // MFMA tile rows: two X panels
constexpr int TC = 128;        // MFMA tile columns: one Y panel
constexpr int NST = 3;         // LDS stages: slabs in flight
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
template <bool VEC>
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
    if (isx && VEC) {
      // X[i][k], k contiguous: lane l reads k0 + 4l .. +3 of a row, 16 bytes
      // (a wave: 1 KB of one row); past K element by element, clamped
      const int l = tid % 64, rg = tid / 64, k = k0 + 4 * l;
#pragma unroll
      for (int j = 0; j < QR / 4; ++j) {
        const float *row = X + (int64_t)min(i0 + rg + 4 * j, n - 1) * ldx;
        if (k + 3 < K) {
          const f32x4 q4 = *reinterpret_cast<const f32x4 *>(row + k);
#pragma unroll
          for (int q = 0; q < 4; ++q) v[j * 4 + q] = q4[q];
        } else {
#pragma unroll
          for (int q = 0; q < 4; ++q) v[j * 4 + q] = row[min(k + q, K - 1)];
        }
      }
    } else if (isx) {
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
      // Y[k][c], c contiguous: half a wave reads 32 consecutive columns; the
      // thread takes 4 consecutive k of its column at a time (one LDS word)
      const int cc = tid % QR, kq = tid / QR;
      const int c = min(i0 + cc, n - 1);
#pragma unroll
      for (int j = 0; j < NV / 4; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q)
          v[j * 4 + q] = Y[(int64_t)min(k0 + 4 * (kq + 8 * j) + q, K - 1) * ldy + c];
    }
  };
  // 4 consecutive k of one row or column, as int8 in one 32-bit word
  auto pack4 = [&](const float (&v)[NV], int at, int k, bool row_ok, uint32_t bad_bit) {
    uint32_t w = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float x = (row_ok && k + q < K) ? v[at + q] : 0.0f;
      const bool ok = small_int(x);
      badmask |= ok ? 0u : bad_bit;
      w |= (uint32_t)(uint8_t)(ok ? (int8_t)(int)x : (int8_t)0) << (8 * q);
    }
    return w;
  };
  auto consume = [&](int k0, const float (&v)[NV]) {
    if (isx && VEC) {
      const int l = tid % 64, rg = tid / 64;
#pragma unroll
      for (int j = 0; j < QR / 4; ++j) {
        const int ri = rg + 4 * j;
        *reinterpret_cast<uint32_t *>(&T[ri][4 * l]) =
            pack4(v, j * 4, k0 + 4 * l, i0 + ri < n, 1u << j);
      }
    } else if (isx) {
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
      const int cc = tid % QR, kq = tid / QR;
#pragma unroll
      for (int j = 0; j < NV / 4; ++j) {
        const int m = kq + 8 * j;
        *reinterpret_cast<uint32_t *>(&T[cc][4 * m]) =
            pack4(v, j * 4, k0 + 4 * m, i0 + cc < n, 1u);
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
// One direct global -> LDS copy of 16 bytes per lane: lane l's bytes land at
// lds + 16 l, lds being the wave's base (the same in every lane).  The copy
// counts as an outstanding vector-memory load until the bytes are in LDS.
__device__ __forceinline__ void lds_copy16(const int8_t *g, int8_t *lds) {
  __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)g,
                                   (__attribute__((address_space(3))) void *)lds, 16, 0, 0);
}

// Wait until at most N of this thread's vector-memory loads are outstanding
// (they complete in issue order) and its LDS reads have returned, then the
// workgroup barrier.  The memory clobber keeps the compiler from moving
// memory accesses across it; no fence is implied, so copies issued earlier
// may still be in flight, as intended.
template <int N>
__device__ __forceinline__ void wait_copies_barrier() {
  asm volatile("s_waitcnt vmcnt(%0) lgkmcnt(0)\n\ts_barrier" ::"n"(N) : "memory");
}

// The wave's four X and four Y fragments of a slab (16 bytes per lane, blocks
// 1 KB apart), waited for in the same statement.  Written out because the
// compiler, seeing LDS reads of memory that direct copies write, would wait
// for every copy in flight, the next slabs' included, before each read.
__device__ __forceinline__ void read_fragments(const int8_t *xs, const int8_t *ys, i32x4 (&a)[4],
                                               i32x4 (&b)[4]) {
  const uint32_t xo = (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) int8_t *)xs;
  const uint32_t yo = (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) int8_t *)ys;
  asm volatile(
      "ds_read_b128 %0, %8\n\t"
      "ds_read_b128 %1, %8 offset:1024\n\t"
      "ds_read_b128 %2, %8 offset:2048\n\t"
      "ds_read_b128 %3, %8 offset:3072\n\t"
      "ds_read_b128 %4, %9\n\t"
      "ds_read_b128 %5, %9 offset:1024\n\t"
      "ds_read_b128 %6, %9 offset:2048\n\t"
      "ds_read_b128 %7, %9 offset:3072\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "=&v"(a[0]), "=&v"(a[1]), "=&v"(a[2]), "=&v"(a[3]), "=&v"(b[0]), "=&v"(b[1]),
        "=&v"(b[2]), "=&v"(b[3])
      : "v"(xo), "v"(yo)
      : "memory");
}

// Tiles of 256 rows x 128 columns, 8 waves of 64x64 quadrants.  The 64-k
// slabs (two X panels' 8 KB runs, one Y panel's) are copied straight into
// LDS, three slabs in flight: slab ks + 2 is issued right after the barrier
// that publishes slab ks, into the stage slab ks - 1 left, so each slab has
// two slabs' products to arrive in.  72 KB of LDS: two workgroups per CU.
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) void msq_mfma_kernel(
    int R, int Cn, int Kp, const int8_t *__restrict__ Xq, const int8_t *__restrict__ Yq,
    const int *__restrict__ nx, const int *__restrict__ ny, float *O, int64_t ldo,
    const float *C, int64_t ldc, int *__restrict__ gate, int gate_ld, int tiles_r,
    int tiles_c, bool vec_out) {
  // stage st: the X slab (two panels) at st * 3 QSLAB, the Y slab after it
  __shared__ __attribute__((aligned(16))) int8_t smem[NST * 3 * QSLAB];
  __shared__ int qbad[6];  // row quadrants 0-3, column quadrants 0-1
  __shared__ int badw[8];  // waves with a sum past 2^24
  const int tid = tid_x();
  const int lane = tid % 64;
  const int wave = tid / 64;
  const int wr = wave / 2, wc = wave % 2;
  int tr, tc;
  map_tile(bid_x(), tiles_r, tiles_c, tr, tc);
  const int row0 = tr * TR, col0 = tc * TC;

  // Quadrants with a row or column that failed the integer check go to the
  // VALU kernel; when all eight do, skip the products.
  if (tid < 6) qbad[tid] = 0;
  __syncthreads();
  if (tid < TR) {
    const int r = row0 + tid;
    if (r < R && nx[r] < 0) qbad[tid / 64] = 1;
  } else if (tid < TR + TC) {
    const int c = col0 + tid - TR;
    if (c < Cn && ny[c] < 0) qbad[4 + (tid - TR) / 64] = 1;
  }
  __syncthreads();
  int *g = gate + (int64_t)(tr * 4 + wr) * gate_ld + tc * 2 + wc;
  const bool mybad = qbad[wr] | qbad[4 + wc];
  bool allbad = true;
#pragma unroll
  for (int q = 0; q < 4; ++q) allbad = allbad && (qbad[q] | qbad[4]) && (qbad[q] | qbad[5]);
  if (allbad) {
    if (lane == 0) *g = 1;
    return;
  }

  // Wave w copies fragment block w of each slab: of X panels 2 tr and
  // 2 tr + 1 (Rq is a multiple of 256, so both exist) and of Y panel tc.
  const int nks = Kp / QK;
  const int8_t *xa = Xq + (int64_t)(2 * tr) * nks * QSLAB + wave * 1024 + lane * 16;
  const int8_t *xb = xa + (int64_t)nks * QSLAB;
  const int8_t *ya = Yq + (int64_t)tc * nks * QSLAB + wave * 1024 + lane * 16;
  auto issue = [&](int ks, int st) {
    const int64_t o = (int64_t)ks * QSLAB;
    int8_t *x = &smem[st * 3 * QSLAB];
    lds_copy16(xa + o, x + wave * 1024);
    lds_copy16(xb + o, x + QSLAB + wave * 1024);
    lds_copy16(ya + o, x + 2 * QSLAB + wave * 1024);
  };

  i32x4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = i32x4{0, 0, 0, 0};

  issue(0, 0);
  if (nks > 1) issue(1, 1);
  int st = 0;
  for (int ks = 0; ks < nks; ++ks) {
    // slab ks is in LDS once at most the next slab's 3 copies are outstanding
    if (ks + 1 < nks)
      wait_copies_barrier<3>();
    else
      wait_copies_barrier<0>();
    // every wave read slab ks - 1 before this barrier: its stage takes ks + 2
    if (ks + 2 < nks) issue(ks + 2, st == 0 ? 2 : st - 1);
    i32x4 a[4], b[4];
    const int8_t *x = &smem[st * 3 * QSLAB];
    read_fragments(x + wr * 4 * 1024 + lane * 16, x + 2 * QSLAB + wc * 4 * 1024 + lane * 16, a, b);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[i], b[j], acc[i][j], 0, 0, 0);
    st = st == NST - 1 ? 0 : st + 1;
  }
  // the epilogue reuses the stages: every wave's reads of the last slab come first
  wait_copies_barrier<0>();
  const int fr = lane % 16;

  // Epilogue of this wave's 64x64 quadrant: first whether every sum is
  // exact in FP32, then the stores.
  const int qr0 = row0 + wr * 64, qc0 = col0 + wc * 64;
  int nyv[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int c = qc0 + j * 16 + fr;
    nyv[j] = c < Cn ? ny[c] : 0;
  }
  // Lane l holds the norm of the quadrant's row l (one load per lane); the
  // accumulator rows take theirs by a lane permute.
  const uint32_t nx_lane = qr0 + lane < R ? (uint32_t)nx[qr0 + lane] : 0u;
  auto row_norm = [&](int i, int v) {
    return (uint32_t)__shfl((int)nx_lane, i * 16 + (lane / 16) * 4 + v);
  };
  // For valid rows and columns the sum s = sum_k (x - y)^2 is below
  // 65536 * 255^2 < 2^32, so it is exact in 32-bit unsigned arithmetic.
  bool bad = mybad;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int r = qr0 + i * 16 + (lane / 16) * 4 + v;
      const uint32_t nxv = row_norm(i, v);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int c = qc0 + j * 16 + fr;
        const uint32_t s = nxv + (uint32_t)nyv[j] - 2u * (uint32_t)acc[i][j][v];
        bad = bad | ((r < R) & (c < Cn) & (s > (uint32_t)EXACT_LIMIT));
      }
    }
  const bool wave_bad = __ballot(bad) != 0;

  // Interior tiles with C (rows 16-byte aligned): C is copied straight into
  // LDS, half the tile at a time (rows wr * 64 + h * 32 .. + 31 of every row
  // quadrant: 128 rows x 512 bytes, two rows per 1 KB copy, 1056 bytes apart
  // so that the 4 row groups of an accumulator-layout access fall in distinct
  // banks); each wave adds its exact sums in place and stores its 32 x 64
  // block from there, 16 bytes per lane along rows.  The whole half's C is in
  // flight at once, where loading it per 16-row block waited for memory 4
  // times in turn.  Taken only when no wave of the tile failed the check.
  if (C && vec_out && row0 + TR <= R && col0 + TC <= Cn) {
    if (lane == 0) badw[wave] = wave_bad ? 1 : 0;
    __syncthreads();
    int anybad = 0;
#pragma unroll
    for (int w = 0; w < 8; ++w) anybad |= badw[w];
    if (!anybad) {
      constexpr int CP = 1056;
      static_assert(64 * CP <= NST * 3 * QSLAB, "half a C tile exceeds the stages");
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (h) wait_copies_barrier<0>();  // every wave has stored the first half
#pragma unroll
        for (int q = 0; q < 8; ++q) {
          const int m = wave * 8 + q;        // copy m: local rows 2m, 2m + 1
          const int lr = 2 * m + lane / 32;  // local row l: tile row (l/32)*64 + h*32 + l%32
          const int trow = (lr / 32) * 64 + h * 32 + lr % 32;
          lds_copy16(reinterpret_cast<const int8_t *>(C + (int64_t)(row0 + trow) * ldc + col0 +
                                                      (lane % 32) * 4),
                     &smem[m * CP]);
        }
        wait_copies_barrier<0>();
#pragma unroll
        for (int ii = 0; ii < 2; ++ii) {
          const int i = 2 * h + ii;
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int rr = (lane / 16) * 4 + v;
            const int lr = wr * 32 + ii * 16 + rr;
            const uint32_t nxv = row_norm(i, v);
            float *pc = reinterpret_cast<float *>(&smem[(lr / 2) * CP + (lr % 2) * 512]) + wc * 64 + fr;
#pragma unroll
            for (int j = 0; j < 4; ++j)  // the gate loop's exact sum, <= 2^24
              pc[j * 16] += (float)(nxv + (uint32_t)nyv[j] - 2u * (uint32_t)acc[i][j][v]);
          }
        }
        // only this wave wrote its block, and its LDS operations complete in order
#pragma unroll
        for (int hh = 0; hh < 8; ++hh) {
          const int idx = lane + 64 * hh;
          const int row = idx / 16, cc = (idx % 16) * 4;
          const int lr = wr * 32 + row;
          const f32x4 d = *reinterpret_cast<const f32x4 *>(
              &smem[(lr / 2) * CP + (lr % 2) * 512 + (wc * 64 + cc) * 4]);
          *reinterpret_cast<f32x4 *>(O + (int64_t)(qr0 + h * 32 + row) * ldo + qc0 + cc) = d;
        }
      }
      if (lane == 0) *g = 0;
      return;
    }
  }
  if (wave_bad) {
    if (lane == 0) *g = 1;
    return;
  }
  // Stores through LDS, one 16-row block of the quadrant at a time: the wave
  // writes the block's exact sums (as floats) to its own LDS region in the
  // accumulator layout and reads them back row-major, so the C loads and the
  // stores are 16 bytes per lane along rows.  The barrier after the main loop
  // has retired every read of the stages; only this wave touches its region
  // (8 x 4352 bytes), and a wave's LDS operations complete in order, so no
  // workgroup barrier is needed (a wave may have left above).
  constexpr int TLD = 68;  // padded row: the 4 row groups of a store fall in distinct banks
  static_assert(8 * 16 * TLD * 4 <= NST * 3 * QSLAB, "store regions exceed the stages");
  float *T = reinterpret_cast<float *>(&smem[0]) + wave * 16 * TLD;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int rr = (lane / 16) * 4 + v;
      const uint32_t nxv = row_norm(i, v);
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
  w.Rq = (R + TR - 1) / TR * TR;
  w.Cq = (Cn + TC - 1) / TC * TC;
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
  // 128-bit loads of X rows when they are 16-byte aligned
  if (ldx % 4 == 0 && (reinterpret_cast<uintptr_t>(X) & 15) == 0)
    msq_quantize_kernel<true><<<dim3(qbx + qby, ksplit), 256, 0, s>>>(
        R, Cn, K, w.Kp, X, ldx, Y, ldy, Xq, Yq, nx, ny, qbx, kchunk);
  else
    msq_quantize_kernel<false><<<dim3(qbx + qby, ksplit), 256, 0, s>>>(
        R, Cn, K, w.Kp, X, ldx, Y, ldy, Xq, Yq, nx, ny, qbx, kchunk);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int tiles_r = w.Rq / TR, tiles_c = w.Cq / TC;
  auto aligned = [](const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
  const bool vec_out = ldo % 4 == 0 && aligned(O) && (!C || (ldc % 4 == 0 && aligned(C)));
  msq_mfma_kernel<<<tiles_r * tiles_c, 512, 0, s>>>(R, Cn, w.Kp, Xq, Yq, nx, ny, O, ldo, C, ldc,
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

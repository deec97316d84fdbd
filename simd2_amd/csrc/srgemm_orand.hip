// or_and semiring on gfx950: D = [C |] OR_k A(i,k) & B(k,j).
//
// Packed kernel (simd2_orand_packed): rows of A are K-bit vectors, rows of
// B and D are N-bit vectors, 32 columns per word.  Row i of D is the OR of
// the rows k of B for which bit A(i,k) is set.  A thread owns TR rows x TW
// words of D; per k it turns its rows' A bits into all-ones / zero masks with
// one v_bfe_i32 each and folds B's words in with one v_and_or_b32 per word,
// i.e. 32 (i,j,k) terms per instruction.  A B slab of 32 rows (one A word per
// row of the tile) sits in LDS; A words come straight from global memory.
// A whole A word that is zero skips its 32 k-steps (wave-uniform test), which
// pays off on sparse graphs.  When the output has too few tiles to fill the
// GPU (4096 x 4096 bits is 128 tiles), K is split over several workgroups per
// tile whose partial words are OR-ed into D with atomics (OR is idempotent
// and commutative, so the result is exact): D is first set to C (or zero) and
// the change flag is taken by a pass of its own afterwards.
//
// int32 kernel (cuasr_orand_srsgemm drop-in): booleans stored one per int.
// The kernel packs on the fly: X rows (k contiguous) are packed along k with
// wave ballots, Y columns are packed along k by each lane walking its own
// column, and a thread then ORs the AND of 32-k words per output element
// (v_and_or_b32 again, 32 terms per instruction).  Its cost is the 4-byte
// reads of 1-bit data, which is why the packed form exists.
#include <algorithm>

#include "internal.hpp"

namespace simd2 {

namespace {

// acc | (a & b) as one instruction; written as asm because the compiler
// otherwise pairs terms into v_and + v_or3 (1.5 instructions per term).
__device__ __forceinline__ uint32_t and_or(uint32_t a, uint32_t b, uint32_t acc) {
  uint32_t r;
  asm("v_and_or_b32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(acc));
  return r;
}

constexpr int PK_TR = 4;     // rows per thread
constexpr int PK_TW = 4;     // words per thread
constexpr int PK_NTR = 16;   // threads along rows
constexpr int PK_NTW = 16;   // threads along words
constexpr int PK_BM = PK_TR * PK_NTR;   // 64 rows per tile
constexpr int PK_BW = PK_TW * PK_NTW;   // 64 words (2048 columns) per tile
constexpr int PK_BK = 32;               // k per slab (one A word)

__global__ __launch_bounds__(256) void orand_packed_kernel(
    int M, int NW, int K, const uint32_t *__restrict__ A, int64_t lda,
    const uint32_t *__restrict__ B, int64_t ldb, const uint32_t *C,
    int64_t ldc, uint32_t *D, int64_t ldd, const uint32_t *P, int64_t ldp,
    int *flag, uint32_t last_mask, int tiles_w, int ksplit, int kw_per) {
  __shared__ __attribute__((aligned(16))) uint32_t Bs[2][PK_BK][PK_BW];
  const int tid = tid_x();
  const int tw = tid % PK_NTW;
  const int trr = tid / PK_NTW;
  const int tile = bid_x() / ksplit, split = bid_x() % ksplit;
  const int tile_r = tile / tiles_w;
  const int tile_w = tile % tiles_w;
  const int row0 = tile_r * PK_BM + trr * PK_TR;
  const int w0 = tile_w * PK_BW;

  uint32_t acc[PK_TR][PK_TW];
#pragma unroll
  for (int i = 0; i < PK_TR; ++i)
#pragma unroll
    for (int j = 0; j < PK_TW; ++j) acc[i][j] = 0u;

  const int nkw = (K + 31) / 32;
  const int kw0 = split * kw_per;
  const int kw_end = min(nkw, kw0 + kw_per);
  // Pipeline over the K-words of the split, one barrier per word: the A words
  // of word kw + 1 are loaded one step ahead (their tile-wide "any bit set"
  // is taken at the barrier), and the B slab of word kw + 1 is loaded during
  // word kw's arithmetic when that word is needed; B slabs alternate between
  // two LDS buffers.  A word whose A bits are zero across the tile is skipped
  // without loading its B slab.
  auto load_a = [&](int kw, uint32_t (&aw)[PK_TR]) {
#pragma unroll
    for (int i = 0; i < PK_TR; ++i) {
      const int r = row0 + i;
      aw[i] = (r < M && kw < kw_end) ? A[(int64_t)r * lda + kw] : 0u;
      if (kw == nkw - 1 && (K & 31)) aw[i] &= (1u << (K & 31)) - 1u;
    }
  };
  constexpr int BPER = PK_BK * PK_BW / 256;  // B words per thread per slab
  uint32_t bpre[BPER];
  auto load_b = [&](int kw) {
#pragma unroll
    for (int h = 0; h < BPER; ++h) {
      const int f = tid + 256 * h;
      const int kk = f / PK_BW, ww = f % PK_BW;
      const int k = kw * 32 + kk, w = w0 + ww;
      bpre[h] = (k < K && w < NW) ? B[(int64_t)k * ldb + w] : 0u;
    }
  };
  auto any_of = [](const uint32_t (&aw)[PK_TR]) {
    uint32_t x = 0;
#pragma unroll
    for (int i = 0; i < PK_TR; ++i) x |= aw[i];
    return x != 0;
  };
  uint32_t aw_cur[PK_TR], aw_nxt[PK_TR];
  load_a(kw0, aw_cur);
  bool need_cur = __syncthreads_or(any_of(aw_cur));
  if (need_cur) load_b(kw0);
  load_a(kw0 + 1, aw_nxt);
  int buf = 0;
  for (int kw = kw0; kw < kw_end; ++kw) {
    if (need_cur) {
#pragma unroll
      for (int h = 0; h < BPER; ++h) {
        const int f = tid + 256 * h;
        Bs[buf][f / PK_BW][f % PK_BW] = bpre[h];
      }
    }
    // publishes Bs[buf]; the buffer written here was last read two words ago,
    // before the previous word's barrier
    const bool need_nxt = __syncthreads_or(any_of(aw_nxt));
    if (need_nxt) load_b(kw + 1);
    uint32_t aw_nn[PK_TR];
    load_a(kw + 2, aw_nn);
    if (need_cur) {
#pragma unroll
      for (int kk = 0; kk < PK_BK; ++kk) {
        uint32_t bw[PK_TW];
#pragma unroll
        for (int j = 0; j < PK_TW; ++j) bw[j] = Bs[buf][kk][tw * PK_TW + j];
#pragma unroll
        for (int i = 0; i < PK_TR; ++i) {
          // bit kk of the A word as an all-ones / all-zeros mask (v_bfe_i32)
          const uint32_t m = (uint32_t)((int32_t)(aw_cur[i] << (31 - kk)) >> 31);
#pragma unroll
          for (int j = 0; j < PK_TW; ++j) acc[i][j] = and_or(m, bw[j], acc[i][j]);
        }
      }
    }
#pragma unroll
    for (int i = 0; i < PK_TR; ++i) {
      aw_cur[i] = aw_nxt[i];
      aw_nxt[i] = aw_nn[i];
    }
    need_cur = need_nxt;
    buf ^= 1;
  }

  if (ksplit > 1) {  // partial words into D (set to C or zero beforehand)
#pragma unroll
    for (int i = 0; i < PK_TR; ++i) {
      const int r = row0 + i;
#pragma unroll
      for (int j = 0; j < PK_TW; ++j) {
        const int w = w0 + tw * PK_TW + j;
        const uint32_t v = acc[i][j] & (w == NW - 1 ? last_mask : ~0u);
        if (r < M && w < NW && v) atomicOr(&D[(int64_t)r * ldd + w], v);
      }
    }
    return;
  }
  bool changed = false;
#pragma unroll
  for (int i = 0; i < PK_TR; ++i) {
    const int r = row0 + i;
    if (r >= M) continue;
#pragma unroll
    for (int j = 0; j < PK_TW; ++j) {
      const int w = w0 + tw * PK_TW + j;
      if (w >= NW) continue;
      uint32_t v = acc[i][j];
      if (C) v |= C[(int64_t)r * ldc + w];
      if (w == NW - 1) v &= last_mask;
      if (P) changed |= (v != (P[(int64_t)r * ldp + w] & (w == NW - 1 ? last_mask : ~0u)));
      D[(int64_t)r * ldd + w] = v;
    }
  }
  if (flag) {
    const bool any = __any(changed);
    if (any && (tid % warpSize) == 0) atomicOr(flag, 1);
  }
}

// D = C (or zero), last word masked: the start of a K-split product.
__global__ __launch_bounds__(256) void orand_init_kernel(int M, int NW, const uint32_t *C,
                                                         int64_t ldc, uint32_t *D, int64_t ldd,
                                                         uint32_t last_mask) {
  const int64_t total = (int64_t)M * NW;
  for (int64_t e = (int64_t)bid_x() * 256 + tid_x(); e < total; e += (int64_t)gridDim.x * 256) {
    const int r = (int)(e / NW), w = (int)(e % NW);
    const uint32_t v = C ? C[(int64_t)r * ldc + w] : 0u;
    D[(int64_t)r * ldd + w] = v & (w == NW - 1 ? last_mask : ~0u);
  }
}

// *flag = 1 when D differs from P (last word masked): the change test of a
// K-split product.  One atomic per workgroup at most (a workgroup-wide OR
// first): when most words changed, one per wave serialised on the flag.
__global__ __launch_bounds__(256) void orand_diff_kernel(int M, int NW, const uint32_t *D,
                                                         int64_t ldd, const uint32_t *P,
                                                         int64_t ldp, uint32_t last_mask,
                                                         int *flag) {
  const int64_t total = (int64_t)M * NW;
  bool changed = false;
  for (int64_t e = (int64_t)bid_x() * 256 + tid_x(); e < total; e += (int64_t)gridDim.x * 256) {
    const int r = (int)(e / NW), w = (int)(e % NW);
    const uint32_t m = w == NW - 1 ? last_mask : ~0u;
    changed |= (D[(int64_t)r * ldd + w] & m) != (P[(int64_t)r * ldp + w] & m);
  }
  if (__syncthreads_or(changed) && tid_x() == 0) atomicOr(flag, 1);
}

// ---- int32 booleans -------------------------------------------------------
constexpr int IB_BM = 64;   // output rows per tile
constexpr int IB_BN = 64;   // output columns per tile
constexpr int IB_BKW = 8;   // 32-bit k-words per slab (256 k)

__global__ __launch_bounds__(256) void orand_i32_kernel(
    int R, int Cn, int K, const int *__restrict__ X, int64_t ldx,
    const int *__restrict__ Y, int64_t ldy, const int *C, int64_t ldc,
    int *O, int64_t ldo, const int *P, int64_t ldp, int *flag, int tiles_c) {
  __shared__ uint32_t Xp[IB_BM][IB_BKW + 1];   // rows packed along k
  __shared__ uint32_t Yp[IB_BN][IB_BKW + 1];   // columns packed along k
  const int tid = tid_x();
  const int lane = tid % 64;
  const int wave = tid / 64;
  const int row0 = (bid_x() / tiles_c) * IB_BM;
  const int col0 = (bid_x() % tiles_c) * IB_BN;
  const int tr = tid / 16;  // 16 x 16 threads, 4 x 4 outputs each
  const int tc = tid % 16;

  uint32_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0u;

  for (int k0 = 0; k0 < K; k0 += IB_BKW * 32) {
    // Loads are issued in batches before any is used (addresses clamped into
    // the matrix, values outside it taken as false), so a batch waits for
    // memory once.
    // X: each wave packs 16 rows, 4 at a time; per row, 4 ballots of 64
    // consecutive k.
    for (int rr = 0; rr < IB_BM / 4; rr += 4) {
      int v[4][IB_BKW / 2];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = min(row0 + wave * (IB_BM / 4) + rr + i, R - 1);
#pragma unroll
        for (int ch = 0; ch < IB_BKW / 2; ++ch)
          v[i][ch] = X[(int64_t)r * ldx + min(k0 + ch * 64 + lane, K - 1)];
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int lr = wave * (IB_BM / 4) + rr + i;
#pragma unroll
        for (int ch = 0; ch < IB_BKW / 2; ++ch) {
          const int k = k0 + ch * 64 + lane;
          const uint64_t m = __ballot(row0 + lr < R && k < K && v[i][ch] != 0);
          if (lane == 0) {
            Xp[lr][ch * 2] = (uint32_t)m;
            Xp[lr][ch * 2 + 1] = (uint32_t)(m >> 32);
          }
        }
      }
    }
    // Y: lane = column, each wave walks 64 k (two words) of its quarter.
    {
      const int c = col0 + lane, cc = min(c, Cn - 1);
#pragma unroll
      for (int half = 0; half < 2; ++half) {
        const int kb = k0 + wave * 64 + half * 32;
        int v[32];
#pragma unroll
        for (int q = 0; q < 32; ++q) v[q] = Y[(int64_t)min(kb + q, K - 1) * ldy + cc];
        uint32_t wv = 0;
#pragma unroll
        for (int q = 0; q < 32; ++q)
          wv |= (uint32_t)(c < Cn && kb + q < K && v[q] != 0) << q;
        Yp[lane][wave * 2 + half] = wv;
      }
    }
    __syncthreads();
#pragma unroll
    for (int w = 0; w < IB_BKW; ++w) {
      uint32_t xa[4], yb[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) xa[i] = Xp[tr * 4 + i][w];
#pragma unroll
      for (int j = 0; j < 4; ++j) yb[j] = Yp[tc * 4 + j][w];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = and_or(xa[i], yb[j], acc[i][j]);
    }
    __syncthreads();
  }

  bool changed = false;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int r = row0 + tr * 4 + i;
    if (r >= R) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int c = col0 + tc * 4 + j;
      if (c >= Cn) continue;
      int v = acc[i][j] != 0u;
      if (C) v = v || (C[(int64_t)r * ldc + c] != 0);
      if (P) changed |= (v != P[(int64_t)r * ldp + c]);
      O[(int64_t)r * ldo + c] = v;
    }
  }
  if (flag) {
    const bool any = __any(changed);
    if (any && (tid % warpSize) == 0) atomicOr(flag, 1);
  }
}

}  // namespace

hipError_t launch_orand_packed(int M, int N, int K, const uint32_t *A,
                               int64_t lda, const uint32_t *B, int64_t ldb,
                               const uint32_t *C, int64_t ldc, uint32_t *D,
                               int64_t ldd, const uint32_t *P, int64_t ldp,
                               int *flag, hipStream_t s) {
  const int NW = (N + 31) / 32;
  const uint32_t last_mask = (N & 31) ? ((1u << (N & 31)) - 1u) : ~0u;
  const int tiles_r = (M + PK_BM - 1) / PK_BM;
  const int tiles_w = (NW + PK_BW - 1) / PK_BW;
  const int tiles = tiles_r * tiles_w;
  const int nkw = (K + 31) / 32;
  // about 1024 workgroups: split K when the tiles alone are fewer
  int ksplit = std::max(1, std::min({(1024 + tiles - 1) / tiles, 16, nkw}));
  const int kw_per = (nkw + ksplit - 1) / std::max(ksplit, 1);
  if (nkw > 0) ksplit = (nkw + kw_per - 1) / kw_per;
  if (M <= 0 || NW <= 0) return hipSuccess;
  const int64_t words = (int64_t)M * NW;
  const int egrid = (int)std::min<int64_t>((words + 255) / 256, 2048);
  if (ksplit > 1)
    orand_init_kernel<<<dim3(egrid), dim3(256), 0, s>>>(M, NW, C, ldc, D, ldd, last_mask);
  orand_packed_kernel<<<dim3(tiles * ksplit), dim3(256), 0, s>>>(
      M, NW, K, A, lda, B, ldb, C, ldc, D, ldd, P, ldp, flag, last_mask, tiles_w, ksplit,
      kw_per);
  if (ksplit > 1 && P && flag)
    orand_diff_kernel<<<dim3(egrid), dim3(256), 0, s>>>(M, NW, D, ldd, P, ldp, last_mask, flag);
  return hipGetLastError();
}

namespace {

// int32 booleans (nonzero = 1) <-> packed rows (bit c % 32 of word c / 32).
// Pack: a wave takes 256 columns of a row -- four 64-column loads in flight,
// four ballots, eight words; a wave that saw a value other than 0 or 1 sets
// *nonbool once.  Unpack: a thread writes 4 consecutive columns.  Grids are
// 2D (column blocks x rows, rows strided past 65535), so no index needs a
// 64-bit division.
constexpr int PACK_COLS = 1024;  // columns per pack workgroup (4 waves x 256)

__global__ __launch_bounds__(256) void pack_i32_kernel(int rows, int cols, const int *__restrict__ X,
                                                       int64_t ldx, uint32_t *P, int64_t ldp,
                                                       int *nonbool) {
  const int lane = tid_x() % 64;
  const int nw = (cols + 31) / 32;
  const int c0 = bid_x() * PACK_COLS + (tid_x() / 64) * 256;
  bool odd = false;
  for (int r = bid_y(); r < rows; r += gridDim.y) {
    const int *x = X + (int64_t)r * ldx;
    int v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int c = c0 + q * 64 + lane;
      v[q] = c < cols ? x[c] : 0;
    }
    uint64_t b[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      odd |= (unsigned)v[q] > 1u;
      b[q] = __ballot(v[q] != 0);
    }
    const int word = c0 / 32 + lane;
    if (lane < 8 && word < nw) {
      const uint64_t bq = (lane >> 1) == 0 ? b[0] : (lane >> 1) == 1 ? b[1] : (lane >> 1) == 2 ? b[2] : b[3];
      P[(int64_t)r * ldp + word] = (lane & 1) ? (uint32_t)(bq >> 32) : (uint32_t)bq;
    }
  }
  if (nonbool && __any(odd) && lane == 0) atomicOr(nonbool, 1);
}

__global__ __launch_bounds__(256) void unpack_i32_kernel(int rows, int cols, const uint32_t *P,
                                                         int64_t ldp, int *X, int64_t ldx) {
  const int c = (bid_x() * 256 + tid_x()) * 4;
  if (c >= cols) return;
  for (int r = bid_y(); r < rows; r += gridDim.y) {
    const uint32_t *p = P + (int64_t)r * ldp;
    int *x = X + (int64_t)r * ldx;
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (c + q < cols) x[c + q] = (int)((p[(c + q) / 32] >> ((c + q) % 32)) & 1u);
  }
}

}  // namespace

hipError_t launch_pack_i32(int rows, int cols, const int *X, int64_t ldx, uint32_t *P, int64_t ldp,
                           int *nonbool, hipStream_t s) {
  if (rows <= 0 || cols <= 0) return hipSuccess;
  const dim3 grid((cols + PACK_COLS - 1) / PACK_COLS, std::min(rows, 65535));
  pack_i32_kernel<<<grid, dim3(256), 0, s>>>(rows, cols, X, ldx, P, ldp, nonbool);
  return hipGetLastError();
}

hipError_t launch_unpack_i32(int rows, int cols, const uint32_t *P, int64_t ldp, int *X,
                             int64_t ldx, hipStream_t s) {
  if (rows <= 0 || cols <= 0) return hipSuccess;
  const dim3 grid((cols + 1023) / 1024, std::min(rows, 65535));
  unpack_i32_kernel<<<grid, dim3(256), 0, s>>>(rows, cols, P, ldp, X, ldx);
  return hipGetLastError();
}

hipError_t launch_orand_i32(int R, int Cn, int K, const int *X, int64_t ldx,
                            const int *Y, int64_t ldy, int *O, int64_t ldo,
                            const int *C, int64_t ldc, const int *P,
                            int64_t ldp, int *flag, hipStream_t s) {
  const int tiles_r = (R + IB_BM - 1) / IB_BM;
  const int tiles_c = (Cn + IB_BN - 1) / IB_BN;
  orand_i32_kernel<<<dim3(tiles_r * tiles_c), dim3(256), 0, s>>>(
      R, Cn, K, X, ldx, Y, ldy, C, ldc, O, ldo, P, ldp, flag, tiles_c);
  return hipGetLastError();
}

}  // namespace simd2

// Scheduled repeated squaring (gfx950): one squaring D' = [D (+)] D (x) D
// that only computes what can change.
//
// Two exact skips, both decided per output tile of the VALU kernel:
//   * clean tiles.  If neither the tile's row panel nor its column panel
//     changed (bit for bit) in the previous squaring, every input of the tile
//     is what it was one squaring earlier, so its result equals the previous
//     result, which equals the tile's current value.  The output buffer of
//     squaring t is the input buffer of squaring t - 1, and by induction it
//     already holds that value: the tile is not touched at all.
//   * empty K slabs.  A 16-deep slab whose X part (the tile's rows) or whose
//     Y part (the tile's columns) holds only the (+)-identity contributes
//     identity (x) y = identity to every output, so the main loop visits only
//     the slabs listed as occupied on both sides.
// The epilogue produces what the next squaring needs: the convergence flag
// (the caller's change test), exact per-panel change flags, and the
// occupancy of the new matrix per (tile row, slab) and (tile column, slab).
#include <float.h>

#include <algorithm>
#include <cstring>

#include "valu_tile.hpp"

namespace simd2 {

namespace {

constexpr int EW = 256;

template <class Op, class Cfg>
__global__ __launch_bounds__((Cfg::BM / Cfg::TM) * (Cfg::BN / Cfg::TN))
__attribute__((amdgpu_waves_per_eu(Cfg::WAVES))) void sched_kernel(SchedArgs a) {
  constexpr int BM = Cfg::BM, BN = Cfg::BN, TM = Cfg::TM, TN = Cfg::TN, BK = Cfg::BK;
  constexpr int NTX = BN / TN;
  constexpr int RS = BM / (TM / 4), CS = BN / (TN / 4);
  constexpr int XS = BN / 16;  // column slabs of a tile
  constexpr int YS = BM / 16;  // row slabs of a tile
  static_assert(BK == 16 && XS <= 32 && YS <= 32 && XS + YS <= 256, "slab geometry");
  __shared__ ValuSmem<BM, BN, BK> sm;
  __shared__ uint16_t list[SCHED_MAX_SLABS];
  __shared__ int s_cnt;
  __shared__ int s_wave[4];
  __shared__ unsigned s_bits[2];

  // A squaring after one that changed nothing computes nothing: the result
  // stays where that squaring left it.
  if (a.conv_prev && *a.conv_prev == 0) return;
  if (a.nnz_in && *a.nnz_in <= a.nnz_limit) return;  // squared row by row instead
  const int tid = tid_x();
  if (a.nnz_out && bid_x() == 0 && tid == 0) *a.nnz_out = 0x7fffffff;  // not sparse any more
  int tr, tc;
  map_tile(bid_x(), a.tiles_r, a.tiles_c, tr, tc);
  if (a.rowchg_prev && !a.rowchg_prev[tr] && !a.colchg_prev[tc]) return;  // clean tile
  const int row0 = tr * BM, col0 = tc * BN;

  // Occupied slabs of this tile, ascending.
  if (tid == 0) {
    s_cnt = 0;
    s_bits[0] = 0;
    s_bits[1] = 0;
  }
  __syncthreads();
  const uint8_t *xo = a.xocc_in + (int64_t)tr * a.slabs;
  const uint8_t *yo = a.yocc_in + (int64_t)tc * a.slabs;
  const int lane = tid % 64, wave = tid / 64;
  for (int base = 0; base < a.slabs; base += 256) {
    const int s = base + tid;
    const bool p = s < a.slabs && xo[s] && yo[s];
    const uint64_t m = __ballot(p);
    if (lane == 0) s_wave[wave] = __popcll(m);
    __syncthreads();
    int off = s_cnt;
    for (int w = 0; w < wave; ++w) off += s_wave[w];
    if (p) list[off + __popcll(m & ((1ull << lane) - 1ull))] = (uint16_t)s;
    __syncthreads();
    if (tid == 0) s_cnt += s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
    __syncthreads();
  }
  const int nlist = s_cnt;

  float acc[TM][TN];
  valu_mainloop<Op, BM, BN, TM, TN, BK, false, true>(a.n, a.n, a.kpad, a.cur, a.ld, a.cur, a.ld,
                                                     row0, col0, list, nlist, sm, acc);

  // Epilogue.  The padded layout keeps the 128-bit reads of the old values in
  // bounds; stores and flags cover only r < n, c < n.
  const int tx = tid % NTX, ty = tid / NTX;
  const uint32_t ident = __float_as_uint(Op::identity);
  bool chg_bits = false, chg_test = false;
  unsigned xb = 0, yb = 0;
#pragma unroll
  for (int gi = 0; gi < TM / 4; ++gi)
#pragma unroll
    for (int ii = 0; ii < 4; ++ii) {
      const int i = gi * 4 + ii;
      const int rl = gi * RS + ty * 4 + ii;
      const int r = row0 + rl;
#pragma unroll
      for (int gj = 0; gj < TN / 4; ++gj) {
        const int cl = gj * CS + tx * 4;
        const int c = col0 + cl;
        const float4 pv = *reinterpret_cast<const float4 *>(a.cur + (int64_t)r * a.ld + c);
        const float p[4] = {pv.x, pv.y, pv.z, pv.w};
        float v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          v[q] = acc[i][gj * 4 + q];
          if (a.epi) v[q] = Op::add(v[q], p[q]);
          if (a.mode == CHECK_APLP && v[q] > -1.0f && v[q] < 1.0f) v[q] = -FLT_MAX;
        }
        if (r < a.n) {
          bool occupied = false;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            if (c + q < a.n) {
              chg_bits |= __float_as_uint(v[q]) != __float_as_uint(p[q]);
              chg_test |= differs(v[q], p[q], a.mode);
              occupied |= __float_as_uint(v[q]) != ident;
            }
          }
          if (occupied) {
            xb |= 1u << (cl / 16);
            yb |= 1u << (rl / 16);
          }
          float *o = a.nxt + (int64_t)r * a.ld + c;
          if (c + 4 <= a.n) {
            *reinterpret_cast<float4 *>(o) = make_float4(v[0], v[1], v[2], v[3]);
          } else {
#pragma unroll
            for (int q = 0; q < 4; ++q)
              if (c + q < a.n) o[q] = v[q];
          }
        }
      }
    }
  if (xb) atomicOr(&s_bits[0], xb);
  if (yb) atomicOr(&s_bits[1], yb);
  if (__any(chg_bits) && lane == 0) {
    atomicOr(a.rowchg + tr, 1);
    atomicOr(a.colchg + tc, 1);
  }
  if (__any(chg_test) && lane == 0) atomicOr(a.conv, 1);
  __syncthreads();
  const int sx = col0 / 16 + tid, sy = row0 / 16 + (tid - XS);
  if (tid < XS) {
    if (sx < a.slabs) a.xocc_out[(int64_t)tr * a.slabs + sx] = (s_bits[0] >> tid) & 1u;
  } else if (tid < XS + YS) {
    if (sy < a.slabs) a.yocc_out[(int64_t)tc * a.slabs + sy] = (s_bits[1] >> (tid - XS)) & 1u;
  }
  if (tid == 0 && a.stats) {
    atomicAdd(a.stats, 1ull);
    atomicAdd(a.stats + 1, (unsigned long long)nlist);
  }
}

// Row-sparse squaring, one workgroup per row i:
//   D'[i][c] = [D[i][c] (+)] (+)_{k : D[i][k] != identity} D[i][k] (x) D[k][c].
// The terms left out are identity (x) y = identity, so the result is the tiled
// kernel's bit for bit; it costs one pass over row k per entry of row i, which
// beats the tiled kernel while rows hold few entries (early squarings of
// sparse graphs).  Row i's entries are compacted into LDS (wave ballots) in
// passes of RL_CAP columns; each thread keeps RL_Q output columns in
// registers, 8192 columns per chunk.
constexpr int RL_Q = 32;
constexpr int RL_CHUNK = EW * RL_Q;
constexpr int RL_CAP = 2048;

template <class Op>
__global__ __launch_bounds__(EW) void sched_rows_kernel(SchedArgs a, int bm, int bn) {
  __shared__ uint16_t lk[RL_CAP];
  __shared__ float lx[RL_CAP];
  __shared__ int s_cnt;
  __shared__ int s_wave[EW / 64];
  if (a.conv_prev && *a.conv_prev == 0) return;
  if (!a.nnz_in || *a.nnz_in > a.nnz_limit) return;  // squared tile by tile instead
  const int i = bid_x();
  const int tid = tid_x(), lane = tid % 64, wave = tid / 64;
  const uint32_t ident = __float_as_uint(Op::identity);
  const float *rowi = a.cur + (int64_t)i * a.ld;
  float *outi = a.nxt + (int64_t)i * a.ld;
  bool chg_bits = false, chg_test = false;
  long long entries = 0;
  for (int c0 = 0; c0 < a.n; c0 += RL_CHUNK) {
    float acc[RL_Q];
#pragma unroll
    for (int q = 0; q < RL_Q; ++q) acc[q] = Op::identity;
    for (int k0 = 0; k0 < a.n; k0 += RL_CAP) {
      if (tid == 0) s_cnt = 0;
      __syncthreads();
      const int k1 = min(a.n, k0 + RL_CAP);
      for (int kb = k0; kb < k1; kb += EW) {
        const int k = kb + tid;
        const float x = k < k1 ? rowi[k] : Op::identity;
        const bool p = k < k1 && __float_as_uint(x) != ident;
        const uint64_t m = __ballot(p);
        if (lane == 0) s_wave[wave] = __popcll(m);
        __syncthreads();
        int off = s_cnt;
        for (int w = 0; w < wave; ++w) off += s_wave[w];
        if (p) {
          const int at = off + __popcll(m & ((1ull << lane) - 1ull));
          lk[at] = (uint16_t)k;
          lx[at] = x;
        }
        __syncthreads();
        if (tid == 0) s_cnt += s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
        __syncthreads();
      }
      const int cnt = s_cnt;
      entries += cnt;
      for (int e = 0; e < cnt; ++e) {
        const float x = lx[e];
        const float *rowk = a.cur + (int64_t)lk[e] * a.ld + c0 + tid;
#pragma unroll
        for (int q = 0; q < RL_Q; ++q)
          if (c0 + tid + EW * q < a.ld) acc[q] = Op::add(acc[q], Op::mul(x, rowk[EW * q]));
      }
      __syncthreads();  // the list is rewritten by the next pass
    }
#pragma unroll
    for (int q = 0; q < RL_Q; ++q) {
      const int c = c0 + tid + EW * q;
      bool col_chg = false;
      if (c < a.n) {
        const float p = rowi[c];
        float v = acc[q];
        if (a.epi) v = Op::add(v, p);
        if (a.mode == CHECK_APLP && v > -1.0f && v < 1.0f) v = -FLT_MAX;
        col_chg = __float_as_uint(v) != __float_as_uint(p);
        chg_bits |= col_chg;
        chg_test |= differs(v, p, a.mode);
        outi[c] = v;
      }
      // a wave covers 64 consecutive columns, inside one column block
      if (__any(col_chg) && lane == 0) a.colchg[(c0 + EW * q + 64 * wave) / bn] = 1;
    }
  }
  if (__any(chg_bits) && lane == 0) a.rowchg[i / bm] = 1;
  if (__any(chg_test) && lane == 0) *a.conv = 1;
  if (tid == 0 && a.stats) {
    atomicAdd(a.stats + 2, 1ull);                               // rows squared row-sparse
    atomicAdd(a.stats + 3, (unsigned long long)entries);        // row entries visited
  }
}

// Occupancy maps of a matrix, one workgroup per bm x bn tile (the maps the
// tiled squaring's epilogue writes), and its count of entries other than the
// identity.  Used for the first squaring's input and after row-sparse
// squarings; after a tiled squaring (or once converged) it does nothing.
__global__ __launch_bounds__(EW) void occupancy_kernel(const float *D, int64_t ld, int n, int bm,
                                                       int bn, int tiles_c, int slabs,
                                                       uint32_t ident, uint8_t *xocc,
                                                       uint8_t *yocc, int *nnz_out,
                                                       const int *conv_prev, const int *nnz_in,
                                                       int nnz_limit) {
  __shared__ unsigned s_bits[2];
  __shared__ int s_count;
  if (conv_prev && *conv_prev == 0) return;
  if (nnz_in && *nnz_in > nnz_limit) return;
  const int tid = tid_x();
  const int tr = bid_x() / tiles_c, tc = bid_x() % tiles_c;
  const int row0 = tr * bm, col0 = tc * bn;
  if (tid == 0) {
    s_bits[0] = 0;
    s_bits[1] = 0;
    s_count = 0;
  }
  __syncthreads();
  unsigned xb = 0, yb = 0;
  int count = 0;
  for (int e = tid; e < bm * bn; e += EW) {
    const int rl = e / bn, cl = e % bn;
    const int r = row0 + rl, c = col0 + cl;
    if (r < n && c < n && __float_as_uint(D[(int64_t)r * ld + c]) != ident) {
      xb |= 1u << (cl / 16);
      yb |= 1u << (rl / 16);
      ++count;
    }
  }
  if (xb) atomicOr(&s_bits[0], xb);
  if (yb) atomicOr(&s_bits[1], yb);
  if (count) atomicAdd(&s_count, count);
  __syncthreads();
  if (tid == 0 && nnz_out && s_count) atomicAdd(nnz_out, s_count);
  const int xs = bn / 16, ys = bm / 16;
  if (tid < xs) {
    const int s = col0 / 16 + tid;
    if (s < slabs) xocc[(int64_t)tr * slabs + s] = (s_bits[0] >> tid) & 1u;
  } else if (tid < xs + ys) {
    const int l = tid - xs, s = row0 / 16 + l;
    if (s < slabs) yocc[(int64_t)tc * slabs + s] = (s_bits[1] >> l) & 1u;
  }
}

__global__ __launch_bounds__(EW) void fill_f32_kernel(float *p, float v, int64_t n) {
  for (int64_t i = (int64_t)bid_x() * EW + tid_x(); i < n; i += (int64_t)gridDim.x * EW) p[i] = v;
}

// inverse == 0: dst[i][j] = src[perm[i]][perm[j]]; inverse == 1:
// dst[perm[i]][perm[j]] = src[i][j].  perm == null: plain copy.
__global__ __launch_bounds__(EW) void permute_copy_kernel(float *dst, int64_t ldd,
                                                          const float *src, int64_t lds, int n,
                                                          const int *perm, int inverse) {
  const int64_t total = (int64_t)n * n;
  for (int64_t e = (int64_t)bid_x() * EW + tid_x(); e < total; e += (int64_t)gridDim.x * EW) {
    const int i = (int)(e / n), j = (int)(e % n);
    const int pi = perm ? perm[i] : i, pj = perm ? perm[j] : j;
    if (!inverse)
      dst[(int64_t)i * ldd + j] = src[(int64_t)pi * lds + pj];
    else
      dst[(int64_t)pi * ldd + pj] = src[(int64_t)i * lds + j];
  }
}

int grid_for(int64_t n) { return (int)std::min<int64_t>((n + EW - 1) / EW, 4096); }

template <class Op>
hipError_t sched_op(int cfg, const SchedArgs &a, hipStream_t s) {
  const dim3 grid(a.tiles_r * a.tiles_c), block(256);
  if (a.nnz_in) {  // the row-sparse form; exactly one of the two kernels works
    const ValuCfgInfo c = valu_cfg_info(cfg);
    sched_rows_kernel<Op><<<dim3(a.n), dim3(EW), 0, s>>>(a, c.bm, c.bn);
  }
  if (cfg == CFG_S64)
    sched_kernel<Op, CfgS><<<grid, block, 0, s>>>(a);
  else if (cfg == CFG_M128x64)
    sched_kernel<Op, CfgM><<<grid, block, 0, s>>>(a);
  else
    sched_kernel<Op, CfgL><<<grid, block, 0, s>>>(a);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_sched_square(int op, int cfg, const SchedArgs &a, hipStream_t s) {
  switch (op) {
  case 0: return sched_op<MinPlus>(cfg, a, s);
  case 2: return sched_op<MaxPlus>(cfg, a, s);
  case 3: return sched_op<MaxMin>(cfg, a, s);
  case 6: return sched_op<MinMax>(cfg, a, s);
  default: return hipErrorInvalidValue;
  }
}

hipError_t launch_sched_occupancy(int op, int cfg, const float *D, int64_t ld, int n,
                                  int tiles_r, int tiles_c, int slabs, uint8_t *xocc,
                                  uint8_t *yocc, int *nnz_out, const SchedArgs *a,
                                  hipStream_t s) {
  const ValuCfgInfo c = valu_cfg_info(cfg);
  const float ident = (op == 0 || op == 6) ? SIMD2_INF : -SIMD2_INF;
  uint32_t bits;
  std::memcpy(&bits, &ident, 4);
  occupancy_kernel<<<dim3(tiles_r * tiles_c), dim3(EW), 0, s>>>(
      D, ld, n, c.bm, c.bn, tiles_c, slabs, bits, xocc, yocc, nnz_out,
      a ? a->conv_prev : nullptr, a ? a->nnz_in : nullptr, a ? a->nnz_limit : 0);
  return hipGetLastError();
}

hipError_t launch_fill_f32(float *p, float v, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  fill_f32_kernel<<<dim3(grid_for(n)), dim3(EW), 0, s>>>(p, v, n);
  return hipGetLastError();
}

hipError_t launch_permute_copy(float *dst, int64_t ldd, const float *src, int64_t lds, int n,
                               const int *perm, int inverse, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  permute_copy_kernel<<<dim3(grid_for((int64_t)n * n)), dim3(EW), 0, s>>>(dst, ldd, src, lds, n,
                                                                          perm, inverse);
  return hipGetLastError();
}

}  // namespace simd2

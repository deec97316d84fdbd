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
constexpr int SCHED_SPLIT_SLABS = 256;  // the split form's limit: n <= 4096

// G > 1 (small problems, 64x64 tiles): the tile's occupied slabs are split
// over G groups of 256 threads, contiguous ranges of the list, and the
// partial tiles combined in group order before the epilogue, as the plain
// kernel's split-K form does (min/max: exact, so regrouping changes no bit).
// Every group runs the same number of slabs: a shorter range repeats its
// last slab (or the list's first), which min/max, being idempotent, ignore.
template <class Op, class Cfg, int G = 1>
__global__ __launch_bounds__((Cfg::BM / Cfg::TM) * (Cfg::BN / Cfg::TN) * G)
__attribute__((amdgpu_waves_per_eu(G > 1 ? 1 : Cfg::WAVES))) void sched_kernel(SchedArgs a) {
  constexpr int BM = Cfg::BM, BN = Cfg::BN, TM = Cfg::TM, TN = Cfg::TN, BK = Cfg::BK;
  constexpr int NTX = BN / TN;
  constexpr int NT = NTX * (BM / TM);  // threads of one group
  constexpr int RS = BM / (TM / 4), CS = BN / (TN / 4);
  constexpr int XS = BN / 16;  // column slabs of a tile
  constexpr int YS = BM / 16;  // row slabs of a tile
  static_assert(BK == 16 && XS <= 32 && YS <= 32 && XS + YS <= 256 && NT == 256,
                "slab geometry");
  constexpr int GL = G > 1 ? SCHED_SPLIT_SLABS / G : 1;  // slabs per group (launcher checks)
  __shared__ ValuSmem<BM, BN, BK> sm_g[G];
  __shared__ uint16_t glist[G][GL];
  __shared__ uint16_t list[SCHED_MAX_SLABS];
  __shared__ int s_cnt;
  __shared__ int s_wave[4];
  __shared__ unsigned s_bits[2];

  // A squaring after one that changed nothing computes nothing: the result
  // stays where that squaring left it.
  if (a.conv_prev && *a.conv_prev == 0) return;
  if (a.sq_mode && *a.sq_mode != 0) return;  // squared row by row instead
  const int tid_all = tid_x();
  const int grp = tid_all / NT, tid = tid_all % NT;
  auto &sm = sm_g[grp];
  int tr, tc;
  map_tile(bid_x(), a.tiles_r, a.tiles_c, tr, tc);
  if (a.rowchg_prev && !a.rowchg_prev[tr] && !a.colchg_prev[tc]) return;  // clean tile
  const int row0 = tr * BM, col0 = tc * BN;

  // Occupied slabs of this tile, ascending.
  // (group 0 builds the list; the other groups only keep the barriers)
  if (tid_all == 0) {
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
    const bool p = grp == 0 && s < a.slabs && xo[s] && yo[s];
    const uint64_t m = __ballot(p);
    if (grp == 0 && lane == 0) s_wave[wave] = __popcll(m);
    __syncthreads();
    if (grp == 0) {
      int off = s_cnt;
      for (int w = 0; w < wave; ++w) off += s_wave[w];
      if (p) list[off + __popcll(m & ((1ull << lane) - 1ull))] = (uint16_t)s;
    }
    __syncthreads();
    if (tid_all == 0) s_cnt += s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
    __syncthreads();
  }
  const int nlist = s_cnt;

  float acc[TM][TN];
  if constexpr (G == 1) {
    valu_mainloop<Op, BM, BN, TM, TN, BK, false, true>(a.n, a.n, a.kpad, a.cur, a.ld, a.cur,
                                                       a.ld, row0, col0, list, nlist, sm, acc);
  } else {
    // group g: entries [b, b + c) of the list, padded to `steps` by repeats
    const int steps = (nlist + G - 1) / G;
    const int b = grp * (nlist / G) + min(grp, nlist % G);
    const int c = nlist / G + (grp < nlist % G ? 1 : 0);
    for (int i = tid; i < steps; i += NT)
      glist[grp][i] = list[c > 0 ? b + min(i, c - 1) : 0];
    __syncthreads();
    valu_mainloop<Op, BM, BN, TM, TN, BK, false, true>(a.n, a.n, a.kpad, a.cur, a.ld, a.cur,
                                                       a.ld, row0, col0, glist[grp], steps, sm,
                                                       acc, tid);
    // the main loop ends with a barrier after every read of the slabs
    if (grp > 0) {
      float *part = reinterpret_cast<float *>(&sm);
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) part[(i * TN + j) * NT + tid] = acc[i][j];
    }
    __syncthreads();
    if (grp == 0) {
#pragma unroll
      for (int q = 1; q < G; ++q) {
        const float *part = reinterpret_cast<const float *>(&sm_g[q]);
#pragma unroll
        for (int i = 0; i < TM; ++i)
#pragma unroll
          for (int j = 0; j < TN; ++j)
            acc[i][j] = Op::add(acc[i][j], part[(i * TN + j) * NT + tid]);
      }
    }
  }
  const bool active = grp == 0;

  // Epilogue.  The padded layout keeps the 128-bit reads of the old values in
  // bounds; stores and flags cover only r < n, c < n.
  const int tx = tid % NTX, ty = tid / NTX;
  const uint32_t ident = __float_as_uint(Op::identity);
  bool chg_bits = false, chg_test = false;
  unsigned xb = 0, yb = 0;
  if (active)
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
  if (active && tid < XS) {
    if (sx < a.slabs) a.xocc_out[(int64_t)tr * a.slabs + sx] = (s_bits[0] >> tid) & 1u;
  } else if (active && tid < XS + YS) {
    if (sy < a.slabs) a.yocc_out[(int64_t)tc * a.slabs + sy] = (s_bits[1] >> (tid - XS)) & 1u;
  }
  if (tid_all == 0 && a.stats) {
    atomicAdd(a.stats, 1ull);
    atomicAdd(a.stats + 1, (unsigned long long)nlist);
  }
}

// Sparse squarings, one workgroup per output row i, over a list of entries
// built before the squaring (by row, ascending columns):
//   * row-sparse: every entry of the input other than the (+)-identity.
//       D'[i][:] = [D[i][:] (+)] (+)_{(i,k) listed} D[i][k] (x) D[k][:]
//     is the tiled result bit for bit: the terms left out are
//     identity (x) y = identity.  It costs one pass over row k per entry.
//   * changed-entry: the entries that changed in the previous squaring.  With
//     the epilogue on, the input is D = D_prev (+) (+)_k D_prev[i][k] (x)
//     D_prev[k][j], so a term D[i][k] (x) D[k][j] whose operands both kept
//     their bits is a term of the previous squaring and already inside
//     D[i][j]; (+) being min or max, taking it again changes no bit.  The
//     remaining terms have (i,k) changed (a pass over row k, as above) or
//     (k,j) changed (scattered into row i's accumulators in LDS).
// Before each squaring: the (tile, slab) pairs the tiled kernel would compute
// (its cost), per-row counts of both kinds of entries, one scan that picks
// the form, then each row's entries compacted with wave ballots (entry =
// (row << 16) | column, and the value).
constexpr int SP_G = 8;                // 4-column groups per thread
constexpr int SP_CH = EW * 4 * SP_G;   // columns per pass of a row
constexpr int SCAN_T = 1024;
constexpr int SP_B = 8;                // changed entries per thread per batch

// (tile, slab) pairs the tiled kernel would compute in this squaring: dirty
// tiles times the slabs occupied on both sides, plus two per dirty tile for
// its epilogue.  One wave per tile, its lanes over the slabs.
__global__ __launch_bounds__(EW) void tiled_work_kernel(SchedArgs a) {
  if (a.conv_prev && *a.conv_prev == 0) return;
  const int t = bid_x() * (EW / 64) + tid_x() / 64, lane = tid_x() % 64;
  if (t >= a.tiles_r * a.tiles_c) return;  // whole waves leave
  const int tr = t / a.tiles_c, tc = t % a.tiles_c;
  if (a.rowchg_prev && !a.rowchg_prev[tr] && !a.colchg_prev[tc]) return;  // clean tile
  const uint8_t *xo = a.xocc_in + (int64_t)tr * a.slabs;
  const uint8_t *yo = a.yocc_in + (int64_t)tc * a.slabs;
  int cnt = 0;
  for (int s = lane; s < a.slabs; s += 64) cnt += xo[s] & yo[s];
  for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o);
  if (lane == 0) atomicAdd(a.work, (unsigned long long)(cnt + 2));
}

// Per row: entries other than the identity (rowptr[i], when the row-sparse
// form is on) and, with `prev`, entries that differ from prev (chgcnt[i]; a
// row of a tile row the previous squaring left unchanged has none).
template <class Op>
__global__ __launch_bounds__(EW) void list_count_kernel(SchedArgs a, int bm) {
  __shared__ int s_wave[2][EW / 64];
  if (a.conv_prev && *a.conv_prev == 0) return;
  const int i = bid_x(), tid = tid_x(), lane = tid % 64, wave = tid / 64;
  const uint32_t ident = __float_as_uint(Op::identity);
  const int64_t r0 = (int64_t)i * a.ld;
  const bool need_nz = a.rows_limit > 0;
  const bool need_ch = a.prev && !(a.rowchg_prev && !a.rowchg_prev[i / bm]);
  int nz = 0, ch = 0;
  // columns n .. ld-1 hold the identity in every buffer, so whole groups of
  // four can be tested (ld is a multiple of 128)
  for (int c = 4 * tid; (need_nz || need_ch) && c < a.n; c += 4 * EW) {
    const vf4 x = *reinterpret_cast<const vf4 *>(a.cur + r0 + c);
    nz += (__float_as_uint(x.x) != ident) + (__float_as_uint(x.y) != ident) +
          (__float_as_uint(x.z) != ident) + (__float_as_uint(x.w) != ident);
    if (need_ch) {
      const vf4 y = *reinterpret_cast<const vf4 *>(a.prev + r0 + c);
      ch += (__float_as_uint(x.x) != __float_as_uint(y.x)) +
            (__float_as_uint(x.y) != __float_as_uint(y.y)) +
            (__float_as_uint(x.z) != __float_as_uint(y.z)) +
            (__float_as_uint(x.w) != __float_as_uint(y.w));
    }
  }
  for (int o = 32; o > 0; o >>= 1) {
    nz += __shfl_xor(nz, o);
    ch += __shfl_xor(ch, o);
  }
  if (lane == 0) {
    s_wave[0][wave] = nz;
    s_wave[1][wave] = ch;
  }
  __syncthreads();
  if (tid == 0) {
    a.rowptr[i] = s_wave[0][0] + s_wave[0][1] + s_wave[0][2] + s_wave[0][3];
    if (a.prev) a.chgcnt[i] = s_wave[1][0] + s_wave[1][1] + s_wave[1][2] + s_wave[1][3];
  }
}

// Totals, the squaring's form, and the exclusive scan of the chosen counts
// into rowptr (one workgroup).
__global__ __launch_bounds__(SCAN_T) void list_scan_kernel(SchedArgs a) {
  __shared__ long long s_nz[SCAN_T], s_ch[SCAN_T];
  __shared__ int s_mode;
  if (a.conv_prev && *a.conv_prev == 0) return;
  const int tid = tid_x();
  const int per = (a.n + SCAN_T - 1) / SCAN_T;
  const int lo = min(a.n, tid * per), hi = min(a.n, lo + per);
  long long nz = 0, ch = 0;
  for (int r = lo; r < hi; ++r) {
    nz += a.rowptr[r];
    if (a.prev) ch += a.chgcnt[r];
  }
  s_nz[tid] = nz;
  s_ch[tid] = ch;
  __syncthreads();
  for (int o = 1; o < SCAN_T; o <<= 1) {
    const long long vn = tid >= o ? s_nz[tid - o] : 0, vc = tid >= o ? s_ch[tid - o] : 0;
    __syncthreads();
    s_nz[tid] += vn;
    s_ch[tid] += vc;
    __syncthreads();
  }
  if (tid == 0) {
    const long long tn = s_nz[SCAN_T - 1], tc = s_ch[SCAN_T - 1];
    // both limits are within the list's capacity
    const bool rows_ok = a.rows_limit > 0 && tn <= a.rows_limit;
    const bool ch_ok = a.prev && a.changed_limit > 0 && tc <= a.changed_limit;
    const bool sym = a.asym && *a.asym == 0;
    int mode = 0;
    if (rows_ok && !a.rows_auto) {
      mode = 1;
    } else if (ch_ok && !a.changed_auto) {
      mode = sym ? 3 : 2;
    } else {
      const double c_rows = (double)tn * a.cost_row_entry + a.cost_list;
      const double c_changed = (double)tc * a.cost_changed_entry + a.cost_list;
      // symmetric: row passes over the changed entries, then a symmetrising pass
      const double c_sym = (double)tc * a.cost_row_entry + 2.0 * a.cost_list;
      double best = (double)*a.work * a.cost_slab_tile;
      if (rows_ok && c_rows < best) {
        mode = 1;
        best = c_rows;
      }
      if (ch_ok && c_changed < best) {
        mode = 2;
        best = c_changed;
      }
      if (ch_ok && sym && c_sym < best) mode = 3;
    }
    *a.sq_mode = mode;
    s_mode = mode;
  }
  __syncthreads();
  const int mode = s_mode;
  if (mode == 0) return;
  const long long *sc = mode == 1 ? s_nz : s_ch;
  const int *cnt = mode == 1 ? a.rowptr : a.chgcnt;
  long long run = sc[tid] - (mode == 1 ? nz : ch);
  for (int r = lo; r < hi; ++r) {
    const int c = cnt[r];  // read before rowptr[r] is overwritten (mode 1)
    a.rowptr[r] = (int)run;
    run += c;
  }
  if (tid == SCAN_T - 1) a.rowptr[a.n] = (int)sc[SCAN_T - 1];
}

template <class Op>
__global__ __launch_bounds__(EW) void list_fill_kernel(SchedArgs a) {
  __shared__ int s_wave[EW / 64];
  if (a.conv_prev && *a.conv_prev == 0) return;
  const int mode = *a.sq_mode;
  if (mode == 0) return;
  const int i = bid_x();
  int at = a.rowptr[i];
  const int end = a.rowptr[i + 1];
  if (at == end) return;
  const int tid = tid_x(), lane = tid % 64, wave = tid / 64;
  const uint32_t ident = __float_as_uint(Op::identity);
  const int64_t r0 = (int64_t)i * a.ld;
  const uint64_t below = (1ull << lane) - 1ull;
  for (int c0 = 0; c0 < a.n && at < end; c0 += 4 * EW) {
    const int c = c0 + 4 * tid;
    bool p[4] = {false, false, false, false};
    if (c < a.n) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const uint32_t x = __float_as_uint(a.cur[r0 + c + q]);
        p[q] = mode == 1 ? x != ident : x != __float_as_uint(a.prev[r0 + c + q]);
      }
    }
    int before = 0, wave_total = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const uint64_t m = __ballot(p[q]);
      before += __popcll(m & below);
      wave_total += __popcll(m);
    }
    if (lane == 0) s_wave[wave] = wave_total;
    __syncthreads();
    int off = at + before;
    for (int w = 0; w < wave; ++w) off += s_wave[w];
    const int total = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
    __syncthreads();  // s_wave is rewritten by the next group
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (p[q] && off < end) {
        a.lkj[off] = ((uint32_t)i << 16) | (uint32_t)(c + q);
        a.lval[off] = a.cur[r0 + c + q];
        ++off;
      }
    at += total;
  }
}

template <class Op>
__device__ __forceinline__ void lds_combine(float *p, float v) {
  if constexpr (Op::identity > 0.0f)
    __hip_atomic_fetch_min(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  else
    __hip_atomic_fetch_max(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <class Op>
__device__ __forceinline__ vf4 relax4(vf4 acc, float x0, vf4 y0, float x1, vf4 y1) {
  acc.x = Op::add3(acc.x, Op::mul(x0, y0.x), Op::mul(x1, y1.x));
  acc.y = Op::add3(acc.y, Op::mul(x0, y0.y), Op::mul(x1, y1.y));
  acc.z = Op::add3(acc.z, Op::mul(x0, y0.z), Op::mul(x1, y1.z));
  acc.w = Op::add3(acc.w, Op::mul(x0, y0.w), Op::mul(x1, y1.w));
  return acc;
}

// One output row per workgroup.  FORM 1: row-sparse; 2: changed-entry;
// 3: changed-entry of a symmetric matrix, row passes only: it writes
// R = D (+) (row passes) and no flags, and symmetrize_kernel finishes
// D' = R (+) R^T (the scatter terms of row i are the row-pass terms of
// column i when D and its changes are symmetric).
template <class Op, int FORM>
__global__ __launch_bounds__(EW) void sparse_square_kernel(SchedArgs a, int bm, int bn) {
  constexpr bool CHANGED = FORM == 2;
  __shared__ float lds_acc[CHANGED ? SP_CH : 1];
  __shared__ float lds_x[CHANGED ? SP_CH : 1];
  if (a.conv_prev && *a.conv_prev == 0) return;
  if (*a.sq_mode != FORM) return;
  const int i = bid_x(), tid = tid_x(), lane = tid % 64;
  const uint32_t ident = __float_as_uint(Op::identity);
  const int e0 = a.rowptr[i], e1 = a.rowptr[i + 1];
  const int64_t r0 = (int64_t)i * a.ld;
  bool chg_bits = false, chg_test = false;
  for (int jc = 0; jc < a.n; jc += SP_CH) {
    if constexpr (CHANGED) {
      // terms with (k, j) changed: lds_acc[j] (+)= D[i][k] (x) D[k][j]
      for (int t = tid; t < SP_CH; t += EW) lds_acc[t] = Op::identity;
      for (int kc = 0; kc < a.n; kc += SP_CH) {
        const int lo = a.rowptr[kc], hi = a.rowptr[min(a.n, kc + SP_CH)];
        if (lo == hi) continue;
        __syncthreads();  // lds_x readers of the previous chunk are done
        for (int t = tid; t < SP_CH; t += EW)
          lds_x[t] = kc + t < a.n ? a.cur[r0 + kc + t] : Op::identity;
        __syncthreads();
        // SP_B entries per thread are loaded before any is used, so that
        // many list reads are in flight
        for (int base = lo; base < hi; base += EW * SP_B) {
          uint32_t u[SP_B];
          float v[SP_B];
#pragma unroll
          for (int q = 0; q < SP_B; ++q) {
            const int e = base + EW * q + tid;
            u[q] = e < hi ? a.lkj[e] : 0xffffffffu;  // column 0xffff: never in range
            v[q] = e < hi ? a.lval[e] : Op::identity;
          }
#pragma unroll
          for (int q = 0; q < SP_B; ++q) {
            const int j = (int)(u[q] & 0xffffu) - jc;
            if ((unsigned)j < (unsigned)SP_CH && (int)(u[q] & 0xffffu) < a.n) {
              const float x = lds_x[(int)(u[q] >> 16) - kc];
              if (__float_as_uint(x) != ident) lds_combine<Op>(&lds_acc[j], Op::mul(x, v[q]));
            }
          }
        }
      }
      __syncthreads();
    }
    // terms with (i, k) listed: a pass over row k per entry, two at a time
    vf4 acc[SP_G];
#pragma unroll
    for (int g = 0; g < SP_G; ++g) acc[g] = vf4{Op::identity, Op::identity, Op::identity, Op::identity};
    const int cbase = jc + 4 * tid;
    int e = e0;
    for (; e + 1 < e1; e += 2) {
      const float x0 = a.lval[e], x1 = a.lval[e + 1];
      const float *y0p = a.cur + (int64_t)(a.lkj[e] & 0xffffu) * a.ld + cbase;
      const float *y1p = a.cur + (int64_t)(a.lkj[e + 1] & 0xffffu) * a.ld + cbase;
      vf4 y0[SP_G], y1[SP_G];
#pragma unroll
      for (int g = 0; g < SP_G; ++g)
        if (cbase + EW * 4 * g < a.n) {
          y0[g] = *reinterpret_cast<const vf4 *>(y0p + EW * 4 * g);
          y1[g] = *reinterpret_cast<const vf4 *>(y1p + EW * 4 * g);
        }
#pragma unroll
      for (int g = 0; g < SP_G; ++g)
        if (cbase + EW * 4 * g < a.n) acc[g] = relax4<Op>(acc[g], x0, y0[g], x1, y1[g]);
    }
    if (e < e1) {
      const float x0 = a.lval[e];
      const float *y0p = a.cur + (int64_t)(a.lkj[e] & 0xffffu) * a.ld + cbase;
#pragma unroll
      for (int g = 0; g < SP_G; ++g)
        if (cbase + EW * 4 * g < a.n) {
          const vf4 y = *reinterpret_cast<const vf4 *>(y0p + EW * 4 * g);
          acc[g] = relax4<Op>(acc[g], x0, y, x0, y);
        }
    }
    // epilogue: combine, compare with the input, store
#pragma unroll
    for (int g = 0; g < SP_G; ++g) {
      const int c = cbase + EW * 4 * g;
      if (c < a.n) {
        const vf4 p = *reinterpret_cast<const vf4 *>(a.cur + r0 + c);
        vf4 v = acc[g];
        bool col = false;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          float t = v[q];
          if constexpr (CHANGED) t = Op::add(t, lds_acc[c - jc + q]);
          if (a.epi) t = Op::add(t, p[q]);
          if (FORM == 1 && a.mode == CHECK_APLP && t > -1.0f && t < 1.0f) t = -FLT_MAX;
          v[q] = t;
          if (c + q < a.n) {
            col |= __float_as_uint(t) != __float_as_uint(p[q]);
            chg_test |= differs(t, p[q], a.mode);
          }
        }
        chg_bits |= col;
        if (FORM != 3 && col) a.colchg[c / bn] = 1;  // c .. c+3 lie in one column block
        float *o = a.nxt + r0 + c;
        if (c + 4 <= a.n) {
          *reinterpret_cast<vf4 *>(o) = v;
        } else {
#pragma unroll
          for (int q = 0; q < 4; ++q)
            if (c + q < a.n) o[q] = v[q];
        }
      }
    }
    if constexpr (CHANGED) __syncthreads();  // lds_acc is reset by the next pass
  }
  if (FORM != 3) {  // the symmetrising pass sets the flags of form 3
    if (__any(chg_bits) && lane == 0) a.rowchg[i / bm] = 1;
    if (__any(chg_test) && lane == 0) *a.conv = 1;
  }
  if (tid == 0 && a.stats) {
    atomicAdd(a.stats + (FORM == 1 ? 2 : 4), 1ull);
    atomicAdd(a.stats + (FORM == 1 ? 3 : 5), (unsigned long long)(e1 - e0));
  }
}

// Pairs of SY x SY tiles (I, J), I <= J, one workgroup each, from the linear
// workgroup index (tile row I holds T - I pairs).
constexpr int SY = 64;
__device__ __forceinline__ void tile_pair(int b, int T, int &I, int &J) {
  I = 0;
  while (b >= T - I) {
    b -= T - I;
    ++I;
  }
  J = I + b;
}

// Loads tile (I, J) of M into ta and tile (J, I) into tb (identity outside n).
template <class Op>
__device__ __forceinline__ void load_pair(const float *M, int64_t ld, int n, int r0, int c0,
                                          float (*ta)[SY + 1], float (*tb)[SY + 1]) {
  for (int e = tid_x(); e < SY * SY; e += EW) {
    const int rl = e / SY, cl = e % SY;
    ta[rl][cl] = r0 + rl < n && c0 + cl < n ? M[(int64_t)(r0 + rl) * ld + c0 + cl] : Op::identity;
    tb[rl][cl] = c0 + rl < n && r0 + cl < n ? M[(int64_t)(c0 + rl) * ld + r0 + cl] : Op::identity;
  }
}

// Form 3, second pass: D'[r][c] = R[r][c] (+) R[c][r] in place, with the
// exact change flags and the convergence flag against the input.  A tile of
// 64 lies inside one row block and one column block (bm, bn are 64 or 128).
template <class Op>
__global__ __launch_bounds__(EW) void symmetrize_kernel(SchedArgs a, int bm, int bn) {
  __shared__ float ta[SY][SY + 1], tb[SY][SY + 1];
  if (a.conv_prev && *a.conv_prev == 0) return;
  if (*a.sq_mode != 3) return;
  const int T = (a.n + SY - 1) / SY;
  int I, J;
  tile_pair(bid_x(), T, I, J);
  const int r0 = I * SY, c0 = J * SY;
  load_pair<Op>(a.nxt, a.ld, a.n, r0, c0, ta, tb);
  __syncthreads();
  bool chg1 = false, chg2 = false, test = false;
  for (int e = tid_x(); e < SY * SY; e += EW) {
    const int rl = e / SY, cl = e % SY;
    if (r0 + rl < a.n && c0 + cl < a.n) {  // element (r0 + rl, c0 + cl) of tile (I, J)
      const int64_t at = (int64_t)(r0 + rl) * a.ld + c0 + cl;
      const float v = Op::add(ta[rl][cl], tb[cl][rl]), p = a.cur[at];
      a.nxt[at] = v;
      chg1 |= __float_as_uint(v) != __float_as_uint(p);
      test |= differs(v, p, a.mode);
    }
    if (I != J && c0 + rl < a.n && r0 + cl < a.n) {  // element (c0 + rl, r0 + cl) of (J, I)
      const int64_t at = (int64_t)(c0 + rl) * a.ld + r0 + cl;
      const float v = Op::add(tb[rl][cl], ta[cl][rl]), p = a.cur[at];
      a.nxt[at] = v;
      chg2 |= __float_as_uint(v) != __float_as_uint(p);
      test |= differs(v, p, a.mode);
    }
  }
  chg1 = __syncthreads_or(chg1);
  chg2 = __syncthreads_or(chg2);
  test = __syncthreads_or(test);
  if (tid_x() == 0) {
    if (chg1) {
      a.rowchg[r0 / bm] = 1;
      a.colchg[c0 / bn] = 1;
    }
    if (chg2) {
      a.rowchg[c0 / bm] = 1;
      a.colchg[r0 / bn] = 1;
    }
    if (test) *a.conv = 1;
  }
}

// *asym = 1 when some D[r][c] differs from D[c][r] in its bits.
__global__ __launch_bounds__(EW) void symmetry_check_kernel(const float *D, int64_t ld, int n,
                                                            int *asym) {
  __shared__ float ta[SY][SY + 1], tb[SY][SY + 1];
  const int T = (n + SY - 1) / SY;
  int I, J;
  tile_pair(bid_x(), T, I, J);
  const int r0 = I * SY, c0 = J * SY;
  load_pair<MinPlus>(D, ld, n, r0, c0, ta, tb);
  __syncthreads();
  bool diff = false;
  for (int e = tid_x(); e < SY * SY; e += EW) {
    const int rl = e / SY, cl = e % SY;
    diff |= __float_as_uint(ta[rl][cl]) != __float_as_uint(tb[cl][rl]);
  }
  if (__syncthreads_or(diff) && tid_x() == 0) *asym = 1;
}

// Occupancy maps of a matrix, one workgroup per bm x bn tile (the maps the
// tiled squaring's epilogue writes), and its count of entries other than the
// identity.  Used for the first squaring's input and after sparse
// squarings; after a tiled squaring (or once converged) it does nothing.
__global__ __launch_bounds__(EW) void occupancy_kernel(const float *D, int64_t ld, int n, int bm,
                                                       int bn, int tiles_c, int slabs,
                                                       uint32_t ident, uint8_t *xocc,
                                                       uint8_t *yocc, const int *conv_prev,
                                                       const int *sq_mode) {
  __shared__ unsigned s_bits[2];
  if (conv_prev && *conv_prev == 0) return;
  if (sq_mode && *sq_mode == 0) return;
  const int tid = tid_x();
  const int tr = bid_x() / tiles_c, tc = bid_x() % tiles_c;
  const int row0 = tr * bm, col0 = tc * bn;
  if (tid == 0) {
    s_bits[0] = 0;
    s_bits[1] = 0;
  }
  __syncthreads();
  // Each thread reads one 16-byte column group of every (EW / (bn/4))-th
  // tile row: rows are 16-byte aligned (ld is a multiple of 128 floats) and
  // a row's padded columns up to the tile edge lie inside its pitch.
  const int vpr = bn / 4;
  const int c4 = tid % vpr;
  const int c = col0 + 4 * c4;
  unsigned xb = 0, yb = 0;
#pragma unroll 4
  for (int rl = tid / vpr; rl < bm; rl += EW / vpr) {
    const int r = row0 + rl;
    if (r >= n) break;
    const uint4 v = *reinterpret_cast<const uint4 *>(D + (int64_t)r * ld + c);
    const bool any = (c < n && v.x != ident) || (c + 1 < n && v.y != ident) ||
                     (c + 2 < n && v.z != ident) || (c + 3 < n && v.w != ident);
    if (any) {
      xb |= 1u << (c4 / 4);
      yb |= 1u << (rl / 16);
    }
  }
  if (xb) atomicOr(&s_bits[0], xb);
  if (yb) atomicOr(&s_bits[1], yb);
  __syncthreads();
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

// Several 32-bit fills in one launch: segment blockIdx.y, 16-byte stores
// from a 16-byte aligned start.
__global__ __launch_bounds__(EW) void fill_multi_kernel(FillSegments f) {
  const int k = bid_y();
  uint32_t *p = f.p[k];
  const int64_t n = f.n[k];
  const uint32_t v = f.v[k];
  const int64_t first = (int64_t)bid_x() * EW + tid_x(), stride = (int64_t)gridDim.x * EW;
  int64_t done = 0;
  if ((reinterpret_cast<uintptr_t>(p) & 15) == 0) {
    const int64_t n4 = n / 4;
    uint4 *p4 = reinterpret_cast<uint4 *>(p);
    for (int64_t i = first; i < n4; i += stride) p4[i] = make_uint4(v, v, v, v);
    done = 4 * n4;
  }
  for (int64_t i = done + first; i < n; i += stride) p[i] = v;
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
  if (a.sq_mode) {  // the list, then whichever of the three forms it chose works
    const ValuCfgInfo c = valu_cfg_info(cfg);
    const int tiles = a.tiles_r * a.tiles_c;
    tiled_work_kernel<<<dim3((tiles + EW / 64 - 1) / (EW / 64)), dim3(EW), 0, s>>>(a);
    list_count_kernel<Op><<<dim3(a.n), dim3(EW), 0, s>>>(a, c.bm);
    list_scan_kernel<<<dim3(1), dim3(SCAN_T), 0, s>>>(a);
    list_fill_kernel<Op><<<dim3(a.n), dim3(EW), 0, s>>>(a);
    if (a.rows_limit > 0)
      sparse_square_kernel<Op, 1><<<dim3(a.n), dim3(EW), 0, s>>>(a, c.bm, c.bn);
    if (a.prev) sparse_square_kernel<Op, 2><<<dim3(a.n), dim3(EW), 0, s>>>(a, c.bm, c.bn);
    if (a.prev && a.asym) {
      const int T = (a.n + SY - 1) / SY;
      sparse_square_kernel<Op, 3><<<dim3(a.n), dim3(EW), 0, s>>>(a, c.bm, c.bn);
      symmetrize_kernel<Op><<<dim3(T * (T + 1) / 2), dim3(EW), 0, s>>>(a, c.bm, c.bn);
    }
  }
  // at most four 64x64 tiles per CU: the tile's slabs split over 4 or 2 groups
  const int tiles = a.tiles_r * a.tiles_c, cus = device_cu_count();
  const bool split = cfg == CFG_S64 && a.slabs <= SCHED_SPLIT_SLABS;
  if (split && tiles < 2 * cus)
    sched_kernel<Op, CfgS, 4><<<grid, dim3(1024), 0, s>>>(a);
  else if (split && tiles <= 4 * cus)
    sched_kernel<Op, CfgS, 2><<<grid, dim3(512), 0, s>>>(a);
  else if (cfg == CFG_S64)
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
                                  uint8_t *yocc, const SchedArgs *a, hipStream_t s) {
  const ValuCfgInfo c = valu_cfg_info(cfg);
  const float ident = (op == 0 || op == 6) ? SIMD2_INF : -SIMD2_INF;
  uint32_t bits;
  std::memcpy(&bits, &ident, 4);
  occupancy_kernel<<<dim3(tiles_r * tiles_c), dim3(EW), 0, s>>>(
      D, ld, n, c.bm, c.bn, tiles_c, slabs, bits, xocc, yocc, a ? a->conv_prev : nullptr,
      a ? a->sq_mode : nullptr);
  return hipGetLastError();
}

hipError_t launch_sched_symmetry(const float *D, int64_t ld, int n, int *asym, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int T = (n + SY - 1) / SY;
  symmetry_check_kernel<<<dim3(T * (T + 1) / 2), dim3(EW), 0, s>>>(D, ld, n, asym);
  return hipGetLastError();
}

hipError_t launch_fill_multi(const FillSegments &f, hipStream_t s) {
  if (f.k <= 0) return hipSuccess;
  int64_t most = 0;
  for (int i = 0; i < f.k; ++i) most = std::max(most, f.n[i]);
  if (most <= 0) return hipSuccess;
  fill_multi_kernel<<<dim3(grid_for((most + 3) / 4), f.k), dim3(EW), 0, s>>>(f);
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

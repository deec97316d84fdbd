// Host driver of the scheduled converge loop (sched.hip): workspace layout,
// the squaring sequence and the deferred flag reads.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <cstdint>
#include <vector>

#include "../../include/simd2/converge.h"
#include "../../include/simd2/schedule.h"
#include "../../include/simd2/srgemm.h"
#include "internal.hpp"

using namespace simd2;

namespace {

int hip_status(hipError_t e) {
  if (e == hipSuccess) return SIMD2_SUCCESS;
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return SIMD2_ERROR_NO_DEVICE;
  return SIMD2_ERROR_INTERNAL;
}

int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

bool sched_op_supported(int op) {
  return op == SIMD2_MIN_PLUS || op == SIMD2_MAX_PLUS || op == SIMD2_MAX_MIN ||
         op == SIMD2_MIN_MAX;
}

// Entries per row (on average) the list of a sparse squaring can hold.
int list_per_row(int n) { return std::min(n, 1024); }

// The automatic sparse forms start at this size: below it the list's kernel
// launches cost about as much as a whole tiled squaring.
constexpr int SPARSE_AUTO_MIN_N = 2048;

// Cost model of one squaring, in picoseconds, measured on MI355X at
// n = 8192 with these kernels (every form is exact; the model only picks the
// fastest):
//   tiled: per (i, j, k) term, 13.3 ms / 8192^3 for min_plus and max_plus
//     (add + min3/max3), 22.6 ms / 8192^3 for max_min and min_max;
//   row-sparse: per (listed entry, column), from the slope of the squaring
//     time over 4 .. 512 entries per row (0.034 ms per entry per row);
//   changed-entry: per (changed entry, n), a row pass plus a scatter into
//     every row's accumulators (ER graphs with 60 and 80 changed per row);
//   either sparse form: two passes over the matrix (list fill, occupancy).
constexpr double PS_TERM_ADD = 0.0242, PS_TERM_MINMAX = 0.0411;
constexpr double PS_ROW_ENTRY_COL = 0.507, PS_CHANGED_ENTRY_N = 1.8, PS_LIST_ELEM = 1.3;

// Workspace carve-up (every piece 256-byte aligned).
struct Layout {
  int P, kpad, tiles_r, tiles_c, slabs;
  size_t buf[2], xocc[2], yocc[2], rowchg, colchg, conv, work, mode, rowptr, chgcnt, lkj, lval,
      asym, stats, total;
};

bool layout(int n, int cfg, int max_itr, Layout *L) {
  if (n <= 0 || n > 16 * SCHED_MAX_SLABS || max_itr < 0) return false;
  if (cfg < 0 || cfg >= CFG_COUNT) return false;
  const ValuCfgInfo c = valu_cfg_info(cfg);
  L->P = (int)round_up(n, 128);
  L->kpad = (int)round_up(n, 16);
  L->tiles_r = (n + c.bm - 1) / c.bm;
  L->tiles_c = (n + c.bn - 1) / c.bn;
  L->slabs = (n + 15) / 16;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t at = off;
    off += (size_t)round_up((int64_t)bytes, 256);
    return at;
  };
  const size_t mat = (size_t)L->P * L->P * sizeof(float);
  L->buf[0] = take(mat);
  L->buf[1] = take(mat);
  for (int k = 0; k < 2; ++k) {
    L->xocc[k] = take((size_t)L->tiles_r * L->slabs);
    L->yocc[k] = take((size_t)L->tiles_c * L->slabs);
  }
  const size_t rounds = (size_t)max_itr + 1;
  L->rowchg = take(rounds * L->tiles_r * sizeof(int));
  L->colchg = take(rounds * L->tiles_c * sizeof(int));
  L->conv = take(rounds * sizeof(int));
  L->work = take(rounds * sizeof(unsigned long long));
  L->mode = take(rounds * sizeof(int));
  L->rowptr = take(((size_t)n + 1) * sizeof(int));
  L->chgcnt = take((size_t)n * sizeof(int));
  const size_t cap = (size_t)list_per_row(n) * n;
  L->lkj = take(cap * sizeof(uint32_t));
  L->lval = take(cap * sizeof(float));
  L->asym = take(sizeof(int));
  L->stats = take(6 * sizeof(unsigned long long));
  L->total = off;
  return true;
}

}  // namespace

extern "C" {

size_t simd2_converge_scheduled_workspace(int n, int tile_config, int max_itr) {
  const int cfg = tile_config >= 0 ? tile_config : valu_pick_cfg(SIMD2_MIN_PLUS, n, n, n);
  Layout L;
  return layout(n, cfg, max_itr, &L) ? L.total : 0;
}

int simd2_converge_scheduled(int op, int n, float *X, int ldx, int do_epilogue,
                             int check_mode, int max_itr, const simd2_schedule *schedule,
                             const int *perm, void *workspace, size_t workspace_bytes,
                             hipStream_t stream, int *num_itr, simd2_sched_report *report) {
  if (num_itr) *num_itr = 0;
  if (report) *report = simd2_sched_report{};
  if (!sched_op_supported(op)) return SIMD2_ERROR_NOT_SUPPORTED;
  if (check_mode < 0 || check_mode > 2) return SIMD2_ERROR_NOT_SUPPORTED;
  if (n < 0 || max_itr < 0) return SIMD2_ERROR_INVALID_PROBLEM;
  if (n == 0) return SIMD2_SUCCESS;
  if (ldx < n || n > 16 * SCHED_MAX_SLABS) return SIMD2_ERROR_INVALID_PROBLEM;
  if (!X || !workspace) return SIMD2_ERROR_NULL_POINTER;
  simd2_schedule sc = schedule ? *schedule : simd2_schedule{-1, 0, -1, -1};
  if (sc.tile_config >= CFG_COUNT || sc.predicted_itr < 0) return SIMD2_ERROR_INVALID_PROBLEM;
  const int cfg = sc.tile_config >= 0 ? sc.tile_config : valu_pick_cfg(op, n, n, n);
  // Sparse forms: a fixed limit (entries per row on average) or, with -1,
  // up to the list's capacity where the cost model favours them.  The
  // changed-entry form needs the epilogue and no APLP mapping: it rests on
  // D = D_prev (+) D_prev (x) D_prev.
  const bool rows_auto = sc.row_sparse < 0, changed_auto = sc.changed_sparse < 0;
  const int per_row = rows_auto ? (n >= SPARSE_AUTO_MIN_N ? list_per_row(n) : 0)
                                : std::min(sc.row_sparse, list_per_row(n));
  const int changed_per_row = changed_auto ? (n >= SPARSE_AUTO_MIN_N ? list_per_row(n) : 0)
                                           : std::min(sc.changed_sparse, list_per_row(n));
  const bool rows_on = per_row > 0;
  const bool changed_on = changed_per_row > 0 && do_epilogue && check_mode != CHECK_APLP;
  const ValuCfgInfo ci = valu_cfg_info(cfg);
  const bool addmul = op == SIMD2_MIN_PLUS || op == SIMD2_MAX_PLUS;
  const double ps_term = addmul ? PS_TERM_ADD : PS_TERM_MINMAX;
  Layout L;
  if (!layout(n, cfg, max_itr, &L)) return SIMD2_ERROR_INVALID_PROBLEM;
  if (workspace_bytes < L.total) return SIMD2_ERROR_WORKSPACE;
  if (!workspace_aligned(workspace)) return SIMD2_ERROR_MISALIGNED_OPERAND;

  char *ws = static_cast<char *>(workspace);
  float *buf[2] = {reinterpret_cast<float *>(ws + L.buf[0]),
                   reinterpret_cast<float *>(ws + L.buf[1])};
  uint8_t *xocc[2] = {reinterpret_cast<uint8_t *>(ws + L.xocc[0]),
                      reinterpret_cast<uint8_t *>(ws + L.xocc[1])};
  uint8_t *yocc[2] = {reinterpret_cast<uint8_t *>(ws + L.yocc[0]),
                      reinterpret_cast<uint8_t *>(ws + L.yocc[1])};
  int *rowchg = reinterpret_cast<int *>(ws + L.rowchg);
  int *colchg = reinterpret_cast<int *>(ws + L.colchg);
  int *conv = reinterpret_cast<int *>(ws + L.conv);
  auto *work = reinterpret_cast<unsigned long long *>(ws + L.work);
  int *mode = reinterpret_cast<int *>(ws + L.mode);
  int *rowptr = reinterpret_cast<int *>(ws + L.rowptr);
  int *chgcnt = reinterpret_cast<int *>(ws + L.chgcnt);
  int *asym = reinterpret_cast<int *>(ws + L.asym);
  uint32_t *lkj = reinterpret_cast<uint32_t *>(ws + L.lkj);
  float *lval = reinterpret_cast<float *>(ws + L.lval);
  unsigned long long *stats = reinterpret_cast<unsigned long long *>(ws + L.stats);
  const float inf = __builtin_inff();
  const float ident = (op == SIMD2_MIN_PLUS || op == SIMD2_MIN_MAX) ? inf : -inf;

  // Padded copies: everything outside the live n x n block is the identity.
  // (the identity in both buffers and every counter cleared: one launch)
  const int64_t P2 = (int64_t)L.P * L.P;
  uint32_t ident_bits;
  std::memcpy(&ident_bits, &ident, 4);
  FillSegments fills;
  fills.add(buf[0], P2, ident_bits);
  fills.add(buf[1], P2, ident_bits);
  fills.add(rowchg, (int64_t)(max_itr + 1) * L.tiles_r, 0);
  fills.add(colchg, (int64_t)(max_itr + 1) * L.tiles_c, 0);
  fills.add(conv, max_itr + 1, 0);
  fills.add(work, 2 * (int64_t)(max_itr + 1), 0);
  fills.add(mode, max_itr + 1, 0);
  fills.add(stats, 12, 0);  // 6 64-bit counters
  hipError_t e = launch_fill_multi(fills, stream);
  if (e == hipSuccess) e = launch_permute_copy(buf[0], L.P, X, ldx, n, perm, 0, stream);
  // a symmetric input stays symmetric (every (x) here commutes bit for bit),
  // which the changed-entry form can use
  if (e == hipSuccess && changed_on) e = launch_fill_i32(asym, 0, 1, stream);
  if (e == hipSuccess && changed_on) e = launch_sched_symmetry(buf[0], L.P, n, asym, stream);
  if (e == hipSuccess)
    e = launch_sched_occupancy(op, cfg, buf[0], L.P, n, L.tiles_r, L.tiles_c, L.slabs, xocc[0],
                               yocc[0], nullptr, stream);
  if (e != hipSuccess) return hip_status(e);

  int t = 0, T = -1, syncs = 0;
  int chunk = sc.predicted_itr > 0 ? sc.predicted_itr : 1;
  std::vector<int> flags;
  while (t < max_itr && T < 0) {
    const int m = std::min(chunk, max_itr - t);
    for (int j = 0; j < m && e == hipSuccess; ++j) {
      const int s = t + j;
      SchedArgs a;
      a.cur = buf[s & 1];
      a.nxt = buf[(s + 1) & 1];
      a.ld = L.P;
      a.n = n;
      a.kpad = L.kpad;
      a.epi = do_epilogue ? 1 : 0;
      a.mode = check_mode;
      a.tiles_r = L.tiles_r;
      a.tiles_c = L.tiles_c;
      a.slabs = L.slabs;
      a.xocc_in = xocc[s & 1];
      a.xocc_out = xocc[(s + 1) & 1];
      a.yocc_in = yocc[s & 1];
      a.yocc_out = yocc[(s + 1) & 1];
      a.rowchg_prev = s > 0 ? rowchg + (int64_t)(s - 1) * L.tiles_r : nullptr;
      a.colchg_prev = s > 0 ? colchg + (int64_t)(s - 1) * L.tiles_c : nullptr;
      a.rowchg = rowchg + (int64_t)s * L.tiles_r;
      a.colchg = colchg + (int64_t)s * L.tiles_c;
      a.conv_prev = s > 0 ? conv + s - 1 : nullptr;
      a.conv = conv + s;
      a.stats = stats;
      // the previous squaring's input is the buffer this one overwrites; it
      // is read (to list the changed entries) before anything is written
      a.prev = changed_on && s > 0 ? a.nxt : nullptr;
      a.asym = changed_on ? asym : nullptr;
      a.rows_limit = rows_on ? per_row * n : 0;
      a.changed_limit = a.prev ? changed_per_row * n : 0;
      a.rows_auto = rows_auto ? 1 : 0;
      a.changed_auto = changed_auto ? 1 : 0;
      a.cost_slab_tile = (float)(16.0 * ci.bm * ci.bn * ps_term);
      a.cost_list = (float)((double)n * n * PS_LIST_ELEM);
      a.cost_row_entry = (float)(n * PS_ROW_ENTRY_COL);
      a.cost_changed_entry = (float)(n * PS_CHANGED_ENTRY_N);
      a.work = work + s;
      a.sq_mode = rows_on || a.prev ? mode + s : nullptr;
      a.rowptr = rowptr;
      a.chgcnt = chgcnt;
      a.lkj = lkj;
      a.lval = lval;
      e = launch_sched_square(op, cfg, a, stream);
      if (e == hipSuccess && a.sq_mode)  // maps after a sparse squaring
        e = launch_sched_occupancy(op, cfg, a.nxt, L.P, n, L.tiles_r, L.tiles_c, L.slabs,
                                   a.xocc_out, a.yocc_out, &a, stream);
    }
    if (e != hipSuccess) return hip_status(e);
    flags.resize(m);
    e = hipMemcpyAsync(flags.data(), conv + t, m * sizeof(int), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip_status(e);
    ++syncs;
    for (int j = 0; j < m; ++j)
      if (flags[j] == 0) {
        T = t + j;
        break;
      }
    if (report) report->squarings_enqueued += m;
    t += m;
    chunk = 1;
  }
  const int done = T >= 0 ? T + 1 : t;  // squarings whose result counts
  float *result = buf[done & 1];
  e = launch_permute_copy(X, ldx, result, L.P, n, perm, 1, stream);
  unsigned long long st[6] = {0, 0, 0, 0, 0, 0};
  if (e == hipSuccess && report)
    e = hipMemcpyAsync(st, stats, sizeof(st), hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) return hip_status(e);
  if (num_itr) *num_itr = done;
  if (report) {
    const unsigned long long tiles = (unsigned long long)L.tiles_r * L.tiles_c;
    report->host_syncs = syncs;
    report->tiles_total = tiles * done;
    report->tiles_computed = st[0];
    report->slab_tiles_total = tiles * L.slabs * done;
    report->slab_tiles_computed = st[1];
    report->rows_sparse = st[2];
    report->row_entries = st[3];
    report->changed_rows = st[4];
    report->changed_entries = st[5];
  }
  return T >= 0 ? SIMD2_SUCCESS : SIMD2_ERROR_NOT_CONVERGED;
}

}  // extern "C"

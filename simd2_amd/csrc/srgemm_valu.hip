// LDS-tiled VALU semiring GEMM for gfx950 (the seven float semirings whose
// (+) is min/max, plus the direct minus_square and a VALU plus_mul).
//
// Work decomposition (row-major problem of internal.hpp):
//   * one workgroup computes a BM x BN output tile; each thread a TM x TN
//     register tile, split into 4x4 sub-blocks spaced BM/(TM/4) rows and
//     BN/(TN/4) columns apart so that every LDS read is one ds_read_b128 and
//     the 16 (or more) lanes reading one row of Y hit distinct banks;
//   * K advances in BK-deep slabs, double-buffered in LDS: the next slab's
//     global loads (128-bit, coalesced along K for X and along columns for Y)
//     are issued before the current slab's arithmetic and written to the
//     other LDS buffer afterwards, one barrier per slab;
//   * X is stored transposed in LDS (k-major, padded by 4 floats so the
//     transposing 32-bit writes of a wave fall in distinct banks);
//   * the inner step handles two k at once: per output, two (x) and one
//     three-input (+) (v_min3_f32 / v_max3_f32), i.e. 3 VALU instructions
//     per 4 semiring operations;
//   * the epilogue fuses C, the convergence change test and per-tile change
//     flags, so the repeated-squaring driver needs no separate pass.
// Tiles are walked XCD-aware: the 8 XCDs receive workgroups round-robin, so
// each XCD is handed one contiguous range of tiles, and inside it tiles are
// visited in groups of GROUP tile-rows so neighbours share panels in L2.
#include <mutex>

#include "valu_tile.hpp"

namespace simd2 {

namespace {

// Launch forms: FAST -- the problem is a whole number of tiles and slabs and
// every operand is 16-byte aligned row by row; SPLIT -- aligned, but the
// shape is off the grid: tiles inside the matrix run the 128-bit path (their
// last, partial K slab checked), the edge tiles the guarded 128-bit path
// (rows past R read as padding, a column group past Cn element by element,
// stores masked); CHECKED -- per-element bounds everywhere.
enum { FORM_FAST = 0, FORM_SPLIT = 1, FORM_CHECKED = 2 };

template <class Op, int BM, int BN, int TM, int TN, int BK, int WAVES, int FORM>
__global__ __launch_bounds__((BM / TM) * (BN / TN))
__attribute__((amdgpu_waves_per_eu(WAVES))) void srgemm_valu_kernel(
    int R, int Cn, int K, const float *__restrict__ X, int64_t ldx,
    const float *__restrict__ Y, int64_t ldy, float *O, int64_t ldo,
    Epilogue e, int tiles_r, int tiles_c) {
  __shared__ ValuSmem<BM, BN, BK> sm;
  int tr, tc;
  map_tile(bid_x(), tiles_r, tiles_c, tr, tc);
  if (e.gate && e.gate[tr * e.gate_ld + tc] == 0) return;
  int *flag = e.tile_flags ? &e.tile_flags[tr * tiles_c + tc] : nullptr;
  const int row0 = tr * BM, col0 = tc * BN;
  if (FORM == FORM_FAST || (FORM == FORM_SPLIT && row0 + BM <= R && col0 + BN <= Cn))
    valu_tile<Op, BM, BN, TM, TN, BK, false, false, FORM == FORM_SPLIT>(
        R, Cn, K, X, ldx, Y, ldy, O, ldo, e, row0, col0, flag, sm);
  else if (FORM == FORM_SPLIT)
    valu_tile<Op, BM, BN, TM, TN, BK, false, true, true, true>(R, Cn, K, X, ldx, Y, ldy, O, ldo, e,
                                                               row0, col0, flag, sm);
  else
    valu_tile<Op, BM, BN, TM, TN, BK, true>(R, Cn, K, X, ldx, Y, ldy, O, ldo, e, row0, col0,
                                            flag, sm);
}

// Small problems (at most four 64x64 tiles per CU; n = 1024 has one): the K
// loop of each tile is split over G groups of 256 threads in one workgroup
// (G = 4 up to two tiles per CU, else 2), each over a contiguous range of
// slabs with its own LDS buffers, and the partial tiles are combined in
// group order, ((p0 (+) p1) (+) p2) (+) p3, before the usual epilogue: G
// times the resident waves.  Only for the semirings whose (+) is min or max,
// which is exact: regrouping the terms changes no bit.  CHECKED: any shape
// and alignment, per-element bounds on loads and stores; every group runs
// ceil(slabs / G) slabs, the ranges past K reading padding, whose terms are
// the (+)-identity.
template <class Op, int SK_GROUPS, bool CHECKED>
__global__ __launch_bounds__(256 * SK_GROUPS) void srgemm_valu_splitk_kernel(
    int R, int Cn, int K, const float *__restrict__ X, int64_t ldx,
    const float *__restrict__ Y, int64_t ldy, float *O, int64_t ldo, Epilogue e,
    int tiles_r, int tiles_c) {
  using S = CfgS;
  __shared__ ValuSmem<S::BM, S::BN, S::BK> sm[SK_GROUPS];
  int tr, tc;
  map_tile(bid_x(), tiles_r, tiles_c, tr, tc);
  int *flag = e.tile_flags ? &e.tile_flags[tr * tiles_c + tc] : nullptr;
  const int row0 = tr * S::BM, col0 = tc * S::BN;
  const int g = tid_x() / 256, t = tid_x() % 256;
  if (tid_x() == 0) sm[0].changed = 0;  // published by the main loop's barriers
  const int per = ((K + S::BK - 1) / S::BK + SK_GROUPS - 1) / SK_GROUPS;
  float acc[S::TM][S::TN];
  valu_mainloop<Op, S::BM, S::BN, S::TM, S::TN, S::BK, CHECKED, false>(
      R, Cn, K, X, ldx, Y, ldy, row0, col0, nullptr, 0, sm[g], acc, t, g * per, (g + 1) * per);
  // The main loop ends with a barrier after every read of the slabs: groups
  // 1.. park their partial tiles in their own LDS (16 KB of its 16.5).
  static_assert(S::TM * S::TN * 256 * sizeof(float) <= sizeof(ValuSmem<S::BM, S::BN, S::BK>) -
                                                           sizeof(int), "partial tile");
  if (g > 0) {
    float *part = reinterpret_cast<float *>(&sm[g]);
#pragma unroll
    for (int i = 0; i < S::TM; ++i)
#pragma unroll
      for (int j = 0; j < S::TN; ++j) part[(i * S::TN + j) * 256 + t] = acc[i][j];
  }
  __syncthreads();
  if (g == 0) {
#pragma unroll
    for (int q = 1; q < SK_GROUPS; ++q) {
      const float *part = reinterpret_cast<const float *>(&sm[q]);
#pragma unroll
      for (int i = 0; i < S::TM; ++i)
#pragma unroll
        for (int j = 0; j < S::TN; ++j)
          acc[i][j] = Op::add(acc[i][j], part[(i * S::TN + j) * 256 + t]);
    }
  }
  valu_epilogue<Op, S::BM, S::BN, S::TM, S::TN, CHECKED, false>(R, Cn, O, ldo, e, row0, col0,
                                                                flag, &sm[0].changed, acc, t,
                                                                g == 0);
}

// groups > 0: the fast form (whole tiles and slab groups, aligned rows);
// groups < 0: the checked form with -groups groups.
template <class Op>
hipError_t launch_splitk(int R, int Cn, int K, const float *X, int64_t ldx, const float *Y,
                         int64_t ldy, float *O, int64_t ldo, const Epilogue &e, int groups,
                         hipStream_t s) {
  const int tiles_r = (R + CfgS::BM - 1) / CfgS::BM, tiles_c = (Cn + CfgS::BN - 1) / CfgS::BN;
  const dim3 grid(tiles_r * tiles_c);
  if (groups == 4)
    srgemm_valu_splitk_kernel<Op, 4, false><<<grid, dim3(1024), 0, s>>>(
        R, Cn, K, X, ldx, Y, ldy, O, ldo, e, tiles_r, tiles_c);
  else if (groups == 2)
    srgemm_valu_splitk_kernel<Op, 2, false><<<grid, dim3(512), 0, s>>>(
        R, Cn, K, X, ldx, Y, ldy, O, ldo, e, tiles_r, tiles_c);
  else if (groups == -4)
    srgemm_valu_splitk_kernel<Op, 4, true><<<grid, dim3(1024), 0, s>>>(
        R, Cn, K, X, ldx, Y, ldy, O, ldo, e, tiles_r, tiles_c);
  else
    srgemm_valu_splitk_kernel<Op, 2, true><<<grid, dim3(512), 0, s>>>(
        R, Cn, K, X, ldx, Y, ldy, O, ldo, e, tiles_r, tiles_c);
  return hipGetLastError();
}

template <class Op, class Cfg>
hipError_t launch_cfg(int R, int Cn, int K, const float *X, int64_t ldx,
                      const float *Y, int64_t ldy, float *O, int64_t ldo,
                      const Epilogue &e, hipStream_t s) {
  const int tiles_r = (R + Cfg::BM - 1) / Cfg::BM;
  const int tiles_c = (Cn + Cfg::BN - 1) / Cfg::BN;
  const int nthreads = (Cfg::BM / Cfg::TM) * (Cfg::BN / Cfg::TN);
  auto aligned = [](const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
  bool vec = ldx % 4 == 0 && ldy % 4 == 0 && ldo % 4 == 0 && aligned(X) && aligned(Y) &&
             aligned(O);
  if (e.C) vec = vec && e.ldc % 4 == 0 && aligned(e.C);
  if (e.P) vec = vec && e.ldp % 4 == 0 && aligned(e.P);
  const bool grid_fit = R % Cfg::BM == 0 && Cn % Cfg::BN == 0 && K % Cfg::BK == 0;
  // the split form pays off once some tile lies wholly inside the matrix
  const bool split = vec && R >= Cfg::BM && Cn >= Cfg::BN;
  dim3 grid(tiles_r * tiles_c), block(nthreads);
  if (vec && grid_fit)
    srgemm_valu_kernel<Op, Cfg::BM, Cfg::BN, Cfg::TM, Cfg::TN, Cfg::BK, Cfg::WAVES, FORM_FAST>
        <<<grid, block, 0, s>>>(R, Cn, K, X, ldx, Y, ldy, O, ldo, e, tiles_r, tiles_c);
  else if (split)
    srgemm_valu_kernel<Op, Cfg::BM, Cfg::BN, Cfg::TM, Cfg::TN, Cfg::BK, Cfg::WAVES, FORM_SPLIT>
        <<<grid, block, 0, s>>>(R, Cn, K, X, ldx, Y, ldy, O, ldo, e, tiles_r, tiles_c);
  else
    srgemm_valu_kernel<Op, Cfg::BM, Cfg::BN, Cfg::TM, Cfg::TN, Cfg::BK, Cfg::WAVES, FORM_CHECKED>
        <<<grid, block, 0, s>>>(R, Cn, K, X, ldx, Y, ldy, O, ldo, e, tiles_r, tiles_c);
  return hipGetLastError();
}

template <class Op>
hipError_t launch_op(int R, int Cn, int K, const float *X, int64_t ldx,
                     const float *Y, int64_t ldy, float *O, int64_t ldo,
                     const Epilogue &e, hipStream_t s, int cfg) {
  if (cfg == CFG_S64) return launch_cfg<Op, CfgS>(R, Cn, K, X, ldx, Y, ldy, O, ldo, e, s);
  if (cfg == CFG_M128x64) return launch_cfg<Op, CfgM>(R, Cn, K, X, ldx, Y, ldy, O, ldo, e, s);
  return launch_cfg<Op, CfgL>(R, Cn, K, X, ldx, Y, ldy, O, ldo, e, s);
}

}  // namespace

int device_cu_count() {
  static int cus = 0;
  if (cus == 0) {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0)
      cus = n;
    else
      cus = 256;
  }
  return cus;
}

ValuCfgInfo valu_cfg_info(int cfg) {
  if (cfg == CFG_S64) return {CfgS::BM, CfgS::BN, CfgS::BK, 256};
  if (cfg == CFG_M128x64) return {CfgM::BM, CfgM::BN, CfgM::BK, 256};
  return {CfgL::BM, CfgL::BN, CfgL::BK, 256};
}

// Tile shape choice.  Rule measured on MI355X for square problems
// (n = 1024 .. 8192, all ops): the VALU semirings need many resident waves to
// dual-issue, so up to ~3072^2 outputs the 64x64 tiles (7-8 waves per SIMD)
// win; at 4096^2 the 128x64 tiles balance 2048 tiles over 1024 slots; from
// ~6144^2 up the 128x128 tiles (fewest LDS reads per op) win.  minus_square
// with a short K (the kNN distance step: 16384^2 x 64..128) keeps the 128x64
// tiles at every larger area (1.38 against 1.52 ms for 128x128 and 1.53 for
// 64x64 at 16384^2 x 128).  Shapes the autotuner has timed (simd2_autotune)
// use the recorded winner instead.
namespace {
struct TunedEntry { int op, R, Cn, K, cfg; };
constexpr int TUNED_MAX = 64;
TunedEntry g_tuned[TUNED_MAX];
int g_tuned_n = 0;
std::mutex g_tuned_mu;
}  // namespace

void valu_record_cfg(int op, int R, int Cn, int K, int cfg) {
  std::lock_guard<std::mutex> lock(g_tuned_mu);
  for (int i = 0; i < g_tuned_n; ++i)
    if (g_tuned[i].op == op && g_tuned[i].R == R && g_tuned[i].Cn == Cn && g_tuned[i].K == K) {
      g_tuned[i].cfg = cfg;
      return;
    }
  const int slot = g_tuned_n < TUNED_MAX ? g_tuned_n++ : (op + R + Cn + K) % TUNED_MAX;
  g_tuned[slot] = {op, R, Cn, K, cfg};
}

int valu_pick_cfg(int op, int R, int Cn, int K) {
  {
    std::lock_guard<std::mutex> lock(g_tuned_mu);
    for (int i = 0; i < g_tuned_n; ++i)
      if (g_tuned[i].op == op && g_tuned[i].R == R && g_tuned[i].Cn == Cn && g_tuned[i].K == K)
        return g_tuned[i].cfg;
  }
  const double area = (double)R * Cn;
  const double scale = 256.0 / device_cu_count();   // rule is for 256 CUs
  const double a = area * scale;
  if (a <= 3072.0 * 3072.0 * 1.01) return CFG_S64;
  if (a <= 4096.0 * 4096.0 * 1.01) return CFG_M128x64;
  if (op == 8 && K <= 128) return CFG_M128x64;
  return CFG_L128;
}

// The split-K form (automatic choice only): a min/max semiring, the 64x64
// tiles chosen, at most four tiles per CU, no gate.  Returns the group count
// G for the fast form (a whole number of tiles and of G-slab groups, every
// operand 16-byte aligned row by row), -G for the checked form (operands
// whose rows are not all 16-byte aligned, where the unsplit kernel takes its
// checked body too, and aligned off-grid shapes of at most two tiles per CU,
// where four groups' waves outweigh the unsplit kernel's 128-bit interior
// loads: 1000x776x300 min_plus 39 -> 31 us), 0 when the split does not apply.
static int splitk_groups(int op, int R, int Cn, int K, const float *X, int64_t ldx,
                         const float *Y, int64_t ldy, const float *O, int64_t ldo,
                         const Epilogue &e) {
  if (op == 1 || op == 8 || op < 0 || op > 6 || e.gate || K <= 0) return 0;
  const int64_t tiles =
      (int64_t)((R + CfgS::BM - 1) / CfgS::BM) * ((Cn + CfgS::BN - 1) / CfgS::BN);
  const int groups = tiles < 2 * device_cu_count() ? 4 : tiles <= 4 * device_cu_count() ? 2 : 0;
  if (groups == 0) return 0;
  auto aligned = [](const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
  bool vec = ldx % 4 == 0 && ldy % 4 == 0 && ldo % 4 == 0 && aligned(X) && aligned(Y) &&
             aligned(O);
  if (e.C) vec = vec && e.ldc % 4 == 0 && aligned(e.C);
  if (e.P) vec = vec && e.ldp % 4 == 0 && aligned(e.P);
  if (!vec) return -groups;
  if (R % CfgS::BM == 0 && Cn % CfgS::BN == 0 && K % (CfgS::BK * groups) == 0) return groups;
  return groups == 4 ? -4 : 0;
}

hipError_t launch_valu_f32(int op, int R, int Cn, int K, const float *X,
                           int64_t ldx, const float *Y, int64_t ldy, float *O,
                           int64_t ldo, const Epilogue &e, hipStream_t s,
                           int cfg) {
  if (cfg == CFG_AUTO) {
    cfg = valu_pick_cfg(op, R, Cn, K);
    const int g = cfg == CFG_S64 ? splitk_groups(op, R, Cn, K, X, ldx, Y, ldy, O, ldo, e) : 0;
    if (g) {
      switch (op) {
      case 0: return launch_splitk<MinPlus>(R, Cn, K, X, ldx, Y, ldy, O, ldo, e, g, s);
      case 2: return launch_splitk<MaxPlus>(R, Cn, K, X, ldx, Y, ldy, O, ldo, e, g, s);
      case 3: return launch_splitk<MaxMin>(R, Cn, K, X, ldx, Y, ldy, O, ldo, e, g, s);
      case 4: return launch_splitk<MaxMul>(R, Cn, K, X, ldx, Y, ldy, O, ldo, e, g, s);
      case 5: return launch_splitk<MinMul>(R, Cn, K, X, ldx, Y, ldy, O, ldo, e, g, s);
      case 6: return launch_splitk<MinMax>(R, Cn, K, X, ldx, Y, ldy, O, ldo, e, g, s);
      default: break;
      }
    }
  }
  switch (op) {
  case 0: return launch_op<MinPlus>(R, Cn, K, X, ldx, Y, ldy, O, ldo, e, s, cfg);
  case 1: return launch_op<PlusMul>(R, Cn, K, X, ldx, Y, ldy, O, ldo, e, s, cfg);
  case 2: return launch_op<MaxPlus>(R, Cn, K, X, ldx, Y, ldy, O, ldo, e, s, cfg);
  case 3: return launch_op<MaxMin>(R, Cn, K, X, ldx, Y, ldy, O, ldo, e, s, cfg);
  case 4: return launch_op<MaxMul>(R, Cn, K, X, ldx, Y, ldy, O, ldo, e, s, cfg);
  case 5: return launch_op<MinMul>(R, Cn, K, X, ldx, Y, ldy, O, ldo, e, s, cfg);
  case 6: return launch_op<MinMax>(R, Cn, K, X, ldx, Y, ldy, O, ldo, e, s, cfg);
  case 8: return launch_op<MinusSquare>(R, Cn, K, X, ldx, Y, ldy, O, ldo, e, s, cfg);
  default: return hipErrorInvalidValue;
  }
}

}  // namespace simd2

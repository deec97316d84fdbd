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
#include <float.h>

#include <mutex>

#include "internal.hpp"
#include "semiring_ops.hpp"

namespace simd2 {

namespace {

constexpr int GROUP = 8;
constexpr int NUM_XCD = 8;

__device__ __forceinline__ void map_tile(int b, int tiles_r, int tiles_c,
                                         int &tr, int &tc) {
  const int nt = tiles_r * tiles_c;
  int t = b;
  if (nt % NUM_XCD == 0) t = (b % NUM_XCD) * (nt / NUM_XCD) + b / NUM_XCD;
  const int width = GROUP * tiles_c;
  const int group = t / width;
  const int first = group * GROUP;
  const int gsz = min(tiles_r - first, GROUP);
  const int in = t - group * width;
  tr = first + in % gsz;
  tc = in / gsz;
}

// Change test of kernel/converge.cuh.  The threshold form only subtracts when
// the bit patterns differ: equal infinities (whose difference is NaN, which
// the reference's comparison treats as "unchanged") never reach the
// subtraction, so the result does not depend on how NaN compares.
__device__ __forceinline__ bool differs(float o, float p, int mode) {
  const bool bits = __float_as_uint(o) != __float_as_uint(p);
  if (mode == CHECK_THRESHOLD) return bits && fabsf(p - o) > 0.0001f;
  if (mode == CHECK_APLP) return o != p;
  return bits;
}

// Load one BK-deep slab of X (BM x BK, k contiguous) and Y (BK x BN) into
// registers.  Fast path: 128-bit loads; checked path: per-element bounds
// with the op's padding values.
template <class Op, int BM, int BN, int BK, int NT, bool CHECKED>
__device__ __forceinline__ void load_slab(float4 (&xv)[BM * BK / 4 / NT],
                                          float4 (&yv)[BK * BN / 4 / NT],
                                          const float *__restrict__ X, int64_t ldx,
                                          const float *__restrict__ Y, int64_t ldy,
                                          int R, int Cn, int K, int row0, int col0,
                                          int k0, int tid) {
  constexpr int XL = BM * BK / 4 / NT, YL = BK * BN / 4 / NT;
#pragma unroll
  for (int h = 0; h < XL; ++h) {
    const int f = tid + NT * h;
    const int r = row0 + f / (BK / 4);
    const int k = k0 + (f % (BK / 4)) * 4;
    if constexpr (!CHECKED) {
      xv[h] = *reinterpret_cast<const float4 *>(X + (int64_t)r * ldx + k);
    } else {
      float t[4];
#pragma unroll
      for (int q = 0; q < 4; ++q)
        t[q] = (r < R && k + q < K) ? X[(int64_t)r * ldx + k + q] : Op::pad_x;
      xv[h] = make_float4(t[0], t[1], t[2], t[3]);
    }
  }
#pragma unroll
  for (int h = 0; h < YL; ++h) {
    const int f = tid + NT * h;
    const int k = k0 + f / (BN / 4);
    const int c = col0 + (f % (BN / 4)) * 4;
    if constexpr (!CHECKED) {
      yv[h] = *reinterpret_cast<const float4 *>(Y + (int64_t)k * ldy + c);
    } else {
      float t[4];
#pragma unroll
      for (int q = 0; q < 4; ++q)
        t[q] = (k < K && c + q < Cn) ? Y[(int64_t)k * ldy + c + q] : Op::pad_y;
      yv[h] = make_float4(t[0], t[1], t[2], t[3]);
    }
  }
}

template <int BM, int BN, int BK, int NT, int XLD>
__device__ __forceinline__ void store_slab(float (*Xs)[XLD], float (*Ys)[BN],
                                           const float4 (&xv)[BM * BK / 4 / NT],
                                           const float4 (&yv)[BK * BN / 4 / NT],
                                           int tid) {
  constexpr int XL = BM * BK / 4 / NT, YL = BK * BN / 4 / NT;
#pragma unroll
  for (int h = 0; h < XL; ++h) {
    const int f = tid + NT * h;
    const int r = f / (BK / 4);
    const int k = (f % (BK / 4)) * 4;
    Xs[k + 0][r] = xv[h].x;
    Xs[k + 1][r] = xv[h].y;
    Xs[k + 2][r] = xv[h].z;
    Xs[k + 3][r] = xv[h].w;
  }
#pragma unroll
  for (int h = 0; h < YL; ++h) {
    const int f = tid + NT * h;
    const int k = f / (BN / 4);
    const int c = (f % (BN / 4)) * 4;
    *reinterpret_cast<float4 *>(&Ys[k][c]) = yv[h];
  }
}

template <class Op, int BM, int BN, int TM, int TN, int BK, int WAVES, bool CHECKED>
__global__ __launch_bounds__((BM / TM) * (BN / TN))
__attribute__((amdgpu_waves_per_eu(WAVES))) void srgemm_valu_kernel(
    int R, int Cn, int K, const float *__restrict__ X, int64_t ldx,
    const float *__restrict__ Y, int64_t ldy, float *O, int64_t ldo,
    Epilogue e, int tiles_r, int tiles_c) {
  constexpr int NTX = BN / TN;          // threads along columns
  constexpr int NTY = BM / TM;          // threads along rows
  constexpr int NT = NTX * NTY;
  constexpr int RS = BM / (TM / 4);     // row stride between 4-row groups
  constexpr int CS = BN / (TN / 4);     // column stride between 4-col groups
  constexpr int XLD = BM + 4;           // padded row of the transposed X slab
  constexpr int XL = BM * BK / 4 / NT;  // float4 loads per thread (X)
  constexpr int YL = BK * BN / 4 / NT;  // float4 loads per thread (Y)
  static_assert((BM * BK / 4) % NT == 0 && (BK * BN / 4) % NT == 0, "slab");
  static_assert(TM % 4 == 0 && TN % 4 == 0 && BK % 2 == 0, "tile shape");
  static_assert(NTY * TM == BM && NTX * TN == BN, "thread grid");

  __shared__ __attribute__((aligned(16))) float Xs[2][BK][XLD];
  __shared__ __attribute__((aligned(16))) float Ys[2][BK][BN];
  __shared__ int s_changed;

  const int tid = threadIdx.x;
  const int tx = tid % NTX;
  const int ty = tid / NTX;
  int tr, tc;
  map_tile(blockIdx.x, tiles_r, tiles_c, tr, tc);
  const int row0 = tr * BM;
  const int col0 = tc * BN;
  if (tid == 0) s_changed = 0;

  float acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = Op::identity;

  float4 xv[XL], yv[YL];
  const int nkt = (K + BK - 1) / BK;
  if (nkt > 0) {
    load_slab<Op, BM, BN, BK, NT, CHECKED>(xv, yv, X, ldx, Y, ldy, R, Cn, K, row0, col0, 0, tid);
    store_slab<BM, BN, BK, NT, XLD>(Xs[0], Ys[0], xv, yv, tid);
  }
  __syncthreads();

  for (int kt = 0; kt < nkt; ++kt) {
    const int buf = kt & 1;
    // The prefetch is unconditional (the last slab is simply loaded again and
    // written to the idle buffer, which nothing reads afterwards): under a
    // branch the compiler keeps the prefetch in scratch and waits for it.
    const int knext = (kt + 1 < nkt ? kt + 1 : kt) * BK;
    load_slab<Op, BM, BN, BK, NT, CHECKED>(xv, yv, X, ldx, Y, ldy, R, Cn, K, row0, col0, knext,
                                           tid);
#pragma unroll
    for (int k = 0; k < BK; k += 2) {
      // Keep each k-pair's LDS reads next to its arithmetic: without this
      // the scheduler hoists the reads of the whole slab and spills.
      __builtin_amdgcn_sched_barrier(0);
      float a0[TM], a1[TM], b0[TN], b1[TN];
#pragma unroll
      for (int g = 0; g < TM / 4; ++g) {
        *reinterpret_cast<float4 *>(&a0[g * 4]) =
            *reinterpret_cast<const float4 *>(&Xs[buf][k][g * RS + ty * 4]);
        *reinterpret_cast<float4 *>(&a1[g * 4]) =
            *reinterpret_cast<const float4 *>(&Xs[buf][k + 1][g * RS + ty * 4]);
      }
#pragma unroll
      for (int g = 0; g < TN / 4; ++g) {
        *reinterpret_cast<float4 *>(&b0[g * 4]) =
            *reinterpret_cast<const float4 *>(&Ys[buf][k][g * CS + tx * 4]);
        *reinterpret_cast<float4 *>(&b1[g * 4]) =
            *reinterpret_cast<const float4 *>(&Ys[buf][k + 1][g * CS + tx * 4]);
      }
#pragma unroll
      for (int i = 0; i < TM; ++i) {
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          if constexpr (is_fused<Op>::value)
            acc[i][j] = Op::fma_step(Op::fma_step(acc[i][j], a0[i], b0[j]), a1[i], b1[j]);
          else
            acc[i][j] = Op::add3(acc[i][j], Op::mul(a0[i], b0[j]), Op::mul(a1[i], b1[j]));
        }
      }
    }
    store_slab<BM, BN, BK, NT, XLD>(Xs[buf ^ 1], Ys[buf ^ 1], xv, yv, tid);
    __syncthreads();
  }

  // Epilogue: (+) C, optional APLP mapping, change test, store.
  bool changed = false;
#pragma unroll
  for (int gi = 0; gi < TM / 4; ++gi)
#pragma unroll
    for (int ii = 0; ii < 4; ++ii) {
      const int i = gi * 4 + ii;
      const int r = row0 + gi * RS + ty * 4 + ii;
#pragma unroll
      for (int gj = 0; gj < TN / 4; ++gj) {
        const int c = col0 + gj * CS + tx * 4;
        float v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = acc[i][gj * 4 + q];
        if constexpr (!CHECKED) {
          if (e.C) {
            const float4 cv = *reinterpret_cast<const float4 *>(e.C + (int64_t)r * e.ldc + c);
            v[0] = Op::add(v[0], cv.x); v[1] = Op::add(v[1], cv.y);
            v[2] = Op::add(v[2], cv.z); v[3] = Op::add(v[3], cv.w);
          }
          if (e.mode == CHECK_APLP) {
#pragma unroll
            for (int q = 0; q < 4; ++q) v[q] = (v[q] > -1.0f && v[q] < 1.0f) ? -FLT_MAX : v[q];
          }
          if (e.P) {
            const float4 pv = *reinterpret_cast<const float4 *>(e.P + (int64_t)r * e.ldp + c);
            changed |= differs(v[0], pv.x, e.mode) | differs(v[1], pv.y, e.mode) |
                       differs(v[2], pv.z, e.mode) | differs(v[3], pv.w, e.mode);
          }
          *reinterpret_cast<float4 *>(O + (int64_t)r * ldo + c) = make_float4(v[0], v[1], v[2], v[3]);
        } else {
          if (r < R) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              if (c + q < Cn) {
                float o = v[q];
                if (e.C) o = Op::add(o, e.C[(int64_t)r * e.ldc + c + q]);
                if (e.mode == CHECK_APLP && o > -1.0f && o < 1.0f) o = -FLT_MAX;
                if (e.P) changed |= differs(o, e.P[(int64_t)r * e.ldp + c + q], e.mode);
                O[(int64_t)r * ldo + c + q] = o;
              }
            }
          }
        }
      }
    }

  if (e.flag || e.tile_flags) {
    const bool any = __any(changed);
    if (e.flag && any && (tid % warpSize) == 0) atomicOr(e.flag, 1);
    if (e.tile_flags) {
      if (any && (tid % warpSize) == 0) atomicOr(&s_changed, 1);
      __syncthreads();
      if (tid == 0) e.tile_flags[tr * tiles_c + tc] = s_changed;
    }
  }
}

// WAVES: waves per SIMD the register allocation must leave room for
// (3 -> at most 168 VGPRs for the 8x8 register tile).
struct CfgL { static constexpr int BM = 128, BN = 128, TM = 8, TN = 8, BK = 16, WAVES = 3; };
struct CfgM { static constexpr int BM = 128, BN = 64, TM = 8, TN = 4, BK = 16, WAVES = 4; };
struct CfgS { static constexpr int BM = 64, BN = 64, TM = 4, TN = 4, BK = 16, WAVES = 4; };

template <class Op, class Cfg>
hipError_t launch_cfg(int R, int Cn, int K, const float *X, int64_t ldx,
                      const float *Y, int64_t ldy, float *O, int64_t ldo,
                      const Epilogue &e, hipStream_t s) {
  const int tiles_r = (R + Cfg::BM - 1) / Cfg::BM;
  const int tiles_c = (Cn + Cfg::BN - 1) / Cfg::BN;
  const int nthreads = (Cfg::BM / Cfg::TM) * (Cfg::BN / Cfg::TN);
  auto aligned = [](const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; };
  bool fast = R % Cfg::BM == 0 && Cn % Cfg::BN == 0 && K % Cfg::BK == 0 &&
              ldx % 4 == 0 && ldy % 4 == 0 && ldo % 4 == 0 && aligned(X) &&
              aligned(Y) && aligned(O);
  if (e.C) fast = fast && e.ldc % 4 == 0 && aligned(e.C);
  if (e.P) fast = fast && e.ldp % 4 == 0 && aligned(e.P);
  dim3 grid(tiles_r * tiles_c), block(nthreads);
  if (fast)
    srgemm_valu_kernel<Op, Cfg::BM, Cfg::BN, Cfg::TM, Cfg::TN, Cfg::BK, Cfg::WAVES, false>
        <<<grid, block, 0, s>>>(R, Cn, K, X, ldx, Y, ldy, O, ldo, e, tiles_r, tiles_c);
  else
    srgemm_valu_kernel<Op, Cfg::BM, Cfg::BN, Cfg::TM, Cfg::TN, Cfg::BK, Cfg::WAVES, true>
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

}  // namespace

ValuCfgInfo valu_cfg_info(int cfg) {
  if (cfg == CFG_S64) return {CfgS::BM, CfgS::BN, CfgS::BK, 256};
  if (cfg == CFG_M128x64) return {CfgM::BM, CfgM::BN, CfgM::BK, 256};
  return {CfgL::BM, CfgL::BN, CfgL::BK, 256};
}

// Tile shape choice.  Rule measured on MI355X for square problems
// (n = 1024 .. 8192, all ops): the VALU semirings need many resident waves to
// dual-issue, so up to ~3072^2 outputs the 64x64 tiles (7-8 waves per SIMD)
// win; at 4096^2 the 128x64 tiles balance 2048 tiles over 1024 slots; from
// ~6144^2 up the 128x128 tiles (fewest LDS reads per op) win, except for
// minus_square whose sub+fma pair keeps preferring occupancy.  Shapes the
// autotuner has timed (simd2_autotune) use the recorded winner instead.
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
  if (op == 8) return CFG_S64;
  if (a <= 4096.0 * 4096.0 * 1.01) return CFG_M128x64;
  return CFG_L128;
}

hipError_t launch_valu_f32(int op, int R, int Cn, int K, const float *X,
                           int64_t ldx, const float *Y, int64_t ldy, float *O,
                           int64_t ldo, const Epilogue &e, hipStream_t s,
                           int cfg) {
  if (cfg == CFG_AUTO) cfg = valu_pick_cfg(op, R, Cn, K);
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

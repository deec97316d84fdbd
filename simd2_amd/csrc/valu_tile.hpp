// Device building blocks of the LDS-tiled VALU semiring GEMM, shared by the
// single-problem kernel (srgemm_valu.hip) and the batched graph solver
// (batch.hip).  See srgemm_valu.hip for the tiling scheme.
#pragma once
#include <float.h>

#include "internal.hpp"
#include "semiring_ops.hpp"

namespace simd2 {

namespace {

constexpr int GROUP = 8;
constexpr int NUM_XCD = 8;

// Native 4-float vector for the slab staging registers: the HIP float4
// struct is copied with memcpy, which scalar replacement does not always
// undo, and the staged slab then lives in scratch.
typedef float vf4 __attribute__((ext_vector_type(4)));

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
template <class Op, int BM, int BN, int BK, int NT, bool CHECKED, bool GUARD = false>
__device__ __forceinline__ void load_slab_form(vf4 (&xv)[BM * BK / 4 / NT],
                                               vf4 (&yv)[BK * BN / 4 / NT],
                                               const float *__restrict__ X, int64_t ldx,
                                               const float *__restrict__ Y, int64_t ldy,
                                               int R, int Cn, int K, int row0, int col0,
                                               int k0, int tid);

// KT: K may end inside a slab -- that slab (uniform over the workgroup) takes
// the checked form, the others the fast one.  GUARD (with the fast form): the
// tile may reach past row R or column Cn; a 4-float load wholly inside the
// matrix stays one 128-bit load, X rows past R read as padding, and a Y group
// reaching past Cn is read element by element.
template <class Op, int BM, int BN, int BK, int NT, bool CHECKED, bool KT = false,
          bool GUARD = false>
__device__ __forceinline__ void load_slab(vf4 (&xv)[BM * BK / 4 / NT],
                                          vf4 (&yv)[BK * BN / 4 / NT],
                                          const float *__restrict__ X, int64_t ldx,
                                          const float *__restrict__ Y, int64_t ldy,
                                          int R, int Cn, int K, int row0, int col0,
                                          int k0, int tid) {
  if (CHECKED || (KT && k0 + BK > K))
    load_slab_form<Op, BM, BN, BK, NT, true>(xv, yv, X, ldx, Y, ldy, R, Cn, K, row0, col0, k0,
                                             tid);
  else
    load_slab_form<Op, BM, BN, BK, NT, false, GUARD>(xv, yv, X, ldx, Y, ldy, R, Cn, K, row0,
                                                     col0, k0, tid);
}

template <class Op, int BM, int BN, int BK, int NT, bool CHECKED, bool GUARD>
__device__ __forceinline__ void load_slab_form(vf4 (&xv)[BM * BK / 4 / NT],
                                               vf4 (&yv)[BK * BN / 4 / NT],
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
      if (GUARD && r >= R)
        xv[h] = vf4{Op::pad_x, Op::pad_x, Op::pad_x, Op::pad_x};
      else
        xv[h] = *reinterpret_cast<const vf4 *>(X + (int64_t)r * ldx + k);
    } else {
      float t[4];
#pragma unroll
      for (int q = 0; q < 4; ++q)
        t[q] = (r < R && k + q < K) ? X[(int64_t)r * ldx + k + q] : Op::pad_x;
      xv[h] = vf4{t[0], t[1], t[2], t[3]};
    }
  }
#pragma unroll
  for (int h = 0; h < YL; ++h) {
    const int f = tid + NT * h;
    const int k = k0 + f / (BN / 4);
    const int c = col0 + (f % (BN / 4)) * 4;
    if constexpr (!CHECKED) {
      if (GUARD && c + 3 >= Cn) {
        float t[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) t[q] = c + q < Cn ? Y[(int64_t)k * ldy + c + q] : Op::pad_y;
        yv[h] = vf4{t[0], t[1], t[2], t[3]};
      } else {
        yv[h] = *reinterpret_cast<const vf4 *>(Y + (int64_t)k * ldy + c);
      }
    } else {
      float t[4];
#pragma unroll
      for (int q = 0; q < 4; ++q)
        t[q] = (k < K && c + q < Cn) ? Y[(int64_t)k * ldy + c + q] : Op::pad_y;
      yv[h] = vf4{t[0], t[1], t[2], t[3]};
    }
  }
}

template <int BM, int BN, int BK, int NT, int XLD>
__device__ __forceinline__ void store_slab(float (*Xs)[XLD], float (*Ys)[BN],
                                           const vf4 (&xv)[BM * BK / 4 / NT],
                                           const vf4 (&yv)[BK * BN / 4 / NT],
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
    *reinterpret_cast<vf4 *>(&Ys[k][c]) = yv[h];
  }
}

// LDS of one workgroup of the VALU kernel (shared by the fast and checked
// tile bodies when a kernel contains both).
template <int BM, int BN, int BK> struct ValuSmem {
  __attribute__((aligned(16))) float Xs[2][BK][BM + 4];
  __attribute__((aligned(16))) float Ys[2][BK][BN];
  int changed;
};

// Main loop of one BM x BN tile: acc[i][j] = (+)_k X[r][k] (x) Y[k][c] over
// the K slabs, double-buffered through LDS.  LIST: only the BK-deep slabs
// whose indices are list[0 .. nlist) (ascending, in LDS) are visited -- the
// others are known to hold only the (+)-identity and contribute nothing.
template <class Op, int BM, int BN, int TM, int TN, int BK, bool CHECKED, bool LIST,
          bool KT = false, bool GUARD = false>
__device__ __forceinline__ void valu_mainloop(int R, int Cn, int K, const float *__restrict__ X,
                                              int64_t ldx, const float *__restrict__ Y,
                                              int64_t ldy, int row0, int col0,
                                              const uint16_t *list, int nlist,
                                              ValuSmem<BM, BN, BK> &sm, float (&acc)[TM][TN],
                                              int tid_in = -1, int kt_begin = 0,
                                              int kt_end = -1) {
  constexpr int NTX = BN / TN;
  constexpr int NT = NTX * (BM / TM);
  constexpr int RS = BM / (TM / 4);
  constexpr int CS = BN / (TN / 4);
  constexpr int XLD = BM + 4;
  constexpr int XL = BM * BK / 4 / NT;
  constexpr int YL = BK * BN / 4 / NT;
  auto &Xs = sm.Xs;
  auto &Ys = sm.Ys;
  // tid_in: the thread's index among the NT threads this tile body uses
  // (default: its workgroup index); [kt_begin, kt_end): the slabs visited
  // (default: all of K)
  const int tid = tid_in >= 0 ? tid_in : tid_x();
  const int tx = tid % NTX;
  const int ty = tid / NTX;

#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = Op::identity;

  vf4 xv[XL], yv[YL];
  const int nkt = LIST ? nlist : kt_end >= 0 ? kt_end - kt_begin : (K + BK - 1) / BK;
  if (nkt > 0) {
    const int k0 = (LIST ? (int)list[0] : kt_begin) * BK;
    load_slab<Op, BM, BN, BK, NT, CHECKED, KT, GUARD>(xv, yv, X, ldx, Y, ldy, R, Cn, K, row0,
                                                      col0, k0, tid);
    store_slab<BM, BN, BK, NT, XLD>(Xs[0], Ys[0], xv, yv, tid);
  }
  __syncthreads();

  for (int kt = 0; kt < nkt; ++kt) {
    const int buf = kt & 1;
    // The prefetch is unconditional (the last slab is simply loaded again and
    // written to the idle buffer, which nothing reads afterwards): under a
    // branch the compiler keeps the prefetch in scratch and waits for it.
    const int kn = kt + 1 < nkt ? kt + 1 : kt;
    const int knext = (LIST ? (int)list[kn] : kt_begin + kn) * BK;
    load_slab<Op, BM, BN, BK, NT, CHECKED, KT, GUARD>(xv, yv, X, ldx, Y, ldy, R, Cn, K, row0,
                                                      col0, knext, tid);
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
}

template <class Op, int BM, int BN, int TM, int TN, bool CHECKED, bool MASKED>
__device__ __forceinline__ void valu_epilogue(int R, int Cn, float *O, int64_t ldo,
                                              const Epilogue &e, int row0, int col0,
                                              int *tile_flag, int *changed_slot,
                                              const float (&acc)[TM][TN], int tid, bool active);

// One BM x BN output tile at (row0, col0) of the row-major problem; the whole
// workgroup calls it.  tile_flag, when set, receives 1 / 0 for "some element
// of this tile changed".
//   CHECKED: per-element bounds on every load (op padding) and store.
//   MASKED (with !CHECKED): the operands are stored padded -- every row and
//     column the tile and the K loop touch exists in memory, and the K
//     padding holds the op's pad values -- so loads stay 128-bit; only the
//     epilogue is limited to r < R, c < Cn.
//   KT (with !CHECKED): K need not be a multiple of BK: the last, partial
//     slab is loaded checked.
//   GUARD (with MASKED and KT): the tile reaches past row R or column Cn of
//     an unpadded matrix whose rows are 16-byte aligned; loads wholly inside
//     stay 128-bit, the rest are guarded (load_slab), and the epilogue is
//     MASKED's.
template <class Op, int BM, int BN, int TM, int TN, int BK, bool CHECKED, bool MASKED = false,
          bool KT = false, bool GUARD = false>
__device__ __forceinline__ void valu_tile(int R, int Cn, int K, const float *__restrict__ X,
                                          int64_t ldx, const float *__restrict__ Y, int64_t ldy,
                                          float *O, int64_t ldo, const Epilogue &e, int row0,
                                          int col0, int *tile_flag,
                                          ValuSmem<BM, BN, BK> &sm) {
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

  const int tid = tid_x();
  if (tid == 0) sm.changed = 0;

  float acc[TM][TN];
  valu_mainloop<Op, BM, BN, TM, TN, BK, CHECKED, false, KT, GUARD>(R, Cn, K, X, ldx, Y, ldy,
                                                                    row0, col0, nullptr, 0, sm,
                                                                    acc);
  (void)XL;
  (void)YL;
  (void)XLD;
  valu_epilogue<Op, BM, BN, TM, TN, CHECKED, MASKED>(R, Cn, O, ldo, e, row0, col0, tile_flag,
                                                     &sm.changed, acc, tid, true);
}

// Epilogue of a BM x BN tile from the thread's acc: (+) C, optional APLP
// mapping, change test, store.  Every thread of the workgroup calls it
// (the tile flag takes a barrier); only threads with `active` set store, tid
// being their index among the tile's threads.  *changed_slot (LDS) must be
// cleared before a barrier that precedes this call.
template <class Op, int BM, int BN, int TM, int TN, bool CHECKED, bool MASKED>
__device__ __forceinline__ void valu_epilogue(int R, int Cn, float *O, int64_t ldo,
                                              const Epilogue &e, int row0, int col0,
                                              int *tile_flag, int *changed_slot,
                                              const float (&acc)[TM][TN], int tid, bool active) {
  constexpr int NTX = BN / TN;
  constexpr int RS = BM / (TM / 4);
  constexpr int CS = BN / (TN / 4);
  const int tx = tid % NTX;
  const int ty = tid / NTX;
  bool changed = false;
  if (active) {
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
        if (!CHECKED && (!MASKED || (r < R && c + 4 <= Cn))) {
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

  }
  if (e.flag || tile_flag) {
    const bool any = __any(changed);
    if (e.flag && any && (tid % warpSize) == 0) atomicOr(e.flag, 1);
    if (tile_flag) {
      if (any && (tid % warpSize) == 0) atomicOr(changed_slot, 1);
      __syncthreads();
      if (active && tid == 0) *tile_flag = *changed_slot;
    }
  }
}

// WAVES: waves per SIMD the register allocation must leave room for
// (3 -> at most 168 VGPRs for the 8x8 register tile).
struct CfgL { static constexpr int BM = 128, BN = 128, TM = 8, TN = 8, BK = 16, WAVES = 3; };
struct CfgM { static constexpr int BM = 128, BN = 64, TM = 8, TN = 4, BK = 16, WAVES = 4; };
struct CfgS { static constexpr int BM = 64, BN = 64, TM = 4, TN = 4, BK = 16, WAVES = 4; };


}  // namespace

}  // namespace simd2

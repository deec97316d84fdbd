// k-nearest neighbours by pairwise squared L2 -- the PLD app's use of the
// minus_square semiring (apps/pld/cuASR/knncuda.cu:399-549) -- in the
// precision form (no counterpart in the reference; the reference's top-k of
// a materialised distance matrix is knn_select.hip).
//
// Candidates come from ||q||^2 + ||r||^2 - 2 q.r with the dot product on BF16
// MFMA (v_mfma_f32_16x16x32_bf16, FP32 accumulate), the answer from exact
// FP32 distances of those candidates, FP64 where FP32 cannot order them:
//
//  1. centre: mu = the references' mean (per dimension, fixed-order double
//     sums).  Distances do not change under a common translation, and
//     centring keeps ||q - mu||^2 and ||r - mu||^2 -- which the expansion's
//     error scales with -- near the data's own spread even for data far from
//     the origin;
//  2. prep: centred points split into BF16 high and low parts (x - hi), dims
//     zero-padded to a multiple of 32, and their FP32 squared norms; then the
//     largest reference norm;
//  3. pass 1: per (query, 64-reference tile) the smallest approximate
//     distance.  q.r ~ qh.rh + qh.rl + ql.rh, three MFMAs per 32-dim step,
//     ~16 significant bits.  Up to 160 dims a workgroup holds its 128
//     queries' fragments in registers and streams references through LDS;
//     beyond, queries and references are both streamed through LDS in 32-dim
//     slabs.  The Q x R matrix is never written;
//  4. tau: per query the CK-th smallest tile minimum (CK = k + 16): CK
//     distinct references have approximate distance <= tau.  With E_q a bound
//     on |approximate - exact| for the query (below), their exact distances
//     are <= tau + E_q, so the k-th nearest's is too, and every one of the k
//     nearest has approximate distance <= thr = tau + 2 E_q;
//  5. pass 2: the same products; every reference with approximate distance
//     <= thr is appended to the query's candidate list (LDS-staged, one
//     global atomic per query and workgroup).  A query whose list overflows
//     (only when more than CAND_CAP references crowd within 2 E_q of its
//     neighbours, e.g. thousands of duplicate points) is listed;
//  6. overflow (listed queries only): the exact FP32 distance to every
//     reference, and the CAND_CAP smallest (distance, index) keys become the
//     query's candidates;
//  7. re-rank, one wave per query: exact FP32 distance of every candidate
//     (eight fixed partial sums over dims j = p mod 8, combined in a fixed
//     tree -- the same value wherever it is computed), sort by (distance,
//     index); runs of candidates whose FP32 distances lie within a relative
//     tol of their neighbour and that reach rank k are recomputed in FP64 and
//     ordered by it.  FP64 is used nowhere else.
//
// E_q.  With x = hi + lo + e, |e| <= 2^-16 |x| (two round-to-nearest BF16
// steps of unit 2^-8), each product the three MFMAs drop (lo.lo and the e
// terms) is below 3.1 2^-16 |x_j y_j|; the FP32 accumulation of the 3 dpad
// products rounds at most 3 dpad times, each by 2^-24 of the running sum of
// magnitudes; so the dot product errs by at most
// (3.1 2^-16 + 3 dpad 2^-24) sum |x_j y_j| <= that / 2 (|x|^2 + |y|^2), and
// the distance (twice the dot) by that times (|x|^2 + |y|^2).  The FP32 norms
// (per-lane sums of d/64 squares and a 6-level tree) and the final two
// additions add (d/64 + 8) 2^-24 (|x|^2 + |y|^2).  E_q = 1.25 x the sum of
// these coefficients x (|q - mu|^2 + max_r |r - mu|^2) bounds the error of
// every approximate distance of query q.
#include <float.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

#include "internal.hpp"

namespace simd2 {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int QG = 2;        // groups of 16 queries per wave
constexpr int CQ = 64 * QG;  // queries per workgroup (4 waves)
constexpr int CR = 64;       // references per LDS step, and per tile of the tile minima
constexpr int CK_MARGIN = 16;
constexpr int KMAX = 256;    // largest k of the precision form
constexpr int CAND_CAP = 1024;
constexpr int FB_GRID = 64;  // workgroups of the overflow pass

// Minimum over the 16 lanes of a DPP row (lanes 16j .. 16j+15), in every lane
// of the row: rotations by 8, 4, 2 and 1 within the row, each one VALU min
// with a DPP operand (no LDS traffic).  The whole wave must be active.
__device__ __forceinline__ float row_min16(float m) {
  m = fminf(m, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(m), 0x128, 0xf, 0xf, false)));
  m = fminf(m, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(m), 0x124, 0xf, 0xf, false)));
  m = fminf(m, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(m), 0x122, 0xf, 0xf, false)));
  m = fminf(m, __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(m), 0x121, 0xf, 0xf, false)));
  return m;
}

// --- 1. centre --------------------------------------------------------------
// partial[split][j] = sum of ref[r][j] over the split's rows, in double: a
// workgroup takes 64 dims (lane = dim) and one of MEAN_SPLITS row ranges;
// wave w sums the w-th quarter of the range in row order, and the four wave
// sums are added in wave order (a fixed order, whatever the scheduling).
constexpr int MEAN_SPLITS = 64;
__global__ __launch_bounds__(256) void mean_partial_kernel(const float *__restrict__ P,
                                                           int64_t ldp, int n, int d,
                                                           double *partial) {
  __shared__ double ws[4][64];
  const int lane = tid_x() % 64, wave = tid_x() / 64;
  const int j = (bid_x() / MEAN_SPLITS) * 64 + lane;
  const int split = bid_x() % MEAN_SPLITS;
  const int per = (n + MEAN_SPLITS - 1) / MEAN_SPLITS;
  const int r0 = split * per, r1 = max(r0, min(n, r0 + per));
  const int q = (r1 - r0 + 3) / 4;
  const int a = r0 + wave * q, b = min(r1, a + q);
  double s = 0.0;
  if (j < d)
    for (int r = a; r < b; ++r) s += (double)P[(int64_t)r * ldp + j];
  ws[wave][lane] = s;
  __syncthreads();
  if (wave == 0 && j < d)
    partial[(int64_t)split * d + j] = ((ws[0][lane] + ws[1][lane]) + ws[2][lane]) + ws[3][lane];
}

// mu[j] = (sum of the partials in split order) / n.
__global__ __launch_bounds__(256) void mean_final_kernel(const double *__restrict__ partial,
                                                         int d, int n, float *mu) {
  const int j = bid_x() * 256 + tid_x();
  if (j >= d) return;
  double s = 0.0;
  for (int b = 0; b < MEAN_SPLITS; ++b) s += partial[(int64_t)b * d + j];
  mu[j] = (float)(s / n);
}

// rnmax = the largest reference norm (one workgroup; norms are >= 0).
__global__ __launch_bounds__(1024) void norm_max_kernel(const float *__restrict__ rn, int n,
                                                        float *rnmax) {
  __shared__ float wm[16];
  float m = 0.f;
  for (int i = tid_x(); i < n; i += 1024) m = fmaxf(m, rn[i]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
  if (tid_x() % 64 == 0) wm[tid_x() / 64] = m;
  __syncthreads();
  if (tid_x() == 0) {
    float r = wm[0];
    for (int w = 1; w < 16; ++w) r = fmaxf(r, wm[w]);
    *rnmax = r;
  }
}

// --- 2. prep ----------------------------------------------------------------
// One wave per point: x - mu split into BF16 high and low parts (dims
// zero-padded to dpad, stored [point][hi | lo]) and the FP32 squared norm of
// x - mu (wave-summed).
__global__ __launch_bounds__(256) void prep_kernel(const float *__restrict__ P, int64_t ldp,
                                                   int n, int d, int dpad,
                                                   const float *__restrict__ mu, __bf16 *Pb,
                                                   float *norm) {
  const int lane = tid_x() % 64;
  const int p = bid_x() * 4 + tid_x() / 64;
  if (p >= n) return;
  float s = 0.f;
  for (int j = lane; j < dpad; j += 64) {
    const float x = j < d ? P[(int64_t)p * ldp + j] - mu[j] : 0.f;
    s += x * x;
    const __bf16 hi = __float2bfloat16(x);
    Pb[(int64_t)p * 2 * dpad + j] = hi;
    Pb[(int64_t)p * 2 * dpad + dpad + j] = __float2bfloat16(x - __bfloat162float(hi));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (lane == 0) norm[p] = s;
}

// --- 3 / 5. the distance passes ----------------------------------------------
struct PassArgs {
  const float *qn, *rn, *thr;
  float *tmin;
  int ntiles;
  float *cand_d;
  int *cand_i, *cand_n, *ovf_n, *ovf_list;
};

// LDS candidate slots of a workgroup's queries (pass 2; one dummy slot in pass 1).
template <int PASS, int SLOTS>
struct CandStage {
  static constexpr int CB = PASS == 2 ? SLOTS : 1;
  float d[CQ][CB];
  int i[CQ][CB];
  int n[CQ];
};

// One candidate straight to the query's global list; the append that
// overflows the list puts the query on the overflow list (once).
__device__ __forceinline__ void append_global(const PassArgs &a, int q, float dd, int r) {
  const int at = atomicAdd(&a.cand_n[q], 1);
  if (at < CAND_CAP) {
    a.cand_d[(int64_t)q * CAND_CAP + at] = dd;
    a.cand_i[(int64_t)q * CAND_CAP + at] = r;
  } else if (at == CAND_CAP) {
    a.ovf_list[atomicAdd(a.ovf_n, 1)] = q;
  }
}

// Pass 1 writes per (query, tile) the smallest approximate distance; a tile is
// 64 references (PASS 1) or, where R / 64 tiles would give tau too few minima
// to be tight, 16 (PASS 3).  PASS 2 collects the candidates.
constexpr int tile_refs(int pass) { return pass == 3 ? 16 : CR; }

// The epilogue of one workgroup step over RB blocks of 16 references
// (r0 .. r0 + 16 RB): acc[g][rb] holds the dot products of this lane's 4
// queries of group g with reference r0 + rb*16 + lane%16.
template <int PASS, int RB, typename Stage>
__device__ __forceinline__ void pass_epilogue(const f32x4 (&acc)[QG][RB], const float (&qnv)[QG][4],
                                              const float (&tv)[QG][4], const int (&qv)[QG][4],
                                              int q0, int Q, int r0, int rb1, const float *Rn,
                                              Stage &cs, const PassArgs &a, int lane) {
  float rnv[RB];
  bool valid[RB];
#pragma unroll
  for (int rb = 0; rb < RB; ++rb) {
    rnv[rb] = Rn[rb * 16 + lane % 16];
    valid[rb] = r0 + rb * 16 + lane % 16 < rb1;
  }
  if constexpr (PASS != 2) {
    constexpr int TB = tile_refs(PASS) / 16;  // 16-reference blocks per tile
#pragma unroll
    for (int g = 0; g < QG; ++g)
#pragma unroll
      for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int t = 0; t < RB / TB; ++t) {  // each tile of the step
          // branch-free: references past the slice enter as +inf
          float m = __builtin_inff();
#pragma unroll
          for (int rb = TB * t; rb < TB * t + TB; ++rb) {
            const float dd = qnv[g][v] + rnv[rb] - 2.0f * acc[g][rb][v];
            m = fminf(m, valid[rb] ? dd : __builtin_inff());
          }
          // the 16 lanes of this query (one DPP row) hold the tile's references
          m = row_min16(m);
          const int tile = r0 / tile_refs(PASS) + t;
          if (lane % 16 == 0 && qv[g][v] < Q && r0 + tile_refs(PASS) * t < rb1)
            a.tmin[(int64_t)qv[g][v] * a.ntiles + tile] = m;
        }
  } else {
#pragma unroll
    for (int g = 0; g < QG; ++g)
#pragma unroll
      for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) {
          const float dd = qnv[g][v] + rnv[rb] - 2.0f * acc[g][rb][v];
          if (valid[rb] && dd <= tv[g][v]) {
            const int q = qv[g][v], ql = q - q0;
            const int at = atomicAdd(&cs.n[ql], 1);
            const int r = r0 + rb * 16 + lane % 16;
            if (at < Stage::CB) {
              cs.d[ql][at] = dd;
              cs.i[ql][at] = r;
            } else {
              append_global(a, q, dd, r);
            }
          }
        }
  }
}

// The LDS candidate slots of a workgroup to the global lists.
template <int PASS, typename Stage>
__device__ __forceinline__ void flush_candidates(Stage &cs, const PassArgs &a, int q0, int Q,
                                                 int tid) {
  if constexpr (PASS == 2) {
    __syncthreads();
    for (int ql = tid; ql < CQ; ql += 256) {
      const int q = q0 + ql, nb = min(cs.n[ql], Stage::CB);
      if (q < Q && nb > 0) {
        const int b = atomicAdd(&a.cand_n[q], nb);
        for (int e = 0; e < nb; ++e) {
          if (b + e < CAND_CAP) {
            a.cand_d[(int64_t)q * CAND_CAP + b + e] = cs.d[ql][e];
            a.cand_i[(int64_t)q * CAND_CAP + b + e] = cs.i[ql][e];
          } else if (b + e == CAND_CAP) {
            a.ovf_list[atomicAdd(a.ovf_n, 1)] = q;
          }
        }
      }
    }
  }
}

// The 4 queries per group whose accumulator rows this lane holds: norms and thresholds.
template <int PASS>
__device__ __forceinline__ void query_rows(const PassArgs &a, int q0, int Q, int wave, int lane,
                                           float (&qnv)[QG][4], float (&tv)[QG][4],
                                           int (&qv)[QG][4]) {
#pragma unroll
  for (int g = 0; g < QG; ++g)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      qv[g][v] = q0 + wave * 16 * QG + g * 16 + 4 * (lane / 16) + v;
      qnv[g][v] = qv[g][v] < Q ? a.qn[qv[g][v]] : 0.f;
      tv[g][v] = PASS == 2 && qv[g][v] < Q ? a.thr[qv[g][v]] : -__builtin_inff();
    }
}

// Up to 160 dims: the workgroup's 128 queries' A fragments stay in registers
// (every reference fragment read from LDS feeds QG x 3 MFMAs); references are
// streamed through LDS in 64-row tiles, the next tile's loads in flight while
// the current one is multiplied.  Workgroup b takes query block b / slices and
// reference slice b % slices (with a multiple of 8 slices, workgroups run on
// XCD b % 8, so each XCD's L2 keeps its slices).
template <int DPAD, int PASS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void knn_pass_kernel(
    const __bf16 *__restrict__ Qb, const __bf16 *__restrict__ Rb, int Q, int R, int slices,
    int slice, PassArgs a) {
  constexpr int KS = DPAD / 32;             // MFMA k-steps
  // LDS row: hi | lo, +32 bytes of padding: a row pitch of DPAD + 8 words
  // (8 mod 16) keeps the 16 rows of a fragment read on distinct banks; the
  // 16-byte padding (4 mod 16) cost twice the LDS cycles in conflicts
  constexpr int RLD = 2 * DPAD + 16;
  constexpr int CHUNKS = CR * 2 * DPAD / 8; // 16-byte chunks per reference tile
  constexpr int PER = CHUNKS / 256;         // chunks per thread
  __shared__ __attribute__((aligned(16))) __bf16 Rt[CR][RLD];
  __shared__ float Rn[CR];
  __shared__ CandStage<PASS, 16> cs;

  const int tid = tid_x();
  const int lane = tid % 64;
  const int wave = tid / 64;
  if (PASS == 2)
    for (int i = tid; i < CQ; i += 256) cs.n[i] = 0;  // published by the loop's first barrier
  const int q0 = (bid_x() / slices) * CQ;
  const int rb0 = (bid_x() % slices) * slice;
  const int rb1 = min(R, rb0 + slice);

  bf16x8 ah[QG][KS], al[QG][KS];
#pragma unroll
  for (int g = 0; g < QG; ++g) {
    const int q = q0 + wave * 16 * QG + g * 16 + lane % 16;
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
      if (q < Q) {
        const __bf16 *base = Qb + (int64_t)q * 2 * DPAD + ks * 32 + 8 * (lane / 16);
        ah[g][ks] = *reinterpret_cast<const bf16x8 *>(base);
        al[g][ks] = *reinterpret_cast<const bf16x8 *>(base + DPAD);
      } else {
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          ah[g][ks][e] = (__bf16)0.0f;
          al[g][ks][e] = (__bf16)0.0f;
        }
      }
    }
  }
  float qnv[QG][4], tv[QG][4];
  int qv[QG][4];
  query_rows<PASS>(a, q0, Q, wave, lane, qnv, tv, qv);

  uint4 pre[PER];
  float pren = 0.f;
  auto fetch = [&](int r0) {
#pragma unroll
    for (int h = 0; h < PER; ++h) {
      const int f = tid + 256 * h;
      const int rr = f / (2 * DPAD / 8);
      const int cc = (f % (2 * DPAD / 8)) * 8;
      const int r = r0 + rr;
      pre[h] = make_uint4(0u, 0u, 0u, 0u);
      if (r < rb1) pre[h] = *reinterpret_cast<const uint4 *>(Rb + (int64_t)r * 2 * DPAD + cc);
    }
    if (tid < CR) pren = (r0 + tid < rb1) ? a.rn[r0 + tid] : 0.f;
  };
  auto stage = [&]() {
#pragma unroll
    for (int h = 0; h < PER; ++h) {
      const int f = tid + 256 * h;
      const int rr = f / (2 * DPAD / 8);
      const int cc = (f % (2 * DPAD / 8)) * 8;
      *reinterpret_cast<uint4 *>(&Rt[rr][cc]) = pre[h];
    }
    if (tid < CR) Rn[tid] = pren;
  };

  if (rb0 < rb1) fetch(rb0);
  for (int r0 = rb0; r0 < rb1; r0 += CR) {
    __syncthreads();  // previous tile fully multiplied
    stage();
    __syncthreads();
    if (r0 + CR < rb1) fetch(r0 + CR);
    f32x4 acc[QG][4];
#pragma unroll
    for (int g = 0; g < QG; ++g)
#pragma unroll
      for (int rb = 0; rb < 4; ++rb) acc[g][rb] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < KS; ++ks) {
#pragma unroll
      for (int rb = 0; rb < 4; ++rb) {
        const __bf16 *row = &Rt[rb * 16 + lane % 16][ks * 32 + 8 * (lane / 16)];
        const bf16x8 bh = *reinterpret_cast<const bf16x8 *>(row);
        const bf16x8 bl = *reinterpret_cast<const bf16x8 *>(row + DPAD);
#pragma unroll
        for (int g = 0; g < QG; ++g) {
          acc[g][rb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[g][ks], bh, acc[g][rb], 0, 0, 0);
          acc[g][rb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[g][ks], bl, acc[g][rb], 0, 0, 0);
          acc[g][rb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[g][ks], bh, acc[g][rb], 0, 0, 0);
        }
      }
    }
    pass_epilogue<PASS, 4>(acc, qnv, tv, qv, q0, Q, r0, rb1, Rn, cs, a, lane);
  }
  flush_candidates<PASS>(cs, a, q0, Q, tid);
}

// Any number of dims: queries and references both streamed through LDS in
// 32-dim slabs (double-buffered, the next slab's loads in flight), 128
// queries x 128 references per workgroup step (each wave 32 x 128).
constexpr int SRB = 8;          // 16-reference blocks per step
constexpr int SREF = 16 * SRB;  // 128 references per step
constexpr int SLD = 72;         // LDS row: 32 hi + 32 lo bf16, +16 bytes of padding

template <int PASS>
__global__ __launch_bounds__(256) void knn_pass_stream_kernel(
    const __bf16 *__restrict__ Qb, const __bf16 *__restrict__ Rb, int Q, int R, int dpad,
    int slices, int slice, PassArgs a) {
  __shared__ __attribute__((aligned(16))) __bf16 Qs[2][CQ][SLD];
  __shared__ __attribute__((aligned(16))) __bf16 Rs[2][SREF][SLD];
  __shared__ float Rn[SREF];
  // 4 slots per query keep two workgroups' LDS within the CU's 160 KB
  __shared__ CandStage<PASS, 4> cs;

  const int tid = tid_x();
  const int lane = tid % 64;
  const int wave = tid / 64;
  if (PASS == 2)
    for (int i = tid; i < CQ; i += 256) cs.n[i] = 0;
  const int q0 = (bid_x() / slices) * CQ;
  const int rb0 = (bid_x() % slices) * slice;
  const int rb1 = min(R, rb0 + slice);
  float qnv[QG][4], tv[QG][4];
  int qv[QG][4];
  query_rows<PASS>(a, q0, Q, wave, lane, qnv, tv, qv);
  const int ksteps = dpad / 32;

  // Slab ks of the query block and of the reference step at r0: each of the
  // 128 + 128 rows is 64 B of hi and 64 B of lo; thread t moves half t % 2
  // (hi, lo) of query row t / 2 and of reference row t / 2, 4 x 16 B each.
  uint4 pq[4], pr[4];
  auto fetch = [&](int r0, int ks) {
    const int row = tid / 2, half = tid % 2;
    const int q = q0 + row, r = r0 + row;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      pq[c] = make_uint4(0u, 0u, 0u, 0u);
      pr[c] = make_uint4(0u, 0u, 0u, 0u);
      if (q < Q)
        pq[c] = *reinterpret_cast<const uint4 *>(Qb + (int64_t)q * 2 * dpad + half * dpad +
                                                 ks * 32 + 8 * c);
      if (r < rb1)
        pr[c] = *reinterpret_cast<const uint4 *>(Rb + (int64_t)r * 2 * dpad + half * dpad +
                                                 ks * 32 + 8 * c);
    }
  };
  auto stage = [&](int buf) {
    const int row = tid / 2, half = tid % 2;
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      *reinterpret_cast<uint4 *>(&Qs[buf][row][half * 32 + 8 * c]) = pq[c];
      *reinterpret_cast<uint4 *>(&Rs[buf][row][half * 32 + 8 * c]) = pr[c];
    }
  };

  for (int r0 = rb0; r0 < rb1; r0 += SREF) {
    __syncthreads();  // the previous step's epilogue has read Rn and the slabs
    if (tid < SREF) Rn[tid] = (r0 + tid < rb1) ? a.rn[r0 + tid] : 0.f;
    f32x4 acc[QG][SRB];
#pragma unroll
    for (int g = 0; g < QG; ++g)
#pragma unroll
      for (int rb = 0; rb < SRB; ++rb) acc[g][rb] = f32x4{0.f, 0.f, 0.f, 0.f};
    fetch(r0, 0);
    stage(0);
    __syncthreads();
    for (int ks = 0; ks < ksteps; ++ks) {
      const int buf = ks & 1;
      if (ks + 1 < ksteps) fetch(r0, ks + 1);
      bf16x8 ah[QG], al[QG];
#pragma unroll
      for (int g = 0; g < QG; ++g) {
        const __bf16 *row = &Qs[buf][wave * 16 * QG + g * 16 + lane % 16][8 * (lane / 16)];
        ah[g] = *reinterpret_cast<const bf16x8 *>(row);
        al[g] = *reinterpret_cast<const bf16x8 *>(row + 32);
      }
#pragma unroll
      for (int rb = 0; rb < SRB; ++rb) {
        const __bf16 *row = &Rs[buf][rb * 16 + lane % 16][8 * (lane / 16)];
        const bf16x8 bh = *reinterpret_cast<const bf16x8 *>(row);
        const bf16x8 bl = *reinterpret_cast<const bf16x8 *>(row + 32);
#pragma unroll
        for (int g = 0; g < QG; ++g) {
          acc[g][rb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[g], bh, acc[g][rb], 0, 0, 0);
          acc[g][rb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[g], bl, acc[g][rb], 0, 0, 0);
          acc[g][rb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[g], bh, acc[g][rb], 0, 0, 0);
        }
      }
      // the other buffer was last read in the previous step, before its barrier
      if (ks + 1 < ksteps) stage(buf ^ 1);
      __syncthreads();
    }
    pass_epilogue<PASS, SRB>(acc, qnv, tv, qv, q0, Q, r0, rb1, Rn, cs, a, lane);
  }
  flush_candidates<PASS>(cs, a, q0, Q, tid);
}

// --- 4. tau -------------------------------------------------------------------
__device__ __forceinline__ uint32_t fkey(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float fkey_inv(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// thr[q] = (the ck-th smallest of the query's tile minima) + 2 E_q, +inf when
// it has fewer than ck tiles; by bisection over the keys, one wave per query.
// Also clears the query's candidate count and exact flag (and, in the first
// wave, the counts of the overflow list and of the re-rank's LDS list).
__global__ __launch_bounds__(256) void knn_tau_kernel(const float *__restrict__ tmin, int ntiles,
                                                      int Q, int ck, float e_coef,
                                                      const float *__restrict__ qn,
                                                      const float *__restrict__ rnmax,
                                                      float *thr, int *cand_n, int *cand_exact,
                                                      int *ovf_n, int *big_n) {
  const int lane = tid_x() % 64;
  const int q = bid_x() * 4 + tid_x() / 64;
  if (q == 0 && lane == 0) {
    *ovf_n = 0;
    *big_n = 0;
  }
  if (q >= Q) return;
  const float *t = tmin + (int64_t)q * ntiles;
  uint32_t lo = 0u, hi = 0xffffffffu;  // smallest key with >= ck minima at or under it
  constexpr int TP = 8;                 // keys held in registers per lane
  uint32_t kv[TP];
#pragma unroll
  for (int i = 0; i < TP; ++i) {
    const int j = lane + 64 * i;
    kv[i] = j < ntiles ? fkey(t[j]) : 0xffffffffu;
  }
  float res = __builtin_inff();
  if (ntiles >= ck) {
    while (lo < hi) {
      const uint32_t mid = lo + (hi - lo) / 2;
      int c = 0;
#pragma unroll
      for (int i = 0; i < TP; ++i)
        c += (int)__popcll(__ballot(lane + 64 * i < ntiles && kv[i] <= mid));
      if (ntiles > 64 * TP) {
        int e = 0;
        for (int j = lane + 64 * TP; j < ntiles; j += 64) e += fkey(t[j]) <= mid ? 1 : 0;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) e += __shfl_xor(e, off);
        c += e;
      }
      if (c >= ck)
        hi = mid;
      else
        lo = mid + 1;
    }
    const float eq = e_coef * (qn[q] + *rnmax);
    res = fkey_inv(lo) + 2.0f * eq;
    res = res + fabsf(res) * 1e-6f;  // the sums above round by less than this
  }
  if (lane == 0) {
    thr[q] = res;
    cand_n[q] = 0;
    cand_exact[q] = 0;
  }
}

// --- canonical exact distances ---------------------------------------------------
// The exact FP32 squared distance: eight partial sums s_p over dims j = p mod 8
// in increasing j, combined as ((s0+s1)+(s2+s3))+((s4+s5)+(s6+s7)).  The wave
// form (8 lanes per candidate, combined by xor shuffles 1, 2, 4) and this
// thread form give the same bits.
__device__ __forceinline__ float exact_d32_thread(const float *__restrict__ q,
                                                  const float *__restrict__ r, int d) {
  float s[8];
#pragma unroll
  for (int p = 0; p < 8; ++p) s[p] = 0.f;
  int j = 0;
  for (; j + 8 <= d; j += 8)
#pragma unroll
    for (int p = 0; p < 8; ++p) {
      const float t = q[j + p] - r[j + p];
      s[p] = s[p] + t * t;
    }
#pragma unroll
  for (int p = 0; p < 8; ++p)
    if (j + p < d) {
      const float t = q[j + p] - r[j + p];
      s[p] = s[p] + t * t;
    }
  return ((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7]));
}

// --- 6. overflow: exact distances to every reference ----------------------------
// One workgroup per listed query (grid-stride): exact distances into the
// workgroup's scratch row, then the min(R, CAND_CAP) smallest (distance,
// index) keys (by bisection on the key) become the query's candidates,
// marked exact.
__global__ __launch_bounds__(256) void knn_overflow_kernel(
    const float *__restrict__ Qf, int64_t ldq, const float *__restrict__ Rf, int64_t ldr, int R,
    int d, const int *__restrict__ ovf_n, const int *__restrict__ ovf_list, float *scratch,
    float *cand_d, int *cand_i, int *cand_n, int *cand_exact) {
  __shared__ int cnt_s;
  const int tid = tid_x();
  const int n_ovf = *ovf_n;
  float *row = scratch + (int64_t)bid_x() * R;
  for (int b = bid_x(); b < n_ovf; b += FB_GRID) {
    const int q = ovf_list[b];
    for (int r = tid; r < R; r += 256)
      row[r] = exact_d32_thread(Qf + (int64_t)q * ldq, Rf + (int64_t)r * ldr, d);
    __threadfence_block();
    __syncthreads();
    const int want = min(R, CAND_CAP);
    // the smallest 64-bit key (fkey(d) << 32 | r) with >= want keys at or under it
    uint64_t lo = 0ull, hi = ~0ull;
    while (lo < hi) {
      const uint64_t mid = lo + (hi - lo) / 2;
      if (tid == 0) cnt_s = 0;
      __syncthreads();
      int c = 0;
      for (int r = tid; r < R; r += 256)
        c += (((uint64_t)fkey(row[r]) << 32) | (uint32_t)r) <= mid ? 1 : 0;
      atomicAdd(&cnt_s, c);
      __syncthreads();
      const int total = cnt_s;
      __syncthreads();
      if (total >= want)
        hi = mid;
      else
        lo = mid + 1;
    }
    if (tid == 0) cnt_s = 0;
    __syncthreads();
    for (int r = tid; r < R; r += 256) {
      if ((((uint64_t)fkey(row[r]) << 32) | (uint32_t)r) <= lo) {
        const int at = atomicAdd(&cnt_s, 1);  // exactly `want` keys pass
        cand_d[(int64_t)q * CAND_CAP + at] = row[r];
        cand_i[(int64_t)q * CAND_CAP + at] = r;
      }
    }
    __syncthreads();
    if (tid == 0) {
      cand_n[q] = cnt_s;
      cand_exact[q] = 1;
    }
    __syncthreads();  // the scratch row is reused by the next listed query
  }
}

// --- 7. re-rank -----------------------------------------------------------------
__device__ __forceinline__ bool key_less(float d, int i, float e, int j) {
  return d < e || (d == e && i < j);
}
__device__ __forceinline__ bool key_less64(double d, int i, double e, int j) {
  return d < e || (d == e && i < j);
}

__device__ __forceinline__ double shfl_d(double v, int src) {
  const uint64_t u = __double_as_longlong(v);
  const int lo = __shfl((int)(uint32_t)u, src), hi = __shfl((int)(uint32_t)(u >> 32), src);
  return __longlong_as_double(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

// Exact FP32 distances of candidates m0 .. m0 + 7 (lanes 8g .. 8g+7 sum dims
// p, p + 8, ... of candidate m0 + g), combined by the xor tree; lane
// 8g + p receives candidate m0 + g's distance.
__device__ __forceinline__ float exact_d32_wave(const float *__restrict__ qrow,
                                                const float *__restrict__ Rf, int64_t ldr, int d,
                                                int r, bool live, int part) {
  float s = 0.f;
  if (live) {
    const float *rrow = Rf + (int64_t)r * ldr;
    for (int j = part; j < d; j += 8) {
      const float t = qrow[j] - rrow[j];
      s = s + t * t;
    }
  }
  s += __shfl_xor(s, 1);
  s += __shfl_xor(s, 2);
  s += __shfl_xor(s, 4);
  return s;
}

// FP64 squared distance of reference r, summed by the whole wave.
__device__ __forceinline__ double exact_d64_wave(const float *__restrict__ qrow,
                                                 const float *__restrict__ Rf, int64_t ldr, int d,
                                                 int r, int lane) {
  const float *rrow = Rf + (int64_t)r * ldr;
  double s = 0.0;
  for (int j = lane; j < d; j += 64) {
    const double t = (double)qrow[j] - (double)rrow[j];
    s += t * t;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  return s;
}

// One wave per query with up to 64 candidates (lane m holds candidate m);
// queries with more go to the list of rerank_big_kernel.
__global__ __launch_bounds__(256) void rerank_small_kernel(
    const float *__restrict__ Qf, int64_t ldq, const float *__restrict__ Rf, int64_t ldr, int Q,
    int d, int k, float tol, const float *__restrict__ cand_d, const int *__restrict__ cand_i,
    const int *__restrict__ cand_n, const int *__restrict__ cand_exact, float *out_d,
    int *out_i, int do_sqrt, int *stats, int *big_n, int *big_list) {
  const int lane = tid_x() % 64;
  const int q = bid_x() * 4 + tid_x() / 64;
  if (q >= Q) return;
  const int c = min(cand_n[q], CAND_CAP);
  if (c > 64) {
    if (lane == 0) big_list[atomicAdd(big_n, 1)] = q;
    return;
  }
  const bool exact = cand_exact[q] != 0;
  const float *qrow = Qf + (int64_t)q * ldq;
  int my_i = lane < c ? cand_i[(int64_t)q * CAND_CAP + lane] : 0x7fffffff;
  float my_d = __builtin_inff();
  if (exact) {
    if (lane < c) my_d = cand_d[(int64_t)q * CAND_CAP + lane];
  } else {
    const int grp = lane / 8, part = lane % 8;
    for (int m0 = 0; m0 < c; m0 += 8) {
      const int r = __shfl(my_i, min(m0 + grp, 63));
      const float s = exact_d32_wave(qrow, Rf, ldr, d, r, m0 + grp < c, part);
      const float got = __shfl(s, 8 * ((lane - m0) & 7));
      if (lane >= m0 && lane < m0 + 8 && lane < c) my_d = got;
    }
  }
  // bitonic sort of the wave by (distance, index); padding sorts last
#pragma unroll
  for (int s = 2; s <= 64; s <<= 1)
#pragma unroll
    for (int j = s >> 1; j > 0; j >>= 1) {
      const float od = __shfl_xor(my_d, j);
      const int oi = __shfl_xor(my_i, j);
      const bool up = (lane & s) == 0, low = (lane & j) == 0;
      const bool take = (low == up) ? key_less(od, oi, my_d, my_i) : key_less(my_d, my_i, od, oi);
      if (take) {
        my_d = od;
        my_i = oi;
      }
    }
  // Near ties: bit i of T says candidate i lies within tol of candidate i - 1.
  const float prev = __shfl(my_d, lane > 0 ? lane - 1 : 0);
  const uint64_t T = __ballot(lane > 0 && lane < c && my_d - prev <= tol * my_d);
  // this lane's run [start, end): start = the last position <= lane whose bit
  // is clear, end = the first position > lane whose bit is clear (or c)
  const uint64_t clear = ~T;
  const uint64_t below = clear & ((lane == 63) ? ~0ull : ((2ull << lane) - 1));
  const int start = 63 - __clzll(below | 1ull);
  const uint64_t above = (lane == 63) ? 0ull : (clear & ~((2ull << lane) - 1));
  const int end = min(c, above ? __ffsll((long long)above) - 1 : 64);
  const bool flag = lane < c && end - start >= 2 && start <= k;
  const uint64_t F = __ballot(flag);
  double my_x = (double)my_d;
  if (F) {
    for (uint64_t f = F; f; f &= f - 1) {
      const int m = __ffsll((long long)f) - 1;
      const double x = exact_d64_wave(qrow, Rf, ldr, d, __shfl(my_i, m), lane);
      if (lane == m) my_x = x;
    }
    // re-sort by (FP64 distance where computed, else the FP32 one; index):
    // runs are separated by gaps above tol that FP32's error cannot cross,
    // so only the order inside the flagged runs can change
#pragma unroll
    for (int s = 2; s <= 64; s <<= 1)
#pragma unroll
      for (int j = s >> 1; j > 0; j >>= 1) {
        const double ox = shfl_d(my_x, lane ^ j);
        const float od = __shfl_xor(my_d, j);
        const int oi = __shfl_xor(my_i, j);
        const bool up = (lane & s) == 0, low = (lane & j) == 0;
        const bool take = (low == up) ? key_less64(ox, oi, my_x, my_i)
                                      : key_less64(my_x, my_i, ox, oi);
        if (take) {
          my_x = ox;
          my_d = od;
          my_i = oi;
        }
      }
    if (stats && lane == 0) {
      atomicAdd(&stats[0], 1);
      atomicAdd(&stats[1], (int)__popcll(F));
    }
  }
  if (lane < k && lane < c) {
    // a flagged candidate reports its FP64 distance rounded to FP32; for the
    // others my_x is their FP32 distance exactly
    const float outd = (float)my_x;
    out_d[(int64_t)q * k + lane] = do_sqrt ? sqrtf(outd) : outd;
    out_i[(int64_t)q * k + lane] = my_i;
  }
  if (stats && lane == 0 && exact) atomicAdd(&stats[2], 1);
}

// Queries with more than 64 candidates: one wave (one workgroup) per listed
// query, grid-stride over the list, the candidates sorted in LDS.
constexpr int BIG_GRID_RR = 512;
__global__ __launch_bounds__(64) void rerank_big_kernel(const float *__restrict__ Qf, int64_t ldq,
                                                        const float *__restrict__ Rf, int64_t ldr,
                                                        int d, int k, float tol,
                                                        const float *__restrict__ cand_d,
                                                        const int *__restrict__ cand_i,
                                                        const int *__restrict__ cand_n,
                                                        const int *__restrict__ cand_exact,
                                                        float *out_d, int *out_i, int do_sqrt,
                                                        int *stats, const int *big_n,
                                                        const int *big_list) {
  __shared__ float sd[CAND_CAP];
  __shared__ int si[CAND_CAP];
  __shared__ double sx[CAND_CAP];
  __shared__ int sflag[CAND_CAP];
  __shared__ int nflag_s;
  const int lane = tid_x();
  const int nbig = *big_n;
  for (int b = bid_x(); b < nbig; b += BIG_GRID_RR) {
  const int q = big_list[b];
  const int c = min(cand_n[q], CAND_CAP);
  const bool exact = cand_exact[q] != 0;
  const float *qrow = Qf + (int64_t)q * ldq;
  const float *cd = cand_d + (int64_t)q * CAND_CAP;
  const int *ci = cand_i + (int64_t)q * CAND_CAP;

  // Exact FP32 distances: lanes 8g .. 8g+7 sum dims p, p + 8, ... of
  // candidate m0 + g (8 candidates in flight), combined by the xor tree.
  if (exact) {
    for (int m = lane; m < c; m += 64) {
      sd[m] = cd[m];
      si[m] = ci[m];
    }
  } else {
    const int grp = lane / 8, part = lane % 8;
    for (int m0 = 0; m0 < c; m0 += 8) {
      const int m = m0 + grp;
      const int r = m < c ? ci[m] : 0;
      float s = 0.f;
      if (m < c) {
        const float *rrow = Rf + (int64_t)r * ldr;
        for (int j = part; j < d; j += 8) {
          const float t = qrow[j] - rrow[j];
          s = s + t * t;
        }
      }
      s += __shfl_xor(s, 1);
      s += __shfl_xor(s, 2);
      s += __shfl_xor(s, 4);
      if (part == 0 && m < c) {
        sd[m] = s;
        si[m] = r;
      }
    }
  }
  int P = 1;
  while (P < c) P <<= 1;
  for (int m = c + lane; m < P; m += 64) {
    sd[m] = __builtin_inff();
    si[m] = 0x7fffffff;
  }
  __syncthreads();
  // bitonic sort of (sd, si) by key
  for (int s = 2; s <= P; s <<= 1)
    for (int j = s >> 1; j > 0; j >>= 1) {
      for (int i = lane; i < P; i += 64) {
        const int p = i ^ j;
        if (p > i) {
          const bool up = (i & s) == 0;
          const bool gt = key_less(sd[p], si[p], sd[i], si[i]);
          if (gt == up) {
            const float td = sd[i];
            const int ti = si[i];
            sd[i] = sd[p];
            si[i] = si[p];
            sd[p] = td;
            si[p] = ti;
          }
        }
      }
      __syncthreads();
    }
  // Near ties: runs of consecutive candidates each within tol (relative) of
  // the previous one; a run that starts at or before rank k (it can move a
  // candidate across the boundary or reorder the answer) is taken to FP64.
  for (int i = lane; i < c; i += 64) {
    sflag[i] = 0;
    sx[i] = (double)sd[i];
  }
  __syncthreads();
  if (lane == 0) {
    int nf = 0, start = 0;
    for (int i = 1; i <= c; ++i) {
      const bool tie = i < c && sd[i] - sd[i - 1] <= tol * sd[i];
      if (!tie) {
        if (i - start >= 2 && start <= k)
          for (int m = start; m < i; ++m) {
            sflag[m] = 1;
            ++nf;
          }
        if (i > k) break;
        start = i;
      }
    }
    nflag_s = nf;
  }
  __syncthreads();
  const int nflag = nflag_s;
  if (nflag > 0) {
    for (int m = 0; m < c; ++m) {
      if (!sflag[m]) continue;  // uniform over the wave: sflag is in LDS
      const float *rrow = Rf + (int64_t)si[m] * ldr;
      double s = 0.0;
      for (int j = lane; j < d; j += 64) {
        const double t = (double)qrow[j] - (double)rrow[j];
        s += t * t;
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
      if (lane == 0) sx[m] = s;
    }
    __syncthreads();
    // each flagged run sorted by (FP64 distance, index): runs are separated
    // by relative gaps above tol, which FP32's error cannot cross
    if (lane == 0) {
      int m = 0;
      while (m < c) {
        if (!sflag[m]) {
          ++m;
          continue;
        }
        int e = m + 1;
        while (e < c && sflag[e] && sd[e] - sd[e - 1] <= tol * sd[e]) ++e;
        for (int x = m + 1; x < e; ++x) {  // insertion sort of [m, e)
          const double xd = sx[x];
          const float fd = sd[x];
          const int xi = si[x];
          int y = x - 1;
          while (y >= m && key_less64(xd, xi, sx[y], si[y])) {
            sx[y + 1] = sx[y];
            sd[y + 1] = sd[y];
            si[y + 1] = si[y];
            --y;
          }
          sx[y + 1] = xd;
          sd[y + 1] = fd;
          si[y + 1] = xi;
        }
        m = e;
      }
      if (stats) {
        atomicAdd(&stats[0], 1);
        atomicAdd(&stats[1], nflag);
      }
    }
    __syncthreads();
  }
  for (int j = lane; j < k && j < c; j += 64) {
    const float outd = sflag[j] ? (float)sx[j] : sd[j];
    out_d[(int64_t)q * k + j] = do_sqrt ? sqrtf(outd) : outd;
    out_i[(int64_t)q * k + j] = si[j];
  }
  if (stats && lane == 0 && exact) atomicAdd(&stats[2], 1);
  __syncthreads();  // the LDS arrays are reused by the next listed query
  }
}

int dpad_of(int d) { return ((d + 31) / 32) * 32; }

struct Bf16Plan {
  int dpad, ntiles, slices, slice, ck;
  bool stream, fine;
  size_t bQb, bRb, bqn, brn, bmu, bpart, bmax, btmin, bthr, bcn, bex, bcd, bci, bovf, bbig,
      bscr;
  size_t bytes() const {
    return bQb + bRb + bqn + brn + bmu + bpart + bmax + btmin + bthr + bcn + bex + bcd + bci +
           bovf + bbig + bscr;
  }
};

size_t a256(size_t b) { return (b + 255) / 256 * 256; }

Bf16Plan bf16_plan(int Q, int R, int d, int k) {
  Bf16Plan p;
  p.dpad = dpad_of(d);
  p.stream = p.dpad > 160;
  p.ck = k + CK_MARGIN;
  // 16-reference tiles while 64-reference ones would number under 4 ck: tau
  // is the ck-th smallest minimum, and from few tiles it lies far out
  p.fine = (R + CR - 1) / CR < 4 * p.ck;
  p.ntiles = (R + tile_refs(p.fine ? 3 : 1) - 1) / tile_refs(p.fine ? 3 : 1);
  const int step = p.stream ? SREF : CR;
  const int steps = (R + step - 1) / step;
  const int qblocks = (Q + CQ - 1) / CQ;
  // reference slices: 8 (one per XCD), doubled until the grid has ~1024
  // workgroups, never more than there are reference steps
  int slices = 8;
  while (slices < 64 && qblocks * slices < 1024) slices *= 2;
  if (slices > steps) slices = steps;
  p.slices = slices;
  p.slice = ((R + slices - 1) / slices + step - 1) / step * step;
  p.bQb = a256((size_t)Q * p.dpad * 4);
  p.bRb = a256((size_t)R * p.dpad * 4);
  p.bqn = a256((size_t)Q * 4);
  p.brn = a256((size_t)R * 4);
  p.bmu = a256((size_t)d * 4);
  p.bpart = a256((size_t)MEAN_SPLITS * d * 8);
  p.bmax = 256;
  p.btmin = a256((size_t)Q * p.ntiles * 4);
  p.bthr = a256((size_t)Q * 4);
  p.bcn = a256((size_t)Q * 4);
  p.bex = a256((size_t)Q * 4);
  p.bcd = a256((size_t)Q * CAND_CAP * 4);
  p.bci = a256((size_t)Q * CAND_CAP * 4);
  p.bovf = a256((size_t)Q * 4 + 4);
  p.bbig = a256((size_t)Q * 4 + 4);
  p.bscr = a256((size_t)FB_GRID * R * 4);
  return p;
}

}  // namespace

int knn_bf16_kmax() { return KMAX; }

// sized for the largest k: a larger k can only choose the finer tiles
size_t knn_bf16_workspace(int Q, int R, int d) { return bf16_plan(Q, R, d, KMAX).bytes(); }

hipError_t launch_knn_bf16(const float *query, int64_t ldq, const float *ref, int64_t ldr, int Q,
                           int R, int d, int k, float *out_d, int *out_i, int do_sqrt,
                           void *workspace, int *stats, hipStream_t s) {
  const Bf16Plan p = bf16_plan(Q, R, d, k);
  char *w = static_cast<char *>(workspace);
  auto take = [&](size_t bytes) {
    char *ptr = w;
    w += bytes;
    return ptr;
  };
  __bf16 *Qb = reinterpret_cast<__bf16 *>(take(p.bQb));
  __bf16 *Rb = reinterpret_cast<__bf16 *>(take(p.bRb));
  float *qn = reinterpret_cast<float *>(take(p.bqn));
  float *rn = reinterpret_cast<float *>(take(p.brn));
  float *mu = reinterpret_cast<float *>(take(p.bmu));
  double *part = reinterpret_cast<double *>(take(p.bpart));
  float *rnmax = reinterpret_cast<float *>(take(p.bmax));
  float *tmin = reinterpret_cast<float *>(take(p.btmin));
  float *thr = reinterpret_cast<float *>(take(p.bthr));
  int *cn = reinterpret_cast<int *>(take(p.bcn));
  int *cex = reinterpret_cast<int *>(take(p.bex));
  float *cd = reinterpret_cast<float *>(take(p.bcd));
  int *ci = reinterpret_cast<int *>(take(p.bci));
  int *ovf = reinterpret_cast<int *>(take(p.bovf));
  int *big = reinterpret_cast<int *>(take(p.bbig));
  float *scr = reinterpret_cast<float *>(take(p.bscr));

  mean_partial_kernel<<<((d + 63) / 64) * MEAN_SPLITS, 256, 0, s>>>(ref, ldr, R, d, part);
  mean_final_kernel<<<(d + 255) / 256, 256, 0, s>>>(part, d, R, mu);
  prep_kernel<<<(Q + 3) / 4, 256, 0, s>>>(query, ldq, Q, d, p.dpad, mu, Qb, qn);
  prep_kernel<<<(R + 3) / 4, 256, 0, s>>>(ref, ldr, R, d, p.dpad, mu, Rb, rn);
  norm_max_kernel<<<1, 1024, 0, s>>>(rn, R, rnmax);
  const PassArgs a{qn, rn, thr, tmin, p.ntiles, cd, ci, cn, ovf, ovf + 1};
  const dim3 grid(((Q + CQ - 1) / CQ) * p.slices);
  // E_q = e_coef (||q - mu||^2 + max ||r - mu||^2): see the bound at the top
  const float e_coef = 1.25f * (3.1f * ldexpf(1.0f, -16) + (float)(3 * p.dpad + d / 64 + 8) *
                                                             ldexpf(1.0f, -24));
  for (int pass = 1; pass <= 2; ++pass) {
    if (p.stream) {
      if (pass == 1 && p.fine)
        knn_pass_stream_kernel<3><<<grid, 256, 0, s>>>(Qb, Rb, Q, R, p.dpad, p.slices, p.slice, a);
      else if (pass == 1)
        knn_pass_stream_kernel<1><<<grid, 256, 0, s>>>(Qb, Rb, Q, R, p.dpad, p.slices, p.slice, a);
      else
        knn_pass_stream_kernel<2><<<grid, 256, 0, s>>>(Qb, Rb, Q, R, p.dpad, p.slices, p.slice, a);
    } else {
#define SIMD2_KNN_PASS(D)                                                                      \
  case D:                                                                                      \
    if (pass == 1 && p.fine)                                                                   \
      knn_pass_kernel<D, 3><<<grid, 256, 0, s>>>(Qb, Rb, Q, R, p.slices, p.slice, a);         \
    else if (pass == 1)                                                                        \
      knn_pass_kernel<D, 1><<<grid, 256, 0, s>>>(Qb, Rb, Q, R, p.slices, p.slice, a);         \
    else                                                                                       \
      knn_pass_kernel<D, 2><<<grid, 256, 0, s>>>(Qb, Rb, Q, R, p.slices, p.slice, a);         \
    break;
      switch (p.dpad) {
        SIMD2_KNN_PASS(32)
        SIMD2_KNN_PASS(64)
        SIMD2_KNN_PASS(96)
        SIMD2_KNN_PASS(128)
        SIMD2_KNN_PASS(160)
      default: return hipErrorInvalidValue;
      }
#undef SIMD2_KNN_PASS
    }
    if (pass == 1)
      knn_tau_kernel<<<(Q + 3) / 4, 256, 0, s>>>(tmin, p.ntiles, Q, p.ck, e_coef, qn, rnmax, thr,
                                                 cn, cex, ovf, big);
  }
  knn_overflow_kernel<<<FB_GRID, 256, 0, s>>>(query, ldq, ref, ldr, R, d, ovf, ovf + 1, scr, cd,
                                              ci, cn, cex);
  // relative gap below which FP32 may order two distances wrongly: the
  // canonical sum's error is at most about (d / 8 + 3) units of 2^-24 each side
  const float tol = fmaxf(1e-5f, (float)(d / 4 + 8) * ldexpf(1.0f, -24));
  rerank_small_kernel<<<(Q + 3) / 4, 256, 0, s>>>(query, ldq, ref, ldr, Q, d, k, tol, cd, ci, cn,
                                                   cex, out_d, out_i, do_sqrt, stats, big, big + 1);
  rerank_big_kernel<<<BIG_GRID_RR, 64, 0, s>>>(query, ldq, ref, ldr, d, k, tol, cd, ci, cn, cex,
                                               out_d, out_i, do_sqrt, stats, big, big + 1);
  return hipGetLastError();
}

}  // namespace simd2

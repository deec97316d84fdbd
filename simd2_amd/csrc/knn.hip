// k-nearest neighbours by pairwise squared L2 -- the PLD app's use of the
// minus_square semiring (apps/pld/cuASR/knncuda.cu:399-549) -- on gfx950.
//
// 1. knn_select_kernel: the reference's post-processing of a materialised
//    distance matrix (modified_insertion_sort_linear + compute_sqrt_linear,
//    knncuda.cu:170-237): per query the k smallest (distance, index) pairs in
//    ascending order, ties kept in index order, optionally square-rooted.
//    Layout as the reference: dist[q + r * Q], outputs [j * Q + q].  Two waves
//    scan two halves of the references for the same 64 queries (coalesced
//    rows), each keeping a sorted register list, and merge through LDS.
//    k <= 32.
//
// 2. The precision form (no counterpart in the reference):
//    * prep_kernel: FP32 points -> BF16 high and low parts (zero-padded to a
//      multiple of 32 dims) and FP32 squared norms;
//    * knn_pass_kernel: ||q||^2 + ||r||^2 - 2 q.r with the dot product on BF16
//      MFMA (v_mfma_f32_16x16x32_bf16, FP32 accumulate; hi.hi + hi.lo + lo.hi,
//      three MFMAs per k-step, for ~16 significant bits).  A workgroup owns 128
//      queries (each wave 32, their A fragments held in registers for the whole
//      scan) and one slice of the references, streamed through LDS in 64-row
//      tiles; the distances stay in the MFMA accumulators (lane l: 4 queries x
//      4 references of a tile).  Two passes over the same products:
//        pass 1 writes, per (query, tile), the smallest distance of the tile;
//        knn_tau_kernel takes each query's CAND_K-th smallest tile minimum,
//          tau: CAND_K distinct references lie at or under it, so it bounds
//          the query's CAND_K-th best distance from above;
//        pass 2 appends every reference at or under tau to the query's
//          candidate buffer (rare: about CAND_K to 2 CAND_K per query).
//      Neither pass keeps lists in the inner loop, and the Q x R matrix is
//      never written;
//    * rerank_kernel: one wave per query picks the 32 best candidates by the
//      approximate distance, recomputes them exactly in FP32 ((q - r)^2 summed
//      by the wave), ranks by (distance, index), and for candidates whose FP32
//      distance lies within a relative 1e-5 of another one near the top k --
//      where cancellation or rounding could order them wrongly -- recomputes the
//      distance in FP64 and ranks those by it.  FP64 is used nowhere else.
#include <float.h>
#include <hip/hip_bf16.h>
#include <stdint.h>

#include "internal.hpp"

namespace simd2 {

namespace {

// ---------------------------------------------------------------------------
// 1. top-k of a materialised distance matrix (reference layout)
// ---------------------------------------------------------------------------
constexpr int SEL_KMAX = 32;
constexpr int SEL_Q = 64;      // queries per workgroup (one per lane)
constexpr int SEL_U = 16;      // distance loads in flight per lane
// Waves splitting the references: as many as the merge lists fit in LDS
// (PARTS x KMAX x 64 x 8 bytes).
template <int KMAX> constexpr int sel_parts() { return KMAX <= 16 ? 16 : 8; }

// The reference's top-k (apps/pld/cuASR/knncuda.cu:170-237) keeps a sorted
// list: while it holds fewer than k entries every element enters; then an
// element enters only with a distance strictly below the k-th, equal
// distances stay in index order, and the last entry drops out.  Here the list
// is kept unsorted with its largest entry cached: an element enters under the
// same rule and replaces the entry that is largest by (distance, index) --
// the one the sorted list drops -- and the k entries are sorted by (distance,
// index) at the end, which gives the same list.  An insertion rescans the k
// entries for the new maximum instead of shifting them.
template <int KMAX>
__device__ __forceinline__ void max_insert(float (&ld)[KMAX], int (&li)[KMAX], int &n,
                                           float &thr, int &tpos, int k, float d, int idx) {
  if (n < k) {
#pragma unroll
    for (int j = 0; j < KMAX; ++j)
      if (j == n) {
        ld[j] = d;
        li[j] = idx;
      }
    if (++n < k) return;
  } else {
    if (!(d < thr)) return;
#pragma unroll
    for (int j = 0; j < KMAX; ++j)
      if (j == tpos) {
        ld[j] = d;
        li[j] = idx;
      }
  }
  float md = ld[0];
  int mi = li[0], mp = 0;
#pragma unroll
  for (int j = 1; j < KMAX; ++j)
    if (j < k && (ld[j] > md || (ld[j] == md && li[j] > mi))) {
      md = ld[j];
      mi = li[j];
      mp = j;
    }
  thr = md;
  tpos = mp;
}

template <int KMAX, int SEL_PARTS = sel_parts<KMAX>()>
__global__ __launch_bounds__(SEL_Q * SEL_PARTS) void knn_select_kernel(
    const float *__restrict__ dist, int Q, int R, int k, float *out_d, int *out_i, int do_sqrt) {
  __shared__ float sd[SEL_PARTS][KMAX][SEL_Q];
  __shared__ int si[SEL_PARTS][KMAX][SEL_Q];
  __shared__ int sn[SEL_PARTS][SEL_Q];
  const int lane = tid_x() % 64;
  const int part = tid_x() / 64;
  const int q = bid_x() * SEL_Q + lane;
  const int chunk = (R + SEL_PARTS - 1) / SEL_PARTS;
  const int r0 = part * chunk;
  const int r1 = min(R, r0 + chunk);
  float ld[KMAX];
  int li[KMAX];
#pragma unroll
  for (int j = 0; j < KMAX; ++j) {  // unfilled entries sort last
    ld[j] = __builtin_inff();
    li[j] = 0x7fffffff;
  }
  int n = 0, tpos = 0;
  float thr = __builtin_inff();
  if (q < Q) {
    int r = r0;
    for (; r + SEL_U <= r1; r += SEL_U) {
      float v[SEL_U];
#pragma unroll
      for (int u = 0; u < SEL_U; ++u) v[u] = dist[(int64_t)(r + u) * Q + q];
#pragma unroll
      for (int u = 0; u < SEL_U; ++u) max_insert<KMAX>(ld, li, n, thr, tpos, k, v[u], r + u);
    }
    for (; r < r1; ++r) max_insert<KMAX>(ld, li, n, thr, tpos, k, dist[(int64_t)r * Q + q], r);
  }
  // sort by (distance, index): odd-even transposition over the registers
#pragma unroll
  for (int pass = 0; pass < KMAX; ++pass)
#pragma unroll
    for (int j = pass & 1; j + 1 < KMAX; j += 2) {
      const bool sw = ld[j + 1] < ld[j] || (ld[j + 1] == ld[j] && li[j + 1] < li[j]);
      const float td = sw ? ld[j + 1] : ld[j];
      const int ti = sw ? li[j + 1] : li[j];
      ld[j + 1] = sw ? ld[j] : ld[j + 1];
      li[j + 1] = sw ? li[j] : li[j + 1];
      ld[j] = td;
      li[j] = ti;
    }
#pragma unroll
  for (int j = 0; j < KMAX; ++j) {
    sd[part][j][lane] = ld[j];
    si[part][j][lane] = li[j];
  }
  sn[part][lane] = n;
  __syncthreads();
  if (part != 0 || q >= Q) return;
  // Merge of the parts' sorted lists: part p holds lower indices than part
  // p + 1, so on equal distances the lowest part wins, which keeps index
  // order.
  int pos[SEL_PARTS];
#pragma unroll
  for (int p = 0; p < SEL_PARTS; ++p) pos[p] = 0;
  for (int j = 0; j < k; ++j) {
    int best = -1;
    float bd = 0.f;
#pragma unroll
    for (int p = 0; p < SEL_PARTS; ++p) {
      if (pos[p] < sn[p][lane]) {
        const float dp = sd[p][pos[p]][lane];
        if (best < 0 || dp < bd) {
          best = p;
          bd = dp;
        }
      }
    }
    if (best < 0) break;
    int bi = 0;
#pragma unroll
    for (int p = 0; p < SEL_PARTS; ++p)
      if (p == best) {
        bi = si[p][pos[p]][lane];
        ++pos[p];
      }
    out_d[(int64_t)j * Q + q] = do_sqrt ? sqrtf(bd) : bd;
    out_i[(int64_t)j * Q + q] = bi;
  }
}

// ---------------------------------------------------------------------------
// 2. precision form
// ---------------------------------------------------------------------------
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

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

// One wave per point: the FP32 coordinates split into a BF16 high part and a
// BF16 low part (x - hi), dims zero-padded to dpad, stored [point][hi | lo],
// and the FP32 squared norm (wave-summed).
__global__ __launch_bounds__(256) void prep_kernel(const float *__restrict__ P, int64_t ldp,
                                                   int n, int d, int dpad, __bf16 *Pb,
                                                   float *norm) {
  const int lane = tid_x() % 64;
  const int p = bid_x() * 4 + tid_x() / 64;
  if (p >= n) return;
  float s = 0.f;
  for (int j = lane; j < dpad; j += 64) {
    const float x = j < d ? P[(int64_t)p * ldp + j] : 0.f;
    s += x * x;
    const __bf16 hi = __float2bfloat16(x);
    Pb[(int64_t)p * 2 * dpad + j] = hi;
    Pb[(int64_t)p * 2 * dpad + dpad + j] = __float2bfloat16(x - __bfloat162float(hi));
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
  if (lane == 0) norm[p] = s;
}

constexpr int QG = 2;       // groups of 16 queries per wave
constexpr int CQ = 64 * QG; // queries per workgroup (4 waves)
constexpr int CR = 64;      // references per tile
constexpr int CAND_K = 32;  // tau bounds the CAND_K-th best distance
// Candidate buffer per query: with no ties at tau at most CAND_K tiles hold
// references at or under it, 64 each; a pathological tie of more tile
// minima at exactly tau keeps the first CAND_CAP arrivals.
constexpr int CAND_CAP = CAND_K * CR;

// q.r ~ qh.rh + qh.rl + ql.rh: three BF16 MFMAs per k-step carry ~16
// significant bits, which the norm expansion needs (centres ~N(0, 10^2) make
// ||q||^2 + ||r||^2 ~ 100x the neighbour distances); a single BF16 product
// leaves errors of ~10 on distances ~250 and loses true neighbours.
// PASS 1: tmin[q][tile] = min over the tile; PASS 2: append r with d <= tau[q].
template <int DPAD, int PASS>
__global__ __launch_bounds__(256) void knn_pass_kernel(
    const __bf16 *__restrict__ Qb, const __bf16 *__restrict__ Rb, const float *__restrict__ qn,
    const float *__restrict__ rn, int Q, int R, int slices, int slice, float *tmin, int ntiles,
    const float *__restrict__ tau, float *cand_d, int *cand_i, int *cand_n) {
  constexpr int KS = DPAD / 32;             // MFMA k-steps
  constexpr int RLD = 2 * DPAD + 8;         // LDS row: hi | lo, +16 bytes of padding
  constexpr int CHUNKS = CR * 2 * DPAD / 8; // 16-byte chunks per reference tile
  constexpr int PER = CHUNKS / 256;         // chunks per thread
  __shared__ __attribute__((aligned(16))) __bf16 Rt[CR][RLD];
  __shared__ float Rn[CR];
  constexpr int CB = PASS == 2 ? 16 : 1;  // LDS candidate slots per query
  __shared__ float Cd[CQ][CB];
  __shared__ int Ci[CQ][CB];
  __shared__ int Cn[CQ];

  const int tid = tid_x();
  const int lane = tid % 64;
  const int wave = tid / 64;
  if (PASS == 2)
    for (int i = tid; i < CQ; i += 256) Cn[i] = 0;  // published by the loop's first barrier
  // Workgroup b takes query block b / slices and reference slice b % slices:
  // with 8 slices, workgroups run on XCD b % 8, so each XCD streams one slice
  // and its L2 keeps it.
  const int q0 = (bid_x() / slices) * CQ;
  const int rb0 = (bid_x() % slices) * slice;
  const int rb1 = min(R, rb0 + slice);

  // A fragments of this wave's QG groups of 16 queries (every reference
  // fragment read from LDS feeds QG x 3 MFMAs): lane l holds query l%16 of
  // each group, dims 8*(l/16) .. +7 of every 32-wide k-step, high and low.
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
  // The 4 queries per group whose accumulator rows this lane holds: norms and tau.
  float qnv[QG][4], tv[QG][4];
  int qv[QG][4];
#pragma unroll
  for (int g = 0; g < QG; ++g)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      qv[g][v] = q0 + wave * 16 * QG + g * 16 + 4 * (lane / 16) + v;
      qnv[g][v] = qv[g][v] < Q ? qn[qv[g][v]] : 0.f;
      tv[g][v] = PASS == 2 && qv[g][v] < Q ? tau[qv[g][v]] : -__builtin_inff();
    }

  // Register prefetch of the next reference tile (global -> registers while
  // the current tile is multiplied, registers -> LDS after).
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
    if (tid < CR) pren = (r0 + tid < rb1) ? rn[r0 + tid] : 0.f;
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
    // Distances of this lane's queries x references rb*16 + lane%16;
    // references past the slice end take no part.
    float rnv[4];
    bool valid[4];
#pragma unroll
    for (int rb = 0; rb < 4; ++rb) {
      rnv[rb] = Rn[rb * 16 + lane % 16];
      valid[rb] = r0 + rb * 16 + lane % 16 < rb1;
    }
#pragma unroll
    for (int g = 0; g < QG; ++g)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        if constexpr (PASS == 1) {  // the tile minimum of each query
          float m = __builtin_inff();
#pragma unroll
          for (int rb = 0; rb < 4; ++rb)
            if (valid[rb]) m = fminf(m, qnv[g][v] + rnv[rb] - 2.0f * acc[g][rb][v]);
          // the 16 lanes of this query (same lane / 16, one DPP row) hold the
          // tile's 64 references
          m = row_min16(m);
          if (lane % 16 == 0 && qv[g][v] < Q) tmin[(int64_t)qv[g][v] * ntiles + r0 / CR] = m;
        }
      }
    if constexpr (PASS == 2) {
      // References at or under tau (about CAND_K per query over all slices)
      // take a slot of the query's LDS buffer; overflow goes straight to the
      // global buffer.
#pragma unroll
      for (int g = 0; g < QG; ++g)
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
          for (int rb = 0; rb < 4; ++rb) {
            const float dd = qnv[g][v] + rnv[rb] - 2.0f * acc[g][rb][v];
            if (valid[rb] && dd <= tv[g][v]) {
              const int q = qv[g][v], ql = q - q0;
              const int at = atomicAdd(&Cn[ql], 1);
              const int r = r0 + rb * 16 + lane % 16;
              if (at < CB) {
                Cd[ql][at] = dd;
                Ci[ql][at] = r;
              } else {
                const int gat = atomicAdd(&cand_n[q], 1);
                if (gat < CAND_CAP) {
                  cand_d[(int64_t)q * CAND_CAP + gat] = dd;
                  cand_i[(int64_t)q * CAND_CAP + gat] = r;
                }
              }
            }
          }
    }
  }
  if constexpr (PASS == 2) {
    // The LDS buffers to the global ones: one atomic per query.
    __syncthreads();
    for (int ql = tid; ql < CQ; ql += 256) {
      const int q = q0 + ql, nb = min(Cn[ql], CB);
      if (q < Q && nb > 0) {
        const int b = atomicAdd(&cand_n[q], nb);
        for (int e = 0; e < nb && b + e < CAND_CAP; ++e) {
          cand_d[(int64_t)q * CAND_CAP + b + e] = Cd[ql][e];
          cand_i[(int64_t)q * CAND_CAP + b + e] = Ci[ql][e];
        }
      }
    }
  }
}

// Order-preserving 32-bit key of a (non-NaN) float.
__device__ __forceinline__ uint32_t fkey(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float fkey_inv(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}

// tau[q] = the CAND_K-th smallest of the query's tile minima (+inf when it has
// fewer tiles), by bisection over the keys; one wave per query.  Also clears
// the query's candidate count.
__global__ __launch_bounds__(256) void knn_tau_kernel(const float *__restrict__ tmin, int ntiles,
                                                      int Q, float *tau, int *cand_n) {
  const int lane = tid_x() % 64;
  const int q = bid_x() * 4 + tid_x() / 64;
  if (q >= Q) return;
  const float *t = tmin + (int64_t)q * ntiles;
  uint32_t lo = 0u, hi = 0xffffffffu;  // smallest key with >= CAND_K minima at or under it
  constexpr int TP = 8;                 // keys held in registers per lane (ntiles <= 512)
  uint32_t kv[TP];
#pragma unroll
  for (int i = 0; i < TP; ++i) {
    const int j = lane + 64 * i;
    kv[i] = j < ntiles ? fkey(t[j]) : 0xffffffffu;
  }
  if (ntiles < CAND_K) {
    lo = fkey(__builtin_inff());
  } else {
    while (lo < hi) {
      const uint32_t mid = lo + (hi - lo) / 2;
      // The wave's count of keys <= mid: one ballot per register of keys,
      // counted in scalar registers (no cross-lane shuffles); keys past the
      // registers (ntiles > 512) are counted per lane and summed.
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
      if (c >= CAND_K)
        hi = mid;
      else
        lo = mid + 1;
    }
  }
  if (lane == 0) {
    tau[q] = fkey_inv(lo);
    cand_n[q] = 0;
  }
}

constexpr int RR_M = 32;        // candidates re-ranked exactly per query
constexpr float RR_TOL = 1e-5f; // relative FP32 gap treated as a near tie

__device__ __forceinline__ bool key_less(float d, int i, float e, int j) {
  return d < e || (d == e && i < j);
}
__device__ __forceinline__ bool key_less64(double d, int i, double e, int j) {
  return d < e || (d == e && i < j);
}

__global__ __launch_bounds__(256) void rerank_kernel(const float *__restrict__ Qf, int64_t ldq,
                                                     const float *__restrict__ Rf, int64_t ldr,
                                                     int Q, int d, int k,
                                                     const float *__restrict__ cand_d,
                                                     const int *__restrict__ cand_i,
                                                     const int *__restrict__ cand_n, int ncand,
                                                     float *out_d, int *out_i, int do_sqrt,
                                                     int *stats) {
  const int lane = tid_x() % 64;
  const int q = bid_x() * 4 + tid_x() / 64;
  if (q >= Q) return;
  const float *cd = cand_d + (int64_t)q * ncand;
  const int *ci = cand_i + (int64_t)q * ncand;
  const int nq = min(cand_n[q], ncand);  // candidates of this query

  // Step 1: the RR_M best candidates by (approximate distance, index).
  int sel = -1;  // lane m < RR_M receives the m-th choice
  if (nq <= 64) {
    // one candidate per lane, bitonic sort of the wave by the key
    float bd = lane < nq ? cd[lane] : __builtin_inff();
    int bi = lane < nq ? ci[lane] : 0x7fffffff;
#pragma unroll
    for (int k = 2; k <= 64; k <<= 1)
#pragma unroll
      for (int j = k >> 1; j > 0; j >>= 1) {
        const float od = __shfl_xor(bd, j);
        const int oi = __shfl_xor(bi, j);
        // the lower lane of the pair keeps the smaller key in an ascending run
        const bool up = (lane & k) == 0, low = (lane & j) == 0;
        const bool take = (low == up) ? key_less(od, oi, bd, bi) : key_less(bd, bi, od, oi);
        if (take) {
          bd = od;
          bi = oi;
        }
      }
    if (lane < RR_M && bi != 0x7fffffff) sel = bi;
  }
  float last_d = -__builtin_inff();
  int last_i = -1;
  for (int m = 0; nq > 64 && m < RR_M; ++m) {
    float bd = __builtin_inff();
    int bi = 0x7fffffff;
    for (int c = lane; c < nq; c += 64) {
      const float dd = cd[c];
      const int ii = ci[c];
      if (ii >= 0 && key_less(last_d, last_i, dd, ii) && key_less(dd, ii, bd, bi)) {
        bd = dd;
        bi = ii;
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      const float od = __shfl_xor(bd, off);
      const int oi = __shfl_xor(bi, off);
      if (key_less(od, oi, bd, bi)) {
        bd = od;
        bi = oi;
      }
    }
    if (bi == 0x7fffffff) break;  // fewer candidates than RR_M
    if (lane == m) sel = bi;
    last_d = bd;
    last_i = bi;
  }

  // Step 2: exact FP32 distances, 8 candidates at a time: lanes 8c .. 8c+7
  // sum dims part, part + 8, ... of candidate m0 + c, so that 8 row loads
  // are in flight instead of one.
  float d32 = __builtin_inff();
  {
    const int grp = lane / 8, part = lane % 8;
    for (int m0 = 0; m0 < RR_M; m0 += 8) {
      const int r = __shfl(sel, m0 + grp);
      float s = 0.f;
      if (r >= 0)
        for (int j = part; j < d; j += 8) {
          const float t = Qf[(int64_t)q * ldq + j] - Rf[(int64_t)r * ldr + j];
          s += t * t;
        }
      s += __shfl_xor(s, 1);
      s += __shfl_xor(s, 2);
      s += __shfl_xor(s, 4);
      const float got = __shfl(s, 8 * ((lane - m0) & 7));
      if (lane >= m0 && lane < m0 + 8 && sel >= 0) d32 = got;
    }
  }
  const int my_i = sel >= 0 ? sel : 0x7fffffff;

  // Step 3: rank by (FP32 distance, index).
  int rank = 0;
  for (int m = 0; m < RR_M; ++m) {
    const float od = __shfl(d32, m);
    const int oi = __shfl(my_i, m);
    if (m != lane && key_less(od, oi, d32, my_i)) ++rank;
  }
  // Step 4: near ties that touch the top k get an FP64 distance.
  bool flag = false;
  for (int m = 0; m < RR_M; ++m) {
    const float od = __shfl(d32, m);
    const int orank = __shfl(rank, m);
    const int oi = __shfl(my_i, m);
    if (lane < RR_M && m != lane && sel >= 0 && oi != 0x7fffffff &&
        fabsf(od - d32) <= RR_TOL * fmaxf(d32, 1e-30f) && (rank <= k || orank <= k))
      flag = true;
  }
  double d64 = (double)d32;
  const uint64_t fl = __ballot(flag);
  if (fl) {
    for (int m = 0; m < RR_M; ++m) {
      if (!((fl >> m) & 1ull)) continue;
      const int r = __shfl(sel, m);
      double s = 0.0;
      for (int j = lane; j < d; j += 64) {
        const double t = (double)Qf[(int64_t)q * ldq + j] - (double)Rf[(int64_t)r * ldr + j];
        s += t * t;
      }
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
      if (lane == m) d64 = s;
    }
    rank = 0;
    for (int m = 0; m < RR_M; ++m) {
      const double od = __shfl(d64, m);
      const int oi = __shfl(my_i, m);
      if (m != lane && key_less64(od, oi, d64, my_i)) ++rank;
    }
    if (stats && lane == 0) {
      atomicAdd(&stats[0], 1);
      atomicAdd(&stats[1], (int)__popcll(fl));
    }
  }
  if (lane < RR_M && sel >= 0 && rank < k) {
    const float outd = flag ? (float)d64 : d32;
    out_d[(int64_t)q * k + rank] = do_sqrt ? sqrtf(outd) : outd;
    out_i[(int64_t)q * k + rank] = sel;
  }
}

int dpad_of(int d) { return ((d + 31) / 32) * 32; }

}  // namespace

hipError_t launch_knn_select(const float *dist, int Q, int R, int k, float *out_d, int *out_i,
                             int do_sqrt, hipStream_t s) {
  const dim3 grid((Q + SEL_Q - 1) / SEL_Q);
  if (k <= 16)
    knn_select_kernel<16><<<grid, SEL_Q * sel_parts<16>(), 0, s>>>(dist, Q, R, k, out_d, out_i,
                                                                   do_sqrt);
  else
    knn_select_kernel<SEL_KMAX><<<grid, SEL_Q * sel_parts<SEL_KMAX>(), 0, s>>>(
        dist, Q, R, k, out_d, out_i, do_sqrt);
  return hipGetLastError();
}

size_t knn_bf16_workspace(int Q, int R, int d, int *slices_out) {
  const int dpad = dpad_of(d);
  // 8 reference slices, one per XCD (see knn_pass_kernel), fewer when the
  // references are fewer than 8 tiles.
  int slices = 8;
  const int max_slices = (R + CR - 1) / CR;
  if (slices > max_slices) slices = max_slices < 1 ? 1 : max_slices;
  if (slices_out) *slices_out = slices;
  const size_t ntiles = ((size_t)R + CR - 1) / CR;
  auto a256 = [](size_t b) { return (b + 255) / 256 * 256; };
  size_t bytes = 0;
  bytes += a256((size_t)Q * dpad * 4);  // hi | lo BF16
  bytes += a256((size_t)R * dpad * 4);
  bytes += a256((size_t)Q * 4);          // query norms
  bytes += a256((size_t)R * 4);          // reference norms
  bytes += a256((size_t)Q * ntiles * 4); // tile minima
  bytes += a256((size_t)Q * 4);          // tau
  bytes += a256((size_t)Q * 4);          // candidate counts
  bytes += 2 * a256((size_t)Q * CAND_CAP * 4);
  return bytes;
}

hipError_t launch_knn_bf16(const float *query, int64_t ldq, const float *ref, int64_t ldr, int Q,
                           int R, int d, int k, float *out_d, int *out_i, int do_sqrt,
                           void *workspace, int *stats, hipStream_t s) {
  int slices = 1;
  knn_bf16_workspace(Q, R, d, &slices);
  const int dpad = dpad_of(d);
  const int ntiles = (R + CR - 1) / CR;
  char *w = static_cast<char *>(workspace);
  auto take = [&](size_t bytes) {
    char *p = w;
    w += (bytes + 255) / 256 * 256;
    return p;
  };
  __bf16 *Qb = reinterpret_cast<__bf16 *>(take((size_t)Q * dpad * 4));
  __bf16 *Rb = reinterpret_cast<__bf16 *>(take((size_t)R * dpad * 4));
  float *qn = reinterpret_cast<float *>(take((size_t)Q * 4));
  float *rn = reinterpret_cast<float *>(take((size_t)R * 4));
  float *tmin = reinterpret_cast<float *>(take((size_t)Q * ntiles * 4));
  float *tau = reinterpret_cast<float *>(take((size_t)Q * 4));
  int *cn = reinterpret_cast<int *>(take((size_t)Q * 4));
  float *cd = reinterpret_cast<float *>(take((size_t)Q * CAND_CAP * 4));
  int *ci = reinterpret_cast<int *>(take((size_t)Q * CAND_CAP * 4));

  prep_kernel<<<(Q + 3) / 4, 256, 0, s>>>(query, ldq, Q, d, dpad, Qb, qn);
  prep_kernel<<<(R + 3) / 4, 256, 0, s>>>(ref, ldr, R, d, dpad, Rb, rn);
  int slice = (R + slices - 1) / slices;
  slice = (slice + CR - 1) / CR * CR;
  const dim3 grid(((Q + CQ - 1) / CQ) * slices);
  for (int pass = 1; pass <= 2; ++pass) {
#define SIMD2_KNN_PASS(D)                                                                      \
  case D:                                                                                      \
    if (pass == 1)                                                                             \
      knn_pass_kernel<D, 1><<<grid, 256, 0, s>>>(Qb, Rb, qn, rn, Q, R, slices, slice, tmin,  \
                                                 ntiles, tau, cd, ci, cn);                     \
    else                                                                                       \
      knn_pass_kernel<D, 2><<<grid, 256, 0, s>>>(Qb, Rb, qn, rn, Q, R, slices, slice, tmin,  \
                                                 ntiles, tau, cd, ci, cn);                     \
    break;
    switch (dpad) {
      SIMD2_KNN_PASS(32)
      SIMD2_KNN_PASS(64)
      SIMD2_KNN_PASS(96)
      SIMD2_KNN_PASS(128)
      SIMD2_KNN_PASS(160)
    default: return hipErrorInvalidValue;
    }
#undef SIMD2_KNN_PASS
    if (pass == 1) knn_tau_kernel<<<(Q + 3) / 4, 256, 0, s>>>(tmin, ntiles, Q, tau, cn);
  }
  rerank_kernel<<<(Q + 3) / 4, 256, 0, s>>>(query, ldq, ref, ldr, Q, d, k, cd, ci, cn, CAND_CAP,
                                           out_d, out_i, do_sqrt, stats);
  return hipGetLastError();
}

}  // namespace simd2

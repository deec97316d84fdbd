// Top-k of a materialised distance matrix -- the reference's PLD
// post-processing (modified_insertion_sort_linear + compute_sqrt_linear,
// apps/pld/cuASR/knncuda.cu:170-237) -- for any k <= R, on gfx950.
//
// Layout as the reference: dist[q + r * Q] (column-major Q x R: a reference's
// distances to consecutive queries are contiguous), outputs [j * Q + q].
//
// Result: per query the k smallest (distance, index) pairs in ascending
// order, equal distances in index order -- the reference's list: its sorted
// insertion admits an element while fewer than k are held, then only with a
// distance strictly below the k-th, and shifts on strict '>', so among equal
// distances the earlier index stays ahead.  That is exactly the order of the
// unique 64-bit key (order-preserving bits of d) << 32 | r.
//
// Method (exact, no per-element lists; two streaming passes over dist):
//   1. group minima: for each query and each group of SG = 32 consecutive
//      references, the smallest key of the group (one wave = 64 queries, one
//      per lane, x 16 groups; every load is 64 consecutive floats);
//   2. tau[q] = the k-th smallest of the query's group minima.  The k group
//      minima at or below tau are k distinct references, so the k-th smallest
//      key of the row is <= tau: every key of the answer is <= tau.  Keys
//      strictly below tau lie in the k - 1 groups whose minimum is below tau
//      (keys are unique), so at most (k - 1) * SG + 1 keys are <= tau;
//   3. collect: a second pass appends every key <= tau to the query's
//      candidate list (typically k to 2k entries; the bound above sizes it);
//   4. final: sort the candidates and write the first k (square-rooted on
//      request), one wave per query for up to 64 candidates, a workgroup
//      with an LDS or global bitonic sort for more.
// Queries are processed in batches that bound the workspace.
#include <stdint.h>

#include <algorithm>

#include "internal.hpp"

namespace simd2 {

namespace {

constexpr int SG = 32;                 // references per group
constexpr int GCH = 16;                // groups per wave task
constexpr int TASK_REFS = SG * GCH;    // 512 references
constexpr int BIG_THREADS = 1024;
constexpr int LDS_SORT = 8192;         // keys sorted in LDS by one workgroup (64 KB)
constexpr int BIG_GRID = 512;
constexpr uint64_t KEY_MAX = ~0ull;

__device__ __forceinline__ uint32_t fkey(float f) {
  const uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float fkey_inv(uint32_t k) {
  return __uint_as_float((k & 0x80000000u) ? (k & 0x7fffffffu) : ~k);
}
__device__ __forceinline__ uint64_t key_of(float d, int r) {
  return ((uint64_t)fkey(d) << 32) | (uint32_t)r;
}

// Loads the SG distances of group g of query q (references clamped into
// [0, R); the caller masks the ones past R).
__device__ __forceinline__ void load_group(const float *__restrict__ dist, int64_t Q, int R,
                                           int q, int g, float (&v)[SG]) {
#pragma unroll
  for (int u = 0; u < SG; ++u) {
    const int r = min(g * SG + u, R - 1);
    v[u] = dist[(int64_t)r * Q + q];
  }
}

// Pass 1.  One wave = one query block (64 queries, one per lane) x GCH groups
// (task); writes gkey[(q - q0) * ngroups + g].
__global__ __launch_bounds__(256) void sel_group_min_kernel(
    const float *__restrict__ dist, int64_t Q, int R, int q0, int qn, int tasks, int ngroups,
    uint64_t *__restrict__ gkey) {
  const int lane = tid_x() % 64;
  // workgroup b: task b % tasks for query blocks 4 (b / tasks) .. + 3, one per
  // wave, so the workgroup reads 1 KB of each row at a time
  const int qblocks = (qn + 63) / 64;
  const int qb = (bid_x() / tasks) * 4 + tid_x() / 64;
  if (qb >= qblocks) return;
  const int ql = qb * 64 + lane;  // query within the batch
  const bool live = ql < qn;
  const int q = q0 + (live ? ql : qn - 1);
  const int g0 = (bid_x() % tasks) * GCH;
  const int g1 = min(ngroups, g0 + GCH);
  uint64_t out[GCH];
  float va[SG], vb[SG];
  load_group(dist, Q, R, q, g0, va);
#pragma unroll
  for (int gi = 0; gi < GCH; gi += 2) {
    // two register sets: the next group's loads are in flight while one is reduced
    if (g0 + gi + 1 < g1) load_group(dist, Q, R, q, g0 + gi + 1, vb);
    if (g0 + gi < g1) {
      const int rb = (g0 + gi) * SG;
      float m = va[0];
      int mi = 0;
#pragma unroll
      for (int u = 1; u < SG; ++u)
        if (rb + u < R && va[u] < m) {  // strict: the earliest index keeps a tie
          m = va[u];
          mi = u;
        }
      out[gi] = key_of(m, rb + mi);
    }
    if (g0 + gi + 2 < g1) load_group(dist, Q, R, q, g0 + gi + 2, va);
    if (g0 + gi + 1 < g1) {
      const int rb = (g0 + gi + 1) * SG;
      float m = vb[0];
      int mi = 0;
#pragma unroll
      for (int u = 1; u < SG; ++u)
        if (rb + u < R && vb[u] < m) {
          m = vb[u];
          mi = u;
        }
      out[gi + 1] = key_of(m, rb + mi);
    }
  }
  if (!live) return;
  uint64_t *dst = gkey + (int64_t)ql * ngroups;
#pragma unroll
  for (int gi = 0; gi < GCH; ++gi)
    if (g0 + gi < g1) dst[g0 + gi] = out[gi];
}

// Number of the query's group-minimum keys satisfying pred, over the wave.
template <int TP, typename Pred>
__device__ __forceinline__ int wave_count(const uint64_t (&kv)[TP], const uint64_t *t,
                                          int ngroups, int lane, Pred pred) {
  int c = 0;
#pragma unroll
  for (int i = 0; i < TP; ++i)
    c += (int)__popcll(__ballot(lane + 64 * i < ngroups && pred(kv[i])));
  if (ngroups > 64 * TP) {
    int e = 0;
    for (int j = lane + 64 * TP; j < ngroups; j += 64) e += pred(t[j]) ? 1 : 0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) e += __shfl_xor(e, off);
    c += e;
  }
  return c;
}

// Pass 2.  One wave per query: tau = the k-th smallest group-minimum key
// (KEY_MAX when there are fewer than k groups), by bisection on the distance
// bits and then, among minima with that distance, on the index.  Also clears
// the query's candidate count (and, in the first wave, the count of queries
// for the workgroup sort).
__global__ __launch_bounds__(256) void sel_tau_kernel(const uint64_t *__restrict__ gkey,
                                                      int ngroups, int qn, int k, uint64_t *tau,
                                                      int *cnt, int *big_n) {
  const int lane = tid_x() % 64;
  const int ql = bid_x() * 4 + tid_x() / 64;
  if (ql == 0 && lane == 0) *big_n = 0;
  if (ql >= qn) return;
  const uint64_t *t = gkey + (int64_t)ql * ngroups;
  constexpr int TP = 8;
  uint64_t kv[TP];
#pragma unroll
  for (int i = 0; i < TP; ++i) {
    const int j = lane + 64 * i;
    kv[i] = j < ngroups ? t[j] : KEY_MAX;
  }
  uint64_t res = KEY_MAX;
  if (ngroups >= k) {
    // the smallest distance key with >= k minima at or under it, between the
    // smallest and the largest of the query's minima
    uint32_t lo = 0xffffffffu, hi = 0u;
#pragma unroll
    for (int i = 0; i < TP; ++i)
      if (lane + 64 * i < ngroups) {
        lo = min(lo, (uint32_t)(kv[i] >> 32));
        hi = max(hi, (uint32_t)(kv[i] >> 32));
      }
    for (int j = lane + 64 * TP; j < ngroups; j += 64) {
      lo = min(lo, (uint32_t)(t[j] >> 32));
      hi = max(hi, (uint32_t)(t[j] >> 32));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      lo = min(lo, (uint32_t)__shfl_xor((int)lo, off));
      hi = max(hi, (uint32_t)__shfl_xor((int)hi, off));
    }
    while (lo < hi) {
      const uint32_t mid = lo + (hi - lo) / 2;
      const int c = wave_count<TP>(kv, t, ngroups, lane,
                                   [&](uint64_t x) { return (uint32_t)(x >> 32) <= mid; });
      if (c >= k)
        hi = mid;
      else
        lo = mid + 1;
    }
    const uint32_t th = lo;
    const int below = wave_count<TP>(kv, t, ngroups, lane,
                                     [&](uint64_t x) { return (uint32_t)(x >> 32) < th; });
    const int need = k - below;  // >= 1: the need-th smallest index among minima with distance th
    const int ties = wave_count<TP>(kv, t, ngroups, lane,
                                    [&](uint64_t x) { return (uint32_t)(x >> 32) == th; });
    uint32_t ilo = 0u, ihi = 0xffffffffu;
    if (need == ties) ilo = ihi = 0xffffffffu;  // all of them: any index bound at or above theirs
    while (ilo < ihi) {
      const uint32_t mid = ilo + (ihi - ilo) / 2;
      const int c = wave_count<TP>(kv, t, ngroups, lane, [&](uint64_t x) {
        return (uint32_t)(x >> 32) == th && (uint32_t)x <= mid;
      });
      if (c >= need)
        ihi = mid;
      else
        ilo = mid + 1;
    }
    res = ((uint64_t)th << 32) | ilo;
  }
  if (lane == 0) {
    tau[ql] = res;
    cnt[ql] = 0;
  }
}

// Pass 3.  Same wave tasks as pass 1: every key <= tau of the query goes to
// its candidate list.
__global__ __launch_bounds__(256) void sel_collect_kernel(
    const float *__restrict__ dist, int64_t Q, int R, int q0, int qn, int tasks, int ngroups,
    const uint64_t *__restrict__ tau, uint64_t *__restrict__ cand, int cap, int *cnt) {
  const int lane = tid_x() % 64;
  const int qblocks = (qn + 63) / 64;
  const int qb = (bid_x() / tasks) * 4 + tid_x() / 64;
  if (qb >= qblocks) return;
  const int ql = qb * 64 + lane;
  const bool live = ql < qn;
  const int q = q0 + (live ? ql : qn - 1);
  const uint64_t tq = live ? tau[ql] : 0ull;
  const uint32_t th = (uint32_t)(tq >> 32), tl = (uint32_t)tq;
  const int g0 = (bid_x() % tasks) * GCH;
  const int g1 = min(ngroups, g0 + GCH);
  auto scan = [&](const float (&v)[SG], int g) {
    const int rb = g * SG;
    uint32_t hit = 0;
#pragma unroll
    for (int u = 0; u < SG; ++u) {
      const uint32_t h = fkey(v[u]);
      const bool in = live && rb + u < R && (h < th || (h == th && (uint32_t)(rb + u) <= tl));
      hit |= in ? 1u << u : 0u;
    }
    if (__ballot(hit != 0) == 0) return;  // the common case: nothing of this group
    while (hit) {
      const int u = __builtin_ctz(hit);
      hit &= hit - 1;
      const int s = atomicAdd(&cnt[ql], 1);
      if (s < cap) cand[(int64_t)ql * cap + s] = key_of(v[u], rb + u);
    }
  };
  float va[SG], vb[SG];
  load_group(dist, Q, R, q, g0, va);
#pragma unroll
  for (int gi = 0; gi < GCH; gi += 2) {
    if (g0 + gi + 1 < g1) load_group(dist, Q, R, q, g0 + gi + 1, vb);
    if (g0 + gi < g1) scan(va, g0 + gi);
    if (g0 + gi + 2 < g1) load_group(dist, Q, R, q, g0 + gi + 2, va);
    if (g0 + gi + 1 < g1) scan(vb, g0 + gi + 1);
  }
}

__device__ __forceinline__ void write_out(uint64_t key, int j, int q, int64_t Q, float *out_d,
                                          int *out_i, int do_sqrt) {
  const float d = fkey_inv((uint32_t)(key >> 32));
  out_d[(int64_t)j * Q + q] = do_sqrt ? sqrtf(d) : d;
  out_i[(int64_t)j * Q + q] = (int)(uint32_t)key;
}

__device__ __forceinline__ uint64_t shfl_xor64(uint64_t v, int m) {
  const int lo = __shfl_xor((int)(uint32_t)v, m);
  const int hi = __shfl_xor((int)(uint32_t)(v >> 32), m);
  return ((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo;
}

// Pass 4a.  One wave per query: up to 64 candidates sorted by a wave bitonic
// network, the first k written; queries with more go to the list of pass 4b.
__global__ __launch_bounds__(256) void sel_final_small_kernel(
    const uint64_t *__restrict__ cand, const int *__restrict__ cnt, int cap, int q0, int qn,
    int64_t Q, int k, float *out_d, int *out_i, int do_sqrt, int *big_n, int *big_list) {
  const int lane = tid_x() % 64;
  const int ql = bid_x() * 4 + tid_x() / 64;
  if (ql >= qn) return;
  const int c = min(cnt[ql], cap);
  if (c > 64) {
    if (lane == 0) big_list[atomicAdd(big_n, 1)] = ql;
    return;
  }
  uint64_t v = lane < c ? cand[(int64_t)ql * cap + lane] : KEY_MAX;
#pragma unroll
  for (int s = 2; s <= 64; s <<= 1)
#pragma unroll
    for (int j = s >> 1; j > 0; j >>= 1) {
      const uint64_t o = shfl_xor64(v, j);
      const bool up = (lane & s) == 0, low = (lane & j) == 0;
      const bool take = (low == up) ? (o < v) : (o > v);
      if (take) v = o;
    }
  if (lane < k && lane < c) write_out(v, lane, q0 + ql, Q, out_d, out_i, do_sqrt);
}

// Pass 4b.  One workgroup per listed query (grid-stride over the list):
// bitonic sort of the candidates padded to a power of two P (the list's
// capacity is one), in LDS when P <= LDS_SORT, else in the list itself in
// global memory; the first k written.
__global__ __launch_bounds__(BIG_THREADS) void sel_final_big_kernel(
    uint64_t *__restrict__ cand, const int *__restrict__ cnt, int cap, int q0, int64_t Q, int k,
    float *out_d, int *out_i, int do_sqrt, const int *big_n, const int *big_list) {
  __shared__ uint64_t S[LDS_SORT];
  const int tid = tid_x();
  const int nbig = *big_n;
  for (int b = bid_x(); b < nbig; b += BIG_GRID) {
    const int ql = big_list[b];
    const int c = min(cnt[ql], cap);
    int P = 1;
    while (P < c) P <<= 1;
    uint64_t *L = cand + (int64_t)ql * cap;
    if (P <= LDS_SORT) {
      for (int i = tid; i < P; i += BIG_THREADS) S[i] = i < c ? L[i] : KEY_MAX;
      __syncthreads();
      for (int s = 2; s <= P; s <<= 1)
        for (int j = s >> 1; j > 0; j >>= 1) {
          for (int i = tid; i < P; i += BIG_THREADS) {
            const int p = i ^ j;
            if (p > i) {
              const uint64_t x = S[i], y = S[p];
              const bool up = (i & s) == 0;
              if ((x > y) == up) {
                S[i] = y;
                S[p] = x;
              }
            }
          }
          __syncthreads();
        }
      for (int j = tid; j < k; j += BIG_THREADS) write_out(S[j], j, q0 + ql, Q, out_d, out_i, do_sqrt);
    } else {
      for (int i = c + tid; i < P; i += BIG_THREADS) L[i] = KEY_MAX;
      __syncthreads();
      for (int s = 2; s <= P; s <<= 1)
        for (int j = s >> 1; j > 0; j >>= 1) {
          for (int i = tid; i < P; i += BIG_THREADS) {
            const int p = i ^ j;
            if (p > i) {
              const uint64_t x = L[i], y = L[p];
              const bool up = (i & s) == 0;
              if ((x > y) == up) {
                L[i] = y;
                L[p] = x;
              }
            }
          }
          __threadfence_block();
          __syncthreads();
        }
      for (int j = tid; j < k; j += BIG_THREADS) write_out(L[j], j, q0 + ql, Q, out_d, out_i, do_sqrt);
    }
    __syncthreads();  // S and L are reused by the next listed query
  }
}

struct SelPlan {
  int ngroups, tasks, cap, qb;
  size_t b_gkey, b_tau, b_cnt, b_cand, b_list;
  size_t bytes() const { return b_gkey + b_tau + b_cnt + b_cand + b_list + 256; }
};

size_t a256(size_t b) { return (b + 255) / 256 * 256; }

// Workspace plan: queries in batches of qb (a multiple of 64) so that the
// batch's group minima and candidate lists stay within SEL_BUDGET bytes.
constexpr size_t SEL_BUDGET = (size_t)256 << 20;

SelPlan sel_plan(int Q, int R, int k) {
  SelPlan p;
  p.ngroups = (R + SG - 1) / SG;
  p.tasks = (p.ngroups + GCH - 1) / GCH;
  const int64_t bound = std::min<int64_t>((int64_t)R, (int64_t)(k - 1) * SG + 1);
  int cap = 1;
  while (cap < bound) cap <<= 1;
  p.cap = cap;
  const size_t per_q = (size_t)p.ngroups * 8 + (size_t)cap * 8 + 8 + 4 + 4;
  size_t qb = SEL_BUDGET / per_q;
  qb = qb / 64 * 64;
  if (qb < 64) qb = 64;
  const size_t qpad = ((size_t)Q + 63) / 64 * 64;
  if (qb > qpad) qb = qpad;
  p.qb = (int)qb;
  p.b_gkey = a256(qb * p.ngroups * 8);
  p.b_tau = a256(qb * 8);
  p.b_cnt = a256(qb * 4);
  p.b_cand = a256(qb * (size_t)cap * 8);
  p.b_list = a256(qb * 4 + 4);
  return p;
}

}  // namespace

size_t knn_select_workspace(int Q, int R, int k) {
  if (Q <= 0 || R <= 0 || k < 1 || k > R) return 0;
  return sel_plan(Q, R, k).bytes();
}

hipError_t launch_knn_select(const float *dist, int Q, int R, int k, float *out_d, int *out_i,
                             int do_sqrt, void *workspace, hipStream_t s) {
  const SelPlan p = sel_plan(Q, R, k);
  char *w = static_cast<char *>(workspace);
  uint64_t *gkey = reinterpret_cast<uint64_t *>(w);
  uint64_t *tau = reinterpret_cast<uint64_t *>(w + p.b_gkey);
  int *cnt = reinterpret_cast<int *>(w + p.b_gkey + p.b_tau);
  uint64_t *cand = reinterpret_cast<uint64_t *>(w + p.b_gkey + p.b_tau + p.b_cnt);
  int *big_n = reinterpret_cast<int *>(w + p.b_gkey + p.b_tau + p.b_cnt + p.b_cand);
  int *big_list = big_n + 1;
  for (int q0 = 0; q0 < Q; q0 += p.qb) {
    const int qn = std::min(p.qb, Q - q0);
    const int wgs = ((qn + 255) / 256) * p.tasks;  // 4 query blocks x one task each
    sel_group_min_kernel<<<wgs, 256, 0, s>>>(dist, Q, R, q0, qn, p.tasks, p.ngroups, gkey);
    sel_tau_kernel<<<(qn + 3) / 4, 256, 0, s>>>(gkey, p.ngroups, qn, k, tau, cnt, big_n);
    sel_collect_kernel<<<wgs, 256, 0, s>>>(dist, Q, R, q0, qn, p.tasks, p.ngroups, tau, cand,
                                           p.cap, cnt);
    sel_final_small_kernel<<<(qn + 3) / 4, 256, 0, s>>>(cand, cnt, p.cap, q0, qn, Q, k, out_d,
                                                        out_i, do_sqrt, big_n, big_list);
    sel_final_big_kernel<<<BIG_GRID, BIG_THREADS, 0, s>>>(cand, cnt, p.cap, q0, Q, k, out_d,
                                                          out_i, do_sqrt, big_n, big_list);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace simd2

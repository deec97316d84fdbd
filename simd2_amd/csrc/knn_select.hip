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
// Method (exact, no per-element lists):
//   1. group minima: for each query and each group of SG = 32 consecutive
//      references, the smallest key of the group (one wave = 64 queries, one
//      per lane, x 16 groups; every load is 64 consecutive floats), and for
//      k <= KF also the group's second-smallest distance;
//   2. tau[q] = the k-th smallest of the query's group minima.  The k group
//      minima at or below tau are k distinct references, so the k-th smallest
//      key of the row is <= tau: every key of the answer is <= tau.  Keys
//      strictly below tau lie in the k - 1 groups whose minimum is below tau
//      (keys are unique), so at most (k - 1) * SG + 1 keys are <= tau;
//   3. collect every key <= tau into the query's candidate list (typically
//      k to 2k entries; the bound above sizes it).  k <= KF: in the same
//      wave as tau -- a group whose minimum is <= tau and whose second
//      distance is above tau's contributes exactly its minimum, read from
//      step 1; only a group holding two or more keys <= tau (rare) is read
//      again, 32 loads by the wave -- so dist is streamed once.  Larger k: a
//      second streaming pass over dist;
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
constexpr int KF = 64;                 // largest k of the one-pass collect

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
// [0, R); the caller masks the ones past R).  The matrix is streamed once:
// nontemporal loads keep it from displacing the small arrays in L2 (7%
// faster than ordinary loads at 16384 x 16384).
__device__ __forceinline__ void load_group(const float *__restrict__ dist, int64_t Q, int R,
                                           int q, int g, float (&v)[SG]) {
#pragma unroll
  for (int u = 0; u < SG; ++u) {
    const int r = min(g * SG + u, R - 1);
    v[u] = __builtin_nontemporal_load(&dist[(int64_t)r * Q + q]);
  }
}

// Pass 1.  One wave = one query block (64 queries, one per lane) x GCH groups
// (task); writes gkey[(q - q0) * ngroups + g] and, with SECOND, the order
// key of the group's second-smallest distance to g2 (same index; all ones
// for a group of one reference).
template <bool SECOND>
__global__ __launch_bounds__(256) void sel_group_min_kernel(
    const float *__restrict__ dist, int64_t Q, int R, int q0, int qn, int tasks, int ngroups,
    uint64_t *__restrict__ gkey, uint32_t *__restrict__ g2) {
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
  uint32_t out2[GCH];
  // smallest key of the group (strict '<': the earliest index keeps a tie)
  // and the second-smallest distance (ties with the minimum included)
  auto reduce = [&](const float (&v)[SG], int g, uint64_t &o, uint32_t &o2) {
    const int rb = g * SG;
    float m = v[0], m2 = __int_as_float(0x7f800000);
    int mi = 0;
    bool two = false;
#pragma unroll
    for (int u = 1; u < SG; ++u)
      if (rb + u < R) {
        two = true;
        if (v[u] < m) {
          m2 = m;
          m = v[u];
          mi = u;
        } else if (SECOND) {
          m2 = fminf(m2, v[u]);
        }
      }
    o = key_of(m, rb + mi);
    o2 = two ? fkey(m2) : 0xffffffffu;
  };
  float va[SG], vb[SG];
  load_group(dist, Q, R, q, g0, va);
#pragma unroll
  for (int gi = 0; gi < GCH; gi += 2) {
    // two register sets: the next group's loads are in flight while one is reduced
    if (g0 + gi + 1 < g1) load_group(dist, Q, R, q, g0 + gi + 1, vb);
    if (g0 + gi < g1) reduce(va, g0 + gi, out[gi], out2[gi]);
    if (g0 + gi + 2 < g1) load_group(dist, Q, R, q, g0 + gi + 2, va);
    if (g0 + gi + 1 < g1) reduce(vb, g0 + gi + 1, out[gi + 1], out2[gi + 1]);
  }
  if (!live) return;
  uint64_t *dst = gkey + (int64_t)ql * ngroups;
#pragma unroll
  for (int gi = 0; gi < GCH; ++gi)
    if (g0 + gi < g1) dst[g0 + gi] = out[gi];
  if (SECOND) {
    uint32_t *dst2 = g2 + (int64_t)ql * ngroups;
#pragma unroll
    for (int gi = 0; gi < GCH; ++gi)
      if (g0 + gi < g1) dst2[g0 + gi] = out2[gi];
  }
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

// tau of one query, by its wave: the k-th smallest group-minimum key
// (KEY_MAX when there are fewer than k groups), by bisection on the distance
// bits and then, among minima with that distance, on the index.
__device__ __forceinline__ uint64_t wave_tau(const uint64_t *t, int ngroups, int k, int lane) {
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
  return res;
}

// Pass 2 for k > KF.  One wave per query: tau, and the query's candidate
// count cleared (and, in the first wave, the count of queries for the
// workgroup sort).
__global__ __launch_bounds__(256) void sel_tau_kernel(const uint64_t *__restrict__ gkey,
                                                      int ngroups, int qn, int k, uint64_t *tau,
                                                      int *cnt, int *big_n) {
  const int lane = tid_x() % 64;
  const int ql = bid_x() * 4 + tid_x() / 64;
  if (ql == 0 && lane == 0) *big_n = 0;
  if (ql >= qn) return;
  const uint64_t res = wave_tau(gkey + (int64_t)ql * ngroups, ngroups, k, lane);
  if (lane == 0) {
    tau[ql] = res;
    cnt[ql] = 0;
  }
}

// Passes 2 and 3 for k <= KF, one wave per query: tau, then the candidates.
// A group whose minimum key is <= tau holds no other key <= tau when its
// second distance's order key is above tau's distance: its minimum is
// appended as read in pass 1.  Otherwise lanes 0 .. SG-1 read the group's
// distances again and append every key <= tau.  Positions come from wave
// ballots in group order, so the list is written without atomics.
__global__ __launch_bounds__(256) void sel_tau_collect_kernel(
    const float *__restrict__ dist, int64_t Q, int R, int q0, const uint64_t *__restrict__ gkey,
    const uint32_t *__restrict__ g2, int ngroups, int qn, int k, uint64_t *__restrict__ cand,
    int cap, int *cnt, int *big_n) {
  const int lane = tid_x() % 64;
  const int ql = bid_x() * 4 + tid_x() / 64;
  if (ql == 0 && lane == 0) *big_n = 0;
  if (ql >= qn) return;
  const uint64_t *t = gkey + (int64_t)ql * ngroups;
  const uint32_t *t2 = g2 + (int64_t)ql * ngroups;
  const uint64_t res = wave_tau(t, ngroups, k, lane);
  const uint32_t th = (uint32_t)(res >> 32);
  const int q = q0 + ql;
  const uint64_t below = (1ull << lane) - 1ull;  // lanes under this one
  uint64_t *list = cand + (int64_t)ql * cap;
  int base = 0;
  for (int j0 = 0; j0 < ngroups; j0 += 64) {
    const int j = j0 + lane;
    const uint64_t key = j < ngroups ? t[j] : KEY_MAX;
    const bool qual = j < ngroups && key <= res;
    const bool again = qual && t2[j] <= th;
    const uint64_t bd = __ballot(qual && !again);
    if (qual && !again) {
      const int pos = base + (int)__popcll(bd & below);
      if (pos < cap) list[pos] = key;
    }
    base += (int)__popcll(bd);
    uint64_t br = __ballot(again);
    while (br) {
      const int g = j0 + (int)__builtin_ctzll(br);
      br &= br - 1;
      const int r = g * SG + lane;
      uint64_t kk = KEY_MAX;
      if (lane < SG && r < R) kk = key_of(dist[(int64_t)r * Q + q], r);
      const bool in = lane < SG && r < R && kk <= res;
      const uint64_t bi = __ballot(in);
      if (in) {
        const int pos = base + (int)__popcll(bi & below);
        if (pos < cap) list[pos] = kk;
      }
      base += (int)__popcll(bi);
    }
  }
  if (lane == 0) cnt[ql] = base;
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
  bool one_pass;
  size_t b_gkey, b_g2, b_tau, b_cnt, b_cand, b_list;
  size_t bytes() const { return b_gkey + b_g2 + b_tau + b_cnt + b_cand + b_list + 256; }
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
  // one pass when k is small and the row has enough groups that tau is a
  // real bound (with fewer than 4k groups most of them would be re-read)
  p.one_pass = k <= KF && p.ngroups >= 4 * k;
  const size_t per_q =
      (size_t)p.ngroups * (p.one_pass ? 12 : 8) + (size_t)cap * 8 + 8 + 4 + 4;
  size_t qb = SEL_BUDGET / per_q;
  qb = qb / 64 * 64;
  if (qb < 64) qb = 64;
  const size_t qpad = ((size_t)Q + 63) / 64 * 64;
  if (qb > qpad) qb = qpad;
  p.qb = (int)qb;
  p.b_gkey = a256(qb * p.ngroups * 8);
  p.b_g2 = p.one_pass ? a256(qb * p.ngroups * 4) : 0;
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
  w += p.b_gkey;
  uint32_t *g2 = reinterpret_cast<uint32_t *>(w);
  w += p.b_g2;
  uint64_t *tau = reinterpret_cast<uint64_t *>(w);
  w += p.b_tau;
  int *cnt = reinterpret_cast<int *>(w);
  w += p.b_cnt;
  uint64_t *cand = reinterpret_cast<uint64_t *>(w);
  w += p.b_cand;
  int *big_n = reinterpret_cast<int *>(w);
  int *big_list = big_n + 1;
  for (int q0 = 0; q0 < Q; q0 += p.qb) {
    const int qn = std::min(p.qb, Q - q0);
    const int wgs = ((qn + 255) / 256) * p.tasks;  // 4 query blocks x one task each
    if (p.one_pass) {
      sel_group_min_kernel<true><<<wgs, 256, 0, s>>>(dist, Q, R, q0, qn, p.tasks, p.ngroups,
                                                     gkey, g2);
      sel_tau_collect_kernel<<<(qn + 3) / 4, 256, 0, s>>>(dist, Q, R, q0, gkey, g2, p.ngroups,
                                                          qn, k, cand, p.cap, cnt, big_n);
    } else {
      sel_group_min_kernel<false><<<wgs, 256, 0, s>>>(dist, Q, R, q0, qn, p.tasks, p.ngroups,
                                                      gkey, nullptr);
      sel_tau_kernel<<<(qn + 3) / 4, 256, 0, s>>>(gkey, p.ngroups, qn, k, tau, cnt, big_n);
      sel_collect_kernel<<<wgs, 256, 0, s>>>(dist, Q, R, q0, qn, p.tasks, p.ngroups, tau, cand,
                                             p.cap, cnt);
    }
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

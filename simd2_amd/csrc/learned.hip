// The learned schedule's device side (gfx950): graph features, breadth-first
// searches, and a 24 -> 64 -> 64 -> 2 MLP with forward, backward and Adam
// kernels.  All reductions run in a fixed order, so every result is
// deterministic.
#include <float.h>
#include <math.h>

#include <algorithm>

#include "internal.hpp"

namespace simd2 {

namespace {

constexpr int NF = 24, NH = 64, NO = 2;
constexpr int OFF_W1 = 0, OFF_B1 = OFF_W1 + NH * NF, OFF_W2 = OFF_B1 + NH,
              OFF_B2 = OFF_W2 + NH * NH, OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NO * NH,
              NPARAM = OFF_B3 + NO;
constexpr int GRAD_CHUNK = 16;  // samples per gradient workgroup
constexpr int RS = 8;           // row statistics per vertex
constexpr int BFS_T = 256;      // threads of a breadth-first search
constexpr int BFS_W = 4;        // adjacency words per thread (n <= 32768)
constexpr int BFS_U = 32;       // frontier vertices expanded together
constexpr int NSRC = 8;         // sampled search sources

// An entry is an edge when it is off the diagonal and a finite weight (not
// the (+)-identity, nor the FLT_MAX "no edge" of the reference apps).
__device__ __forceinline__ bool is_edge(float v, int i, int j) {
  return i != j && fabsf(v) < FLT_MAX;
}

// Per row i: out-degree, sum |w|, sum w^2, min |w|, max |w|, sum |i-j|,
// max |i-j|, edges with |i-j| <= 16 (doubles, fixed-order tree reduction).
// When adj is set, the row is also packed there (bit (j & 31) of word j / 32:
// edge (i, j)), one ballot per 64 entries.
__global__ __launch_bounds__(256) void row_stats_kernel(const float *X, int64_t ld, int n,
                                                        double *stats, uint32_t *adj, int nw,
                                                        int *indeg, int *sym) {
  __shared__ double sh[RS][256];
  const int i = bid_x(), t = tid_x();
  // clears the in-degree and symmetry counts the next two kernels add into
  // (a launch each otherwise, and launches, not work, bound small graphs)
  if (t == 0) indeg[i] = 0;
  if (i == 0 && t < 2) sym[t] = 0;
  const int lane = t % 64;
  double deg = 0, ws = 0, wq = 0, wmin = 1e300, wmax = 0, ds = 0, dmax = 0, near = 0;
  for (int j0 = 0; j0 < n; j0 += 256) {
    const int j = j0 + t;
    const float v = j < n ? X[(int64_t)i * ld + j] : 0.0f;
    const bool edge = j < n && is_edge(v, i, j);
    if (adj) {
      const uint64_t m = __ballot(edge);
      const int w0 = (j0 + (t - lane)) / 32;
      if (lane == 0 && w0 < nw) adj[(int64_t)i * nw + w0] = (uint32_t)m;
      if (lane == 32 && w0 + 1 < nw) adj[(int64_t)i * nw + w0 + 1] = (uint32_t)(m >> 32);
    }
    if (edge) {
      const double w = v, aw = fabs(w), d = abs(i - j);
      deg += 1;
      ws += aw;
      wq += w * w;
      wmin = fmin(wmin, aw);
      wmax = fmax(wmax, aw);
      ds += d;
      dmax = fmax(dmax, d);
      near += d <= 16 ? 1 : 0;
    }
  }
  sh[0][t] = deg; sh[1][t] = ws; sh[2][t] = wq; sh[3][t] = wmin;
  sh[4][t] = wmax; sh[5][t] = ds; sh[6][t] = dmax; sh[7][t] = near;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
      for (int k = 0; k < RS; ++k) {
        const double a = sh[k][t], b = sh[k][t + s];
        sh[k][t] = k == 3 ? fmin(a, b) : (k == 4 || k == 6) ? fmax(a, b) : a + b;
      }
    }
    __syncthreads();
  }
  if (t < RS) stats[(int64_t)i * RS + t] = sh[t][0];
}

// In-degrees: workgroup (column block, row chunk) counts its edges.
__global__ __launch_bounds__(256) void in_degree_kernel(const float *X, int64_t ld, int n,
                                                        int rows_per_chunk, int *indeg) {
  const int j = bid_x() * 256 + tid_x();
  const int r0 = bid_y() * rows_per_chunk;
  if (j >= n) return;
  int c = 0;
  const int r1 = min(n, r0 + rows_per_chunk);
  for (int i = r0; i < r1; ++i) c += is_edge(X[(int64_t)i * ld + j], i, j) ? 1 : 0;
  if (c) atomicAdd(indeg + j, c);
}

// Reverse-edge count over the sampled rows i = k * step, from the packed
// adjacency: out[0] += edges (i, j), out[1] += those with (j, i) an edge too.
// Thread t takes words t, t + 256, ... of row i and tests bit i of each
// listed row j.
__global__ __launch_bounds__(256) void symmetry_kernel(const uint32_t *__restrict__ adj, int n,
                                                       int nw, int step, int *out) {
  const int i = bid_x() * step;
  if (i >= n) return;
  int e = 0, s = 0;
  for (int w = tid_x(); w < nw; w += 256) {
    uint32_t bits = adj[(int64_t)i * nw + w];
    e += __popc(bits);
    for (; bits; bits &= bits - 1) {
      const int j = w * 32 + __ffs(bits) - 1;
      s += (adj[(int64_t)j * nw + i / 32] >> (i % 32)) & 1u;
    }
  }
  if (e) atomicAdd(out, e);
  if (s) atomicAdd(out + 1, s);
}

// Bit-packed adjacency: bit (j & 31) of word adj[i][j / 32] is edge (i, j).
// A wave turns 64 entries of a row into two words with one ballot.
__global__ __launch_bounds__(256) void pack_adj_kernel(const float *X, int64_t ld, int n, int nw,
                                                       uint32_t *adj) {
  const int i = bid_y();
  const int j = bid_x() * 256 + tid_x();
  const bool e = j < n && is_edge(X[(int64_t)i * ld + j], i, j);
  const uint64_t m = __ballot(e);
  const int lane = tid_x() % 64;
  const int w0 = (bid_x() * 256 + (tid_x() - lane)) / 32;
  if (lane == 0 && w0 < nw) adj[(int64_t)i * nw + w0] = (uint32_t)m;
  if (lane == 32 && w0 + 1 < nw) adj[(int64_t)i * nw + w0 + 1] = (uint32_t)(m >> 32);
}

// Breadth-first search from src over the packed adjacency, by one
// workgroup.  The frontier is a vertex list in global memory (cur / nxt, n
// ints each); thread t owns adjacency words t + 256 q.  Writes depth (the
// last level reached), reached (vertex count) and the lowest-numbered vertex
// of the last level to res[0..2] and, when level is set, each vertex's level
// (-1: unreached).  Every level adds at least one vertex or ends the search,
// so it stops within n levels.  s_cnt: 3 ints of LDS.
__device__ void bfs(const uint32_t *adj, int n, int nw, int src, int *cur, int *nxt, int *level,
                    int *res, uint32_t *visited, uint32_t *red, int *s_cnt) {
  const int t = tid_x();
  // Short adjacency rows (nw <= 128 words, n <= 4096) would leave most
  // threads idle: the threads form G groups of wpg words, each group expands
  // its share of the frontier, and the groups' words are OR-ed through LDS.
  const int wpg = nw <= 64 ? 64 : (nw <= 128 ? 128 : BFS_T);
  const int G = BFS_T / wpg, g = t / wpg, tw = t % wpg;
  for (int w = t; w < nw; w += BFS_T) visited[w] = 0;
  if (level)
    for (int v = t; v < n; v += BFS_T) level[v] = v == src ? 0 : -1;
  __syncthreads();
  if (t == 0) {
    visited[src / 32] |= 1u << (src % 32);
    cur[0] = src;
    s_cnt[0] = 1;      // frontier size
    s_cnt[1] = 0;      // next frontier size
    s_cnt[2] = 0x7fffffff;  // lowest vertex of the next frontier
  }
  __syncthreads();
  int depth = 0, reached = 1, far = src;
  for (;;) {
    const int cnt = s_cnt[0];
    uint32_t acc[BFS_W] = {0u, 0u, 0u, 0u};
    // BFS_U frontier vertices per group at a time, so that many row loads
    // are in flight
    for (int base = 0; base < cnt; base += BFS_U * G) {
      int u[BFS_U];
#pragma unroll
      for (int k = 0; k < BFS_U; ++k) {
        const int at = base + g + G * k;
        u[k] = at < cnt ? cur[at] : -1;
      }
#pragma unroll
      for (int q = 0; q < BFS_W; ++q) {
        const int w = tw + BFS_T * q;
        if (w < nw) {
#pragma unroll
          for (int k = 0; k < BFS_U; ++k)
            if (u[k] >= 0) acc[q] |= adj[(int64_t)u[k] * nw + w];
        }
      }
    }
    if (G > 1) {  // nw <= wpg < BFS_T: one word per thread
      red[t] = acc[0];
      __syncthreads();
      if (g == 0)
        for (int h = 1; h < G; ++h) acc[0] |= red[h * wpg + tw];
    }
#pragma unroll
    for (int q = 0; q < BFS_W; ++q) {
      const int w = tw + BFS_T * q;
      if (g == 0 && w < nw) {
        uint32_t fresh = acc[q] & ~visited[w];
        if (fresh) {
          visited[w] |= fresh;
          atomicMin(&s_cnt[2], w * 32 + __ffs(fresh) - 1);
          int at = atomicAdd(&s_cnt[1], __popc(fresh));
          while (fresh) {
            const int b = __ffs(fresh) - 1;
            fresh &= fresh - 1;
            const int v = w * 32 + b;
            nxt[at++] = v;
            if (level) level[v] = depth + 1;
          }
        }
      }
    }
    __syncthreads();
    const int ncnt = s_cnt[1];
    const int nfar = s_cnt[2];
    __syncthreads();
    if (ncnt == 0) break;
    ++depth;
    reached += ncnt;
    far = nfar;
    int *tmp = cur;
    cur = nxt;
    nxt = tmp;
    if (t == 0) {
      s_cnt[0] = ncnt;
      s_cnt[1] = 0;
      s_cnt[2] = 0x7fffffff;
    }
    __syncthreads();
  }
  if (t == 0) {
    res[0] = depth;
    res[1] = reached;
    res[2] = far;
  }
  __syncthreads();
}

// Searches of graphs with n <= SB_N: the whole packed adjacency (128 KB), the
// frontier lists and the visited set live in LDS, so a level costs two
// barriers and LDS latency instead of two dependent global loads.  The
// frontier's rows are OR-ed into a level bitset by LDS atomics (thread t owns
// word t % 32 for every eighth frontier vertex); wave 0 then takes the fresh
// bits, numbers them by a prefix sum over the words and writes the next
// list.  Same results as bfs(): depth, reached count, lowest vertex of the
// last level, levels.  The LDS (136 KB) allows one workgroup per CU, so it
// takes 1024 threads: 32 frontier vertices expanded at once.
constexpr int SB_NW = 32;
constexpr int SB_N = SB_NW * 32;
constexpr int SB_T = 1024;

struct SmallBfsSmem {
  alignas(16) uint32_t adj[SB_N * SB_NW];
  int list[2][SB_N];
  uint32_t visited[SB_NW];
  uint32_t nextb[SB_NW];
  int cnt, far;
};

__device__ void load_adj_small(const uint32_t *adj, int n, int nw, SmallBfsSmem &sm) {
  const int words = n * nw;
  int done = 0;
  if ((reinterpret_cast<uintptr_t>(adj) & 15) == 0) {  // 16-byte copies
    const int w4 = words / 4;
    const uint4 *src = reinterpret_cast<const uint4 *>(adj);
    uint4 *dst = reinterpret_cast<uint4 *>(sm.adj);
    for (int i = tid_x(); i < w4; i += SB_T) dst[i] = src[i];
    done = 4 * w4;
  }
  for (int i = done + tid_x(); i < words; i += SB_T) sm.adj[i] = adj[i];
  __syncthreads();
}

__device__ void bfs_small(int n, int nw, int src, int *level, int *res, SmallBfsSmem &sm) {
  const int t = tid_x();
  constexpr int G = SB_T / SB_NW;
  const int g = t / SB_NW, w = t % SB_NW;
  if (t < SB_NW) {
    sm.visited[t] = 0;
    sm.nextb[t] = 0;
  }
  if (level)
    for (int v = t; v < n; v += SB_T) level[v] = v == src ? 0 : -1;
  __syncthreads();
  if (t == 0) {
    sm.visited[src / 32] |= 1u << (src % 32);
    sm.list[0][0] = src;
    sm.cnt = 1;
  }
  __syncthreads();
  int depth = 0, reached = 1, far = src, p = 0;
  for (;;) {
    const int cnt = sm.cnt;
    const int *cur = sm.list[p];
    uint32_t acc = 0;
    if (w < nw)
      for (int i = g; i < cnt; i += G) acc |= sm.adj[cur[i] * nw + w];
    if (acc) atomicOr(&sm.nextb[w], acc);
    __syncthreads();
    if (t < 64) {
      const int lane = t;
      uint32_t fresh = 0;
      if (lane < nw) {
        fresh = sm.nextb[lane] & ~sm.visited[lane];
        sm.visited[lane] |= fresh;
        sm.nextb[lane] = 0;
      }
      const int c = __popc(fresh);
      int incl = c;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int o = __shfl_up(incl, off);
        if (lane >= off) incl += o;
      }
      int at = incl - c;
      int *nxt = sm.list[p ^ 1];
      for (uint32_t f = fresh; f; f &= f - 1) {
        const int v = lane * 32 + __ffs(f) - 1;
        nxt[at++] = v;
        if (level) level[v] = depth + 1;
      }
      const uint64_t nz = __ballot(fresh != 0);
      if (lane == 63) sm.cnt = incl;
      if (nz && lane == __ffsll((unsigned long long)nz) - 1) sm.far = lane * 32 + __ffs(fresh) - 1;
    }
    __syncthreads();
    const int ncnt = sm.cnt;
    if (ncnt == 0) break;
    ++depth;
    reached += ncnt;
    far = sm.far;
    p ^= 1;
  }
  if (t == 0) {
    res[0] = depth;
    res[1] = reached;
    res[2] = far;
  }
  __syncthreads();
}

__global__ __launch_bounds__(SB_T) void bfs_sources_small_kernel(const uint32_t *adj, int n,
                                                                  int nw, int *res) {
  __shared__ SmallBfsSmem sm;
  load_adj_small(adj, n, nw, sm);
  const int src = (int)((int64_t)bid_x() * n / NSRC);
  bfs_small(n, nw, src, nullptr, res + 3 * bid_x(), sm);
}

__global__ __launch_bounds__(SB_T) void bfs_levels_small_kernel(const uint32_t *adj, int n,
                                                                 int nw, int *level, int *res,
                                                                 const int *far) {
  __shared__ SmallBfsSmem sm;
  load_adj_small(adj, n, nw, sm);
  int src;
  if (far) {
    src = *far;
  } else {
    bfs_small(n, nw, 0, nullptr, res, sm);
    src = res[2];
  }
  bfs_small(n, nw, src, level, res + 3, sm);
}

// Feature searches: workgroup k searches from vertex k * n / 8.
__global__ __launch_bounds__(BFS_T) void bfs_sources_kernel(const uint32_t *adj, int n, int nw,
                                                            int *lists, int *res) {
  __shared__ uint32_t visited[BFS_T * BFS_W];
  __shared__ uint32_t red[BFS_T];
  __shared__ int s_cnt[3];
  const int k = bid_x();
  const int src = (int)((int64_t)k * n / NSRC);
  int *cur = lists + (int64_t)k * 2 * n;
  bfs(adj, n, nw, src, cur, cur + n, nullptr, res + 3 * k, visited, red, s_cnt);
}

// Vertex levels for the order: the search from the lowest-numbered vertex of
// the last level of a search from vertex 0 (a pseudo-peripheral vertex).
// far: that vertex when already known (the feature searches start one at
// vertex 0), else null and the first search runs here.
__global__ __launch_bounds__(BFS_T) void bfs_levels_kernel(const uint32_t *adj, int n, int nw,
                                                           int *lists, int *level, int *res,
                                                           const int *far) {
  __shared__ uint32_t visited[BFS_T * BFS_W];
  __shared__ uint32_t red[BFS_T];
  __shared__ int s_cnt[3];
  int src;
  if (far) {
    src = *far;
  } else {
    bfs(adj, n, nw, 0, lists, lists + n, nullptr, res, visited, red, s_cnt);
    src = res[2];
  }
  bfs(adj, n, nw, src, lists, lists + n, level, res + 3, visited, red, s_cnt);
}

// The 24 features (one workgroup; fixed-order reductions in double).
constexpr int NRED = 13;
// The MLP's layers on one normalised sample xs (LDS): thread j < NH computes
// hidden unit j; every thread of the workgroup calls it (it has barriers).
__device__ __forceinline__ void mlp_layers(const float *P, const float *xs, float *h1, float *h2,
                                           float *y, const float *norm, int j) {
  if (j < NH) {
    float a = P[OFF_B1 + j];
    for (int f = 0; f < NF; ++f) a += P[OFF_W1 + j * NF + f] * xs[f];
    h1[j] = a > 0.f ? a : 0.f;
  }
  __syncthreads();
  if (j < NH) {
    float a = P[OFF_B2 + j];
    for (int k = 0; k < NH; ++k) a += P[OFF_W2 + j * NH + k] * h1[k];
    h2[j] = a > 0.f ? a : 0.f;
  }
  __syncthreads();
  if (j < NO) {
    float o = P[OFF_B3 + j];
    for (int k = 0; k < NH; ++k) o += P[OFF_W3 + j * NH + k] * h2[k];
    y[j] = norm ? o * norm[2 * NF + j] + norm[2 * NF + NO + j] : o;
  }
}

__global__ __launch_bounds__(256) void features_kernel(const double *rstats, const int *indeg,
                                                       const int *sym, const int *bfs_res, int n,
                                                       int max_based, int epi, float *f,
                                                       const float *P, float *y,
                                                       const float *norm) {
  __shared__ double sh[NRED][256];
  __shared__ float fx[NF], h1[NH], h2[NH];
  const int t = tid_x();
  double v[NRED] = {0, 0, 0, 0, 0, 0, 1e300, 0, 0, 0, 0, 0, 0};
  for (int i = t; i < n; i += 256) {
    const double *r = rstats + (int64_t)i * RS;
    const double din = indeg[i];
    v[0] += r[0];                      // edges
    v[1] += r[0] * r[0];               // sum out-degree^2
    v[2] = fmax(v[2], r[0]);           // max out-degree
    v[3] += r[0] == 0 ? 1 : 0;         // vertices without out-edges
    v[4] += r[1];                      // sum |w|
    v[5] += r[2];                      // sum w^2
    if (r[0] > 0) v[6] = fmin(v[6], r[3]);  // min |w|
    v[7] = fmax(v[7], r[4]);           // max |w|
    v[8] += r[5];                      // sum |i - j|
    v[9] = fmax(v[9], r[6]);           // max |i - j|
    v[10] += r[7];                     // edges with |i - j| <= 16
    v[11] += din == 0 ? 1 : 0;         // vertices without in-edges
    v[12] += din * din;                // sum in-degree^2
  }
  for (int k = 0; k < NRED; ++k) sh[k][t] = v[k];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s)
      for (int k = 0; k < NRED; ++k) {
        const double a = sh[k][t], b = sh[k][t + s];
        sh[k][t] = (k == 2 || k == 7 || k == 9) ? fmax(a, b) : k == 6 ? fmin(a, b) : a + b;
      }
    __syncthreads();
  }
  if (t == 0) {
    const double N = n, m = sh[0][0];
    const double mean = m / N;
    const double sd_out = sqrt(fmax(sh[1][0] / N - mean * mean, 0.0));
    const double sd_in = sqrt(fmax(sh[12][0] / N - mean * mean, 0.0));
    const double density = n > 1 ? m / (N * (N - 1)) : 0.0;
    const double wmax = sh[7][0];
    const double wmean = m > 0 ? sh[4][0] / m : 0.0;
    const double wsd = m > 0 ? sqrt(fmax(sh[5][0] / m - wmean * wmean, 0.0)) : 0.0;
    double dmean = 0, dmx = 0, reach = 0, width = 0;
    for (int k = 0; k < NSRC; ++k) {
      const double d = bfs_res[3 * k], r = bfs_res[3 * k + 1];
      dmean += d / NSRC;
      dmx = fmax(dmx, d);
      reach += r / N / NSRC;
      width += r / (d + 1) / N / NSRC;
  }
    double out[NF];
    out[0] = log2(N) / 16;
    out[1] = density;
    out[2] = (log10(density + 1e-9) + 9) / 9;
    out[3] = log2(1 + mean) / 16;
    out[4] = sd_out / (1 + mean);
    out[5] = log2(1 + sh[2][0]) / 16;
    out[6] = sh[3][0] / N;
    out[7] = sh[11][0] / N;
    out[8] = sd_in / (1 + mean);
    out[9] = wmax > 0 ? wmean / wmax : 0;
    out[10] = m > 0 ? fmin(wsd / (wmean + 1e-12), 10.0) : 0;
    out[11] = (m > 0 && wmax > 0) ? sh[6][0] / wmax : 0;
    out[12] = log10(1 + wmax) / 10;
    out[13] = sym[0] > 0 ? (double)sym[1] / sym[0] : 0;
    out[14] = m > 0 ? sh[8][0] / m / N : 0;
    out[15] = m > 0 ? sh[9][0] / N : 0;
    out[16] = m > 0 ? sh[10][0] / m : 0;
    out[17] = log2(1 + dmean) / 16;
    out[18] = log2(1 + dmx) / 16;
    out[19] = reach;
    out[20] = width;
    out[21] = max_based ? 1 : 0;
    out[22] = epi ? 1 : 0;
    out[23] = (ceil(log2(fmax(dmx, 1.0))) + 1) / 16;
    for (int k = 0; k < NF; ++k) {
      f[k] = (float)out[k];
      fx[k] = (float)out[k];
    }
  }
  // With a model: its prediction from the same values, as mlp_forward_kernel
  // would compute it from f (one launch fewer).
  if (!P) return;
  __syncthreads();
  if (t < NF && norm) fx[t] = (fx[t] - norm[t]) * norm[NF + t];
  __syncthreads();
  mlp_layers(P, fx, h1, h2, y, norm, t);
}

// ---- MLP -------------------------------------------------------------------

// y = MLP(x) for one sample per 64-thread workgroup.  norm (optional):
// feature mean[24], scale[24], output scale[2], offset[2].
__global__ __launch_bounds__(64) void mlp_forward_kernel(const float *P, const float *x, int B,
                                                         float *y, const float *norm) {
  __shared__ float xs[NF], h1[NH], h2[NH];
  const int b = bid_x(), j = tid_x();
  if (b >= B) return;
  if (j < NF) {
    const float v = x[(int64_t)b * NF + j];
    xs[j] = norm ? (v - norm[j]) * norm[NF + j] : v;
  }
  __syncthreads();
  mlp_layers(P, xs, h1, h2, y + (int64_t)b * NO, norm, j);
}

// Gradient of L = (1/B) sum_b sum_o w_o (y_bo - t_bo)^2 over GRAD_CHUNK
// samples per 64-thread workgroup: thread j accumulates the gradient of
// hidden unit j's parameters (W1 row j, b1[j], W2 row j, b2[j], W3[.][j],
// and b3[j] for j < 2) in registers, sample after sample, and writes the
// chunk's partial gradient (and loss, last entry) to partial[chunk].
__global__ __launch_bounds__(64) void mlp_grad_kernel(const float *P, const float *x,
                                                      const float *t, const float *w, int B,
                                                      float *partial) {
  __shared__ float xs[NF], h1[NH], h2[NH], dh2[NH], dh1[NH], dy[NO], df[NO];
  const int c = bid_x(), j = tid_x();
  float g1[NF], g2[NH], g3[NO], gb1 = 0.f, gb2 = 0.f, gb3 = 0.f, loss = 0.f;
  _Pragma("unroll") for (int f = 0; f < NF; ++f) g1[f] = 0.f;
  _Pragma("unroll") for (int k = 0; k < NH; ++k) g2[k] = 0.f;
  g3[0] = g3[1] = 0.f;
  const float invB = 1.0f / (float)B;
  const int s1 = min(B, (c + 1) * GRAD_CHUNK);
  for (int s = c * GRAD_CHUNK; s < s1; ++s) {
    if (j < NF) xs[j] = x[(int64_t)s * NF + j];
    __syncthreads();
    float a = P[OFF_B1 + j];
    for (int f = 0; f < NF; ++f) a += P[OFF_W1 + j * NF + f] * xs[f];
    h1[j] = a > 0.f ? a : 0.f;
    __syncthreads();
    a = P[OFF_B2 + j];
    for (int k = 0; k < NH; ++k) a += P[OFF_W2 + j * NH + k] * h1[k];
    h2[j] = a > 0.f ? a : 0.f;
    __syncthreads();
    if (j < NO) {
      float o = P[OFF_B3 + j];
      for (int k = 0; k < NH; ++k) o += P[OFF_W3 + j * NH + k] * h2[k];
      const float d = o - t[(int64_t)s * NO + j];
      df[j] = d;
      dy[j] = 2.f * w[j] * d * invB;
    }
    __syncthreads();
    if (j == 0) loss += (w[0] * df[0] * df[0] + w[1] * df[1] * df[1]) * invB;
    const float h2j = h2[j];
    g3[0] += dy[0] * h2j;
    g3[1] += dy[1] * h2j;
    if (j < NO) gb3 += dy[j];
    const float back2 = dy[0] * P[OFF_W3 + j] + dy[1] * P[OFF_W3 + NH + j];
    const float d2 = h2j > 0.f ? back2 : 0.f;
    dh2[j] = d2;
    __syncthreads();
    float back1 = 0.f;
    for (int jj = 0; jj < NH; ++jj) back1 += dh2[jj] * P[OFF_W2 + jj * NH + j];
    dh1[j] = h1[j] > 0.f ? back1 : 0.f;
    _Pragma("unroll") for (int k = 0; k < NH; ++k) g2[k] += d2 * h1[k];
    gb2 += d2;
    __syncthreads();
    const float d1 = dh1[j];
    _Pragma("unroll") for (int f = 0; f < NF; ++f) g1[f] += d1 * xs[f];
    gb1 += d1;
    __syncthreads();
  }
  float *out = partial + (int64_t)c * (NPARAM + 1);
  _Pragma("unroll") for (int f = 0; f < NF; ++f) out[OFF_W1 + j * NF + f] = g1[f];
  out[OFF_B1 + j] = gb1;
  _Pragma("unroll") for (int k = 0; k < NH; ++k) out[OFF_W2 + j * NH + k] = g2[k];
  out[OFF_B2 + j] = gb2;
  out[OFF_W3 + j] = g3[0];
  out[OFF_W3 + NH + j] = g3[1];
  if (j < NO) out[OFF_B3 + j] = gb3;
  if (j == 0) out[NPARAM] = loss;
}

// grad[p] = sum over chunks, in chunk order (p == NPARAM: the loss).
__global__ __launch_bounds__(256) void mlp_reduce_kernel(const float *partial, int chunks,
                                                         float *grad, float *loss) {
  const int p = bid_x() * 256 + tid_x();
  if (p > NPARAM) return;
  float s = 0.f;
  for (int c = 0; c < chunks; ++c) s += partial[(int64_t)c * (NPARAM + 1) + p];
  if (p < NPARAM) {
    if (grad) grad[p] = s;
  } else if (loss) {
    *loss = s;
  }
}

// Adam with bias corrections c1 = 1 / (1 - beta1^t), c2 = 1 / (1 - beta2^t).
__global__ __launch_bounds__(256) void adam_kernel(float *P, float *m, float *v, const float *g,
                                                   float lr, float c1, float c2) {
  const int p = bid_x() * 256 + tid_x();
  if (p >= NPARAM) return;
  const float gp = g[p];
  const float mp = 0.9f * m[p] + 0.1f * gp;
  const float vp = 0.999f * v[p] + 0.001f * gp * gp;
  m[p] = mp;
  v[p] = vp;
  P[p] -= lr * (mp * c1) / (sqrtf(vp * c2) + 1e-8f);
}

struct FeatLayout {
  size_t rstats, indeg, sym, adj, lists, res, level, total;
};

FeatLayout feat_layout(int n) {
  FeatLayout L;
  size_t off = 0;
  auto take = [&](size_t b) {
    const size_t at = off;
    off += (b + 255) / 256 * 256;
    return at;
  };
  const int nw = (n + 31) / 32;
  L.rstats = take((size_t)n * RS * sizeof(double));
  L.indeg = take((size_t)n * sizeof(int));
  L.sym = take(2 * sizeof(int));
  L.adj = take((size_t)n * nw * sizeof(uint32_t));
  L.lists = take((size_t)NSRC * 2 * n * sizeof(int));
  L.res = take((3 * NSRC + 6) * sizeof(int));
  L.level = take((size_t)n * sizeof(int));
  L.total = off;
  return L;
}

hipError_t pack_adjacency(const float *X, int64_t ld, int n, uint32_t *adj, hipStream_t s) {
  const int nw = (n + 31) / 32;
  pack_adj_kernel<<<dim3((n + 255) / 256, n), dim3(256), 0, s>>>(X, ld, n, nw, adj);
  return hipGetLastError();
}

}  // namespace

size_t graph_features_workspace(int n) { return n > 0 ? feat_layout(n).total : 0; }

hipError_t launch_graph_features(int op, int n, const float *X, int64_t ld, int epi, float *f,
                                 void *ws, hipStream_t s, const float *model, float *pred) {
  const FeatLayout L = feat_layout(n);
  char *w = static_cast<char *>(ws);
  double *rstats = reinterpret_cast<double *>(w + L.rstats);
  int *indeg = reinterpret_cast<int *>(w + L.indeg);
  int *sym = reinterpret_cast<int *>(w + L.sym);
  uint32_t *adj = reinterpret_cast<uint32_t *>(w + L.adj);
  int *lists = reinterpret_cast<int *>(w + L.lists);
  int *res = reinterpret_cast<int *>(w + L.res);
  const int nw = (n + 31) / 32;
  row_stats_kernel<<<dim3(n), dim3(256), 0, s>>>(X, ld, n, rstats, adj, nw, indeg, sym);
  // about 2048 workgroups whatever n: each thread walks `chunk` rows of one column
  const int chunk = std::max(8, (int)(((int64_t)n * n) / (256 * 2048)));
  in_degree_kernel<<<dim3((n + 255) / 256, (n + chunk - 1) / chunk), dim3(256), 0, s>>>(
      X, ld, n, chunk, indeg);
  const int step = std::max(1, n / 256);
  symmetry_kernel<<<dim3((n + step - 1) / step), dim3(256), 0, s>>>(adj, n, nw, step, sym);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (nw <= SB_NW)
    bfs_sources_small_kernel<<<dim3(NSRC), dim3(SB_T), 0, s>>>(adj, n, nw, res);
  else
    bfs_sources_kernel<<<dim3(NSRC), dim3(BFS_T), 0, s>>>(adj, n, nw, lists, res);
  const int max_based = (op == 2 || op == 3 || op == 4) ? 1 : 0;
  features_kernel<<<dim3(1), dim3(256), 0, s>>>(rstats, indeg, sym, res, n, max_based, epi, f,
                                                 model, pred, model ? model + NPARAM : nullptr);
  return hipGetLastError();
}

hipError_t launch_graph_levels(int n, const float *X, int64_t ld, int *level, void *ws,
                               bool after_features, hipStream_t s) {
  const FeatLayout L = feat_layout(n);
  char *w = static_cast<char *>(ws);
  uint32_t *adj = reinterpret_cast<uint32_t *>(w + L.adj);
  int *lists = reinterpret_cast<int *>(w + L.lists);
  int *res = reinterpret_cast<int *>(w + L.res);
  // After launch_graph_features on the same workspace the adjacency is packed
  // and the search from vertex 0 (source 0) has found the far vertex.
  if (!after_features) {
    hipError_t e = pack_adjacency(X, ld, n, adj, s);
    if (e != hipSuccess) return e;
  }
  const int nw = (n + 31) / 32;
  if (nw <= SB_NW)
    bfs_levels_small_kernel<<<dim3(1), dim3(SB_T), 0, s>>>(adj, n, nw, level, res + 3 * NSRC,
                                                            after_features ? res + 2 : nullptr);
  else
    bfs_levels_kernel<<<dim3(1), dim3(BFS_T), 0, s>>>(adj, n, nw, lists, level, res + 3 * NSRC,
                                                      after_features ? res + 2 : nullptr);
  return hipGetLastError();
}

int mlp_param_count() { return NPARAM; }

hipError_t launch_mlp_forward(const float *P, const float *x, int B, float *y, const float *norm,
                              hipStream_t s) {
  if (B <= 0) return hipSuccess;
  mlp_forward_kernel<<<dim3(B), dim3(64), 0, s>>>(P, x, B, y, norm);
  return hipGetLastError();
}

size_t mlp_workspace(int B) {
  const size_t chunks = (size_t)(B + GRAD_CHUNK - 1) / GRAD_CHUNK;
  return chunks * (NPARAM + 1) * sizeof(float) + 256 + (NPARAM + 1) * sizeof(float);
}

hipError_t launch_mlp_gradient(const float *P, const float *x, const float *t, const float *w,
                               int B, float *grad, float *loss, void *ws, hipStream_t s) {
  const int chunks = (B + GRAD_CHUNK - 1) / GRAD_CHUNK;
  float *partial = static_cast<float *>(ws);
  mlp_grad_kernel<<<dim3(chunks), dim3(64), 0, s>>>(P, x, t, w, B, partial);
  mlp_reduce_kernel<<<dim3((NPARAM + 1 + 255) / 256), dim3(256), 0, s>>>(partial, chunks, grad,
                                                                        loss);
  return hipGetLastError();
}

hipError_t launch_mlp_adam(float *P, float *m, float *v, const float *g, float lr, float c1,
                           float c2, hipStream_t s) {
  adam_kernel<<<dim3((NPARAM + 255) / 256), dim3(256), 0, s>>>(P, m, v, g, lr, c1, c2);
  return hipGetLastError();
}

}  // namespace simd2

// Batched repeated squaring of many independent graphs (gfx950).
//
// Every graph lives in a pool of two padded buffers (row stride ld, at least
// as many rows as its tiles cover, K padding holding the op's pad values), so
// the VALU tile body runs with 128-bit loads everywhere and only its stores
// are masked to the live n x n block (valu_tile MASKED mode).
//
// A round is a handful of launches, one per (op, tile configuration) class,
// over a static table of (graph, tile) work items built on the host: one list
// per XCD (all tiles of a graph in one list, so its panels stay in one L2;
// graphs balanced across the lists, largest first), and workgroup b takes
// item b / 8 of list b % 8 -- the hardware dispatches workgroups to the XCDs
// round-robin and in order.  Items of graphs that have converged return at
// once; a launch whose whole class has converged returns at once.  After the
// launches a tiny update kernel retires every graph whose fused change flag
// stayed clear.  Rounds are level synchronous (kernel boundaries order them)
// and no workgroup ever waits on another, so the whole solve is captured as
// one hipGraph.  (Workgroups that stay resident and claim items through
// atomic counters measured about 10% slower on the 4096-graph workload: the
// counters serialise in L2.)
#include <algorithm>
#include <cstring>

#include "valu_tile.hpp"

namespace simd2 {

namespace {

constexpr int EW = 256;

// Work item of workgroup b: the (b / 8)-th item of the list of XCD b % 8
// (workgroups are dispatched to the XCDs round-robin and in order, so each
// XCD works through its own list, largest graphs first); -1 past its end.
__device__ __forceinline__ int item_of_group(const BatchLaunchDev *__restrict__ L) {
  const int b = bid_x();
  const int x = b % BATCH_XCDS, j = b / BATCH_XCDS;
  return j < L->len[x] ? L->off[x] + j : -1;
}

template <class Op, class Cfg>
__global__ __launch_bounds__((Cfg::BM / Cfg::TM) * (Cfg::BN / Cfg::TN))
__attribute__((amdgpu_waves_per_eu(Cfg::WAVES))) void batch_valu_kernel(
    const BatchItem *__restrict__ items, const BatchLaunchDev *__restrict__ L,
    const BatchGraphDev *__restrict__ graphs, float *pool, const int *__restrict__ active,
    int *changed, int parity, const int *class_active) {
  __shared__ ValuSmem<Cfg::BM, Cfg::BN, Cfg::BK> sm;
  if (class_active && *class_active == 0) return;  // every graph of the class has converged
  const int idx = __builtin_amdgcn_readfirstlane(item_of_group(L));
  if (idx < 0) return;
  const int g = __builtin_amdgcn_readfirstlane(items[idx].g);
  const int tile = __builtin_amdgcn_readfirstlane(items[idx].tile);
  if (!active[g]) return;  // uniform over the workgroup
  const BatchGraphDev G = graphs[g];
  const float *cur = pool + (parity ? G.off[1] : G.off[0]);
  float *nxt = pool + (parity ? G.off[0] : G.off[1]);
  const Epilogue e{G.epi ? cur : nullptr, G.ld, cur, G.ld, changed + g, nullptr, G.mode};
  const int tr = tile >> 16, tc = tile & 0xffff;
  valu_tile<Op, Cfg::BM, Cfg::BN, Cfg::TM, Cfg::TN, Cfg::BK, false, true>(
      G.n, G.n, G.kpad, cur, G.ld, cur, G.ld, nxt, G.ld, e, tr * Cfg::BM, tc * Cfg::BN,
      nullptr, sm);
}

// acc | (a & b) in one instruction.
__device__ __forceinline__ uint32_t and_or(uint32_t a, uint32_t b, uint32_t acc) {
  uint32_t r;
  asm("v_and_or_b32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(acc));
  return r;
}

// Packed or_and squaring of one 64-row x 16-word (512-column) tile: row i of
// the result is the OR of the rows k of the current matrix whose bit (i, k)
// is set.  256 threads: 16 row groups of 4 rows x 16 words.
constexpr int BO_TR = 4, BO_NTR = 16, BO_NTW = 16;
constexpr int BO_BM = BO_TR * BO_NTR;  // 64 rows
constexpr int BO_BW = BO_NTW;          // 16 words

__device__ __forceinline__ void orand_item(const BatchGraphDev &G, int g, int tile,
                                           uint32_t *pool, int *changed, int parity,
                                           uint32_t (&Bs)[32][BO_BW]);

__global__ __launch_bounds__(256) void batch_orand_kernel(
    const BatchItem *__restrict__ items, const BatchLaunchDev *__restrict__ L,
    const BatchGraphDev *__restrict__ graphs, uint32_t *pool, const int *__restrict__ active,
    int *changed, int parity, const int *class_active) {
  __shared__ uint32_t Bs[32][BO_BW];
  if (class_active && *class_active == 0) return;
  const int idx = __builtin_amdgcn_readfirstlane(item_of_group(L));
  if (idx < 0) return;
  const int g = __builtin_amdgcn_readfirstlane(items[idx].g);
  const int tile = __builtin_amdgcn_readfirstlane(items[idx].tile);
  if (!active[g]) return;
  orand_item(graphs[g], g, tile, pool, changed, parity, Bs);
}

// One 64-row x 16-word tile of a packed or_and squaring.
__device__ __forceinline__ void orand_item(const BatchGraphDev &G, int g, int tile,
                                           uint32_t *pool, int *changed, int parity,
                                           uint32_t (&Bs)[32][BO_BW]) {
  const uint32_t *cur = pool + (parity ? G.off[1] : G.off[0]);
  uint32_t *nxt = pool + (parity ? G.off[0] : G.off[1]);
  const int n = G.n, nw = G.cols;
  const int64_t ld = G.ld;
  const int tid = tid_x();
  const int tw = tid % BO_NTW;
  const int row0 = (tile >> 16) * BO_BM + (tid / BO_NTW) * BO_TR;
  const int wb = (tile & 0xffff) * BO_BW;
  const int w = wb + tw;
  const uint32_t last_mask = (n & 31) ? (1u << (n & 31)) - 1u : ~0u;

  uint32_t acc[BO_TR];
#pragma unroll
  for (int i = 0; i < BO_TR; ++i) acc[i] = 0u;
  for (int kw = 0; kw < nw; ++kw) {
    uint32_t aw[BO_TR];
    uint32_t any = 0;
#pragma unroll
    for (int i = 0; i < BO_TR; ++i) {
      const int r = row0 + i;
      aw[i] = r < n ? cur[(int64_t)r * ld + kw] : 0u;
      if (kw == nw - 1) aw[i] &= last_mask;
      any |= aw[i];
    }
    if (!__syncthreads_or(any != 0)) continue;  // no k of this slab is used
    for (int f = tid; f < 32 * BO_BW; f += 256) {
      const int kk = f / BO_BW, ww = f % BO_BW;
      const int k = kw * 32 + kk, wd = wb + ww;
      Bs[kk][ww] = (k < n && wd < nw) ? cur[(int64_t)k * ld + wd] : 0u;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 32; ++kk) {
      const uint32_t b = Bs[kk][tw];
#pragma unroll
      for (int i = 0; i < BO_TR; ++i) {
        const uint32_t m = (uint32_t)((int32_t)(aw[i] << (31 - kk)) >> 31);
        acc[i] = and_or(m, b, acc[i]);
      }
    }
    __syncthreads();
  }
  bool ch = false;
  if (w < nw) {
    const uint32_t mask = (w == nw - 1) ? last_mask : ~0u;
#pragma unroll
    for (int i = 0; i < BO_TR; ++i) {
      const int r = row0 + i;
      if (r >= n) continue;
      const uint32_t p = cur[(int64_t)r * ld + w];
      uint32_t v = acc[i];
      if (G.epi) v |= p;
      v &= mask;
      ch |= v != (p & mask);
      nxt[(int64_t)r * ld + w] = v;
    }
  }
  if (__any(ch) && (tid % warpSize) == 0) atomicOr(changed + g, 1);
}

__global__ __launch_bounds__(EW) void batch_reset_kernel(const BatchGraphDev *graphs,
                                                         int *active, int *changed, int *iters,
                                                         int count) {
  const int g = bid_x() * EW + tid_x();
  if (g >= count) return;
  active[g] = graphs[g].n > 0;
  changed[g] = 0;
  iters[g] = 0;
}

// End of round `round` (0-based): a graph whose squaring changed nothing is
// done after round + 1 squarings; on the last round a graph that still
// changed keeps active = 1, which the host reports as not converged.
__global__ __launch_bounds__(EW) void batch_update_kernel(const BatchGraphDev *graphs,
                                                          int *active, int *changed, int *iters,
                                                          int *class_active_next, int count,
                                                          int round, int last) {
  const int g = bid_x() * EW + tid_x();
  if (g >= count) return;
  if (active[g]) {
    if (!changed[g]) {
      active[g] = 0;
      iters[g] = round + 1;
    } else {
      if (last) iters[g] = round + 1;
      atomicAdd(class_active_next + graphs[g].cls, 1);
    }
  }
  changed[g] = 0;
}

// Results of graphs that stopped after an odd number of squarings sit in
// buffer 1: copy their live block back to buffer 0.  Grid (graphs, chunks).
__global__ __launch_bounds__(EW) void batch_finalize_kernel(const BatchGraphDev *graphs,
                                                            uint32_t *pool, const int *iters) {
  const int g = bid_x();
  if (!(iters[g] & 1)) return;
  const BatchGraphDev G = graphs[g];
  const uint32_t *src = pool + G.off[1];
  uint32_t *dst = pool + G.off[0];
  // Rows 0 .. n-1 are contiguous in both buffers and their padding is equal,
  // so the copy is one contiguous range.
  const int64_t total = (int64_t)G.n * G.ld;
  if (G.ld % 4 == 0) {
    const uint4 *s4 = reinterpret_cast<const uint4 *>(src);
    uint4 *d4 = reinterpret_cast<uint4 *>(dst);
    for (int64_t i = (int64_t)bid_y() * EW + tid_x(); i < total / 4; i += (int64_t)gridDim.y * EW)
      d4[i] = s4[i];
  } else {
    for (int64_t i = (int64_t)bid_y() * EW + tid_x(); i < total; i += (int64_t)gridDim.y * EW)
      dst[i] = src[i];
  }
}

// Pad values of one buffer: rows >= n get pad_y (they are read as Y rows of
// the K padding), everything else pad_x (read as X columns of the K padding);
// the live block is overwritten by the upload.
__global__ __launch_bounds__(EW) void batch_pad_kernel(uint32_t *buf, int rows, int64_t ld,
                                                       int n, uint32_t pad_x, uint32_t pad_y) {
  const int64_t total = (int64_t)rows * ld;
  for (int64_t i = (int64_t)bid_x() * EW + tid_x(); i < total;
       i += (int64_t)gridDim.x * EW)
    buf[i] = (i / ld) >= n ? pad_y : pad_x;
}

template <class Op>
hipError_t launch_op_cfg(int cfg, int grid, const BatchItem *items, const BatchLaunchDev *L,
                         const BatchGraphDev *graphs, float *pool, const int *active,
                         int *changed, int parity, const int *class_active, hipStream_t s) {
  const dim3 g(grid), block(256);
  if (cfg == CFG_S64)
    batch_valu_kernel<Op, CfgS><<<g, block, 0, s>>>(items, L, graphs, pool, active, changed,
                                                    parity, class_active);
  else if (cfg == CFG_M128x64)
    batch_valu_kernel<Op, CfgM><<<g, block, 0, s>>>(items, L, graphs, pool, active, changed,
                                                    parity, class_active);
  else
    batch_valu_kernel<Op, CfgL><<<g, block, 0, s>>>(items, L, graphs, pool, active, changed,
                                                    parity, class_active);
  return hipGetLastError();
}

}  // namespace

bool batch_op_supported(int op) {
  return op == 0 || op == 2 || op == 3 || op == 4 || op == 5 || op == 6 || op == 7;
}

void batch_pad_values(int op, float *pad_x, float *pad_y) {
  switch (op) {
  case 0: *pad_x = MinPlus::pad_x; *pad_y = MinPlus::pad_y; break;
  case 2: *pad_x = MaxPlus::pad_x; *pad_y = MaxPlus::pad_y; break;
  case 3: *pad_x = MaxMin::pad_x; *pad_y = MaxMin::pad_y; break;
  case 4: *pad_x = MaxMul::pad_x; *pad_y = MaxMul::pad_y; break;
  case 5: *pad_x = MinMul::pad_x; *pad_y = MinMul::pad_y; break;
  case 6: *pad_x = MinMax::pad_x; *pad_y = MinMax::pad_y; break;
  default: *pad_x = 0.0f; *pad_y = 0.0f; break;  // or_and: zero bits
  }
}

hipError_t launch_batch_round(int op, int cfg, int grid, const BatchItem *items,
                              const BatchLaunchDev *L, const BatchGraphDev *graphs, void *pool,
                              const int *active, int *changed, int parity,
                              const int *class_active, hipStream_t s) {
  if (grid <= 0) return hipSuccess;
  float *fp = static_cast<float *>(pool);
#define SIMD2_BATCH_OP(code, Op)                                                             \
  case code:                                                                                 \
    return launch_op_cfg<Op>(cfg, grid, items, L, graphs, fp, active, changed, parity,      \
                             class_active, s);
  switch (op) {
    SIMD2_BATCH_OP(0, MinPlus)
    SIMD2_BATCH_OP(2, MaxPlus)
    SIMD2_BATCH_OP(3, MaxMin)
    SIMD2_BATCH_OP(4, MaxMul)
    SIMD2_BATCH_OP(5, MinMul)
    SIMD2_BATCH_OP(6, MinMax)
  case 7:
    batch_orand_kernel<<<dim3(grid), dim3(256), 0, s>>>(items, L, graphs,
                                                        static_cast<uint32_t *>(pool), active,
                                                        changed, parity, class_active);
    return hipGetLastError();
  default: return hipErrorInvalidValue;
  }
#undef SIMD2_BATCH_OP
}

void batch_tile_shape(int op, int cfg, int *bm, int *bn) {
  if (op == 7) {
    *bm = BO_BM;
    *bn = BO_BW;  // words
    return;
  }
  const ValuCfgInfo c = valu_cfg_info(cfg);
  *bm = c.bm;
  *bn = c.bn;
}

hipError_t launch_batch_reset(const BatchGraphDev *graphs, int *active, int *changed, int *iters,
                              int count, hipStream_t s) {
  batch_reset_kernel<<<dim3((count + EW - 1) / EW), dim3(EW), 0, s>>>(graphs, active, changed,
                                                                       iters, count);
  return hipGetLastError();
}

hipError_t launch_batch_update(const BatchGraphDev *graphs, int *active, int *changed,
                               int *iters, int *class_active_next, int count, int round,
                               int last, hipStream_t s) {
  batch_update_kernel<<<dim3((count + EW - 1) / EW), dim3(EW), 0, s>>>(
      graphs, active, changed, iters, class_active_next, count, round, last);
  return hipGetLastError();
}

hipError_t launch_batch_finalize(const BatchGraphDev *graphs, void *pool, const int *iters,
                                 int count, hipStream_t s) {
  batch_finalize_kernel<<<dim3(count, 8), dim3(EW), 0, s>>>(
      graphs, static_cast<uint32_t *>(pool), iters);
  return hipGetLastError();
}

hipError_t launch_batch_pad(void *buf, int rows, int64_t ld, int n, float pad_x, float pad_y,
                            hipStream_t s) {
  const int64_t total = (int64_t)rows * ld;
  const int blocks = (int)std::min<int64_t>((total + EW - 1) / EW, 1024);
  uint32_t px, py;
  std::memcpy(&px, &pad_x, 4);
  std::memcpy(&py, &pad_y, 4);
  batch_pad_kernel<<<dim3(blocks > 0 ? blocks : 1), dim3(EW), 0, s>>>(
      static_cast<uint32_t *>(buf), rows, ld, n, px, py);
  return hipGetLastError();
}

}  // namespace simd2

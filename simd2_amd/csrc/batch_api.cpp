// Host side of the batched solver: pool layout, the size-binned schedule
// (work-item tables per (op, tile configuration)), its hipGraph capture and
// the autotuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <vector>

#include "../../include/simd2/batch.h"
#include "../../include/simd2/converge.h"
#include "../../include/simd2/srgemm.h"
#include "internal.hpp"

using namespace simd2;

namespace {

constexpr int NUM_XCD = 8;
constexpr int BIN = 32;  // size bins of the schedule: n in (32 (b-1), 32 b]
constexpr int BRANCHES = 4;  // parallel branches of a captured round

int hip_status(hipError_t e) {
  if (e == hipSuccess) return SIMD2_SUCCESS;
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return SIMD2_ERROR_NO_DEVICE;
  return SIMD2_ERROR_INTERNAL;
}

int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

struct Launch {
  int op, cfg;
  int branch;  // parallel branch of the captured round that runs it
  int items;  // work items (tiles) of the launch
  int grid;   // workgroups: 8 x the longest per-XCD list
};

}  // namespace

struct simd2_batch {
  int count = 0;
  int max_rounds = 0;
  std::vector<simd2_batch_graph> graphs;
  std::vector<BatchGraphDev> desc;
  std::vector<int> rows;  // allocated rows per buffer
  std::vector<int> cfg;   // VALU tile configuration per graph (-1: or_and / empty)
  std::vector<Launch> launches;
  void *pool = nullptr;
  size_t pool_bytes = 0;
  BatchGraphDev *d_desc = nullptr;
  int *d_state = nullptr;  // active | changed | iters, count each
  BatchItem *d_items = nullptr;
  size_t items_cap = 0;
  BatchLaunchDev *d_launch = nullptr;  // per launch: its per-XCD item lists
  int *d_class = nullptr;              // [max_rounds + 1][launches]: active graphs per class
  size_t launch_cap = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // Recorded after every solve: the tables and the pool are replaced or freed
  // only after the last solve, on whatever stream it ran, has finished.
  hipEvent_t solved = nullptr;

  int *active() { return d_state; }
  int *changed() { return d_state + count; }
  int *iters() { return d_state + 2 * count; }
};

namespace {

bool is_float_op(int op) { return op != SIMD2_OR_AND; }

// Time of one semiring term relative to min_plus's in the VALU kernels (the
// single-problem kernels' measured rates at 4096^3); or_and's terms are bits.
double op_weight(int op) {
  switch (op) {
    case SIMD2_MAX_MIN:
    case SIMD2_MIN_MAX: return 1.6;
    case SIMD2_MAX_MUL:
    case SIMD2_MIN_MUL: return 1.15;
    case SIMD2_OR_AND: return 0.07;
    default: return 1.0;
  }
}

// Work of one tile row (or word block) of graph g under `cfg`: proportional
// to the tile area the kernel sweeps times K.
double tile_cost(const simd2_batch &b, int g, int cfg) {
  const int op = b.graphs[g].op;
  int bm, bn;
  batch_tile_shape(op, cfg, &bm, &bn);
  if (op == SIMD2_OR_AND) return (double)bm * bn * 32.0 * b.desc[g].cols / 32.0;
  return (double)bm * bn * b.desc[g].kpad;
}

int tiles_of(const simd2_batch &b, int g, int cfg, int *tr, int *tc) {
  const int op = b.graphs[g].op;
  int bm, bn;
  batch_tile_shape(op, cfg, &bm, &bn);
  const int n = b.graphs[g].n;
  *tr = (n + bm - 1) / bm;
  *tc = op == SIMD2_OR_AND ? (b.desc[g].cols + bn - 1) / bn : (n + bn - 1) / bn;
  return *tr * *tc;
}

// Work items of one launch as 8 per-XCD lists appended to `out`: graphs
// largest first, all tiles of a graph in one list (workgroup i runs on XCD
// i % 8 and takes item i / 8 of list i % 8, so a graph's panels stay in one
// L2), lists balanced by work.
BatchLaunchDev build_items(const simd2_batch &b, const std::vector<int> &gs, int cfg,
                           std::vector<BatchItem> &out) {
  std::vector<int> order(gs);
  std::stable_sort(order.begin(), order.end(),
                   [&](int x, int y) { return b.graphs[x].n > b.graphs[y].n; });
  std::vector<std::vector<BatchItem>> lists(NUM_XCD);
  double load[NUM_XCD] = {0};
  for (int g : order) {
    int tr, tc;
    const int nt = tiles_of(b, g, cfg, &tr, &tc);
    const int x = (int)(std::min_element(load, load + NUM_XCD) - load);
    for (int r = 0; r < tr; ++r)
      for (int c = 0; c < tc; ++c) lists[x].push_back({g, (r << 16) | c});
    load[x] += nt * tile_cost(b, g, cfg);
  }
  BatchLaunchDev L;
  for (int x = 0; x < NUM_XCD; ++x) {
    L.off[x] = (int)out.size();
    L.len[x] = (int)lists[x].size();
    out.insert(out.end(), lists[x].begin(), lists[x].end());
  }
  return L;
}

// Groups graphs by (op, cfg) into launches (largest estimated time first),
// fills `items` and the launch descriptors, and records each graph's launch
// in cls[g].  Launches go to the round's parallel branches longest first,
// each to the branch with the least estimated time so far.  FP32 graphs with cfg < 0 are left out (the autotuner plans one size
// bin at a time), or_and graphs unless with_orand.
void plan(const simd2_batch &b, const std::vector<int> &cfg, bool with_orand,
          std::vector<Launch> &launches, std::vector<BatchLaunchDev> &ldev,
          std::vector<BatchItem> &items, std::vector<int> &cls) {
  std::map<std::pair<int, int>, std::vector<int>> classes;
  for (int g = 0; g < b.count; ++g) {
    if (b.graphs[g].n == 0) continue;
    const bool orand = b.graphs[g].op == SIMD2_OR_AND;
    if (orand ? with_orand : cfg[g] >= 0) classes[{b.graphs[g].op, cfg[g]}].push_back(g);
  }
  std::vector<std::pair<double, std::pair<int, int>>> order;
  for (auto &kv : classes) {
    double w = 0;
    for (int g : kv.second) {
      int tr, tc;
      w += tiles_of(b, g, kv.first.second, &tr, &tc) * tile_cost(b, g, kv.first.second);
    }
    order.push_back({-w * op_weight(kv.first.first), kv.first});
  }
  std::sort(order.begin(), order.end());
  launches.clear();
  ldev.clear();
  items.clear();
  cls.assign(b.count, 0);
  const int nb = std::max(1, std::min<int>(BRANCHES, (int)order.size()));
  std::vector<double> branch_load(nb, 0.0);
  for (auto &o : order) {
    Launch L;
    L.op = o.second.first;
    L.cfg = o.second.second;
    L.branch = (int)(std::min_element(branch_load.begin(), branch_load.end()) -
                     branch_load.begin());
    branch_load[L.branch] -= o.first;
    const int first = (int)items.size();
    for (int g : classes[o.second]) cls[g] = (int)launches.size();
    const BatchLaunchDev d = build_items(b, classes[o.second], L.cfg, items);
    ldev.push_back(d);
    L.items = (int)items.size() - first;
    L.grid = NUM_XCD * *std::max_element(d.len, d.len + NUM_XCD);
    launches.push_back(L);
  }
}

// Copies a plan to the device (growing the buffers when needed) and makes it
// the one launch_round runs.
int upload_plan(simd2_batch &b, const std::vector<Launch> &launches,
                const std::vector<BatchLaunchDev> &ldev, const std::vector<BatchItem> &items,
                const std::vector<int> &cls) {
  hipError_t e = hipSuccess;
  if (items.size() > b.items_cap) {
    if (b.d_items) (void)hipFree(b.d_items);
    b.d_items = nullptr;
    b.items_cap = 0;
    e = hipMalloc(&b.d_items, items.size() * sizeof(BatchItem));
    if (e != hipSuccess) return hip_status(e);
    b.items_cap = items.size();
  }
  if (launches.size() > b.launch_cap) {
    if (b.d_launch) (void)hipFree(b.d_launch);
    if (b.d_class) (void)hipFree(b.d_class);
    b.d_launch = nullptr;
    b.d_class = nullptr;
    b.launch_cap = 0;
    const size_t nl = launches.size();
    e = hipMalloc(&b.d_launch, nl * sizeof(BatchLaunchDev));
    if (e == hipSuccess) e = hipMalloc(&b.d_class, (b.max_rounds + 1) * nl * sizeof(int));
    if (e != hipSuccess) return hip_status(e);
    b.launch_cap = nl;
  }
  if (!items.empty())
    e = hipMemcpy(b.d_items, items.data(), items.size() * sizeof(BatchItem),
                  hipMemcpyHostToDevice);
  if (e == hipSuccess && !ldev.empty())
    e = hipMemcpy(b.d_launch, ldev.data(), ldev.size() * sizeof(BatchLaunchDev),
                  hipMemcpyHostToDevice);
  for (int g = 0; g < b.count; ++g) b.desc[g].cls = cls[g];
  if (e == hipSuccess && b.count > 0)
    e = hipMemcpy(b.d_desc, b.desc.data(), b.desc.size() * sizeof(BatchGraphDev),
                  hipMemcpyHostToDevice);
  if (e != hipSuccess) return hip_status(e);
  b.launches = launches;
  return SIMD2_SUCCESS;
}

// Round r of the current plan on stream s: clear the next round's class
// counts, then one launch per class (classes with no active graph left
// return at once).
hipError_t launch_round(simd2_batch &b, int r, hipStream_t s) {
  const int nl = (int)b.launches.size();
  if (nl == 0) return hipSuccess;
  hipError_t e = launch_fill_i32(b.d_class + (int64_t)(r + 1) * nl, 0, nl, s);
  for (int i = 0; i < nl && e == hipSuccess; ++i) {
    const Launch &L = b.launches[i];
    e = launch_batch_round(L.op, L.cfg, L.grid, b.d_items, b.d_launch + i, b.d_desc, b.pool,
                           b.active(), b.changed(), r & 1,
                           r > 0 ? b.d_class + (int64_t)r * nl + i : nullptr, s);
  }
  return e;
}

// Round r with its class launches as parallel branches: the class counts are
// cleared on s, then each launch runs on its branch's stream after an event
// of s, and s waits for every branch before the round's update.  Graphs of
// different classes share no memory, so the classes may overlap: the tail of
// one launch fills with the next one's workgroups, and in the late rounds the
// classes whose graphs take more squarings run side by side.  Used while s
// is being captured (the branches join the capture through the events).
hipError_t launch_round_branched(simd2_batch &b, int r, hipStream_t s,
                                 const std::vector<hipStream_t> &br,
                                 std::vector<hipEvent_t> &events) {
  const int nl = (int)b.launches.size();
  if (nl == 0) return hipSuccess;
  auto event = [&](hipEvent_t *ev) {
    hipError_t e = hipEventCreateWithFlags(ev, hipEventDisableTiming);
    if (e == hipSuccess) events.push_back(*ev);
    return e;
  };
  hipError_t e = launch_fill_i32(b.d_class + (int64_t)(r + 1) * nl, 0, nl, s);
  hipEvent_t fork;
  if (e == hipSuccess) e = event(&fork);
  if (e == hipSuccess) e = hipEventRecord(fork, s);
  const int nb = std::min<int>((int)br.size(), nl);
  for (int k = 0; k < nb && e == hipSuccess; ++k) e = hipStreamWaitEvent(br[k], fork, 0);
  for (int i = 0; i < nl && e == hipSuccess; ++i) {
    const Launch &L = b.launches[i];
    e = launch_batch_round(L.op, L.cfg, L.grid, b.d_items, b.d_launch + i, b.d_desc, b.pool,
                           b.active(), b.changed(), r & 1,
                           r > 0 ? b.d_class + (int64_t)r * nl + i : nullptr, br[L.branch % nb]);
  }
  for (int k = 0; k < nb && e == hipSuccess; ++k) {
    hipEvent_t join;
    e = event(&join);
    if (e == hipSuccess) e = hipEventRecord(join, br[k]);
    if (e == hipSuccess) e = hipStreamWaitEvent(s, join, 0);
  }
  return e;
}

// Waits for the last solve (any stream) before its tables change.
void wait_solved(simd2_batch &b) {
  if (b.solved) (void)hipEventSynchronize(b.solved);
}

void destroy_graph(simd2_batch &b) {
  wait_solved(b);
  if (b.exec) (void)hipGraphExecDestroy(b.exec);
  if (b.graph) (void)hipGraphDestroy(b.graph);
  b.exec = nullptr;
  b.graph = nullptr;
}

// Installs cfg as the schedule and captures reset, max_rounds rounds and the
// final copy-back as one graph.
int install(simd2_batch &b, const std::vector<int> &cfg) {
  std::vector<Launch> launches;
  std::vector<BatchLaunchDev> ldev;
  std::vector<BatchItem> items;
  std::vector<int> cls;
  destroy_graph(b);  // the old graph refers to buffers the upload may replace
  plan(b, cfg, true, launches, ldev, items, cls);
  int st = upload_plan(b, launches, ldev, items, cls);
  if (st != SIMD2_SUCCESS) return st;
  b.cfg = cfg;
  if (b.count == 0) return SIMD2_SUCCESS;
  hipStream_t cs;
  hipError_t e = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
  if (e != hipSuccess) return hip_status(e);
  // branch streams of the rounds (default priority, like every stream here)
  std::vector<hipStream_t> br;
  std::vector<hipEvent_t> events;
  for (int k = 0; k < BRANCHES && e == hipSuccess; ++k) {
    hipStream_t t;
    e = hipStreamCreateWithFlags(&t, hipStreamNonBlocking);
    if (e == hipSuccess) br.push_back(t);
  }
  if (e == hipSuccess) e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    hipError_t le = launch_batch_reset(b.d_desc, b.active(), b.changed(), b.iters(), b.count, cs);
    for (int r = 0; r < b.max_rounds && le == hipSuccess; ++r) {
      le = launch_round_branched(b, r, cs, br, events);
      if (le == hipSuccess)
        le = launch_batch_update(b.d_desc, b.active(), b.changed(), b.iters(),
                                 b.d_class + (int64_t)(r + 1) * b.launches.size(), b.count, r,
                                 r == b.max_rounds - 1, cs);
    }
    if (le == hipSuccess) le = launch_batch_finalize(b.d_desc, b.pool, b.iters(), b.count, cs);
    e = hipStreamEndCapture(cs, &b.graph);
    if (le != hipSuccess) e = le;
  }
  if (e == hipSuccess) e = hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0);
  (void)hipStreamDestroy(cs);
  for (hipStream_t t : br) (void)hipStreamDestroy(t);
  for (hipEvent_t ev : events) (void)hipEventDestroy(ev);
  if (e != hipSuccess) {
    destroy_graph(b);
    return hip_status(e);
  }
  return SIMD2_SUCCESS;
}

// Default tile configuration for an n-vertex graph: the shape that wastes the
// least padded work, preferring larger tiles (fewer LDS reads per op) on ties
// within 10%.
int default_cfg(int n) {
  if (n <= 64) return CFG_S64;
  const double w64 = (double)round_up(n, 64) * round_up(n, 64);
  const double w128 = (double)round_up(n, 128) * round_up(n, 128);
  const double w12864 = (double)round_up(n, 128) * round_up(n, 64);
  if (w128 <= 1.10 * w64) return CFG_L128;
  if (w12864 <= 1.05 * w64) return CFG_M128x64;
  return CFG_S64;
}

bool cfg_allowed(int n, int cfg) { return n > 64 || cfg == CFG_S64; }

}  // namespace

extern "C" {

int simd2_batch_create(const simd2_batch_graph *graphs, int count, int max_rounds,
                       simd2_batch **out) {
  if (!out) return SIMD2_ERROR_NULL_POINTER;
  *out = nullptr;
  if (count < 0 || max_rounds < 0) return SIMD2_ERROR_INVALID_PROBLEM;
  if (count > 0 && !graphs) return SIMD2_ERROR_NULL_POINTER;
  int nmax = 0;
  for (int g = 0; g < count; ++g) {
    const simd2_batch_graph &G = graphs[g];
    if (G.n < 0 || G.n > (1 << 20)) return SIMD2_ERROR_INVALID_PROBLEM;
    if (!batch_op_supported(G.op)) return SIMD2_ERROR_NOT_SUPPORTED;
    if (G.op != SIMD2_OR_AND && (G.check_mode < 0 || G.check_mode > 2))
      return SIMD2_ERROR_NOT_SUPPORTED;
    nmax = std::max(nmax, G.n);
  }
  simd2_batch *b = new simd2_batch;
  b->count = count;
  b->graphs.assign(graphs, graphs + count);
  if (max_rounds == 0) {
    int lg = 0;
    while ((1 << lg) < std::max(nmax - 1, 1)) ++lg;
    max_rounds = lg + 3;
  }
  b->max_rounds = max_rounds;
  b->desc.resize(count);
  b->rows.resize(count);
  b->cfg.assign(count, -1);
  int64_t total = 0;
  for (int g = 0; g < count; ++g) {
    const simd2_batch_graph &G = graphs[g];
    BatchGraphDev &D = b->desc[g];
    D.n = G.n;
    D.epi = G.do_epilogue ? 1 : 0;
    if (G.op == SIMD2_OR_AND) {
      D.cols = (G.n + 31) / 32;
      D.ld = D.cols;
      D.kpad = G.n;
      D.mode = CHECK_EXACT;
      b->rows[g] = G.n;
    } else {
      // Rows and columns padded to the largest tile any configuration may
      // use, so every tile and every K step the kernel touches exists.
      const int P = G.n <= 64 ? (int)round_up(G.n, 64) : (int)round_up(G.n, 128);
      D.cols = G.n;
      D.ld = P;
      D.kpad = (int)round_up(G.n, 16);
      D.mode = G.check_mode;
      b->rows[g] = P;
      if (G.n > 0) b->cfg[g] = default_cfg(G.n);
    }
    if (G.n == 0) b->rows[g] = 0;
    const int64_t sz = round_up((int64_t)b->rows[g] * D.ld, 64);  // 256-byte aligned
    D.off[0] = total;
    D.off[1] = total + sz;
    total += 2 * sz;
  }
  hipError_t e = hipSuccess;
  b->pool_bytes = (size_t)total * 4;
  if (total > 0) e = hipMalloc(&b->pool, b->pool_bytes);
  if (e == hipSuccess && count > 0)
    e = hipMalloc(&b->d_desc, sizeof(BatchGraphDev) * count);
  if (e == hipSuccess && count > 0) e = hipMalloc(&b->d_state, sizeof(int) * 3 * count);
  if (e == hipSuccess && count > 0)
    e = hipMemcpy(b->d_desc, b->desc.data(), sizeof(BatchGraphDev) * count,
                  hipMemcpyHostToDevice);
  for (int g = 0; g < count && e == hipSuccess; ++g) {
    if (graphs[g].n == 0) continue;
    float px, py;
    batch_pad_values(graphs[g].op, &px, &py);
    for (int k = 0; k < 2 && e == hipSuccess; ++k)
      e = launch_batch_pad(static_cast<uint32_t *>(b->pool) + b->desc[g].off[k], b->rows[g],
                           b->desc[g].ld, graphs[g].n, px, py, nullptr);
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  int st = hip_status(e);
  if (st == SIMD2_SUCCESS) st = install(*b, b->cfg);
  if (st != SIMD2_SUCCESS) {
    simd2_batch_destroy(b);
    return st;
  }
  *out = b;
  return SIMD2_SUCCESS;
}

int simd2_batch_destroy(simd2_batch *b) {
  if (!b) return SIMD2_SUCCESS;
  destroy_graph(*b);
  if (b->solved) (void)hipEventDestroy(b->solved);
  if (b->pool) (void)hipFree(b->pool);
  if (b->d_desc) (void)hipFree(b->d_desc);
  if (b->d_state) (void)hipFree(b->d_state);
  if (b->d_items) (void)hipFree(b->d_items);
  if (b->d_launch) (void)hipFree(b->d_launch);
  if (b->d_class) (void)hipFree(b->d_class);
  delete b;
  return SIMD2_SUCCESS;
}

static int copy_graph(simd2_batch *b, int g, void *host_side, int64_t ld, bool in,
                      hipStream_t stream) {
  if (!b) return SIMD2_ERROR_NULL_POINTER;
  if (g < 0 || g >= b->count) return SIMD2_ERROR_INVALID_PROBLEM;
  const BatchGraphDev &D = b->desc[g];
  if (D.n == 0) return SIMD2_SUCCESS;
  if (!host_side) return SIMD2_ERROR_NULL_POINTER;
  if (ld < D.cols) return SIMD2_ERROR_INVALID_PROBLEM;
  void *dev = static_cast<uint32_t *>(b->pool) + D.off[0];
  const size_t row = (size_t)D.cols * 4;
  hipError_t e = in ? hipMemcpy2DAsync(dev, (size_t)D.ld * 4, host_side, (size_t)ld * 4, row,
                                       D.n, hipMemcpyDefault, stream)
                    : hipMemcpy2DAsync(host_side, (size_t)ld * 4, dev, (size_t)D.ld * 4, row,
                                       D.n, hipMemcpyDefault, stream);
  return hip_status(e);
}

int simd2_batch_upload(simd2_batch *b, int g, const void *src, int64_t ld, hipStream_t stream) {
  return copy_graph(b, g, const_cast<void *>(src), ld, true, stream);
}

int simd2_batch_download(simd2_batch *b, int g, void *dst, int64_t ld, hipStream_t stream) {
  return copy_graph(b, g, dst, ld, false, stream);
}

int simd2_batch_buffer(simd2_batch *b, int g, void **ptr, int64_t *ld) {
  if (!b || !ptr || !ld) return SIMD2_ERROR_NULL_POINTER;
  if (g < 0 || g >= b->count) return SIMD2_ERROR_INVALID_PROBLEM;
  *ptr = b->desc[g].n ? static_cast<uint32_t *>(b->pool) + b->desc[g].off[0] : nullptr;
  *ld = b->desc[g].ld;
  return SIMD2_SUCCESS;
}

int simd2_batch_solve(simd2_batch *b, hipStream_t stream) {
  if (!b) return SIMD2_ERROR_NULL_POINTER;
  if (b->count == 0) return SIMD2_SUCCESS;
  if (!b->exec) return SIMD2_ERROR_INTERNAL;
  if (!b->solved) {
    const hipError_t e = hipEventCreateWithFlags(&b->solved, hipEventDisableTiming);
    if (e != hipSuccess) {
      b->solved = nullptr;
      return hip_status(e);
    }
  }
  hipError_t e = hipGraphLaunch(b->exec, stream);
  if (e == hipSuccess) e = hipEventRecord(b->solved, stream);
  return hip_status(e);
}

int simd2_batch_iterations(simd2_batch *b, int *iters, int *unconverged, hipStream_t stream) {
  if (!b) return SIMD2_ERROR_NULL_POINTER;
  if (unconverged) *unconverged = 0;
  if (b->count == 0) return SIMD2_SUCCESS;
  std::vector<int> state(3 * (size_t)b->count);
  hipError_t e = hipMemcpyAsync(state.data(), b->d_state, state.size() * sizeof(int),
                                hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) return hip_status(e);
  int bad = 0;
  for (int g = 0; g < b->count; ++g) {
    if (iters) iters[g] = state[2 * (size_t)b->count + g];
    bad += state[g] != 0;  // still active after the last round
  }
  if (unconverged) *unconverged = bad;
  return SIMD2_SUCCESS;
}

int simd2_batch_schedule(simd2_batch *b, int *cfg, int *launches) {
  if (!b) return SIMD2_ERROR_NULL_POINTER;
  if (cfg)
    for (int g = 0; g < b->count; ++g) cfg[g] = b->cfg[g];
  if (launches) *launches = (int)b->launches.size();
  return SIMD2_SUCCESS;
}

// Autotuner.  Step 1: for every (op, size bin) with more than one admissible
// tile shape, time one squaring of the bin's graphs per shape (round 0:
// reads buffer 0, writes buffer 1) and keep the fastest.  Step 2: time a
// whole first round for that per-bin schedule and for each uniform schedule
// (one shape for every graph that admits it), which also prices the launch
// count and tails; install the fastest.
int simd2_batch_autotune(simd2_batch *b, hipStream_t stream) {
  if (!b) return SIMD2_ERROR_NULL_POINTER;
  if (b->count == 0) return SIMD2_SUCCESS;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return SIMD2_ERROR_INTERNAL;
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return SIMD2_ERROR_INTERNAL;
  }
  const std::vector<int> saved_cfg = b->cfg;
  int st = SIMD2_SUCCESS;
  destroy_graph(*b);  // trial plans replace the device tables it refers to

  // Times `reps` runs of round 0 of the plan for cfg (all graphs active).
  auto time_plan = [&](const std::vector<int> &cfg, bool with_orand, int reps,
                       float *ms) -> int {
    std::vector<Launch> launches;
    std::vector<BatchLaunchDev> ldev;
    std::vector<BatchItem> items;
    std::vector<int> cls;
    plan(*b, cfg, with_orand, launches, ldev, items, cls);
    int s = upload_plan(*b, launches, ldev, items, cls);
    if (s != SIMD2_SUCCESS) return s;
    hipError_t e = launch_batch_reset(b->d_desc, b->active(), b->changed(), b->iters(),
                                      b->count, stream);
    if (e == hipSuccess) e = launch_round(*b, 0, stream);  // warm-up
    if (e == hipSuccess) e = hipEventRecord(e0, stream);
    for (int r = 0; r < reps && e == hipSuccess; ++r) e = launch_round(*b, 0, stream);
    if (e == hipSuccess) e = hipEventRecord(e1, stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    *ms = t / reps;
    return hip_status(e);
  };

  // Step 1: per (op, bin).
  std::map<std::pair<int, int>, std::vector<int>> bins;
  for (int g = 0; g < b->count; ++g)
    if (b->graphs[g].n > 0 && is_float_op(b->graphs[g].op))
      bins[{b->graphs[g].op, (b->graphs[g].n + BIN - 1) / BIN}].push_back(g);
  std::vector<int> best = b->cfg;
  for (auto &kv : bins) {
    if (st != SIMD2_SUCCESS) break;
    const std::vector<int> &gs = kv.second;
    int nmin = 1 << 30;
    for (int g : gs) nmin = std::min(nmin, b->graphs[g].n);
    float best_ms = 1e30f;
    int arg = -1;
    for (int cfg = 0; cfg < CFG_COUNT && st == SIMD2_SUCCESS; ++cfg) {
      if (!cfg_allowed(nmin, cfg)) continue;
      std::vector<int> trial(b->count, -1);  // this bin's graphs only
      for (int g : gs) trial[g] = cfg;
      float ms = 0.f;
      st = time_plan(trial, false, 3, &ms);
      if (st == SIMD2_SUCCESS && ms < best_ms) {
        best_ms = ms;
        arg = cfg;
      }
    }
    if (arg >= 0)
      for (int g : gs) best[g] = arg;
  }

  // Step 2: whole-round comparison of the per-bin schedule and the uniform ones.
  std::vector<std::vector<int>> candidates = {best};
  for (int cfg = 0; cfg < CFG_COUNT; ++cfg) {
    std::vector<int> c = b->cfg;
    for (int g = 0; g < b->count; ++g)
      if (c[g] >= 0) c[g] = cfg_allowed(b->graphs[g].n, cfg) ? cfg : CFG_S64;
    candidates.push_back(c);
  }
  std::vector<int> chosen = best;
  float chosen_ms = 1e30f;
  for (auto &c : candidates) {
    if (st != SIMD2_SUCCESS) break;
    float ms = 0.f;
    st = time_plan(c, true, 3, &ms);
    if (st == SIMD2_SUCCESS && ms < chosen_ms) {
      chosen_ms = ms;
      chosen = c;
    }
  }

  // Step 3: whole solves.  Round 0 alone misjudges the later rounds, where
  // fewer graphs remain and the class launches run side by side, so the
  // schedule in use, the per-bin one and the round-0 winner are each timed
  // on full solves (the better of two) of a copy of the uploaded matrices,
  // and the fastest is kept.  Without the memory for the copy, the round-0
  // choice stands.
  std::vector<std::vector<int>> finalists = {saved_cfg};
  if (best != saved_cfg) finalists.push_back(best);
  if (chosen != saved_cfg && chosen != best) finalists.push_back(chosen);
  void *backup = nullptr;
  if (st == SIMD2_SUCCESS && finalists.size() > 1 && b->pool_bytes > 0) {
    if (hipMalloc(&backup, b->pool_bytes) != hipSuccess) {
      backup = nullptr;
      (void)hipGetLastError();  // the failed allocation is not an error of the call
    }
  }
  if (backup) {
    hipError_t e = hipMemcpyAsync(backup, b->pool, b->pool_bytes, hipMemcpyDeviceToDevice, stream);
    float fastest = 1e30f;
    size_t arg = 0;
    for (size_t f = 0; f < finalists.size() && e == hipSuccess && st == SIMD2_SUCCESS; ++f) {
      st = install(*b, finalists[f]);
      for (int rep = 0; rep < 2 && st == SIMD2_SUCCESS && e == hipSuccess; ++rep) {
        e = hipMemcpyAsync(b->pool, backup, b->pool_bytes, hipMemcpyDeviceToDevice, stream);
        if (e == hipSuccess) e = hipEventRecord(e0, stream);
        if (e == hipSuccess) e = hipGraphLaunch(b->exec, stream);
        if (e == hipSuccess) e = hipEventRecord(e1, stream);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float t = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
        if (e == hipSuccess && t < fastest) {
          fastest = t;
          arg = f;
        }
      }
    }
    // the uploaded matrices back in place, whatever happened above
    const hipError_t r = hipMemcpyAsync(b->pool, backup, b->pool_bytes, hipMemcpyDeviceToDevice,
                                        stream);
    const hipError_t w = hipStreamSynchronize(stream);
    (void)hipFree(backup);
    if (e == hipSuccess) e = r != hipSuccess ? r : w;
    if (e != hipSuccess && st == SIMD2_SUCCESS) st = hip_status(e);
    if (st == SIMD2_SUCCESS) chosen = finalists[arg];
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st != SIMD2_SUCCESS) {
    (void)install(*b, saved_cfg);
    return st;
  }
  return install(*b, chosen);
}

}  // extern "C"

// Internal launcher interface shared by the kernels and the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#ifdef __HIP__
#include "hw.hpp"  // kernel sources only
#endif

namespace simd2 {

// Row-major problem every kernel solves:
//   O[r][c] = (+)_{k<K} X[r][k] (x) Y[k][c]   (r < R, c < Cn)
// then, per element, the epilogue below.  The column-major C ABI maps onto
// it by swapping operands (a column-major MxN product is the row-major NxM
// product of the transposes), which is exact because every (x) here is
// commutative bit for bit.
struct Epilogue {
  const float *C = nullptr;  // O = O (+) C when set (may alias O)
  int64_t ldc = 0;
  const float *P = nullptr;  // previous values: change detection when set
  int64_t ldp = 0;
  int *flag = nullptr;       // set to 1 when any element changed
  int *tile_flags = nullptr; // per output tile: 1 changed / 0 unchanged
  int mode = 0;              // change test: 0 bits differ, 1 |p-o| > 1e-4,
                             // 2 APLP: o in (-1,1) -> -FLT_MAX, then o != p
  const int *gate = nullptr; // when set, tile (tr, tc) is computed only if
  int gate_ld = 0;           //   gate[tr * gate_ld + tc] != 0
};

enum CheckMode { CHECK_EXACT = 0, CHECK_THRESHOLD = 1, CHECK_APLP = 2 };

// Tile configurations of the VALU kernel.
enum ValuCfg { CFG_AUTO = -1, CFG_L128 = 0, CFG_S64 = 1, CFG_M128x64 = 2, CFG_COUNT = 3 };
struct ValuCfgInfo { int bm, bn, bk, threads; };
ValuCfgInfo valu_cfg_info(int cfg);
int valu_pick_cfg(int op, int R, int Cn, int K);
int device_cu_count();  // compute units of the current device (first call cached)
// Caller workspaces are carved into 256-byte aligned pieces that the kernels
// read with 16-byte vector loads: their base must be 256-byte aligned.
inline bool workspace_aligned(const void *p) {
  return (reinterpret_cast<uintptr_t>(p) & 255) == 0;
}
void valu_record_cfg(int op, int R, int Cn, int K, int cfg);

hipError_t launch_valu_f32(int op, int R, int Cn, int K, const float *X,
                           int64_t ldx, const float *Y, int64_t ldy, float *O,
                           int64_t ldo, const Epilogue &e, hipStream_t s,
                           int cfg = CFG_AUTO);

// minus_square, same row-major problem, with C added when set: tiles whose
// operands are integers in [-128, 127] and whose sums stay <= 2^24 on the
// int8 matrix cores (exact), the rest on the VALU kernel (msq_exact.hip).
// The workspace is the caller's; 0 bytes (K > 65536 or an empty problem)
// means the VALU kernel does every tile.
size_t minus_square_exact_workspace(int R, int Cn, int K);
hipError_t launch_minus_square_exact(int R, int Cn, int K, const float *X, int64_t ldx,
                                     const float *Y, int64_t ldy, float *O, int64_t ldo,
                                     const float *C, int64_t ldc, void *ws, size_t ws_bytes,
                                     hipStream_t s);
// The same on the library's own scratch (a device array of MSQ_NOALLOC_BYTES
// per device, no allocation); the caller has checked that the problem's
// workspace fits.  Calls that overlap on one device race on the scratch.
constexpr size_t MSQ_NOALLOC_BYTES = (size_t)40 << 20;
hipError_t minus_square_exact_noalloc(int R, int Cn, int K, const float *X, int64_t ldx,
                                      const float *Y, int64_t ldy, float *O, int64_t ldo,
                                      const float *C, int64_t ldc, hipStream_t s);
// 64x64 tiles of the last launch_minus_square_exact on this workspace and
// problem computed on the matrix cores / on the VALU kernel; synchronises s.
hipError_t minus_square_exact_tile_counts(const void *ws, int R, int Cn, int K, hipStream_t s,
                                          int *mfma_tiles, int *valu_tiles);

// plus_mul on MFMA (v_mfma_f32_16x16x4_f32), same row-major problem.
hipError_t launch_plusmul_mfma(int R, int Cn, int K, const float *X,
                               int64_t ldx, const float *Y, int64_t ldy,
                               float *O, int64_t ldo, const float *C,
                               int64_t ldc, hipStream_t s);

// or_and on int32 booleans (row-major problem as above).
hipError_t launch_orand_i32(int R, int Cn, int K, const int *X, int64_t ldx,
                            const int *Y, int64_t ldy, int *O, int64_t ldo,
                            const int *C, int64_t ldc, const int *P,
                            int64_t ldp, int *flag, hipStream_t s);

// or_and on bit-packed rows.
// int32 booleans (nonzero = 1) to packed rows and back (srgemm_orand.hip);
// the packing sets *nonbool (when not null) if X holds a value other than 0/1.
hipError_t launch_pack_i32(int rows, int cols, const int *X, int64_t ldx, uint32_t *P, int64_t ldp,
                           int *nonbool, hipStream_t s);
hipError_t launch_unpack_i32(int rows, int cols, const uint32_t *P, int64_t ldp, int *X,
                             int64_t ldx, hipStream_t s);
hipError_t launch_orand_packed(int M, int N, int K, const uint32_t *A,
                               int64_t lda, const uint32_t *B, int64_t ldb,
                               const uint32_t *C, int64_t ldc, uint32_t *D,
                               int64_t ldd, const uint32_t *P, int64_t ldp,
                               int *flag, hipStream_t s);

// FP16-input emulation kernel and conversions (tensor_emu.hip).
hipError_t launch_tensor_emu(int op, const void *a, const void *b, const float *c, float *d,
                             int m, int n, int k, int do_epilogue, hipStream_t s);
hipError_t launch_f2h(void *out, const float *in, int64_t n, hipStream_t s);
hipError_t launch_h2f(float *out, const void *in, int64_t n, hipStream_t s);

// Top-k of a materialised distance matrix (knn_select.hip): any 1 <= k <= R,
// the workspace holds knn_select_workspace(Q, R, k) bytes.
size_t knn_select_workspace(int Q, int R, int k);
hipError_t launch_knn_select(const float *dist, int Q, int R, int k, float *out_d, int *out_i,
                             int do_sqrt, void *workspace, hipStream_t s);
// k-nearest neighbours, precision form (knn.hip).
size_t knn_bf16_workspace(int Q, int R, int d);
int knn_bf16_kmax();
hipError_t launch_knn_bf16(const float *query, int64_t ldq, const float *ref, int64_t ldr, int Q,
                           int R, int d, int k, float *out_d, int *out_i, int do_sqrt,
                           void *workspace, int *stats, hipStream_t s);

// Batched repeated squaring (batch.hip).  Buffers are 4-byte elements of one
// pool (floats, or packed 32-bit words for or_and).
struct BatchGraphDev {
  int64_t off[2];  // element offsets of the two buffers in the pool
  int n;           // vertices
  int cols;        // live elements per row: n, or ceil(n/32) words for or_and
  int ld;          // row stride in elements
  int kpad;        // K loop bound of the VALU kernel (multiple of 16)
  int epi;         // 1: result (+)= current matrix
  int mode;        // CheckMode of the change test
  int cls;         // launch (op, tile configuration class) the graph belongs to
  int pad_;
};
struct BatchItem {
  int g;     // graph index
  int tile;  // tile row << 16 | tile column (or word-block column)
};
// One launch: its items as 8 per-XCD lists (items[off[x] .. off[x] +
// len[x])); workgroup b runs item b / 8 of list b % 8.
struct BatchLaunchDev {
  int off[8];
  int len[8];
};
constexpr int BATCH_XCDS = 8;
bool batch_op_supported(int op);
void batch_pad_values(int op, float *pad_x, float *pad_y);
void batch_tile_shape(int op, int cfg, int *bm, int *bn);
// L: the launch's descriptor in device memory; grid: 8 x the longest list;
// class_active: graphs of this launch's class still active (null: all are).
hipError_t launch_batch_round(int op, int cfg, int grid, const BatchItem *items,
                              const BatchLaunchDev *L, const BatchGraphDev *graphs, void *pool,
                              const int *active, int *changed, int parity,
                              const int *class_active, hipStream_t s);
hipError_t launch_batch_reset(const BatchGraphDev *graphs, int *active, int *changed, int *iters,
                              int count, hipStream_t s);
// End of round: retire converged graphs; count the still-active graphs of
// each class into class_active_next[cls].
hipError_t launch_batch_update(const BatchGraphDev *graphs, int *active, int *changed,
                               int *iters, int *class_active_next, int count, int round,
                               int last, hipStream_t s);
hipError_t launch_batch_finalize(const BatchGraphDev *graphs, void *pool, const int *iters,
                                 int count, hipStream_t s);
hipError_t launch_batch_pad(void *buf, int rows, int64_t ld, int n, float pad_x, float pad_y,
                            hipStream_t s);

// Scheduled repeated squaring (sched.hip): one squaring D' = [D (+)] D (x) D
// over padded ld x ld buffers that skips (a) tiles whose row panel and column
// panel did not change in the previous squaring (their result is bit-for-bit
// the previous one, already in the output buffer) and (b) K slabs that hold
// only the (+)-identity in the tile's rows or columns.
constexpr int SCHED_MAX_SLABS = 2048;  // n <= 32768
struct SchedArgs {
  const float *cur;
  float *nxt;
  int64_t ld;
  int n, kpad, epi, mode;
  int tiles_r, tiles_c, slabs;  // slab = 16 consecutive k
  const uint8_t *xocc_in;       // [tiles_r][slabs]: D rows of tile row, 16 columns
  uint8_t *xocc_out;
  const uint8_t *yocc_in;       // [tiles_c][slabs]: 16 D rows, columns of tile column
  uint8_t *yocc_out;
  const int *rowchg_prev;       // tile rows / columns changed by the previous squaring
  const int *colchg_prev;       //   (null: first squaring, every tile is computed)
  int *rowchg;                  // written by this squaring (exact bit test)
  int *colchg;
  const int *conv_prev;         // change-test flag of the previous squaring (null: first)
  int *conv;                    // change-test flag (check mode) of this squaring
  unsigned long long *stats;    // optional: [0] tiles computed, [1] slab-tiles computed,
                                //   [2] / [3] rows / entries of row-sparse squarings,
                                //   [4] / [5] rows / entries of changed-entry squarings
  // Sparse squarings, one workgroup per row over a list of entries built
  // before the squaring (by row, ascending columns; entry = (row << 16) |
  // column, and its value):
  //   row-sparse (*sq_mode 1): the input's entries other than the
  //     (+)-identity, when there are at most rows_limit;
  //   changed-entry (*sq_mode 2): the input's entries that differ from
  //     `prev` (the previous squaring's input), when at most changed_limit;
  //     *sq_mode 3: the same list for a symmetric matrix (*asym == 0): row
  //     passes only, then the result symmetrised in place.
  //   *sq_mode 0: the tiled kernel squares.  sq_mode null: tiled only.
  // A form whose *_auto flag is set is taken only when the cost model says
  // it beats the other forms: cost_slab_tile per (tile, slab) pair the tiled
  // kernel would compute (*work, counted before the squaring), cost_row_entry
  // per listed entry, cost_changed_entry per changed entry, plus cost_list
  // for either sparse form.
  const float *prev;  // null: no changed-entry form for this squaring
  const int *asym;    // 0: the input of the loop is symmetric (null: unknown)
  int rows_limit, changed_limit;
  int rows_auto, changed_auto;
  float cost_slab_tile, cost_row_entry, cost_changed_entry, cost_list;
  unsigned long long *work;
  int *sq_mode;
  int *rowptr;        // n + 1
  int *chgcnt;        // n
  uint32_t *lkj;
  float *lval;
};
hipError_t launch_sched_square(int op, int cfg, const SchedArgs &a, hipStream_t s);
// *asym = 1 unless the n x n matrix D equals its transpose bit for bit
// (*asym must be 0 beforehand).
hipError_t launch_sched_symmetry(const float *D, int64_t ld, int n, int *asym, hipStream_t s);
// After a sparse squaring (nothing after a tiled one): occupancy maps of
// its output.  With a == null: of `D` for the first squaring.
hipError_t launch_sched_occupancy(int op, int cfg, const float *D, int64_t ld, int n,
                                  int tiles_r, int tiles_c, int slabs, uint8_t *xocc,
                                  uint8_t *yocc, const SchedArgs *a, hipStream_t s);
hipError_t launch_fill_f32(float *p, float v, int64_t n, hipStream_t s);
// Up to 8 fills of 32-bit words in one launch (one kernel instead of one per
// buffer: for small problems the launches, not the stores, take the time).
struct FillSegments {
  uint32_t *p[8];
  int64_t n[8];
  uint32_t v[8];
  int k = 0;
  void add(void *ptr, int64_t count, uint32_t value) {
    p[k] = static_cast<uint32_t *>(ptr);
    n[k] = count;
    v[k] = value;
    ++k;
  }
};
hipError_t launch_fill_multi(const FillSegments &f, hipStream_t s);
hipError_t launch_permute_copy(float *dst, int64_t ldd, const float *src, int64_t lds, int n,
                               const int *perm, int inverse, hipStream_t s);

// Learned schedule (learned.hip).
size_t graph_features_workspace(int n);
// With model (params, then the normalisation: simd2_converge_learned's
// layout), pred receives its prediction from the same launch.
hipError_t launch_graph_features(int op, int n, const float *X, int64_t ld, int epi, float *f,
                                 void *ws, hipStream_t s, const float *model = nullptr,
                                 float *pred = nullptr);
// after_features: launch_graph_features already ran on this workspace (its
// packed adjacency and search from vertex 0 are reused).
hipError_t launch_graph_levels(int n, const float *X, int64_t ld, int *level, void *ws,
                               bool after_features, hipStream_t s);
int mlp_param_count();
size_t mlp_workspace(int B);
hipError_t launch_mlp_forward(const float *P, const float *x, int B, float *y, const float *norm,
                              hipStream_t s);
hipError_t launch_mlp_gradient(const float *P, const float *x, const float *t, const float *w,
                               int B, float *grad, float *loss, void *ws, hipStream_t s);
hipError_t launch_mlp_adam(float *P, float *m, float *v, const float *g, float lr, float c1,
                           float c2, hipStream_t s);

// Element-wise helpers (converge.hip).
hipError_t launch_comp_update_f32(float *prev, float *cur, int *flag,
                                  int64_t n, int mode, hipStream_t s);
hipError_t launch_comp_update_i32(int *prev, const int *cur, int *flag,
                                  int64_t n, hipStream_t s);
hipError_t launch_fill_i32(int *p, int v, int64_t n, hipStream_t s);
hipError_t launch_copy_2d(void *dst, int64_t ldd, const void *src,
                          int64_t lds, int rows, int row_bytes, hipStream_t s);

}  // namespace simd2

/*
 * simd2 -- batched repeated squaring: thousands of independent graphs of
 * uneven size solved together (C ABI of libsimd2.so).
 *
 * The reference solves one graph per process with the apps' loop
 * (apps/apsp/cuASR/apsp_cuASR.cu:37-54, apps/mcp/cuASR/mcp_cuASR.cu,
 * apps/mst/cuASR/mst_cuASR.cu, apps/gtc/cuASR/gtc_cuASR.cu:40-51: square,
 * compare, repeat until unchanged).  This interface has no counterpart
 * there; per graph it computes exactly what simd2_converge_f32 /
 * simd2_converge_orand_packed compute (same squarings, same change test,
 * same iteration count, bit-identical matrices), for a whole batch in one
 * hipGraph launch.
 *
 * Life cycle:
 *   simd2_batch_create     layout, device memory, default schedule (allocates)
 *   simd2_batch_upload     copy each graph's matrix in (host or device memory)
 *   simd2_batch_autotune   optional: time tile shapes per (op, size bin) on
 *                          the uploaded data and fix the schedule
 *   simd2_batch_solve      one hipGraphLaunch on `stream` (asynchronous)
 *   simd2_batch_iterations squarings per graph (synchronises `stream`)
 *   simd2_batch_download   copy results out
 *   simd2_batch_destroy
 * One batch object must not be solved concurrently from two streams.
 */
#ifndef SIMD2_BATCH_H
#define SIMD2_BATCH_H

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct simd2_batch simd2_batch;

/* One graph of the batch.
 *   n            vertices (0 allowed: nothing to do, 0 iterations)
 *   op           SIMD2_MIN_PLUS, SIMD2_MAX_PLUS, SIMD2_MAX_MIN, SIMD2_MAX_MUL,
 *                SIMD2_MIN_MUL, SIMD2_MIN_MAX (FP32, n x n row-major) or
 *                SIMD2_OR_AND (bit-packed rows: ceil(n/32) uint32 words per
 *                row, bit j of word w is column 32 w + j)
 *   do_epilogue  1: X <- X (+) X (x) X  (APSP, APLP, MCP, maxrp, minrp, GTC)
 *                0: X <- X (x) X        (MST)
 *   check_mode   SIMD2_CHECK_EXACT / _THRESHOLD / _APLP (converge.h);
 *                ignored for SIMD2_OR_AND, whose test is exact. */
typedef struct simd2_batch_graph {
  int n;
  int op;
  int do_epilogue;
  int check_mode;
} simd2_batch_graph;

/* Creates a batch of `count` graphs.  max_rounds caps the squarings per
 * graph (0 = ceil(log2(max n - 1)) + 3, enough for any graph whose values
 * settle once all shortest/widest paths are found).  Allocates the
 * matrix pool (two padded buffers per graph) and state on the current
 * device and captures the default schedule.  Returns SIMD2_SUCCESS and the
 * batch in *out. */
int simd2_batch_create(const simd2_batch_graph *graphs, int count, int max_rounds,
                       simd2_batch **out);

/* Releases everything simd2_batch_create allocated (NULL is a no-op). */
int simd2_batch_destroy(simd2_batch *batch);

/* Copies graph g's matrix in from `src` (host or device memory; leading
 * dimension ld elements, or words for or_and), asynchronously on `stream`. */
int simd2_batch_upload(simd2_batch *batch, int g, const void *src, int64_t ld,
                       hipStream_t stream);

/* Copies graph g's matrix (after a solve: the result) out to `dst`. */
int simd2_batch_download(simd2_batch *batch, int g, void *dst, int64_t ld,
                         hipStream_t stream);

/* Device address and row stride (elements / words) of graph g's matrix
 * inside the pool, for callers that fill or read it in place.  Only the
 * live n x n block (n x ceil(n/32) words) may be written. */
int simd2_batch_buffer(simd2_batch *batch, int g, void **ptr, int64_t *ld);

/* Times the tile shapes of every (op, size bin) on the uploaded graphs and
 * whole first rounds of the candidate schedules, then whole solves of the
 * schedule in use and the best candidates on a temporary device copy of
 * the uploaded matrices (allocated for the call; skipped when that memory
 * is not available), keeps the fastest and re-captures the solve graph.
 * Leaves the uploaded matrices unchanged; synchronises `stream`. */
int simd2_batch_autotune(simd2_batch *batch, hipStream_t stream);

/* Solves every graph in place: one hipGraphLaunch on `stream`. */
int simd2_batch_solve(simd2_batch *batch, hipStream_t stream);

/* After a solve: squarings per graph into iters[count] (counting the final
 * one that changed nothing, as the reference's loop counts) and, if
 * `unconverged` is not NULL, the number of graphs that still changed at
 * max_rounds.  Synchronises `stream`. */
int simd2_batch_iterations(simd2_batch *batch, int *iters, int *unconverged,
                           hipStream_t stream);

/* The schedule in use: tile configuration per graph into cfg[count] (the
 * simd2_srgemm_ex numbering; -1 for or_and and empty graphs) and the number
 * of kernel launches per round into *launches. */
int simd2_batch_schedule(simd2_batch *batch, int *cfg, int *launches);

#ifdef __cplusplus
}
#endif

#endif /* SIMD2_BATCH_H */

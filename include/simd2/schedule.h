/*
 * simd2 -- scheduled repeated squaring (C ABI of libsimd2.so).
 *
 * The apps' converge loop (apps/apsp/cuASR/apsp_cuASR.cu:37-54, the same
 * shape in aplp / mcp / mst: square, compare, read the flag, repeat) with
 * work left out that provably has no effect on the result:
 *   - tiles whose row panel and column panel did not change in the previous
 *     squaring (their result is the value they already hold);
 *   - 16-deep K slabs that hold only the (+)-identity in a tile's rows or
 *     columns (their terms are identity (x) y = identity);
 *   - while the matrix is sparse, every term with an identity operand: the
 *     squaring runs row by row over the non-identity entries;
 *   - with the epilogue on, every term whose two operands kept their bits in
 *     the previous squaring (it is already inside the entry it would update:
 *     D = D_prev (+) D_prev (x) D_prev and (+) is min or max), when few
 *     entries changed; that squaring runs row by row over the changed ones;
 * and with the flag read deferred: `predicted_itr` squarings are enqueued
 * before the host first reads the change flags.  Squarings enqueued after
 * the one that changed nothing return at once on the device, so the matrix
 * and the iteration count are exactly those of simd2_converge_f32.
 *
 * The schedule (tile shape, squarings to enqueue) is an input: a fixed
 * choice, or the learned one of learned.h.
 */
#ifndef SIMD2_SCHEDULE_H
#define SIMD2_SCHEDULE_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct simd2_schedule {
  int tile_config;   /* simd2_srgemm_ex numbering (0 128x128, 1 64x64, 2 128x64);
                        -1: the automatic choice of simd2_pick_tile_config */
  int predicted_itr; /* squarings enqueued before the first flag read;
                        0: read the flag after every squaring */
  int row_sparse;    /* a squaring whose input holds at most row_sparse * n
                        entries other than the (+)-identity runs row by row
                        over those entries (one pass over row k per entry
                        (i, k)) instead of tile by tile; 0: never; -1: the
                        default, 128; at most 1024 */
  int changed_sparse; /* a squaring after one that changed at most
                        changed_sparse * n entries runs over those entries
                        only (needs do_epilogue and check_mode != APLP);
                        0: never; -1: the default, 64; at most 1024 */
} simd2_schedule;

typedef struct simd2_sched_report {
  int host_syncs;                          /* flag reads (host round trips) */
  int squarings_enqueued;                  /* including ones that returned at once */
  unsigned long long tiles_total;          /* tiles the plain loop computes */
  unsigned long long tiles_computed;       /* tiles this run computed */
  unsigned long long slab_tiles_total;     /* (tile, 16-deep K slab) pairs, plain loop */
  unsigned long long slab_tiles_computed;  /* pairs this run computed */
  unsigned long long rows_sparse;          /* rows squared row by row */
  unsigned long long row_entries;          /* row entries those passes visited */
  unsigned long long changed_rows;         /* rows squared over changed entries */
  unsigned long long changed_entries;      /* changed entries of their own rows */
} simd2_sched_report;

/* Bytes of workspace simd2_converge_scheduled needs (two padded copies of
 * the matrix plus per-tile maps and flags); 0 for invalid arguments. */
size_t simd2_converge_scheduled_workspace(int n, int tile_config, int max_itr);

/* Repeated squaring of the n x n row-major X (leading dimension ldx) in place
 * until no element changes (check_mode of converge.h) or max_itr squarings.
 * op: SIMD2_MIN_PLUS, SIMD2_MAX_PLUS, SIMD2_MAX_MIN or SIMD2_MIN_MAX (entries
 * must not make (x) produce NaN: no -inf with min_plus, no +inf with
 * max_plus).  perm: optional device array of n vertex indices; the loop runs
 * on X[perm][perm] and the result is mapped back (a vertex order with
 * locality lets more tiles and slabs be skipped; the result does not depend
 * on it).  schedule: NULL for { -1, 0, -1, -1 }.  Returns like simd2_converge_f32;
 * *num_itr and *report (both optional) describe the run.  Synchronises
 * `stream`. */
int simd2_converge_scheduled(int op, int n, float *X, int ldx, int do_epilogue,
                             int check_mode, int max_itr, const simd2_schedule *schedule,
                             const int *perm, void *workspace, size_t workspace_bytes,
                             hipStream_t stream, int *num_itr, simd2_sched_report *report);

#ifdef __cplusplus
}
#endif

#endif /* SIMD2_SCHEDULE_H */

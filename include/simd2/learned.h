/*
 * simd2 -- the learned schedule of the converge loop (C ABI of libsimd2.so).
 *
 * No counterpart in the reference, which always runs the plain loop
 * (apps/apsp/cuASR/apsp_cuASR.cu:37-54).  Pieces:
 *   - simd2_graph_features: 24 cheap features of an adjacency matrix
 *     (size, density, degree and weight statistics, locality, depths of
 *     breadth-first searches from 8 sampled sources), on the GPU;
 *   - simd2_graph_order: a vertex order with locality (breadth-first levels
 *     from a pseudo-peripheral vertex) for simd2_converge_scheduled;
 *   - a 24 -> 64 -> 64 -> 2 MLP (ReLU) with hand-written forward, backward
 *     and Adam kernels, predicting (squarings to convergence, tile shape);
 *   - simd2_converge_learned: features -> MLP -> schedule -> scheduled loop.
 * The loop's own change test still decides when to stop, so results are
 * those of simd2_converge_f32 bit for bit whatever the network predicts.
 */
#ifndef SIMD2_LEARNED_H
#define SIMD2_LEARNED_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#include "schedule.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SIMD2_FEATURES 24
#define SIMD2_MLP_HIDDEN 64
#define SIMD2_MLP_OUT 2
/* Parameter vector layout: W1[64][24] b1[64] W2[64][64] b2[64] W3[2][64] b3[2]. */
#define SIMD2_MLP_PARAMS (64 * 24 + 64 + 64 * 64 + 64 + 2 * 64 + 2)
/* A model: the parameters, then feature mean[24] and scale[24] (inputs are
 * (f - mean) * scale), then output scale[2] and offset[2] (prediction =
 * network output * scale + offset). */
#define SIMD2_MODEL_FLOATS (SIMD2_MLP_PARAMS + 2 * SIMD2_FEATURES + 2 * SIMD2_MLP_OUT)

/* Workspace bytes of simd2_graph_features / simd2_graph_order for n vertices. */
size_t simd2_graph_features_workspace(int n);

/* features[24] (device) of the n x n row-major adjacency X (ldx): an entry
 * is an edge when it is off the diagonal and differs from the op's
 * (+)-identity.  Feature list: see the oracle restatement
 * (oracle/learned.py) and learned.hip. */
int simd2_graph_features(int op, int n, const float *X, int ldx, int do_epilogue,
                         float *features, void *workspace, size_t workspace_bytes,
                         hipStream_t stream);

/* perm[n] (device): vertices sorted by breadth-first level from a
 * pseudo-peripheral vertex (unreached ones last), ties by index. */
int simd2_graph_order(int op, int n, const float *X, int ldx, int *perm, void *workspace,
                      size_t workspace_bytes, hipStream_t stream);

/* y[B][2] = MLP(x[B][24]) with the parameters `params` (device). */
int simd2_mlp_forward(const float *params, const float *x, int B, float *y, hipStream_t stream);

/* Workspace bytes of simd2_mlp_gradient / simd2_mlp_train_step for batch B. */
size_t simd2_mlp_workspace(int B);

/* Loss L = (1/B) sum_b sum_o w[o] (y[b][o] - t[b][o])^2 and its gradient
 * grad[SIMD2_MLP_PARAMS] with respect to params (all device pointers;
 * loss: one float).  Deterministic (fixed reduction order). */
int simd2_mlp_gradient(const float *params, const float *x, const float *t, const float *w,
                       int B, float *grad, float *loss, void *workspace, size_t workspace_bytes,
                       hipStream_t stream);

/* One Adam step (beta1 0.9, beta2 0.999, eps 1e-8, bias-corrected with the
 * 1-based step number) on params with moments m, v, all device arrays of
 * SIMD2_MLP_PARAMS floats; loss as in simd2_mlp_gradient. */
int simd2_mlp_train_step(float *params, float *m, float *v, int step, float lr, const float *x,
                         const float *t, const float *w, int B, float *loss, void *workspace,
                         size_t workspace_bytes, hipStream_t stream);

/* Workspace bytes of simd2_converge_learned. */
size_t simd2_converge_learned_workspace(int n, int max_itr);

/* The converge loop under the learned schedule: features and vertex order of
 * X, the model's prediction (one 8-byte read), then
 * simd2_converge_scheduled with that order and schedule (the predicted tile
 * shape up to n = 4096, the sizes the shipped model was trained on; the
 * automatic tile choice above).  model: device array of
 * SIMD2_MODEL_FLOATS.  *used (optional) receives the schedule.
 * Returns like simd2_converge_f32. */
int simd2_converge_learned(int op, int n, float *X, int ldx, int do_epilogue, int check_mode,
                           int max_itr, const float *model, void *workspace,
                           size_t workspace_bytes, hipStream_t stream, int *num_itr,
                           simd2_schedule *used, simd2_sched_report *report);

#ifdef __cplusplus
}
#endif

#endif /* SIMD2_LEARNED_H */

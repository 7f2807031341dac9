/*
 * sbc_costmodel.h -- learned cost model for scheduling the windowed update
 * (no counterpart in the reference; its BALANCE estimate, src/streamingbc.c:
 * 98-122, plays the same role with a hand-written formula).
 *
 * A 12-64-64-1 MLP (ReLU) predicts log(1 + affected vertices) of one
 * (root, insertion) item from 12 cheap features read from the forest before
 * the window runs.  It is trained on the GPU by the library's own FP32
 * kernels (mini-batch Adam on mean squared error) and evaluated with BF16
 * weights on MFMA (FP32 accumulation).  Predictions only order and bin roots
 * (sbcBatchConfig.root_order / heavy_roots), so results never depend on it.
 *
 * Features of item (root r, insertion (u,v)), with s the deeper and p the
 * shallower endpoint in r's tree (all zero when dist[u] == dist[v]):
 *   0 1 (nonzero gap)          1 |gap| >= 2 (movement case)
 *   2 min(|gap|,16)/16         3 min(dist[s],16)/16      4 min(dist[p],16)/16
 *   5 log1p(deg s)/10          6 log1p(deg p)/10
 *   7 log1p(sigma[p])/20       8 log1p(sigma[s])/20
 *   9 log1p(delta[s])/10       10 log1p(delta[p])/10
 *   11 root history (a per-root value, e.g. the decayed mean label of the
 *      root's recent items)/10: caller-supplied, or kept on the device by
 *      streamingBCSetRootHistory / streamingBCUpdateRootHistory
 */
#ifndef SBC_COSTMODEL_H
#define SBC_COSTMODEL_H

#include <stdint.h>

#include "streamingbc_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SBC_COST_FEATURES 12

typedef struct sbcCostModel sbcCostModel;

/* Parameter count of the flat FP32 parameter vector:
 * W1[16][64] (rows 12..15 unused, kept 0), b1[64], W2[64][64], b2[64], w3[64], b3. */
int32_t sbcCostModelNumParams(void);

sbcStatus sbcCostModelCreate(uint64_t seed, sbcCostModel **out);
void sbcCostModelFree(sbcCostModel *m);
sbcStatus sbcCostModelGetParams(sbcCostModel *m, float *params);
sbcStatus sbcCostModelSetParams(sbcCostModel *m, const float *params);

/* Mini-batch Adam on mean squared error over (X[n][12], y[n]); samples are
 * visited in a seeded random order each epoch.  final_loss (optional): mean
 * squared error over all n after training. */
sbcStatus sbcCostModelTrain(sbcCostModel *m, const float *X, const float *y, int64_t n, int32_t epochs,
                            int32_t batch, float lr, uint64_t seed, float *final_loss);

/* Full-batch gradient of the mean squared error (for gradient checks). */
sbcStatus sbcCostModelGradients(sbcCostModel *m, const float *X, const float *y, int64_t n, float *grad,
                                float *loss);

/* Predictions: BF16 weights on MFMA with FP32 accumulation, or the FP32
 * reference path. */
sbcStatus sbcCostModelPredict(sbcCostModel *m, const float *X, int64_t n, float *out);
sbcStatus sbcCostModelPredictFP32(sbcCostModel *m, const float *X, int64_t n, float *out);

/* Features of every (insertion i, root index r) item of a window, computed on
 * the device from the forest's current state: X[(i * K + r) * 12 + f].
 * root_hist: K per-root values for feature 11 (NULL: zeros). */
sbcStatus streamingBCWindowFeatures(bcForest *forest, struct stinger *G, extraArraysPerThread **eAPT,
                                    const int64_t *u, const int64_t *v, int64_t count, const float *root_hist,
                                    float *X);

/* Schedules the next window on the device: features of every (insertion,
 * root) item, BF16-MFMA predictions, per-root window cost
 * sum(expm1(max(pred, 0)) + item_overhead), roots sorted by descending cost
 * into the context's root order (used by the next batched call with
 * use_context_order = 1; any batched call consumes it, so schedule before
 * each call that uses it).  Ties keep root-array order.  heavy_out (optional): roots with cost >=
 * light_max (light_max <= 0: -1, every root heavy).  pred_out (optional):
 * the predictions, pred_out[i * K + r].  Needs K <= 4096. */
sbcStatus streamingBCScheduleWindow(bcForest *forest, struct stinger *G, extraArraysPerThread **eAPT,
                                    sbcCostModel *model, const int64_t *u, const int64_t *v, int64_t count,
                                    const float *root_hist, float light_max, float item_overhead,
                                    int32_t *heavy_out, float *pred_out);
/* (root_hist NULL in streamingBCScheduleWindow: the context's device history
 * below when it holds K roots, else zeros.) */

/* The history feature kept on the device, so scheduling a window copies no
 * item records: SetRootHistory uploads K starting values (e.g. those learned
 * while collecting training data); UpdateRootHistory folds the last batched
 * call's last window in: hist[r] = decay * hist[r] + (1 - decay) * mean over
 * the window's items of log1p(affected). */
sbcStatus streamingBCSetRootHistory(extraArraysPerThread **eAPT, int64_t K, const float *hist);
sbcStatus streamingBCUpdateRootHistory(extraArraysPerThread **eAPT, int64_t K, float decay);

#ifdef __cplusplus
}
#endif

#endif

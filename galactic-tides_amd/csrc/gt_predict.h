// Learned list lengths for the one-pass walk (see gt_predict.hip).
#pragma once

#include "gt_internal.h"

namespace gth {

constexpr int kPredFeatures = 8;
constexpr int kPredHidden = 64;
constexpr int kPredOutputs = 2;  // log(1 + cells), log(1 + bodies)
constexpr int kPredParams = kPredHidden * kPredFeatures + kPredHidden + kPredHidden * kPredHidden +
                            kPredHidden + kPredOutputs * kPredHidden + kPredOutputs;
constexpr int kPredStats = 2 * (kPredFeatures + kPredOutputs);  // means, then deviations
constexpr int kPredSampleFloats = kPredFeatures + kPredOutputs;

// feat: kPredFeatures floats per group (raw, not standardised); group_node:
// the node each group was selected from, node_depth: every node's depth.
void launch_group_features(const gt_group *groups, const unsigned long long *frame,
                           uint32_t group_cap, uint32_t node_cap, const double *tpos,
                           const int4 *lnk, const double *geom, const uint32_t *group_root,
                           const uint32_t *group_node, const uint32_t *node_depth, float *feat,
                           hipStream_t s);
// Per-group slots (multiples of 4; 0 past the group count) and the predicted
// log(1 + length) per group and output (pred, 2 floats per group).  stats
// holds kPredStats + 1 floats: the standardisation, then the log-space margin.
void launch_predict_slots(const float *feat, const unsigned long long *frame, uint32_t group_cap,
                          const float *params, const float *stats, uint32_t *slot_c,
                          uint32_t *slot_b, float *pred, hipStream_t s);
// *out = bit pattern of the largest |log(1 + true) - predicted| over the samples.
void launch_sample_residual_max(const float *samples, uint32_t n, const float *params,
                                const float *stats, unsigned int *out, hipStream_t s);
// Sets kAbortSlots when frame[kPredCells] / frame[kPredBodies] exceed the buffers.
void launch_check_slot_totals(unsigned long long *frame, unsigned long long cap_c,
                              unsigned long long cap_b, hipStream_t s);
// err += 1e6 * mean over outputs of |log(1 + true) - predicted| per group.
void launch_prediction_error(const unsigned long long *frame, uint32_t group_cap, const float *pred,
                             const uint32_t *cc, const uint32_t *bc, unsigned long long *err,
                             hipStream_t s);
void launch_collect_samples(const unsigned long long *frame, uint32_t group_cap, uint32_t stride,
                            const float *feat, const uint32_t *cc, const uint32_t *bc,
                            uint32_t base, uint32_t capacity, uint32_t count, float *samples,
                            hipStream_t s);
void launch_sample_stats(const float *samples, uint32_t n, float *stats, hipStream_t s);
// The network alone: y (n x kPredOutputs, standardised) from normalised x
// (n x kPredFeatures).
void launch_mlp_forward(const float *x, uint32_t n, const float *params, float *y, hipStream_t s);
// One Adam iteration over blocks x 64 samples; grad holds blocks x kPredParams
// floats, loss blocks floats (sums of squared standardised errors).
void launch_train_iteration(const float *samples, uint32_t n, const float *stats, float *params,
                            float *m, float *v, float *grad, float *loss, uint32_t blocks,
                            uint32_t iter, float lr, hipStream_t s);

}  // namespace gth

// Learned interaction-list lengths (the one-pass walk's per-group slots).
//
// A small FP32 MLP, 8 -> 64 -> 64 -> 2 with ReLU, predicts log(1 + cells) and
// log(1 + bodies) of every group's lists from eight features of the group
// and its place in the tree.  The predictions size each group's slots in the
// one-pass walk (scanned to offsets, so the lists stay packed); a group whose
// lists outgrow their slots is caught by the walk and the lists are rewritten
// at exact offsets, as with uniform slots.
//
// Training (Adam, hand-written kernels): samples (features, true lengths) of
// a spread of groups are gathered from the first evaluations; features and
// targets are standardised with their means and deviations; each iteration
// takes a minibatch of 64 samples per workgroup, every workgroup writes the
// gradient of its samples, and the gradients are added in workgroup order
// (deterministic) before the Adam update.
#include "gt_predict.h"

#include <algorithm>

namespace gth {
namespace {

constexpr int kBlock = 256;
constexpr int kF = kPredFeatures, kH = kPredHidden, kO = kPredOutputs;
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kH * kF, kOffW2 = kOffB1 + kH,
              kOffB2 = kOffW2 + kH * kH, kOffW3 = kOffB2 + kH, kOffB3 = kOffW3 + kO * kH;
static_assert(kOffB3 + kO == kPredParams, "parameter layout");
constexpr int kTrainBlock = 64;  // samples per workgroup in training (one wave)

inline uint32_t blocks_for(size_t n, int b = kBlock) {
  return static_cast<uint32_t>((n + b - 1) / b);
}

// ---- features ------------------------------------------------------------------------

// f0 log size of the group's bounding box, f1 log size of the node the group
// was selected from (the chunked node, a lone packed child, or the parent of
// packed siblings: group_node), f2 that node's depth, f3 targets, f4 log
// distance of the group centre from its root's centre, f5 log bodies per
// volume of that node, f6 f7 the bounding box's middle and smallest extent
// over its largest.
__global__ __launch_bounds__(kBlock) void group_features(
    const gt_group *__restrict__ groups, const unsigned long long *__restrict__ frame,
    uint32_t group_cap, uint32_t node_cap, const double *__restrict__ tpos,
    const int4 *__restrict__ lnk, const double *__restrict__ geom,
    const uint32_t *__restrict__ group_root, const uint32_t *__restrict__ group_node,
    const uint32_t *__restrict__ node_depth, float *__restrict__ feat) {
  const uint32_t ng = bounded_count(frame + frame::kGroups, group_cap);
  const uint32_t nn = bounded_count(frame + frame::kNodes, node_cap);
  const uint32_t g = blockIdx.x * kBlock + threadIdx.x;
  if (g >= ng || nn == 0) return;
  const gt_group G = groups[g];
  double lo[3], hi[3];
  for (int a = 0; a < 3; ++a) lo[a] = hi[a] = tpos[3 * static_cast<size_t>(G.begin) + a];
  for (uint32_t k = G.begin; k < G.end; ++k)
    for (int a = 0; a < 3; ++a) {
      const double v = tpos[3 * static_cast<size_t>(k) + a];
      lo[a] = fmin(lo[a], v);
      hi[a] = fmax(hi[a], v);
    }
  double e[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
  const uint32_t root = group_root ? group_root[g] : 0u;
  const uint32_t gn = group_node[g];
  const uint32_t x = gn < nn ? gn : 0u;
  const uint32_t depth = node_depth[x];
  const double *gx = geom + static_cast<size_t>(kGeomStride) * x;
  const double *gr = geom + static_cast<size_t>(kGeomStride) * (root < nn ? root : 0u);
  const double cell = 2.0 * gx[3], root_size = 2.0 * gr[3];
  const int4 LX = lnk[x];
  const double in_cell = static_cast<double>(LX.w - LX.z);
  const double emax = fmax(e[0], fmax(e[1], e[2])), emin = fmin(e[0], fmin(e[1], e[2]));
  const double emid = e[0] + e[1] + e[2] - emax - emin;
  double d2 = 0.0;
  for (int a = 0; a < 3; ++a) {
    const double d = 0.5 * (lo[a] + hi[a]) - gr[a];
    d2 += d * d;
  }
  const double tiny = 1e-6 * root_size + 1e-300;
  float *f = feat + static_cast<size_t>(kF) * g;
  f[0] = static_cast<float>(log(emax + tiny));
  f[1] = static_cast<float>(log(cell + tiny));
  f[2] = static_cast<float>(depth);
  f[3] = static_cast<float>(G.end - G.begin);
  f[4] = static_cast<float>(log(sqrt(d2) + tiny));
  f[5] = static_cast<float>(log(in_cell / ((cell + tiny) * (cell + tiny) * (cell + tiny))));
  f[6] = static_cast<float>(emid / (emax + tiny));
  f[7] = static_cast<float>(emin / (emax + tiny));
}

// ---- inference -------------------------------------------------------------------------

// Normalised features -> the two standardised outputs, for one sample.  P is
// the parameter vector in global memory: every lane reads the same weights at
// wave-uniform addresses, so they come through the scalar cache into SGPRs
// and are scalar operands of the FMAs (no LDS read per weight).
__device__ inline void mlp_forward(const float *__restrict__ P, const float *x, float *y) {
  float h1[kH];
#pragma unroll
  for (int i = 0; i < kH; ++i) {
    float a = P[kOffB1 + i];
#pragma unroll
    for (int k = 0; k < kF; ++k) a = fmaf(P[kOffW1 + i * kF + k], x[k], a);
    h1[i] = fmaxf(a, 0.0f);
  }
  float out[kO];
#pragma unroll
  for (int o = 0; o < kO; ++o) out[o] = P[kOffB3 + o];
  for (int i = 0; i < kH; ++i) {
    float a = P[kOffB2 + i];
#pragma unroll
    for (int j = 0; j < kH; ++j) a = fmaf(P[kOffW2 + i * kH + j], h1[j], a);
    const float h2 = fmaxf(a, 0.0f);
#pragma unroll
    for (int o = 0; o < kO; ++o) out[o] = fmaf(P[kOffW3 + o * kH + i], h2, out[o]);
  }
#pragma unroll
  for (int o = 0; o < kO; ++o) y[o] = out[o];
}

// stats: mean[kF + kO], then deviation[kF + kO] (features, then targets).
__device__ inline void normalise(const float *__restrict__ f, const float *__restrict__ stats,
                                 float *x) {
#pragma unroll
  for (int k = 0; k < kF; ++k) x[k] = (f[k] - stats[k]) / stats[kF + kO + k];
}

// Slots of every group from the predicted lengths: expm1(prediction +
// margin), at least 4, rounded up to a multiple of 4 (slots start 16-byte
// aligned for the walk's vector stores).  Groups past the count get 0, so the
// scans over group_cap entries add up.  pred keeps the predicted log(1 + n)
// for the error statistics.
__global__ __launch_bounds__(kBlock) void predict_slots(
    const float *__restrict__ feat, const unsigned long long *__restrict__ frame,
    uint32_t group_cap, const float *__restrict__ params, const float *__restrict__ stats,
    uint32_t *__restrict__ slot_c, uint32_t *__restrict__ slot_b, float *__restrict__ pred) {
  __shared__ float S[kPredStats + 1];
  if (threadIdx.x < kPredStats + 1) S[threadIdx.x] = stats[threadIdx.x];
  __syncthreads();
  const float margin = S[kPredStats];  // in device memory: captured graphs see its updates
  const uint32_t ng = bounded_count(frame + frame::kGroups, group_cap);
  const uint32_t g = blockIdx.x * kBlock + threadIdx.x;
  if (g >= group_cap) return;
  if (g >= ng) {
    slot_c[g] = slot_b[g] = 0u;
    return;
  }
  float x[kF], y[kO];
  normalise(feat + static_cast<size_t>(kF) * g, S, x);
  mlp_forward(params, x, y);
  uint32_t slots[kO];
#pragma unroll
  for (int o = 0; o < kO; ++o) {
    const float lg = y[o] * S[kF + kO + kF + o] + S[kF + o];  // log(1 + length)
    pred[static_cast<size_t>(kO) * g + o] = lg;
    const float len = expm1f(fminf(lg + margin, 30.0f));
    const uint32_t sl = static_cast<uint32_t>(fminf(fmaxf(ceilf(len), 4.0f), 1.0e9f));
    slots[o] = (sl + 3u) & ~3u;
  }
  slot_c[g] = slots[0];
  slot_b[g] = slots[1];
}

// |log(1 + true length) - prediction| summed over the groups (atomics on
// integers: micro-units, so the sum does not depend on the order).  A few
// hundred blocks stride over the groups and each adds its total with one
// atomic (one per wave made the atomics on the one word the kernel's cost).
constexpr uint32_t kErrorBlocks = 256;

__global__ __launch_bounds__(kBlock) void prediction_error(
    const unsigned long long *__restrict__ frame, uint32_t group_cap,
    const float *__restrict__ pred, const uint32_t *__restrict__ cc,
    const uint32_t *__restrict__ bc, unsigned long long *__restrict__ err) {
  __shared__ unsigned long long part[kBlock / 64];
  const uint32_t ng = bounded_count(frame + frame::kGroups, group_cap);
  unsigned long long e = 0;
  for (uint32_t g = blockIdx.x * kBlock + threadIdx.x; g < ng; g += kErrorBlocks * kBlock) {
    const float ec = fabsf(log1pf(static_cast<float>(cc[g])) - pred[2 * static_cast<size_t>(g)]);
    const float eb =
        fabsf(log1pf(static_cast<float>(bc[g])) - pred[2 * static_cast<size_t>(g) + 1]);
    e += static_cast<unsigned long long>(1e6f * (ec + eb) * 0.5f);
  }
  for (int off = 32; off > 0; off >>= 1) e += __shfl_xor(e, off, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x / 64] = e;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < kBlock / 64; ++w) t += part[w];
    if (t) atomicAdd(err, t);
  }
}

// Walk ahead: the scanned slot totals must fit the list buffers.
__global__ void check_slot_totals(unsigned long long *frame, unsigned long long cap_c,
                                  unsigned long long cap_b) {
  if (frame[frame::kPredCells] > cap_c || frame[frame::kPredBodies] > cap_b)
    frame[frame::kStatus] |= kAbortSlots;
}

// Largest |log(1 + true) - prediction| over the samples (non-negative floats
// compare like their bit patterns, so an integer max is exact and
// order-independent).
__global__ __launch_bounds__(kBlock) void sample_residual_max(const float *__restrict__ samples,
                                                              uint32_t n,
                                                              const float *__restrict__ params,
                                                              const float *__restrict__ stats,
                                                              unsigned int *__restrict__ out) {
  __shared__ float S[kPredStats];
  if (threadIdx.x < kPredStats) S[threadIdx.x] = stats[threadIdx.x];
  __syncthreads();
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const float *smp = samples + static_cast<size_t>(kF + kO) * i;
  float x[kF], y[kO];
  normalise(smp, S, x);
  mlp_forward(params, x, y);
  float r = 0.0f;
  for (int o = 0; o < kO; ++o) {
    const float lg = y[o] * S[kF + kO + kF + o] + S[kF + o];
    r = fmaxf(r, fabsf(lg - smp[kF + o]));
  }
  atomicMax(out, __float_as_uint(r));
}

// ---- training ---------------------------------------------------------------------------

// Samples of this evaluation: every `stride`-th group, written from `base`
// on (modulo the buffer): raw features, then log(1 + cells), log(1 + bodies).
__global__ __launch_bounds__(kBlock) void collect_samples(
    const unsigned long long *__restrict__ frame, uint32_t group_cap, uint32_t stride,
    const float *__restrict__ feat, const uint32_t *__restrict__ cc,
    const uint32_t *__restrict__ bc, uint32_t base, uint32_t capacity,
    float *__restrict__ samples) {
  const uint32_t ng = bounded_count(frame + frame::kGroups, group_cap);
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t g = i * stride;
  if (g >= ng) return;
  float *s = samples + static_cast<size_t>(kF + kO) * ((base + i) % capacity);
  for (int k = 0; k < kF; ++k) s[k] = feat[static_cast<size_t>(kF) * g + k];
  s[kF] = log1pf(static_cast<float>(cc[g]));
  s[kF + 1] = log1pf(static_cast<float>(bc[g]));
}

// One workgroup: mean and deviation of each column over the n samples, sums
// in FP64 in a fixed order.
__global__ __launch_bounds__(kBlock) void sample_stats(const float *__restrict__ samples,
                                                       uint32_t n, float *__restrict__ stats) {
  constexpr int kC = kF + kO;
  __shared__ double sum[kBlock], sq[kBlock];
  for (int col = 0; col < kC; ++col) {
    double s = 0.0, q = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += kBlock) {
      const double v = samples[static_cast<size_t>(kC) * i + col];
      s += v;
      q += v * v;
    }
    sum[threadIdx.x] = s;
    sq[threadIdx.x] = q;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
      if (threadIdx.x < static_cast<unsigned>(w)) {
        sum[threadIdx.x] += sum[threadIdx.x + w];
        sq[threadIdx.x] += sq[threadIdx.x + w];
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      const double mean = n ? sum[0] / n : 0.0;
      const double var = n ? sq[0] / n - mean * mean : 0.0;
      stats[col] = static_cast<float>(mean);
      stats[kC + col] = static_cast<float>(var > 1e-12 ? sqrt(var) : 1.0);
    }
    __syncthreads();
  }
}

// Gradient of the mean squared error (standardised targets) over one
// minibatch: workgroup b takes samples (iter * batch + b * 64 + t) spread
// over the buffer by a fixed odd stride, and writes its gradient to grad[b].
__global__ __launch_bounds__(kTrainBlock) void mlp_gradients(
    const float *__restrict__ samples, uint32_t n, const float *__restrict__ stats,
    const float *__restrict__ params, uint32_t iter, uint32_t batch, float *__restrict__ grad,
    float *__restrict__ loss) {
  // rows padded by one float: a lane's row reads hit distinct banks
  __shared__ float xs[kTrainBlock][kF + 1];
  __shared__ float h1s[kTrainBlock][kH + 1];
  __shared__ float h2s[kTrainBlock][kH + 1];  // then d1
  __shared__ float d2s[kTrainBlock][kH + 1];
  __shared__ float d3s[kTrainBlock][kO];
  constexpr int kC = kF + kO;
  const uint32_t t = threadIdx.x;
  const uint64_t pick = (static_cast<uint64_t>(iter) * batch + blockIdx.x * kTrainBlock + t) *
                        2654435761ull;
  const uint32_t idx = static_cast<uint32_t>(pick % n);
  const float *s = samples + static_cast<size_t>(kC) * idx;
  float x[kF];
  for (int k = 0; k < kF; ++k) x[k] = (s[k] - stats[k]) / stats[kC + k];
  for (int k = 0; k < kF; ++k) xs[t][k] = x[k];
  // forward
  const float *P = params;
  for (int i = 0; i < kH; ++i) {
    float a = P[kOffB1 + i];
    for (int k = 0; k < kF; ++k) a = fmaf(P[kOffW1 + i * kF + k], x[k], a);
    h1s[t][i] = fmaxf(a, 0.0f);
  }
  float out[kO];
  for (int o = 0; o < kO; ++o) out[o] = P[kOffB3 + o];
  for (int i = 0; i < kH; ++i) {
    float a = P[kOffB2 + i];
    for (int j = 0; j < kH; ++j) a = fmaf(P[kOffW2 + i * kH + j], h1s[t][j], a);
    const float h2 = fmaxf(a, 0.0f);
    h2s[t][i] = h2;
    for (int o = 0; o < kO; ++o) out[o] = fmaf(P[kOffW3 + o * kH + i], h2, out[o]);
  }
  // output error (d loss / d out, loss = mean over batch and outputs of err^2)
  float l = 0.0f;
  for (int o = 0; o < kO; ++o) {
    const float target = (s[kF + o] - stats[kF + o]) / stats[kC + kF + o];
    const float err = out[o] - target;
    l += err * err;
    d3s[t][o] = 2.0f * err / static_cast<float>(batch * kO);
  }
  // back through layer 3, then 2
  for (int i = 0; i < kH; ++i) {
    float a = 0.0f;
    for (int o = 0; o < kO; ++o) a = fmaf(P[kOffW3 + o * kH + i], d3s[t][o], a);
    d2s[t][i] = h2s[t][i] > 0.0f ? a : 0.0f;
  }
  __syncthreads();
  // layer-3 gradient (needs h2 before it is overwritten by d1)
  float *gb = grad + static_cast<size_t>(blockIdx.x) * kPredParams;
  for (int p = t; p < kO * kH + kO; p += kTrainBlock) {
    float a = 0.0f;
    if (p < kO * kH) {
      const int o = p / kH, i = p % kH;
      for (int r = 0; r < kTrainBlock; ++r) a = fmaf(d3s[r][o], h2s[r][i], a);
      gb[kOffW3 + p] = a;
    } else {
      const int o = p - kO * kH;
      for (int r = 0; r < kTrainBlock; ++r) a += d3s[r][o];
      gb[kOffB3 + o] = a;
    }
  }
  __syncthreads();
  for (int j = 0; j < kH; ++j) {
    float a = 0.0f;
    for (int i = 0; i < kH; ++i) a = fmaf(P[kOffW2 + i * kH + j], d2s[t][i], a);
    h2s[t][j] = h1s[t][j] > 0.0f ? a : 0.0f;  // d1
  }
  __syncthreads();
  for (int p = t; p < kH * kH + kH; p += kTrainBlock) {
    float a = 0.0f;
    if (p < kH * kH) {
      const int i = p / kH, j = p % kH;
      for (int r = 0; r < kTrainBlock; ++r) a = fmaf(d2s[r][i], h1s[r][j], a);
      gb[kOffW2 + p] = a;
    } else {
      const int i = p - kH * kH;
      for (int r = 0; r < kTrainBlock; ++r) a += d2s[r][i];
      gb[kOffB2 + i] = a;
    }
  }
  for (int p = t; p < kH * kF + kH; p += kTrainBlock) {
    float a = 0.0f;
    if (p < kH * kF) {
      const int i = p / kF, k = p % kF;
      for (int r = 0; r < kTrainBlock; ++r) a = fmaf(h2s[r][i], xs[r][k], a);
      gb[kOffW1 + p] = a;
    } else {
      const int i = p - kH * kF;
      for (int r = 0; r < kTrainBlock; ++r) a += h2s[r][i];
      gb[kOffB1 + i] = a;
    }
  }
  // per-workgroup loss (sum over its samples), for monitoring
  __shared__ float ls[kTrainBlock];
  ls[t] = l;
  __syncthreads();
  if (t == 0) {
    float a = 0.0f;
    for (int r = 0; r < kTrainBlock; ++r) a += ls[r];
    loss[blockIdx.x] = a;
  }
}

// Gradients added in workgroup order, then one Adam step (bias-corrected).
__global__ __launch_bounds__(kBlock) void adam_step(const float *__restrict__ grad,
                                                    uint32_t blocks, float *__restrict__ params,
                                                    float *__restrict__ m, float *__restrict__ v,
                                                    float lr, float c1, float c2) {
  const uint32_t p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= static_cast<uint32_t>(kPredParams)) return;
  float g = 0.0f;
  for (uint32_t b = 0; b < blocks; ++b) g += grad[static_cast<size_t>(b) * kPredParams + p];
  const float mm = 0.9f * m[p] + 0.1f * g;
  const float vv = 0.999f * v[p] + 0.001f * g * g;
  m[p] = mm;
  v[p] = vv;
  params[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + 1e-8f);
}

// The network on normalised inputs, one row per thread (x: n x kF, y: n x kO,
// standardised outputs): the inference of predict_slots without the features
// and slots around it.
__global__ __launch_bounds__(kBlock) void forward_rows(const float *__restrict__ x, uint32_t n,
                                                       const float *__restrict__ params,
                                                       float *__restrict__ y) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  float xi[kF], yi[kO];
#pragma unroll
  for (int k = 0; k < kF; ++k) xi[k] = x[static_cast<size_t>(kF) * i + k];
  mlp_forward(params, xi, yi);
#pragma unroll
  for (int o = 0; o < kO; ++o) y[static_cast<size_t>(kO) * i + o] = yi[o];
}

}  // namespace

void launch_mlp_forward(const float *x, uint32_t n, const float *params, float *y,
                        hipStream_t s) {
  if (!n) return;
  hipLaunchKernelGGL(forward_rows, dim3(blocks_for(n)), dim3(kBlock), 0, s, x, n, params, y);
  GTH_LAUNCH_CHECK();
}

void launch_group_features(const gt_group *groups, const unsigned long long *frame,
                           uint32_t group_cap, uint32_t node_cap, const double *tpos,
                           const int4 *lnk, const double *geom, const uint32_t *group_root,
                           const uint32_t *group_node, const uint32_t *node_depth, float *feat,
                           hipStream_t s) {
  if (!group_cap) return;
  hipLaunchKernelGGL(group_features, dim3(blocks_for(group_cap)), dim3(kBlock), 0, s, groups, frame,
                     group_cap, node_cap, tpos, lnk, geom, group_root, group_node, node_depth,
                     feat);
  GTH_LAUNCH_CHECK();
}

void launch_predict_slots(const float *feat, const unsigned long long *frame, uint32_t group_cap,
                          const float *params, const float *stats, uint32_t *slot_c,
                          uint32_t *slot_b, float *pred, hipStream_t s) {
  if (!group_cap) return;
  hipLaunchKernelGGL(predict_slots, dim3(blocks_for(group_cap)), dim3(kBlock), 0, s, feat, frame,
                     group_cap, params, stats, slot_c, slot_b, pred);
  GTH_LAUNCH_CHECK();
}

void launch_sample_residual_max(const float *samples, uint32_t n, const float *params,
                                const float *stats, unsigned int *out, hipStream_t s) {
  launch_fill_u32(out, 1, 0u, s);
  if (!n) return;
  hipLaunchKernelGGL(sample_residual_max, dim3(blocks_for(n)), dim3(kBlock), 0, s, samples, n,
                     params, stats, out);
  GTH_LAUNCH_CHECK();
}

void launch_check_slot_totals(unsigned long long *frame, unsigned long long cap_c,
                              unsigned long long cap_b, hipStream_t s) {
  hipLaunchKernelGGL(check_slot_totals, dim3(1), dim3(1), 0, s, frame, cap_c, cap_b);
  GTH_LAUNCH_CHECK();
}

void launch_prediction_error(const unsigned long long *frame, uint32_t group_cap, const float *pred,
                             const uint32_t *cc, const uint32_t *bc, unsigned long long *err,
                             hipStream_t s) {
  if (!group_cap) return;
  const uint32_t blocks = std::min<uint32_t>(blocks_for(group_cap), kErrorBlocks);
  hipLaunchKernelGGL(prediction_error, dim3(blocks), dim3(kBlock), 0, s, frame, group_cap, pred,
                     cc, bc, err);
  GTH_LAUNCH_CHECK();
}

void launch_collect_samples(const unsigned long long *frame, uint32_t group_cap, uint32_t stride,
                            const float *feat, const uint32_t *cc, const uint32_t *bc,
                            uint32_t base, uint32_t capacity, uint32_t count, float *samples,
                            hipStream_t s) {
  if (!count) return;
  hipLaunchKernelGGL(collect_samples, dim3(blocks_for(count)), dim3(kBlock), 0, s, frame,
                     group_cap, stride, feat, cc, bc, base, capacity, samples);
  GTH_LAUNCH_CHECK();
}

void launch_sample_stats(const float *samples, uint32_t n, float *stats, hipStream_t s) {
  hipLaunchKernelGGL(sample_stats, dim3(1), dim3(kBlock), 0, s, samples, n, stats);
  GTH_LAUNCH_CHECK();
}

void launch_train_iteration(const float *samples, uint32_t n, const float *stats, float *params,
                            float *m, float *v, float *grad, float *loss, uint32_t blocks,
                            uint32_t iter, float lr, hipStream_t s) {
  const uint32_t batch = blocks * kTrainBlock;
  hipLaunchKernelGGL(mlp_gradients, dim3(blocks), dim3(kTrainBlock), 0, s, samples, n, stats,
                     params, iter, batch, grad, loss);
  GTH_LAUNCH_CHECK();
  const float c1 = 1.0f - powf(0.9f, static_cast<float>(iter + 1));
  const float c2 = 1.0f - powf(0.999f, static_cast<float>(iter + 1));
  hipLaunchKernelGGL(adam_step, dim3(blocks_for(kPredParams)), dim3(kBlock), 0, s, grad, blocks,
                     params, m, v, lr, c1, c2);
  GTH_LAUNCH_CHECK();
}

}  // namespace gth

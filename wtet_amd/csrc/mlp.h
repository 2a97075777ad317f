// Learned buffer presizing: per-cloud features and a small FP32 MLP.
//
// Features of a cloud (kFeat = 522): the 512-bin occupancy histogram of the
// top 9 bits of the 30-bit Morton code (10-bit floor quantization of each
// axis over the cloud's box, src/morton.rs:22-35), as fractions of the
// cloud's points; the mean (3) and covariance (6) of the box-normalized
// coordinates; log(n).  The network (kFeat -> 128 -> 128 -> 3, ReLU)
// predicts, in standardized log space, the tet slots the triangulation
// uses, the peak active-list length and the number of stuck faces, so that
// device buffers can be sized once.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wtet {

constexpr int kHistBins = 512;
constexpr int kFeat = kHistBins + 3 + 6 + 1;
constexpr int kHidden = 128;
constexpr int kOut = 3;

// Parameter blob layout (floats): feature mean[kFeat], feature std[kFeat],
// target mean[kOut], target std[kOut], W1[kFeat*kHidden] (row-major, input
// major), b1[kHidden], W2[kHidden*kHidden], b2[kHidden], W3[kHidden*kOut],
// b3[kOut], margin[kOut] (log-space safety margin added to predictions).
constexpr size_t kOffFeatMean = 0;
constexpr size_t kOffFeatStd = kOffFeatMean + kFeat;
constexpr size_t kOffTgtMean = kOffFeatStd + kFeat;
constexpr size_t kOffTgtStd = kOffTgtMean + kOut;
constexpr size_t kOffW1 = kOffTgtStd + kOut;
constexpr size_t kOffB1 = kOffW1 + (size_t)kFeat * kHidden;
constexpr size_t kOffW2 = kOffB1 + kHidden;
constexpr size_t kOffB2 = kOffW2 + (size_t)kHidden * kHidden;
constexpr size_t kOffW3 = kOffB2 + kHidden;
constexpr size_t kOffB3 = kOffW3 + (size_t)kHidden * kOut;
constexpr size_t kOffMargin = kOffB3 + kOut;
constexpr size_t kParamCount = kOffMargin + kOut;
// trainable part: W1 .. b3
constexpr size_t kTrainBegin = kOffW1;
constexpr size_t kTrainEnd = kOffMargin;

// Per-cloud features from raw points (cloud c = points off[c]..off[c+1]).
// mm: per-cloud ordered-u32 min/max (as launch_prepare leaves it).
// acc: device scratch of C * (kHistBins + 9) doubles; feat: C * kFeat.
void launch_features(const float* xyz, uint32_t n, const uint32_t* off, uint32_t C,
                     const uint32_t* mm, double* acc, float* feat, hipStream_t s);

// Inference: pred[C * kOut] (un-standardized log targets + margin) from
// features feat[C * kFeat] with the parameter blob on the device.
void launch_mlp_predict(const float* params, const float* feat, uint32_t C, float* pred,
                        hipStream_t s);

// Training pieces (minibatch of B rows, standardized inputs and targets).
struct MlpWork {
  float* x;    // B * kFeat (standardized features)
  float* y;    // B * kOut  (standardized targets)
  float* h1;   // B * kHidden (post-ReLU)
  float* h2;   // B * kHidden
  float* out;  // B * kOut
  float* d3;   // B * kOut
  float* d2;   // B * kHidden
  float* d1;   // B * kHidden
  float* grad; // kParamCount (only the trainable range is written)
  float* loss; // 1
};
// forward + MSE loss + backward into w.grad (trainable range)
void launch_mlp_forward_backward(const float* params, const MlpWork& w, uint32_t B, hipStream_t s);
// Adam step on the trainable range
void launch_adam(float* params, const float* grad, float* m, float* v, float lr, float b1,
                 float b2, float eps, uint32_t step, hipStream_t s);
// standardize rows: x = (raw - mean) / std, gathering rows idx[0..B)
void launch_gather_standardize(const float* raw, const uint32_t* idx, uint32_t B, int width,
                               const float* mean, const float* stdv, float* out, hipStream_t s);

}  // namespace wtet

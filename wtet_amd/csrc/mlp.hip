// Learned presizing kernels: per-cloud features, MLP inference, and the
// forward / backward / Adam kernels that train the network on the GPU.
// All FP32 except the feature moments (f64 accumulation).
#include <algorithm>
#include <cstring>

#include "kernels.h"
#include "mlp.h"

namespace wtet {

namespace {

__device__ __forceinline__ uint32_t gtid() { return blockIdx.x * blockDim.x + threadIdx.x; }

// spread the low 10 bits of v to every third bit
__device__ __forceinline__ uint32_t expand10(uint32_t v) {
  v &= 0x3FFu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

constexpr int kMom = 9;

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

}  // namespace

// Histogram counts (u32) and coordinate moments (f64) per cloud.  Each
// block accumulates the points of its first cloud in LDS (histogram and
// moments) and adds them to the global arrays once at the end; points of
// other clouds go straight to the global arrays.  One cloud: a grid-stride
// loop over a few hundred blocks, so that the global adds are per block
// and not per point (a global atomic per point on 512 words serializes at
// the memory side).  Many clouds: one point per thread, so that a block's
// points mostly share a cloud.
__global__ void __launch_bounds__(kBlock) k_feat_accum(const float* xyz, uint32_t n, const uint32_t* off,
                                                       uint32_t C, const uint32_t* mm, uint32_t* hist,
                                                       double* mom) {
  __shared__ uint32_t lh[kHistBins];
  __shared__ double lm[kMom];
  for (int b = threadIdx.x; b < kHistBins; b += blockDim.x) lh[b] = 0u;
  if (threadIdx.x < kMom) lm[threadIdx.x] = 0.0;
  const uint32_t first = blockIdx.x * blockDim.x;
  const uint32_t cb = (C > 1 && first < n) ? cloud_of_point(off, C, first) : 0u;
  __syncthreads();
  double m[kMom] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = gtid(); i < n; i += stride) {
    uint32_t c = C > 1 ? cloud_of_point(off, C, i) : 0u;
    const uint32_t* b = mm + 6 * (size_t)c;
    float u[3];
    uint32_t q[3];
    for (int k = 0; k < 3; ++k) {
      float lo = unord_f32(b[k]), hi = unord_f32(b[3 + k]);
      float ext = hi - lo;
      float v = ext > 0.0f ? (xyz[3 * (size_t)i + k] - lo) / ext : 0.0f;
      v = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
      u[k] = v;
      q[k] = (uint32_t)(v * 1023.0f);
    }
    uint32_t code = (expand10(q[2]) << 2) | (expand10(q[1]) << 1) | expand10(q[0]);
    double pm[kMom] = {u[0], u[1], u[2], (double)u[0] * u[0], (double)u[1] * u[1], (double)u[2] * u[2],
                       (double)u[0] * u[1], (double)u[0] * u[2], (double)u[1] * u[2]};
    if (c == cb) {
      atomicAdd(&lh[code >> 21], 1u);
      for (int k = 0; k < kMom; ++k) m[k] += pm[k];
    } else {
      atomicAdd(&hist[(size_t)c * kHistBins + (code >> 21)], 1u);
      for (int k = 0; k < kMom; ++k) atomicAdd(&mom[(size_t)c * kMom + k], pm[k]);
    }
  }
  // the block's own cloud: wave sums into LDS, then one global add each
  for (int k = 0; k < kMom; ++k) {
    double w = wave_sum(m[k]);
    if (__lane_id() == 0 && w != 0.0) atomicAdd(&lm[k], w);
  }
  __syncthreads();
  if (first < n) {
    for (int b = threadIdx.x; b < kHistBins; b += blockDim.x)
      if (lh[b]) atomicAdd(&hist[(size_t)cb * kHistBins + b], lh[b]);
    if (threadIdx.x < kMom && lm[threadIdx.x] != 0.0)
      atomicAdd(&mom[(size_t)cb * kMom + threadIdx.x], lm[threadIdx.x]);
  }
}

__global__ void k_feat_finish(const uint32_t* off, uint32_t C, const uint32_t* hist,
                              const double* mom, float* feat) {
  uint32_t g = gtid();
  if (g >= (uint64_t)C * kFeat) return;
  uint32_t c = g / kFeat, k = g - c * kFeat;
  double n = (double)(off[c + 1] - off[c]);
  const double* M = mom + (size_t)c * kMom;
  float v;
  if (k < (uint32_t)kHistBins) {
    v = (float)(hist[(size_t)c * kHistBins + k] / n);
  } else if (k < (uint32_t)kHistBins + 3) {
    v = (float)(M[k - kHistBins] / n);
  } else if (k < (uint32_t)kHistBins + 9) {
    // covariance: xx, yy, zz, xy, xz, yz
    const int a[6] = {0, 1, 2, 0, 0, 1}, b[6] = {0, 1, 2, 1, 2, 2};
    int j = (int)k - kHistBins - 3;
    v = (float)(M[3 + j] / n - (M[a[j]] / n) * (M[b[j]] / n));
  } else {
    v = (float)log(n);
  }
  feat[g] = v;
}

void launch_features(const float* xyz, uint32_t n, const uint32_t* off, uint32_t C,
                     const uint32_t* mm, double* acc, float* feat, hipStream_t s) {
  uint32_t* hist = reinterpret_cast<uint32_t*>(acc);  // C * kHistBins u32 at the front
  double* mom = acc + ((size_t)C * kHistBins + 1) / 2;
  launch_fill_u32(hist, 0u, (uint64_t)C * kHistBins, s);
  launch_fill_u64(reinterpret_cast<uint64_t*>(mom), 0ull, (uint64_t)C * kMom, s);
  uint32_t grid = C > 1 ? nblocks(n) : std::min(nblocks(n), 512u);
  if (n) k_feat_accum<<<grid, kBlock, 0, s>>>(xyz, n, off, C, mm, hist, mom);
  k_feat_finish<<<nblocks((uint64_t)C * kFeat), kBlock, 0, s>>>(off, C, hist, mom, feat);
}

// one block of kHidden threads per cloud
__global__ void __launch_bounds__(kHidden) k_mlp_predict(const float* P, const float* feat,
                                                         uint32_t C, float* pred) {
  __shared__ float x[kFeat];
  __shared__ float h1[kHidden];
  __shared__ float h2[kHidden];
  uint32_t c = blockIdx.x;
  int t = threadIdx.x;
  if (c >= C) return;
  for (int k = t; k < kFeat; k += kHidden)
    x[k] = (feat[(size_t)c * kFeat + k] - P[kOffFeatMean + k]) / P[kOffFeatStd + k];
  __syncthreads();
  float a = P[kOffB1 + t];
  for (int k = 0; k < kFeat; ++k) a += x[k] * P[kOffW1 + (size_t)k * kHidden + t];
  h1[t] = a > 0.0f ? a : 0.0f;
  __syncthreads();
  a = P[kOffB2 + t];
  for (int k = 0; k < kHidden; ++k) a += h1[k] * P[kOffW2 + (size_t)k * kHidden + t];
  h2[t] = a > 0.0f ? a : 0.0f;
  __syncthreads();
  if (t < kOut) {
    a = P[kOffB3 + t];
    for (int k = 0; k < kHidden; ++k) a += h2[k] * P[kOffW3 + (size_t)k * kOut + t];
    pred[(size_t)c * kOut + t] = a * P[kOffTgtStd + t] + P[kOffTgtMean + t] + P[kOffMargin + t];
  }
}

void launch_mlp_predict(const float* params, const float* feat, uint32_t C, float* pred,
                        hipStream_t s) {
  if (C) k_mlp_predict<<<C, kHidden, 0, s>>>(params, feat, C, pred);
}

// ---------------------------------------------------------------------------
// training
// ---------------------------------------------------------------------------
constexpr int kTile = 16;

// C[M x N] = op(A) op(B) (+ bias[N]); op(A) is M x K, op(B) is K x N.
// tA: A stored K x M (use its transpose); tB: B stored N x K.
// post: 0 none, 1 ReLU, 2 multiply by (mask > 0) (mask shaped like C).
__global__ void __launch_bounds__(kTile * kTile)
k_gemm(int M, int N, int K, const float* A, int lda, int tA, const float* B, int ldb, int tB,
       float* Cm, int ldc, const float* bias, int post, const float* mask) {
  __shared__ float As[kTile][kTile + 1];
  __shared__ float Bs[kTile][kTile + 1];
  int tx = threadIdx.x % kTile, ty = threadIdx.x / kTile;
  int row = blockIdx.y * kTile + ty, col = blockIdx.x * kTile + tx;
  float acc = 0.0f;
  for (int k0 = 0; k0 < K; k0 += kTile) {
    int ka = k0 + tx;  // A tile: row, k
    float av = 0.0f;
    if (row < M && ka < K) av = tA ? A[(size_t)ka * lda + row] : A[(size_t)row * lda + ka];
    As[ty][tx] = av;
    int kb = k0 + ty;  // B tile: k, col
    float bv = 0.0f;
    if (kb < K && col < N) bv = tB ? B[(size_t)col * ldb + kb] : B[(size_t)kb * ldb + col];
    Bs[ty][tx] = bv;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kTile; ++k) acc += As[ty][k] * Bs[k][tx];
    __syncthreads();
  }
  if (row < M && col < N) {
    if (bias) acc += bias[col];
    if (post == 1) acc = acc > 0.0f ? acc : 0.0f;
    else if (post == 2) acc = mask[(size_t)row * ldc + col] > 0.0f ? acc : 0.0f;
    Cm[(size_t)row * ldc + col] = acc;
  }
}

static void gemm(int M, int N, int K, const float* A, int lda, int tA, const float* B, int ldb,
                 int tB, float* Cm, int ldc, const float* bias, int post, const float* mask,
                 hipStream_t s) {
  dim3 grid((N + kTile - 1) / kTile, (M + kTile - 1) / kTile);
  k_gemm<<<grid, kTile * kTile, 0, s>>>(M, N, K, A, lda, tA, B, ldb, tB, Cm, ldc, bias, post, mask);
}

// column sums of d (rows x cols) into g
__global__ void k_colsum(const float* d, int rows, int cols, float* g) {
  int j = gtid();
  if (j >= cols) return;
  float a = 0.0f;
  for (int r = 0; r < rows; ++r) a += d[(size_t)r * cols + j];
  g[j] = a;
}

// MSE over B x kOut: loss = mean (out - y)^2, d = 2 (out - y) / (B kOut)
__global__ void __launch_bounds__(256) k_mse(const float* out, const float* y, int count, float* d,
                                             float* loss) {
  __shared__ float part[256];
  float a = 0.0f;
  for (int i = threadIdx.x; i < count; i += 256) {
    float e = out[i] - y[i];
    d[i] = 2.0f * e / (float)count;
    a += e * e;
  }
  part[threadIdx.x] = a;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss[0] = part[0] / (float)count;
}

void launch_mlp_forward_backward(const float* P, const MlpWork& w, uint32_t B, hipStream_t s) {
  const int b = (int)B, F = kFeat, H = kHidden, O = kOut;
  const float* W1 = P + kOffW1;
  const float* W2 = P + kOffW2;
  const float* W3 = P + kOffW3;
  // forward
  gemm(b, H, F, w.x, F, 0, W1, H, 0, w.h1, H, P + kOffB1, 1, nullptr, s);
  gemm(b, H, H, w.h1, H, 0, W2, H, 0, w.h2, H, P + kOffB2, 1, nullptr, s);
  gemm(b, O, H, w.h2, H, 0, W3, O, 0, w.out, O, P + kOffB3, 0, nullptr, s);
  k_mse<<<1, 256, 0, s>>>(w.out, w.y, b * O, w.d3, w.loss);
  // backward
  float* G = w.grad;
  gemm(H, O, b, w.h2, H, 1, w.d3, O, 0, G + kOffW3, O, nullptr, 0, nullptr, s);  // h2^T d3
  k_colsum<<<1, 64, 0, s>>>(w.d3, b, O, G + kOffB3);
  gemm(b, H, O, w.d3, O, 0, W3, O, 1, w.d2, H, nullptr, 2, w.h2, s);  // (d3 W3^T) * [h2>0]
  gemm(H, H, b, w.h1, H, 1, w.d2, H, 0, G + kOffW2, H, nullptr, 0, nullptr, s);
  k_colsum<<<(H + 63) / 64, 64, 0, s>>>(w.d2, b, H, G + kOffB2);
  gemm(b, H, H, w.d2, H, 0, W2, H, 1, w.d1, H, nullptr, 2, w.h1, s);  // (d2 W2^T) * [h1>0]
  gemm(F, H, b, w.x, F, 1, w.d1, H, 0, G + kOffW1, H, nullptr, 0, nullptr, s);
  k_colsum<<<(H + 63) / 64, 64, 0, s>>>(w.d1, b, H, G + kOffB1);
}

__global__ void k_adam(float* p, const float* g, float* m, float* v, uint32_t n, float lr, float b1,
                       float b2, float eps, float c1, float c2) {
  uint32_t i = gtid();
  if (i >= n) return;
  float gi = g[i];
  float mi = b1 * m[i] + (1.0f - b1) * gi;
  float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

void launch_adam(float* params, const float* grad, float* m, float* v, float lr, float b1, float b2,
                 float eps, uint32_t step, hipStream_t s) {
  uint32_t n = (uint32_t)(kTrainEnd - kTrainBegin);
  float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
  k_adam<<<nblocks(n), kBlock, 0, s>>>(params + kTrainBegin, grad + kTrainBegin, m + kTrainBegin,
                                       v + kTrainBegin, n, lr, b1, b2, eps, c1, c2);
}

__global__ void k_gather_std(const float* raw, const uint32_t* idx, uint32_t B, int width,
                             const float* mean, const float* stdv, float* out) {
  uint32_t g = gtid();
  if (g >= (uint64_t)B * width) return;
  uint32_t r = g / width, k = g - r * width;
  out[g] = (raw[(size_t)idx[r] * width + k] - mean[k]) / stdv[k];
}

void launch_gather_standardize(const float* raw, const uint32_t* idx, uint32_t B, int width,
                               const float* mean, const float* stdv, float* out, hipStream_t s) {
  if (B) k_gather_std<<<nblocks((uint64_t)B * width), kBlock, 0, s>>>(raw, idx, B, width, mean, stdv, out);
}

}  // namespace wtet

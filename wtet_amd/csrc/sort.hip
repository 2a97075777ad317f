// Stable LSD radix sort of (key, value) pairs, 8-bit digits (sort.h).
//
// Per pass, over tiles of 2048 pairs (256 threads x 8):
//  * k_radix_up: each tile's digit counts, from wave64 ballots (the lanes
//    with this lane's digit) so that one LDS atomic per distinct digit per
//    wave is issued, stored digit-major (counts[d * tiles + tile]);
//  * k_radix_scan: one workgroup per digit turns its row into global
//    offsets: the pairs of smaller digits (from the all-pass histogram that
//    k_radix_hist builds once, before the first pass) plus the exclusive
//    prefix over earlier tiles;
//  * k_radix_down: each tile ranks its pairs stably (slab by slab in index
//    order: the wave's ballot rank, the earlier waves' counts, the earlier
//    slabs' counts) and scatters them to their offsets.
// Equal digits keep their order in every pass, so the whole sort is stable.
#include "sort.h"

#include <algorithm>

namespace wtet {

namespace {

constexpr int kSortThreads = 256;
constexpr int kSortItems = 8;
constexpr uint32_t kTile = kSortThreads * kSortItems;
constexpr int kDigitBits = 8;
constexpr int kRadix = 1 << kDigitBits;
constexpr int kMaxPasses = 8;
constexpr int kWaves = kSortThreads / 64;
static_assert(kSortThreads == kRadix, "one thread per digit in the per-digit steps");

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
inline uint32_t tiles_of(uint32_t n) { return (n + kTile - 1) / kTile; }
__host__ __device__ inline int passes_of(int end_bit) { return (end_bit + kDigitBits - 1) / kDigitBits; }

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

template <typename K>
__device__ __forceinline__ uint32_t digit_of(K key, int shift, uint32_t mask) {
  return (uint32_t)(key >> shift) & mask;
}

// lanes of this wave that hold a valid pair with digit d (this lane's own
// when it is valid); every lane of the wave must call it
__device__ __forceinline__ uint64_t peer_mask(uint32_t d, bool valid) {
  uint64_t m = __ballot(valid);
#pragma unroll
  for (int b = 0; b < kDigitBits; ++b) {
    bool bit = (d >> b) & 1u;
    uint64_t bm = __ballot(bit);
    m &= bit ? bm : ~bm;
  }
  return m;
}

// inclusive sum over the lanes of the wave up to this one
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t x) {
  uint32_t lane = lane_id();
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t y = __shfl_up(x, off);
    if (lane >= (uint32_t)off) x += y;
  }
  return x;
}

__global__ void k_radix_zero(uint32_t* p, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// digit histograms of every pass in one read of the keys
template <typename K>
__global__ void __launch_bounds__(kSortThreads) k_radix_hist(const K* keys, uint32_t n, int end_bit,
                                                             uint32_t* ghist) {
  __shared__ uint32_t lh[kMaxPasses * kRadix];
  const int P = passes_of(end_bit);
  for (int i = threadIdx.x; i < P * kRadix; i += kSortThreads) lh[i] = 0u;
  __syncthreads();
  const uint32_t lane = lane_id();
  // the loop bound depends on the block only, so every lane of a wave takes
  // part in each ballot
  for (uint32_t base = blockIdx.x * kSortThreads; base < n; base += gridDim.x * kSortThreads) {
    uint32_t i = base + threadIdx.x;
    bool valid = i < n;
    K key = valid ? keys[i] : (K)0;
    for (int p = 0; p < P; ++p) {
      int shift = p * kDigitBits;
      int bits = min(kDigitBits, end_bit - shift);
      uint32_t d = digit_of(key, shift, (1u << bits) - 1u);
      uint64_t peers = peer_mask(d, valid);
      if (valid && (uint32_t)__builtin_ctzll(peers) == lane)
        atomicAdd(&lh[p * kRadix + d], (uint32_t)__popcll(peers));
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < P * kRadix; i += kSortThreads)
    if (lh[i]) atomicAdd(&ghist[i], lh[i]);
}

template <typename K>
__global__ void __launch_bounds__(kSortThreads) k_radix_up(const K* keys, uint32_t n, int shift,
                                                           uint32_t mask, uint32_t tiles,
                                                           uint32_t* counts) {
  __shared__ uint32_t cnt[kRadix];
  const uint32_t t = threadIdx.x, lane = lane_id();
  const uint32_t base = blockIdx.x * kTile;
  cnt[t] = 0u;
  K kk[kSortItems];
#pragma unroll
  for (int k = 0; k < kSortItems; ++k) {
    uint32_t i = base + (uint32_t)k * kSortThreads + t;
    kk[k] = i < n ? keys[i] : (K)0;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < kSortItems; ++k) {
    uint32_t i = base + (uint32_t)k * kSortThreads + t;
    bool valid = i < n;
    uint32_t d = valid ? digit_of(kk[k], shift, mask) : 0u;
    uint64_t peers = peer_mask(d, valid);
    if (valid && (uint32_t)__builtin_ctzll(peers) == lane) atomicAdd(&cnt[d], (uint32_t)__popcll(peers));
  }
  __syncthreads();
  counts[(size_t)t * tiles + blockIdx.x] = cnt[t];
}

// block d: counts[d * tiles + b] -> pairs before tile b's first pair of digit d
__global__ void __launch_bounds__(kSortThreads) k_radix_scan(const uint32_t* hist, uint32_t tiles,
                                                             uint32_t* counts) {
  __shared__ uint32_t ws[kWaves];
  const uint32_t d = blockIdx.x, t = threadIdx.x, lane = lane_id(), wave = t >> 6;
  // pairs of smaller digits
  uint32_t x = wave_incl_scan(t < d ? hist[t] : 0u);
  if (lane == 63) ws[wave] = x;
  __syncthreads();
  uint32_t running = 0;
#pragma unroll
  for (int w = 0; w < kWaves; ++w) running += ws[w];
  __syncthreads();
  uint32_t* row = counts + (size_t)d * tiles;
  for (uint32_t c = 0; c < tiles; c += kSortThreads * 4) {
    uint32_t j = c + 4 * t;
    uint32_t a[4], sum = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      a[q] = j + q < tiles ? row[j + q] : 0u;
      sum += a[q];
    }
    uint32_t incl = wave_incl_scan(sum);
    if (lane == 63) ws[wave] = incl;
    __syncthreads();
    uint32_t before = 0, total = 0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) {
      before += (uint32_t)w < wave ? ws[w] : 0u;
      total += ws[w];
    }
    uint32_t e = running + before + incl - sum;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (j + q < tiles) row[j + q] = e;
      e += a[q];
    }
    running += total;
    __syncthreads();
  }
}

template <typename K>
__global__ void __launch_bounds__(kSortThreads) k_radix_down(const K* kin, const uint32_t* vin, K* kout,
                                                             uint32_t* vout, uint32_t n, int shift,
                                                             uint32_t mask, uint32_t tiles,
                                                             const uint32_t* offs) {
  __shared__ uint32_t goff[kRadix];
  __shared__ uint32_t run[kRadix];
  __shared__ uint32_t wc[kWaves][kRadix];
  const uint32_t t = threadIdx.x, lane = lane_id(), wave = t >> 6;
  const uint32_t base = blockIdx.x * kTile;
  goff[t] = offs[(size_t)t * tiles + blockIdx.x];
  run[t] = 0u;
  K kk[kSortItems];
  uint32_t vv[kSortItems];
#pragma unroll
  for (int k = 0; k < kSortItems; ++k) {
    uint32_t i = base + (uint32_t)k * kSortThreads + t;
    bool valid = i < n;
    kk[k] = valid ? kin[i] : (K)0;
    vv[k] = valid ? vin[i] : 0u;
  }
  const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
  for (int k = 0; k < kSortItems; ++k) {
#pragma unroll
    for (int w = 0; w < kWaves; ++w) wc[w][t] = 0u;
    __syncthreads();
    uint32_t i = base + (uint32_t)k * kSortThreads + t;
    bool valid = i < n;
    uint32_t d = valid ? digit_of(kk[k], shift, mask) : 0u;
    uint64_t peers = peer_mask(d, valid);
    uint32_t wr = (uint32_t)__popcll(peers & lt);
    if (valid && (uint32_t)__builtin_ctzll(peers) == lane) wc[wave][d] = (uint32_t)__popcll(peers);
    __syncthreads();
    // thread t: digit t's slab counts -> the offsets of each wave's run
    {
      uint32_t x = run[t];
#pragma unroll
      for (int w = 0; w < kWaves; ++w) {
        uint32_t y = wc[w][t];
        wc[w][t] = x;
        x += y;
      }
      run[t] = x;
    }
    __syncthreads();
    if (valid) {
      uint32_t pos = goff[d] + wc[wave][d] + wr;
      kout[pos] = kk[k];
      vout[pos] = vv[k];
    }
    __syncthreads();  // wc is cleared by the next slab
  }
}

template <typename K>
size_t temp_bytes_t(uint32_t n, int end_bit) {
  (void)end_bit;
  uint32_t tiles = std::max<uint32_t>(1u, tiles_of(n));
  return align256((size_t)kMaxPasses * kRadix * 4) + align256((size_t)kRadix * tiles * 4) +
         align256((size_t)n * sizeof(K)) + align256((size_t)n * 4);
}

template <typename K>
hipError_t sort_t(const K* kin, const uint32_t* vin, K* kout, uint32_t* vout, uint32_t n, int end_bit,
                  void* temp, size_t temp_bytes, hipStream_t s) {
  if (n == 0) return hipSuccess;
  end_bit = std::max(0, std::min(end_bit, (int)(8 * sizeof(K))));
  if (temp_bytes < temp_bytes_t<K>(n, end_bit) || !temp) return hipErrorInvalidValue;
  const int P = passes_of(end_bit);
  if (P == 0) {
    hipError_t e = hipMemcpyAsync(kout, kin, (size_t)n * sizeof(K), hipMemcpyDeviceToDevice, s);
    if (e != hipSuccess) return e;
    return hipMemcpyAsync(vout, vin, (size_t)n * 4, hipMemcpyDeviceToDevice, s);
  }
  const uint32_t tiles = tiles_of(n);
  char* tp = static_cast<char*>(temp);
  uint32_t* hist = reinterpret_cast<uint32_t*>(tp);
  tp += align256((size_t)kMaxPasses * kRadix * 4);
  uint32_t* counts = reinterpret_cast<uint32_t*>(tp);
  tp += align256((size_t)kRadix * tiles * 4);
  K* kalt = reinterpret_cast<K*>(tp);
  tp += align256((size_t)n * sizeof(K));
  uint32_t* valt = reinterpret_cast<uint32_t*>(tp);
  k_radix_zero<<<(P * kRadix + 255) / 256, 256, 0, s>>>(hist, (uint32_t)(P * kRadix));
  uint32_t hgrid = std::min<uint32_t>(1024u, (n + kSortThreads - 1) / kSortThreads);
  k_radix_hist<K><<<hgrid, kSortThreads, 0, s>>>(kin, n, end_bit, hist);
  const K* ksrc = kin;
  const uint32_t* vsrc = vin;
  for (int p = 0; p < P; ++p) {
    // the last pass writes the caller's output; earlier ones alternate
    bool to_out = ((P - 1 - p) & 1) == 0;
    K* kd = to_out ? kout : kalt;
    uint32_t* vd = to_out ? vout : valt;
    int shift = p * kDigitBits;
    int bits = std::min(kDigitBits, end_bit - shift);
    uint32_t mask = (1u << bits) - 1u;
    k_radix_up<K><<<tiles, kSortThreads, 0, s>>>(ksrc, n, shift, mask, tiles, counts);
    k_radix_scan<<<kRadix, kSortThreads, 0, s>>>(hist + p * kRadix, tiles, counts);
    k_radix_down<K><<<tiles, kSortThreads, 0, s>>>(ksrc, vsrc, kd, vd, n, shift, mask, tiles, counts);
    ksrc = kd;
    vsrc = vd;
  }
  return hipGetLastError();
}

}  // namespace

size_t radix_temp_bytes_u64(uint32_t n, int end_bit) { return temp_bytes_t<uint64_t>(n, end_bit); }
size_t radix_temp_bytes_u32(uint32_t n, int end_bit) { return temp_bytes_t<uint32_t>(n, end_bit); }

hipError_t radix_sort_u64(const uint64_t* keys_in, const uint32_t* vals_in, uint64_t* keys_out,
                          uint32_t* vals_out, uint32_t n, int end_bit, void* temp, size_t temp_bytes,
                          hipStream_t s) {
  return sort_t<uint64_t>(keys_in, vals_in, keys_out, vals_out, n, end_bit, temp, temp_bytes, s);
}

hipError_t radix_sort_u32(const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
                          uint32_t* vals_out, uint32_t n, int end_bit, void* temp, size_t temp_bytes,
                          hipStream_t s) {
  return sort_t<uint32_t>(keys_in, vals_in, keys_out, vals_out, n, end_bit, temp, temp_bytes, s);
}

}  // namespace wtet

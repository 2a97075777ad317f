// HIP kernels of the triangulation loop (gfx950, wave64).
//
// Element-wise kernels wrap the bodies of ops.h.  List compaction uses
// wave ballots: each wavefront ranks its lanes with popcount(ballot & lt),
// wave totals are combined in LDS, block totals are scanned by one block,
// and a scatter pass writes every element once (stable order).
#include <algorithm>
#include <cstring>


#include "kernels.h"
#include "prep.h"
#include "sort.h"

namespace wtet {

__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint64_t n);

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t gid() { return blockIdx.x * blockDim.x + threadIdx.x; }

// Adds this thread's local counters to the global u64 counters (one atomic
// per counter per wavefront).  Must be reached by every lane of the wave.
__device__ __forceinline__ void flush_stats(const PredStats& ls, uint64_t* st) {
  if (!st) return;
  unsigned long long* g = reinterpret_cast<unsigned long long*>(st);
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    uint32_t v = ls.c[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if (__lane_id() == 0 && v) atomicAdd(g + k, (unsigned long long)v);
  }
}

// exclusive rank of `flag` within the block; *total = block count.
// Every thread of the block must call it.
__device__ __forceinline__ uint32_t block_rank(bool flag, uint32_t* total) {
  __shared__ uint32_t wsum[kBlock / 64];
  uint64_t m = __ballot(flag);
  int lane = __lane_id();
  int wave = threadIdx.x >> 6;
  uint32_t r = __popcll(m & ((1ull << lane) - 1ull));
  if (lane == 0) wsum[wave] = __popcll(m);
  __syncthreads();
  uint32_t off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < kBlock / 64; ++w) {
    off += w < wave ? wsum[w] : 0u;
    tot += wsum[w];
  }
  *total = tot;
  return off + r;
}

// in-place exclusive scan of blk[0..nb) by one 1024-thread block
// mode 0: insertion round -> ctr[C_WIN] = total, reserve 3*total slots
// mode 1: output          -> ctr[C_ALIVE] = total
// mode 2: point list      -> ctr[C_NPTS] = total
__global__ void __launch_bounds__(1024) k_scan_blocks(uint32_t* blk, uint32_t nb, uint32_t* ctr,
                                                     int mode) {
  __shared__ uint32_t part[1024];
  uint32_t tid = threadIdx.x;
  uint32_t chunk = (nb + 1023u) / 1024u;
  uint32_t b0 = tid * chunk, b1 = min(nb, b0 + chunk);
  uint32_t s = 0;
  for (uint32_t i = b0; i < b1; ++i) s += blk[i];
  part[tid] = s;
  __syncthreads();
  for (uint32_t off = 1; off < 1024u; off <<= 1) {
    uint32_t v = tid >= off ? part[tid - off] : 0u;
    __syncthreads();
    part[tid] += v;
    __syncthreads();
  }
  uint32_t run = part[tid] - s;  // exclusive prefix of this chunk
  for (uint32_t i = b0; i < b1; ++i) {
    uint32_t v = blk[i];
    blk[i] = run;
    run += v;
  }
  if (tid == 1023u) {
    uint32_t total = part[1023];
    if (mode == 0) {
      ctr[C_WIN] = total;
      ctr[C_SURV] = ctr[C_ULEN] - total;
      ctr[C_BASE] = ctr[C_TETS];
      ctr[C_TETS] += 3u * total;
    } else if (mode == 1) {
      ctr[C_ALIVE] = total;
    } else {
      ctr[C_NPTS] = total;
    }
  }
}

// ---------------------------------------------------------------------------
// point preparation (layout of one or many clouds: prep.h)
// ---------------------------------------------------------------------------
__global__ void k_minmax_init(uint32_t* mm, uint32_t C) {
  uint32_t i = gid();
  if (i < 6u * C) mm[i] = (i % 6u) < 3u ? 0xFFFFFFFFu : 0u;
}

// Per-cloud min/max of each axis.  Lanes in the cloud of the wave's first
// lane combine with a wave reduction (one atomic per wave); lanes of other
// clouds (at most a few per cloud boundary) add their own atomics.
__global__ void k_minmax(const float* xyz, uint32_t n, Clouds K, uint32_t* mm) {
  uint32_t i = gid();
  uint32_t c = (i < n && K.C > 1) ? cloud_of_point(K.off, K.C, i) : 0u;
  uint32_t lo[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, hi[3] = {0u, 0u, 0u};
  if (i < n)
    for (int k = 0; k < 3; ++k) {
      float v = xyz[3 * (size_t)i + k];
      if (v == v) {
        uint32_t o = ord_f32(v);
        lo[k] = o;
        hi[k] = o;
      }
    }
  uint32_t c0 = __shfl(c, 0);
  bool same = c == c0;
  for (int k = 0; k < 3; ++k) {
    uint32_t l = same ? lo[k] : 0xFFFFFFFFu, h = same ? hi[k] : 0u;
    for (int off = 32; off > 0; off >>= 1) {
      l = min(l, (uint32_t)__shfl_xor(l, off));
      h = max(h, (uint32_t)__shfl_xor(h, off));
    }
    if (__lane_id() == 0) {
      atomicMin(&mm[6 * (size_t)c0 + k], l);
      atomicMax(&mm[6 * (size_t)c0 + 3 + k], h);
    }
    if (!same && i < n) {
      atomicMin(&mm[6 * (size_t)c + k], lo[k]);
      atomicMax(&mm[6 * (size_t)c + 3 + k], hi[k]);
    }
  }
}

__global__ void k_normalize(const float* xyz, uint32_t n, Clouds K, const uint32_t* mm,
                            float* xyzn, uint64_t* keys, uint32_t* ids) {
  uint32_t i = gid();
  if (i < n) normalize_body(xyz, i, K, mm, xyzn, keys, ids);
}

__global__ void k_gather_points(const float* xyzn, const uint32_t* ids, const uint64_t* keys,
                                uint32_t n, Clouds K, float4* pts) {
  uint32_t j = gid();
  if (j < n) gather_body(xyzn, ids, keys, j, K, pts);
}

__global__ void k_tail_points(Clouds K, float4* pts) {
  uint32_t g = gid();
  if (g <= 4u * K.C) tail_body(g, K, pts);
}

__global__ void k_start_tets(Mesh M, StartTets S, Clouds K) {
  uint32_t g = gid();
  if (g < 5u * K.C) start_tet_body(M, S, g, K);
}

// One cloud: every thread folds many points (grid-stride, coalesced), the
// block combines in LDS, and one lane per block issues the 6 atomics.
__global__ void __launch_bounds__(kBlock) k_minmax_single(const float* xyz, uint32_t n, uint32_t* mm) {
  __shared__ uint32_t red[6][kBlock / 64];
  uint32_t lo[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}, hi[3] = {0u, 0u, 0u};
  uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = gid(); i < n; i += stride)
    for (int k = 0; k < 3; ++k) {
      float v = xyz[3 * (size_t)i + k];
      if (v == v) {
        uint32_t o = ord_f32(v);
        lo[k] = min(lo[k], o);
        hi[k] = max(hi[k], o);
      }
    }
  int wave = threadIdx.x >> 6;
  for (int k = 0; k < 3; ++k) {
    for (int off = 32; off > 0; off >>= 1) {
      lo[k] = min(lo[k], (uint32_t)__shfl_xor(lo[k], off));
      hi[k] = max(hi[k], (uint32_t)__shfl_xor(hi[k], off));
    }
    if (__lane_id() == 0) {
      red[k][wave] = lo[k];
      red[3 + k][wave] = hi[k];
    }
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    int k = threadIdx.x;
    uint32_t v = k < 3 ? 0xFFFFFFFFu : 0u;
    for (int w = 0; w < kBlock / 64; ++w) v = k < 3 ? min(v, red[k][w]) : max(v, red[k][w]);
    if (k < 3) atomicMin(&mm[k], v);
    else atomicMax(&mm[k], v);
  }
}

void launch_minmax(const float* xyz, uint32_t n, const Clouds& K, uint32_t* mm, hipStream_t s) {
  k_minmax_init<<<nblocks(6u * K.C), kBlock, 0, s>>>(mm, K.C);
  if (K.C <= 1)
    k_minmax_single<<<std::max(1u, std::min(nblocks(n) / 16u, 1024u)), kBlock, 0, s>>>(xyz, n, mm);
  else
    k_minmax<<<nblocks(n), kBlock, 0, s>>>(xyz, n, K, mm);
}

void launch_prepare(const float* xyz, uint32_t n, const Clouds& K, uint32_t* mm, float* xyzn,
                    uint64_t* keys, uint32_t* ids, hipStream_t s) {
  launch_minmax(xyz, n, K, mm, s);
  k_normalize<<<nblocks(n), kBlock, 0, s>>>(xyz, n, K, mm, xyzn, keys, ids);
}

void launch_gather_points(const float* xyzn, const uint32_t* ids, const uint64_t* keys,
                          uint32_t n, const Clouds& K, float4* pts, hipStream_t s) {
  k_gather_points<<<nblocks(n), kBlock, 0, s>>>(xyzn, ids, keys, n, K, pts);
  k_tail_points<<<nblocks(4u * K.C + 1u), kBlock, 0, s>>>(K, pts);
}

void launch_start_tets(const Mesh& M, const Clouds& K, hipStream_t s) {
  static const StartTets S = make_start_tets();
  k_start_tets<<<nblocks(5u * K.C), kBlock, 0, s>>>(M, S, K);
}

size_t sort_temp_bytes(uint32_t n, int end_bit) { return radix_temp_bytes_u64(n, end_bit); }

hipError_t sort_pairs(uint64_t* keys, uint32_t* ids, uint64_t* keys_tmp, uint32_t* ids_tmp, uint32_t n,
                      int end_bit, void* temp, size_t temp_bytes, hipStream_t s) {
  return radix_sort_u64(keys, ids, keys_tmp, ids_tmp, n, end_bit, temp, temp_bytes, s);
}

size_t sort_temp_bytes_u32(uint32_t n, int end_bit) { return radix_temp_bytes_u32(n, end_bit); }

hipError_t sort_pairs_u32(uint32_t* keys, uint32_t* vals, uint32_t* keys_tmp, uint32_t* vals_tmp,
                          uint32_t n, int end_bit, void* temp, size_t temp_bytes, hipStream_t s) {
  return radix_sort_u32(keys, vals, keys_tmp, vals_tmp, n, end_bit, temp, temp_bytes, s);
}

// sorted position j of cloud c: kept unless it is a redundant duplicate;
// the uninserted list holds (vertex j + 4c, start super tet 5c)
__device__ __forceinline__ bool keep_point(const float4* pts, const uint64_t* keys, uint32_t n,
                                           const Clouds& K, uint32_t j, uint32_t* c,
                                           uint32_t* capped) {
  if (j >= n) return false;
  *c = cloud_of_key(K, keys[j]);
  return !redundant_dup(pts + 4u * *c, keys, n, j, capped);
}

__global__ void k_uninit_count(const float4* pts, const uint64_t* keys, uint32_t n, Clouds K,
                               uint32_t* blk, uint32_t* ctr) {
  uint32_t j = gid();
  uint32_t capped = 0, c = 0;
  bool keep = keep_point(pts, keys, n, K, j, &c, &capped);
  if (capped) atomicAdd(&ctr[C_DUPSCAN], 1u);
  uint32_t tot;
  block_rank(keep, &tot);
  if (threadIdx.x == 0) blk[blockIdx.x] = tot;
}

__global__ void k_uninit_scatter(const float4* pts, const uint64_t* keys, uint32_t n, Clouds K,
                                 const uint32_t* blk, uint32_t* uvert, uint32_t* utet) {
  uint32_t j = gid();
  uint32_t capped = 0, c = 0;
  bool keep = keep_point(pts, keys, n, K, j, &c, &capped);
  uint32_t tot;
  uint32_t r = blk[blockIdx.x] + block_rank(keep, &tot);
  if (keep) {
    uvert[r] = j + 4u * c;
    utet[r] = 5u * c;  // every point starts in its cloud's super tet
  }
}

void launch_init_uninserted(const float4* pts, const uint64_t* keys, uint32_t n, const Clouds& K,
                            uint32_t* blk, uint32_t* uvert, uint32_t* utet, uint32_t* ctr,
                            hipStream_t s) {
  uint32_t nb = nblocks(n);
  k_uninit_count<<<nb, kBlock, 0, s>>>(pts, keys, n, K, blk, ctr);
  k_scan_blocks<<<1, 1024, 0, s>>>(blk, nb, ctr, 2);
  k_uninit_scatter<<<nb, kBlock, 0, s>>>(pts, keys, n, K, blk, uvert, utet);
}

// Reference insertion order of the Hilbert mode: within each cloud the
// uninserted list follows the stratified order of the sorted points
// (src/gpu/mod.rs:473-600).  key = (cloud, rank among all the cloud's points);
// after the sort each entry's start tet is its cloud's super tet again.
__global__ void k_stratify_keys(const uint32_t* uvert, uint32_t U, Clouds K, uint64_t* key) {
  uint32_t r = gid();
  if (r >= U) return;
  uint32_t v = uvert[r];
  uint32_t c = K.C > 1 ? cloud_of_vertex(K.off, K.C, v) : 0u;
  uint32_t nc = K.off[c + 1] - K.off[c];
  uint32_t jl = v - 4u * c - K.off[c];
  key[r] = ((uint64_t)c << 32) | stratified_rank(jl, nc, initial_stride(nc));
}

__global__ void k_list_from_sorted(const uint32_t* sv, uint32_t U, Clouds K, uint32_t* uvert,
                                   uint32_t* utet) {
  uint32_t r = gid();
  if (r >= U) return;
  uint32_t v = sv[r];
  uvert[r] = v;
  utet[r] = 5u * (K.C > 1 ? cloud_of_vertex(K.off, K.C, v) : 0u);
}

hipError_t launch_stratify(uint32_t* uvert, uint32_t* utet, uint32_t U, const Clouds& K, uint64_t* key,
                           uint64_t* key_tmp, uint32_t* val_tmp, void* temp, size_t temp_bytes,
                           hipStream_t s) {
  if (!U) return hipSuccess;
  k_stratify_keys<<<nblocks(U), kBlock, 0, s>>>(uvert, U, K, key);
  int bits = 32 + (K.C > 1 ? cloud_key_bits(K.C) : 0);
  hipError_t e = radix_sort_u64(key, uvert, key_tmp, val_tmp, U, bits, temp, temp_bytes, s);
  if (e != hipSuccess) return e;
  k_list_from_sorted<<<nblocks(U), kBlock, 0, s>>>(val_tmp, U, K, uvert, utet);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// insertion rounds
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t shfl_u64(uint64_t v, int src) {
  uint32_t lo = __shfl((uint32_t)v, src), hi = __shfl((uint32_t)(v >> 32), src);
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int m) {
  uint32_t lo = __shfl_xor((uint32_t)v, m), hi = __shfl_xor((uint32_t)(v >> 32), m);
  return ((uint64_t)hi << 32) | lo;
}

// Lanes voting for the same tet as the wave's first pending lane combine
// their keys with a wave min first (a few passes), so early rounds -- when
// most points still share a few tets -- issue one atomic per wave instead of
// one per point; the remaining lanes vote directly.
__global__ void __launch_bounds__(kBlock) k_vote(Mesh M, const uint32_t* uvert, const uint32_t* utet) {
  const uint32_t U = M.ctr[C_ULEN];
  uint32_t i = gid();
  uint64_t key = kVoteNone;
  uint32_t t = 0;
  if (i < U) {
    t = utet[i];
    key = vote_key(M, uvert, utet, i);
  }
  bool pending = key != kVoteNone;
  for (int pass = 0; pass < 4; ++pass) {
    uint64_t act = __ballot(pending);
    if (!act) break;
    int leader = __ffsll((unsigned long long)act) - 1;
    uint32_t lt = __shfl(t, leader);
    bool mine = pending && t == lt;
    uint64_t k = mine ? key : kVoteNone;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      uint64_t o = shfl_xor_u64(k, off);
      k = o < k ? o : k;
    }
    if (__lane_id() == leader) at_min_u64(&M.tvote[lt], k);
    if (mine) pending = false;
  }
  if (pending) at_min_u64(&M.tvote[t], key);
}

__global__ void __launch_bounds__(kBlock) k_pick_count(Mesh M, const uint32_t* utet, uint32_t* blk, int ring) {
  const uint32_t U = M.ctr[C_ULEN];
  uint32_t i = gid();
  bool w = i < U && is_winner(M, utet, i, ring);
  uint32_t tot;
  block_rank(w, &tot);
  if (threadIdx.x == 0) blk[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kBlock) k_pick_scatter(Mesh M, const uint32_t* uvert, const uint32_t* utet,
                               uint2* wlist, uint32_t* uvert2, uint32_t* utet2, const uint32_t* blk,
                               int ring) {
  const uint32_t U = M.ctr[C_ULEN], round = M.ctr[C_ROUND];
  uint32_t i = gid();
  bool w = i < U && is_winner(M, utet, i, ring);
  uint32_t tot;
  uint32_t r = blk[blockIdx.x] + block_rank(w, &tot);
  if (i >= U) return;
  uint32_t t = utet[i], v = uvert[i];
  if (w) {
    wlist[r] = make_uint2(t, v);
    M.tsplit[t] = make_uint2(round, M.ctr[C_BASE] + 3u * r);
    M.tpoint[t] = v;
  } else {
    uvert2[i - r] = v;
    utet2[i - r] = t;
  }
}

// Launch grids of the round kernels: a power of two of workgroups covering
// the host's bound, so that a captured round can be replayed for any count
// up to that bound (the kernels read the counts from ctr).
uint32_t round_grid(uint64_t bound) {
  uint32_t want = std::max<uint32_t>(1u, nblocks(bound)), g = 1;
  while (g < want) g <<= 1;
  return g;
}

void launch_vote(const Mesh& M, const uint32_t* uvert, const uint32_t* utet, uint32_t U,
                 hipStream_t s) {
  k_vote<<<round_grid(U), kBlock, 0, s>>>(M, uvert, utet);
}

// votes of tets whose winner was deferred must not survive into the next
// round (list positions change); split resets the votes of winning tets
__global__ void k_clear_votes(Mesh M, const uint32_t* utet2, const uint32_t* ctr) {
  uint32_t i = gid();
  if (i < ctr[C_SURV]) M.tvote[utet2[i]] = kVoteNone;
}

void launch_pick(const Mesh& M, const uint32_t* uvert, const uint32_t* utet, uint32_t U,
                 uint2* wlist, uint32_t* uvert2, uint32_t* utet2, uint32_t* blk, int ring,
                 hipStream_t s) {
  // blocks past the count write zero block totals, so the scan may cover
  // the whole grid
  uint32_t nb = round_grid(U);
  k_pick_count<<<nb, kBlock, 0, s>>>(M, utet, blk, ring);
  k_scan_blocks<<<1, 1024, 0, s>>>(blk, nb, M.ctr, 0);
  k_pick_scatter<<<nb, kBlock, 0, s>>>(M, uvert, utet, wlist, uvert2, utet2, blk, ring);
  if (ring > 0) k_clear_votes<<<nb, kBlock, 0, s>>>(M, utet2, M.ctr);
}

// The children join the active list with one counter atomic per wave, in
// child-major order within the wave (the wave's child-0 slots, then its
// child-1 slots, ...).  The order matters: the fresh slots of the flips
// that follow are numbered in list order, and lower slot ids win flip
// votes.  On co-spherical inputs a tet-major order left about four times
// as many faces stuck (1M points on a sphere: 3,150 against 866).
__global__ void __launch_bounds__(kBlock) k_split(Mesh M, const uint2* wlist, FlipLists L) {
  const uint32_t W = M.ctr[C_WIN], round = M.ctr[C_ROUND];
  uint32_t k = gid();
  uint32_t c[4], m = 0u;
  if (k < W) m = split_body(M, wlist, k, round, c);
  uint64_t lt = (1ull << __lane_id()) - 1ull;
  uint64_t b0 = __ballot(m & 1u), bv = __ballot(m != 0u);
  uint32_t n0 = __popcll(b0), nv = __popcll(bv);
  uint32_t tot = n0 + 3u * nv;
  if (tot == 0u) return;
  uint32_t base = 0u;
  if (__lane_id() == 0) base = atomicAdd(&M.ctr[C_ACT_NEXT], tot);
  base = __shfl(base, 0);
  uint32_t* act_next = L.act[M.ctr[C_PAR] & 1u];
  if (m & 1u) act_next[base + __popcll(b0 & lt)] = c[0];
  if (m) {
    uint32_t r = __popcll(bv & lt);
    for (uint32_t j = 1; j < 4; ++j) act_next[base + n0 + (j - 1u) * nv + r] = c[j];
  }
}

// ctr[dst] = ctr[src]
__global__ void k_ctr_copy(uint32_t* ctr, int dst, int src) {
  if (threadIdx.x == 0) ctr[dst] = ctr[src];
}

void launch_split(const Mesh& M, const uint2* wlist, uint32_t W, const FlipLists& L, hipStream_t s) {
  // the children are appended after the active entries carried over
  k_ctr_copy<<<1, 64, 0, s>>>(M.ctr, C_ACT_NEXT, C_A);
  k_split<<<round_grid(W), kBlock, 0, s>>>(M, wlist, L);
  k_ctr_copy<<<1, 64, 0, s>>>(M.ctr, C_A, C_ACT_NEXT);
}

// the next round's uninserted count and number, on the device
__global__ void k_round_next(uint32_t* ctr) {
  if (threadIdx.x == 0) {
    ctr[C_ULEN] = ctr[C_SURV];
    ctr[C_ROUND] += 1u;
  }
}

void launch_round_next(uint32_t* ctr, hipStream_t s) { k_round_next<<<1, 64, 0, s>>>(ctr); }

// Point kernels in three forms: LEAN runs the FP32 filter alone and defers
// the points it cannot decide to `slow` (ctr[C_SLOW_P]); SLOW runs the full
// staged predicates on those; FULL runs them on every point (exact mode).
enum KForm : int { LEAN = 0, FULL = 1, SLOW = 2 };

// Grid of the point kernels' SLOW form, for the points the LEAN form
// deferred (their count is on the device): a quarter of the LEAN grid, at
// most 1,024 blocks.  Each deferred point keeps its thread in the full
// predicates for microseconds; degenerate input defers many (a 128^3
// lattice's relocation: 26.0 ms at a sixteenth and 256 blocks, 18.5 ms).
static uint32_t slow_grid(uint32_t g) { return std::min(g / 4u + 16u, 1024u); }

// The survivors of the round's pick (ctr[C_SURV]) in the round ctr[C_ROUND].
template <int FORM>
__global__ void __launch_bounds__(kBlock) k_split_points(Mesh M, const uint32_t* uvert,
                                                         uint32_t* utet, uint64_t* st, uint32_t* slow) {
  const uint32_t U = M.ctr[C_SURV], round = M.ctr[C_ROUND];
  PredStats ls = {};
  if (FORM == SLOW) {
    uint32_t n = M.ctr[C_SLOW_P], stride = gridDim.x * blockDim.x;
    for (uint32_t k = gid(); k < n; k += stride)
      (void)split_points_body_t<false>(M, uvert, utet, slow[k], round, &ls);
  } else {
    uint32_t i = gid();
    bool defer = false;
    if (i < U) defer = split_points_body_t<FORM == LEAN>(M, uvert, utet, i, round, &ls);
    if (defer) slow[append_slot(&M.ctr[C_SLOW_P])] = i;
  }
  flush_stats(ls, st);
}

void launch_split_points(const Mesh& M, const uint32_t* uvert, uint32_t* utet, uint32_t U,
                         uint64_t* st, uint32_t* slow, hipStream_t s) {
  if (!U) return;
  uint32_t g = round_grid(U);
  if (M.mode == PRED_FILTERED) {
    launch_set_ctr(M.ctr, C_SLOW_P, 0u, -1, 0u, s);
    k_split_points<LEAN><<<g, kBlock, 0, s>>>(M, uvert, utet, st, slow);
    k_split_points<SLOW><<<slow_grid(g), kBlock, 0, s>>>(M, uvert, utet, st, slow);
  } else {
    k_split_points<FULL><<<g, kBlock, 0, s>>>(M, uvert, utet, st, slow);
  }
}

// ---------------------------------------------------------------------------
// flips
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// flip iterations, driven from the device: every count, the iteration tag,
// the round and the list parity are read from ctr, so a sequence of
// iterations runs (and can be captured in a graph) without the host.  The
// kernels are persistent: a fixed grid strides over the current list.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void iter_begin(uint32_t* ctr) {
  uint32_t A = ctr[C_A];
  uint32_t run = 0u;
  if (!ctr[C_STOP] && A > 0u) {
    // worst case every proposal wins a 2-3 flip and takes a fresh slot
    if ((uint64_t)ctr[C_TETS] + A + 64u > (uint64_t)ctr[C_CAP]) ctr[C_STOP] = 1u;
    else run = 1u;
  }
  ctr[C_RUN] = run;
  if (run) {
    ctr[C_IT] += 1u;
    ctr[C_ACT_NEXT] = 0u;
    ctr[C_FLIPS] = 0u;
    ctr[C_SLOW_T] = 0u;
    ctr[C_TETS_IT] = ctr[C_TETS];
    if (A > ctr[C_PEAKA]) ctr[C_PEAKA] = A;
  }
}

__global__ void k_iter_begin(uint32_t* ctr) {
  if (threadIdx.x == 0) iter_begin(ctr);
}

// sums (IterSum): checked tets, flips, iterations, 2-3 flips (a 2-3 flip
// takes one fresh slot, a 3-2 none), then the checked tets and flips of the
// iterations k_flip_small ran.
__device__ __forceinline__ void iter_end(uint32_t* ctr, unsigned long long* sums) {
  if (!ctr[C_RUN]) return;
  sums[S_CHECKED] += ctr[C_A];
  sums[S_FLIPS] += ctr[C_FLIPS];
  sums[S_ITERS] += 1ull;
  sums[S_FLIPS23] += ctr[C_TETS] - ctr[C_TETS_IT];
  ctr[C_SLOW_SUM] += ctr[C_SLOW_T];
  ctr[C_A] = ctr[C_ACT_NEXT];
  ctr[C_PAR] ^= 1u;
}

__global__ void k_iter_end(uint32_t* ctr, unsigned long long* sums) {
  if (threadIdx.x == 0) iter_end(ctr, sums);
}

// end of one iteration and begin of the next in one launch
__global__ void k_iter_step(uint32_t* ctr, unsigned long long* sums) {
  if (threadIdx.x) return;
  iter_end(ctr, sums);
  iter_begin(ctr);
}

template <int FORM>
__global__ void __launch_bounds__(kBlock) k_flip_check(Mesh M, FlipLists L, uint64_t* st) {
  if (!M.ctr[C_RUN]) return;
  uint32_t it = M.ctr[C_IT];
  const uint32_t* cur = L.act[M.ctr[C_PAR] & 1u];
  uint32_t stride = gridDim.x * blockDim.x;
  PredStats ls = {};
  // LEAN defers what its filters leave to k_flip_check_deferred; FULL runs
  // the full predicates on every entry
  uint32_t A = M.ctr[C_A];
  for (uint32_t j = gid(); j < A; j += stride) {
    bool defer = flip_check_body_t<FORM == LEAN>(M, cur, L.prop, j, it, &ls);
    if (defer) L.slow[append_slot(&M.ctr[C_SLOW_T])] = j;
  }
  flush_stats(ls, st);
}

// The full check of the tets the lean check deferred, four lanes per tet,
// one per face: each lane runs its face's staged insphere and, where the
// face's apex lies inside, the three orient3d that classify the flip; the
// lowest face that proposes is the tet's proposal, as flip_check_body_t's
// loop over the faces would choose it.  A tet's faces then reach the exact
// and SoS stages together rather than one after another in divergent
// passes.  A tet's lanes are consecutive in one wave (k = 4 * entry + face,
// and a wave's first k is a multiple of 64).
__global__ void __launch_bounds__(kBlock) k_flip_check_deferred(Mesh M, FlipLists L, uint64_t* st) {
  if (!M.ctr[C_RUN]) return;
  const uint32_t it = M.ctr[C_IT];
  const uint32_t* cur = L.act[M.ctr[C_PAR] & 1u];
  const uint32_t n4 = 4u * M.ctr[C_SLOW_T], stride = gridDim.x * blockDim.x;
  PredStats ls = {};
  for (uint32_t k = gid(); k < n4; k += stride) {
    const uint32_t j = L.slow[k >> 2];
    const int f = (int)(k & 3u);
    const uint32_t t = cur[j];
    const uint32_t info = INFO(M, t);
    const uint4 tv = TV(M, t);
    const uint4 tov = TOV(M, t);
    const bool live = (info & kAlive) && tv.x != M.inf && tv.y != M.inf && tv.z != M.inf &&
                      tv.w != M.inf;
    bool qual = false;
    uint32_t kind = 0u, bad = 0u, third = kInvalid;
    if (live) {
      float P[4][3], Q[4][3];
      uint32_t R[4], RQ[4];
      for (int q = 0; q < 4; ++q) load_pt(M, get4(tv, q), P[q], &R[q]);
      for (int q = 0; q < 4; ++q) load_pt(M, get4(tov, q), Q[q], &RQ[q]);
      qual = flip_face_full(M, t, tov, P, R, Q, RQ, f, &ls, &kind, &bad, &third);
    }
    const uint64_t qb = __ballot(qual);
    const uint32_t g0 = __lane_id() & ~3u;
    const uint32_t grp = (uint32_t)(qb >> g0) & 0xFu;
    const uint32_t cf = grp ? (uint32_t)__builtin_ctz(grp) : 0u;
    const uint32_t pk = (uint32_t)f | kind << 2 | bad << 3;
    const uint32_t pk_c = (uint32_t)__shfl((int)pk, (int)(g0 + cf));
    const uint32_t third_c = (uint32_t)__shfl((int)third, (int)(g0 + cf));
    if (f == 0) {
      // each entry is in the list once and only its lanes touch its info
      INFO(M, t) = info & ~kQueued;
      if (grp) {
        uint32_t nt = opp_tet(*opp_word(M, t, cf));
        uint64_t key = flip_key(it, t);
        at_min_u64(&M.fvote[t], key);
        at_min_u64(&M.fvote[nt], key);
        if ((pk_c >> 2) & 1u) at_min_u64(&M.fvote[third_c], key);
        L.prop[j] = make_uint4(t, pk_c, nt, third_c);
      } else {
        L.prop[j].x = kInvalid;
      }
    }
  }
  flush_stats(ls, st);
}

// Exclusive prefixes of three small counts in one pass (a, b in {0, 1},
// c in 0..3), packed 10, 10 and 12 bits (block totals of a and b stay
// below 1024, of c below 4096, for blocks of at most 512 threads).  Every
// thread of the block of BS threads must call it; w is LDS of BS/64.
template <int BS>
__device__ __forceinline__ uint32_t block_prefix3(uint32_t a, uint32_t b, uint32_t c, uint32_t* tot,
                                                  uint32_t* w) {
  uint64_t lt = (1ull << __lane_id()) - 1ull;
  uint64_t ba = __ballot(a), bb = __ballot(b), c0 = __ballot(c & 1u), c1 = __ballot(c & 2u);
  auto pack = [](uint32_t x, uint32_t y, uint32_t z) { return x | (y << 10) | (z << 20); };
  uint32_t r = pack(__popcll(ba & lt), __popcll(bb & lt), __popcll(c0 & lt) + 2u * __popcll(c1 & lt));
  int wave = threadIdx.x >> 6;
  if (__lane_id() == 0) w[wave] = pack(__popcll(ba), __popcll(bb), __popcll(c0) + 2u * __popcll(c1));
  __syncthreads();
  uint32_t off = 0, t = 0;
#pragma unroll
  for (int k = 0; k < BS / 64; ++k) {
    off += k < wave ? w[k] : 0u;
    t += w[k];
  }
  __syncthreads();
  *tot = t;
  return off + r;
}

// Decide every proposal; a winner loads its flip's tets at once, so those
// loads overlap its queue claims and the block's counter claims.  Then the
// block claims list space with one atomic per counter (flips, fresh slots,
// next active entries) -- the counters share a cache line, so per-wave
// atomics from every wave would serialize -- and every winner executes its
// own flip from registers.
// Blocks of 512 threads: each tile of proposals claims its flips, fresh
// slots and queue entries with one atomic per counter, and those
// same-address atomics complete one at a time at the memory side; 512
// halves them against 256 (1M uniform device 26.0-26.1 -> 25.7-26.0 ms).
constexpr uint32_t kDecBlock = 512;
static_assert(kDecBlock <= 512, "block_prefix3 packs totals of at most 512 threads");

__global__ void __launch_bounds__(kDecBlock) k_flip_decide_exec(Mesh M, FlipLists L) {
  constexpr uint32_t kBS = kDecBlock;
  __shared__ uint32_t wsum[kBS / 64];
  __shared__ uint32_t base[3];
  if (!M.ctr[C_RUN]) return;
  uint32_t A = M.ctr[C_A], it = M.ctr[C_IT], round = M.ctr[C_ROUND];
  uint32_t* act_next = L.act[(M.ctr[C_PAR] & 1u) ^ 1u];
  for (uint32_t b0 = blockIdx.x * kBS; b0 < A; b0 += gridDim.x * kBS) {
    uint32_t j = b0 + threadIdx.x;
    uint4 pr = make_uint4(kInvalid, 0u, 0u, 0u);
    if (j < A) pr = L.prop[j];
    uint32_t qid[2], nq = 0, nslot = 0;
    FlipTets F;
    bool win = flip_decide_load(M, pr, it, qid, &nq, &nslot, F);
    uint32_t nqa = nq + nslot;  // the fresh slot is queued too
    uint32_t tot;
    uint32_t r = block_prefix3<kBS>(win ? 1u : 0u, nslot, nqa, &tot, wsum);
    uint32_t rf = r & 1023u, rs = (r >> 10) & 1023u, rq = r >> 20;
    uint32_t tf = tot & 1023u, ts = (tot >> 10) & 1023u, tq = tot >> 20;
    if (threadIdx.x == 0) {
      base[0] = tf ? atomicAdd(&M.ctr[C_FLIPS], tf) : 0u;
      base[1] = ts ? atomicAdd(&M.ctr[C_TETS], ts) : 0u;
      base[2] = tq ? atomicAdd(&M.ctr[C_ACT_NEXT], tq) : 0u;
    }
    __syncthreads();
    uint32_t q = base[2] + rq;
    for (uint32_t k = 0; k < nq; ++k) act_next[q + k] = qid[k];
    if (win) {
      uint32_t sl = nslot ? base[1] + rs : kInvalid;
      if (nslot) act_next[q + nq] = sl;
      flip_exec_loaded(M, pr, F, it, round, sl, &L.out[2 * (size_t)(base[0] + rf)]);
    }
    // the next tile's prefix sums synchronize before base is rewritten
  }
}

__global__ void __launch_bounds__(kBlock) k_flip_update(Mesh M, FlipLists L) {
  if (!M.ctr[C_RUN]) return;
  uint32_t F = M.ctr[C_FLIPS], it = M.ctr[C_IT];
  uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t k = gid(); k < F; k += stride) flip_update_body(M, L.out, k, it);
}

// Up to k flip iterations of a short list in one workgroup: the phases of
// an iteration (check with the full predicates, decide + execute, adjacency
// update, end) are separated by workgroup barriers instead of launches.
// Stops at the first iteration whose list is empty or longer than `limit`;
// a longer list stays for the next batch of iterations.  Counters that
// other threads of the block change are read with atomic loads.
__global__ void __launch_bounds__(kBlock) k_flip_small(Mesh M, FlipLists L, uint64_t* st,
                                                       unsigned long long* sums, int k, uint32_t limit) {
  __shared__ uint32_t run;
  PredStats ls = {};
  uint32_t* ctr = M.ctr;
  auto ld = [&](int i) { return __atomic_load_n(&ctr[i], __ATOMIC_RELAXED); };
  const uint32_t round = ld(C_ROUND);
  for (int i = 0; i < k; ++i) {
    if (threadIdx.x == 0) {
      uint32_t A = ld(C_A);
      uint32_t go = 0u;
      if (A > 0u && A <= limit && !ld(C_STOP)) {
        if ((uint64_t)ld(C_TETS) + A + 64u > (uint64_t)ld(C_CAP)) {
          __atomic_store_n(&ctr[C_STOP], 1u, __ATOMIC_RELAXED);
        } else {
          go = 1u;
          __atomic_store_n(&ctr[C_IT], ld(C_IT) + 1u, __ATOMIC_RELAXED);
          __atomic_store_n(&ctr[C_ACT_NEXT], 0u, __ATOMIC_RELAXED);
          __atomic_store_n(&ctr[C_FLIPS], 0u, __ATOMIC_RELAXED);
          __atomic_store_n(&ctr[C_SLOW_T], 0u, __ATOMIC_RELAXED);
          __atomic_store_n(&ctr[C_TETS_IT], ld(C_TETS), __ATOMIC_RELAXED);
          if (A > ld(C_PEAKA)) __atomic_store_n(&ctr[C_PEAKA], A, __ATOMIC_RELAXED);
        }
      }
      __atomic_store_n(&ctr[C_RUN], go, __ATOMIC_RELAXED);
      run = go;
    }
    __syncthreads();
    if (!run) break;
    uint32_t A = ld(C_A), it = ld(C_IT), par = ld(C_PAR) & 1u;
    const uint32_t* cur = L.act[par];
    uint32_t* nxt = L.act[par ^ 1u];
    for (uint32_t j = threadIdx.x; j < A; j += blockDim.x)
      (void)flip_check_body_t<false>(M, cur, L.prop, j, it, &ls);
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < A; j += blockDim.x)
      flip_decide_exec_body(M, L.prop, j, it, round, L.out, nxt);
    __syncthreads();
    uint32_t F = ld(C_FLIPS);
    for (uint32_t f = threadIdx.x; f < F; f += blockDim.x) flip_update_body(M, L.out, f, it);
    __syncthreads();
    if (threadIdx.x == 0) {
      sums[S_CHECKED] += A;
      sums[S_FLIPS] += F;
      sums[S_ITERS] += 1ull;
      sums[S_FLIPS23] += ld(C_TETS) - ld(C_TETS_IT);
      sums[S_SMALL_CHECKED] += A;
      sums[S_SMALL_FLIPS] += F;
      __atomic_store_n(&ctr[C_A], ld(C_ACT_NEXT), __ATOMIC_RELAXED);
      __atomic_store_n(&ctr[C_PAR], par ^ 1u, __ATOMIC_RELAXED);
    }
    __syncthreads();
  }
  flush_stats(ls, st);
}

void launch_flip_small(const Mesh& M, const FlipLists& L, uint64_t* st, uint64_t* sums, int k,
                       uint32_t limit, hipStream_t s) {
  k_flip_small<<<1, kBlock, 0, s>>>(M, L, st, reinterpret_cast<unsigned long long*>(sums), k, limit);
}

__global__ void k_set_ctr(uint32_t* ctr, int idx, uint32_t v, int idx2, uint32_t v2) {
  if (threadIdx.x == 0) {
    ctr[idx] = v;
    if (idx2 >= 0) ctr[idx2] = v2;
  }
}

void launch_set_ctr(uint32_t* ctr, int idx, uint32_t v, int idx2, uint32_t v2, hipStream_t s) {
  k_set_ctr<<<1, 64, 0, s>>>(ctr, idx, v, idx2, v2);
}

void launch_iter_begin(uint32_t* ctr, hipStream_t s) { k_iter_begin<<<1, 64, 0, s>>>(ctr); }
void launch_iter_end(uint32_t* ctr, uint64_t* sums, hipStream_t s) {
  k_iter_end<<<1, 64, 0, s>>>(ctr, reinterpret_cast<unsigned long long*>(sums));
}
void launch_iter_step(uint32_t* ctr, uint64_t* sums, hipStream_t s) {
  k_iter_step<<<1, 64, 0, s>>>(ctr, reinterpret_cast<unsigned long long*>(sums));
}
void launch_flip_check(const Mesh& M, const FlipLists& L, uint32_t grid, uint64_t* st, hipStream_t s) {
  if (M.mode == PRED_FILTERED) k_flip_check<LEAN><<<grid, kBlock, 0, s>>>(M, L, st);
  else k_flip_check<FULL><<<grid, kBlock, 0, s>>>(M, L, st);
}
void launch_flip_check_deferred(const Mesh& M, const FlipLists& L, uint32_t grid, uint64_t* st,
                                hipStream_t s) {
  if (M.mode == PRED_FILTERED) k_flip_check_deferred<<<grid, kBlock, 0, s>>>(M, L, st);
}
// grid: blocks of kBlock threads, as the other list kernels; the same
// threads in blocks of kDecBlock
void launch_flip_decide_exec(const Mesh& M, const FlipLists& L, uint32_t grid, hipStream_t s) {
  k_flip_decide_exec<<<std::max<uint32_t>(grid * kBlock / kDecBlock, 1u), kDecBlock, 0, s>>>(M, L);
}
void launch_flip_update(const Mesh& M, const FlipLists& L, uint32_t grid, hipStream_t s) {
  k_flip_update<<<grid, kBlock, 0, s>>>(M, L);
}

// The uninserted points after the round's pick (ctr[C_SURV]), round ctr[C_ROUND].
template <int FORM>
__global__ void __launch_bounds__(kBlock) k_relocate(Mesh M, const uint32_t* uvert, uint32_t* utet,
                                                     uint64_t* st, uint32_t* slow) {
  const uint32_t U = M.ctr[C_SURV], round = M.ctr[C_ROUND];
  PredStats ls = {};
  if (FORM == SLOW) {
    uint32_t n = M.ctr[C_SLOW_P], stride = gridDim.x * blockDim.x;
    for (uint32_t k = gid(); k < n; k += stride)
      (void)relocate_body_t<false>(M, uvert, utet, slow[k], round, &ls);
  } else {
    uint32_t i = gid();
    bool defer = false;
    if (i < U) defer = relocate_body_t<FORM == LEAN>(M, uvert, utet, i, round, &ls);
    if (defer) slow[append_slot(&M.ctr[C_SLOW_P])] = i;
  }
  flush_stats(ls, st);
}

void launch_relocate(const Mesh& M, const uint32_t* uvert, uint32_t* utet, uint32_t U,
                     uint64_t* st, uint32_t* slow, hipStream_t s) {
  if (!U) return;
  uint32_t g = round_grid(U);
  if (M.mode == PRED_FILTERED) {
    launch_set_ctr(M.ctr, C_SLOW_P, 0u, -1, 0u, s);
    k_relocate<LEAN><<<g, kBlock, 0, s>>>(M, uvert, utet, st, slow);
    k_relocate<SLOW><<<slow_grid(g), kBlock, 0, s>>>(M, uvert, utet, st, slow);
  } else {
    k_relocate<FULL><<<g, kBlock, 0, s>>>(M, uvert, utet, st, slow);
  }
}

// LEAN: every slot, FP32 filter, undecided tets to slow[ctr[C_SLOW_T]];
// SLOW: the full predicates on those; FULL: the full predicates on every slot
template <int FORM>
__global__ void __launch_bounds__(kBlock) k_audit(Mesh M, uint32_t T, uint32_t* rec, uint32_t cap,
                                                  uint64_t* st, uint32_t* slow) {
  PredStats ls = {};
  if (FORM == SLOW) {
    uint32_t n = M.ctr[C_SLOW_T], stride = gridDim.x * blockDim.x;
    for (uint32_t k = gid(); k < n; k += stride) (void)audit_body_t<false>(M, slow[k], rec, cap, &ls);
  } else {
    uint32_t t = gid();
    if (t < T && audit_body_t<FORM == LEAN>(M, t, rec, cap, &ls))
      slow[append_slot(&M.ctr[C_SLOW_T])] = t;
  }
  flush_stats(ls, st);
}

void launch_audit(const Mesh& M, uint32_t T, uint32_t* rec, uint32_t cap, uint64_t* st,
                  uint32_t* slow, hipStream_t s) {
  if (!T) return;
  if (M.mode == PRED_FILTERED) {
    k_set_ctr<<<1, 64, 0, s>>>(M.ctr, C_SLOW_T, 0u, -1, 0u);
    k_audit<LEAN><<<nblocks(T), kBlock, 0, s>>>(M, T, rec, cap, st, slow);
    k_audit<SLOW><<<std::min<uint32_t>(nblocks(T), 1024u), kBlock, 0, s>>>(M, T, rec, cap, st, slow);
  } else {
    k_audit<FULL><<<nblocks(T), kBlock, 0, s>>>(M, T, rec, cap, st, slow);
  }
}

// ---------------------------------------------------------------------------
// output
// ---------------------------------------------------------------------------
__global__ void k_alive_count(Mesh M, uint32_t T, uint32_t* blk) {
  uint32_t t = gid();
  bool a = t < T && (INFO(M, t) & kAlive);
  uint32_t tot;
  block_rank(a, &tot);
  if (threadIdx.x == 0) blk[blockIdx.x] = tot;
}

// one cloud: live slots in slot order are the rows
__global__ void k_alive_scatter(Mesh M, uint32_t T, const uint32_t* blk, uint32_t* olist,
                                uint32_t* slot2pos) {
  uint32_t t = gid();
  bool a = t < T && (INFO(M, t) & kAlive);
  uint32_t tot;
  uint32_t r = blk[blockIdx.x] + block_rank(a, &tot);
  if (t < T) {
    slot2pos[t] = a ? r : kInvalid;
    if (a) olist[r] = t;
  }
}

// many clouds: sort key per slot = cloud of a live tet, or C for a free slot
__global__ void k_slot_keys(Mesh M, uint32_t T, Clouds K, uint32_t* okey, uint32_t* oslot) {
  uint32_t t = gid();
  if (t >= T) return;
  okey[t] = (INFO(M, t) & kAlive) ? cloud_of_tet(M, t, K) : K.C;
  oslot[t] = t;
}

// after the sort: row j holds slot olist[j]; cstart[c] = first row of
// cloud c (every cloud has live tets), cstart[C] = number of live tets
__global__ void k_row_positions(const uint32_t* okey, const uint32_t* olist, uint32_t T, uint32_t C,
                                uint32_t* slot2pos, uint32_t* cstart) {
  uint32_t j = gid();
  if (j >= T) return;
  uint32_t c = okey[j];
  slot2pos[olist[j]] = c < C ? j : kInvalid;
  if (c < C && (j == 0 || okey[j - 1] != c)) cstart[c] = j;
}

__global__ void k_cstart_tail(uint32_t* cstart, uint32_t C, const uint32_t* ctr) {
  if (threadIdx.x == 0) cstart[C] = ctr[C_ALIVE];
}

__global__ void k_output_rows(Mesh M, uint32_t T, const uint32_t* olist, const uint32_t* okey,
                              const uint32_t* cstart, const uint32_t* slot2pos, const uint32_t* vrank,
                              Clouds K, uint32_t* out_tets, uint32_t* out_adj) {
  uint32_t j = gid();
  if (j >= T || j >= M.ctr[C_ALIVE]) return;
  output_body(M, j, olist, okey, cstart, slot2pos, vrank, K, out_tets, out_adj);
}

__global__ void k_vertex_ranks(const float4* pts, uint32_t nv, uint32_t* vrank) {
  uint32_t v = gid();
  if (v < nv) vrank[v] = pt_rank(pts[v]);
}

hipError_t launch_output(const Mesh& M, uint32_t T, const Clouds& K, const OutputBufs& B, hipStream_t s) {
  uint32_t nb = nblocks(T);
  k_alive_count<<<nb, kBlock, 0, s>>>(M, T, B.blk);
  k_scan_blocks<<<1, 1024, 0, s>>>(B.blk, nb, M.ctr, 1);
  const uint32_t* olist;
  const uint32_t* okey = nullptr;
  if (K.C <= 1) {
    k_alive_scatter<<<nb, kBlock, 0, s>>>(M, T, B.blk, B.olist, B.slot2pos);
    k_fill_u32<<<1, kBlock, 0, s>>>(B.cstart, 0u, 1);
    k_cstart_tail<<<1, 64, 0, s>>>(B.cstart, 1u, M.ctr);
    olist = B.olist;
  } else {
    k_slot_keys<<<nb, kBlock, 0, s>>>(M, T, K, B.okey, B.olist);
    hipError_t e = sort_pairs_u32(B.okey, B.olist, B.okey2, B.olist2, T, cloud_key_bits(K.C), B.temp,
                                  B.temp_bytes, s);
    if (e != hipSuccess) return e;
    k_row_positions<<<nb, kBlock, 0, s>>>(B.okey2, B.olist2, T, K.C, B.slot2pos, B.cstart);
    k_cstart_tail<<<1, 64, 0, s>>>(B.cstart, K.C, M.ctr);
    olist = B.olist2;
    okey = B.okey2;
  }
  if (B.out_tets) {
    k_vertex_ranks<<<nblocks((uint64_t)K.V + 1u), kBlock, 0, s>>>(M.pts, K.V + 1u, B.vrank);
    k_output_rows<<<nb, kBlock, 0, s>>>(M, T, olist, okey, B.cstart, B.slot2pos, B.vrank, K, B.out_tets,
                                        B.out_adj);
  }
  return hipGetLastError();
}

__global__ void k_map_records(uint32_t* rec, uint32_t n, const uint32_t* slot2idx) {
  uint32_t i = gid();
  if (i < n) rec[i] = (slot2idx[rec[i] >> 2] << 2) | (rec[i] & 3u);
}

void launch_map_records(uint32_t* rec, uint32_t n, const uint32_t* slot2idx, hipStream_t s) {
  if (n) k_map_records<<<nblocks(n), kBlock, 0, s>>>(rec, n, slot2idx);
}

// ---------------------------------------------------------------------------
// neighbourhood of the rows the host repair will touch: the rows of the
// recorded faces and the last `tail` rows (moved into holes), grown ring by
// ring over the output adjacency.  list[0, cnt) holds each marked row once;
// ring k reads list[bounds[k], bounds[k+1]).
// ---------------------------------------------------------------------------
__global__ void k_nbr_seed(const uint32_t* rec, uint32_t nrec, uint32_t M, uint32_t tail,
                           uint32_t* mark, uint32_t* list, uint32_t* cnt) {
  uint32_t i = gid();
  uint32_t row = kInvalid;
  if (i < nrec) row = rec[i] >> 2;
  else if (i < nrec + tail) row = M - tail + (i - nrec);
  bool add = row < M && atomicOr(&mark[row], 1u) == 0u;
  if (add) list[append_slot(cnt)] = row;
}

__global__ void k_nbr_snap(const uint32_t* cnt, uint32_t* bounds, int k) {
  if (threadIdx.x == 0) bounds[k] = cnt[0];
}

__global__ void k_nbr_ring(const uint32_t* adj, uint32_t M, uint32_t* mark, uint32_t* list,
                           uint32_t* cnt, const uint32_t* bounds, int k) {
  uint32_t b = bounds[k], e = bounds[k + 1];
  uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = b + gid(); i < e; i += stride) {
    uint32_t r = list[i];
    for (int f = 0; f < 4; ++f) {
      uint32_t a = adj[4 * (size_t)r + f];
      if (a == kInvalid) continue;
      uint32_t nb = a >> 5;
      if (nb < M && atomicOr(&mark[nb], 1u) == 0u) list[append_slot(cnt)] = nb;
    }
  }
}

// rows list[0, n) of the output arrays -> packed (row, tet[4], adj[4])
__global__ void k_nbr_gather(const uint32_t* list, uint32_t n, const uint32_t* tets,
                             const uint32_t* adj, uint32_t* out) {
  uint32_t i = gid();
  if (i >= n) return;
  uint32_t r = list[i];
  uint32_t* q = out + 9 * (size_t)i;
  q[0] = r;
  uint4 t = reinterpret_cast<const uint4*>(tets)[r];
  uint4 a = reinterpret_cast<const uint4*>(adj)[r];
  q[1] = t.x; q[2] = t.y; q[3] = t.z; q[4] = t.w;
  q[5] = a.x; q[6] = a.y; q[7] = a.z; q[8] = a.w;
}

void launch_neighbourhood(const uint32_t* rec, uint32_t nrec, uint32_t M, uint32_t tail, int rings,
                          const uint32_t* adj, uint32_t* mark, uint32_t* list, uint32_t* cnt,
                          uint32_t* bounds, bool more, hipStream_t s) {
  if (!more) {
    k_fill_u32<<<nblocks(M), kBlock, 0, s>>>(mark, 0u, M);
    k_fill_u32<<<1, kBlock, 0, s>>>(cnt, 0u, 1);
  }
  k_nbr_snap<<<1, 64, 0, s>>>(cnt, bounds, 0);
  if (nrec + tail)
    k_nbr_seed<<<nblocks((uint64_t)nrec + tail), kBlock, 0, s>>>(rec, nrec, M, tail, mark, list, cnt);
  for (int k = 0; k < rings; ++k) {
    k_nbr_snap<<<1, 64, 0, s>>>(cnt, bounds, k + 1);
    k_nbr_ring<<<256, kBlock, 0, s>>>(adj, M, mark, list, cnt, bounds, k);
  }
}

void launch_nbr_gather(const uint32_t* list, uint32_t n, const uint32_t* tets, const uint32_t* adj,
                       uint32_t* out, hipStream_t s) {
  if (n) k_nbr_gather<<<nblocks(n), kBlock, 0, s>>>(list, n, tets, adj, out);
}

__global__ void k_apply_patch(const uint32_t* patch, uint32_t n, uint32_t* tets, uint32_t* adj) {
  uint32_t i = gid();
  if (i >= n) return;
  const uint32_t* q = patch + 9 * (size_t)i;
  uint32_t sl = q[0];
  reinterpret_cast<uint4*>(tets)[sl] = make_uint4(q[1], q[2], q[3], q[4]);
  reinterpret_cast<uint4*>(adj)[sl] = make_uint4(q[5], q[6], q[7], q[8]);
}

void launch_apply_patch(const uint32_t* patch, uint32_t n, uint32_t* tets, uint32_t* adj,
                        hipStream_t s) {
  if (n) k_apply_patch<<<nblocks(n), kBlock, 0, s>>>(patch, n, tets, adj);
}

// ---------------------------------------------------------------------------
// fills and predicate batches
// ---------------------------------------------------------------------------
__global__ void k_fill_u32(uint32_t* p, uint32_t v, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_u64(uint64_t* p, uint64_t v, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_copy16(uint4* dst, const uint4* src, uint64_t n) {
  uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] = src[i];
}
void launch_copy16(uint4* dst, const uint4* src, uint64_t n, hipStream_t s) {
  if (n) k_copy16<<<(uint32_t)std::min<uint64_t>(nblocks(n), 8192), kBlock, 0, s>>>(dst, src, n);
}

void launch_fill_u32(uint32_t* p, uint32_t v, uint64_t n, hipStream_t s) {
  if (n) k_fill_u32<<<nblocks(n), kBlock, 0, s>>>(p, v, n);
}
void launch_fill_u64(uint64_t* p, uint64_t v, uint64_t n, hipStream_t s) {
  if (n) k_fill_u64<<<nblocks(n), kBlock, 0, s>>>(p, v, n);
}

__global__ void k_pred_batch(int kind, const float* pts, const uint32_t* idx, uint32_t n, int mode,
                             int32_t* out, uint64_t* st) {
  uint32_t i = gid();
  PredStats ls = {};
  if (i < n) {
    if (kind == 0) {
      const float* q = pts + 12 * (size_t)i;
      const uint32_t* k = idx + 4 * (size_t)i;
      out[i] = orient3d_sos(q, q + 3, q + 6, q + 9, k[0], k[1], k[2], k[3], mode, &ls);
    } else if (kind == 1) {
      const float* q = pts + 15 * (size_t)i;
      const uint32_t* k = idx + 5 * (size_t)i;
      out[i] = orient4d(q, q + 3, q + 6, q + 9, q + 12, k[0], k[1], k[2], k[3], k[4], mode, &ls);
    } else if (kind == 2) {
      const float* q = pts + 12 * (size_t)i;
      out[i] = orient3d_raw(q, q + 3, q + 6, q + 9, mode, &ls);
    } else {
      const float* q = pts + 15 * (size_t)i;
      out[i] = insphere_raw(q, q + 3, q + 6, q + 9, q + 12, mode, &ls);
    }
  }
  flush_stats(ls, st);
}

void launch_pred_batch(int kind, const float* pts, const uint32_t* idx, uint32_t n, int mode,
                       int32_t* out, uint64_t* st, hipStream_t s) {
  if (n) k_pred_batch<<<nblocks(n), kBlock, 0, s>>>(kind, pts, idx, n, mode, out, st);
}

}  // namespace wtet

// HIP kernels of the triangulation loop (gfx950, wave64).
//
// Element-wise kernels wrap the bodies of ops.h.  List compaction uses
// wave ballots: each wavefront ranks its lanes with popcount(ballot & lt),
// wave totals are combined in LDS, block totals are scanned by one block,
// and a scatter pass writes every element once (stable order).
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "kernels.h"

namespace wtet {

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
// point preparation
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t ord_f32(float f) {
  uint32_t u = f32_bits(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float unord_f32(uint32_t o) {
  uint32_t u = (o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

__global__ void k_minmax(const float* xyz, uint32_t n, uint32_t* mm6) {
  uint32_t i = gid();
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
  for (int k = 0; k < 3; ++k) {
    for (int off = 32; off > 0; off >>= 1) {
      lo[k] = min(lo[k], (uint32_t)__shfl_xor(lo[k], off));
      hi[k] = max(hi[k], (uint32_t)__shfl_xor(hi[k], off));
    }
    if (__lane_id() == 0) {
      atomicMin(&mm6[k], lo[k]);
      atomicMax(&mm6[3 + k], hi[k]);
    }
  }
}

// spread the low 21 bits of v to every third bit of a 64-bit word
__device__ __forceinline__ uint64_t expand21(uint64_t v) {
  v &= 0x1FFFFFull;
  v = (v | (v << 32)) & 0x001F00000000FFFFull;
  v = (v | (v << 16)) & 0x001F0000FF0000FFull;
  v = (v | (v << 8)) & 0x100F00F00F00F00Full;
  v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}

#pragma clang fp contract(off)
__global__ void k_normalize(const float* xyz, uint32_t n, const uint32_t* mm6, float* xyzn,
                            uint64_t* keys, uint32_t* ids) {
  uint32_t i = gid();
  if (i >= n) return;
  float mn[3], ext[3];
  for (int k = 0; k < 3; ++k) {
    mn[k] = unord_f32(mm6[k]);
    ext[k] = unord_f32(mm6[3 + k]) - mn[k];
  }
  float s = ext[0];
  if (ext[1] > s) s = ext[1];
  if (ext[2] > s) s = ext[2];
  if (!(s >= 1e-10f)) s = 1e-10f;
  uint32_t q[3];
  for (int k = 0; k < 3; ++k) {
    float v = (xyz[3 * (size_t)i + k] - mn[k]) / s;  // IEEE f32 sub and div
    xyzn[3 * (size_t)i + k] = v;
    float c = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
    q[k] = (uint32_t)(c * 2097151.0f);
  }
  // 63-bit Morton key: spatial order for memory locality (the insertion
  // order cannot change the triangulation)
  keys[i] = (expand21(q[2]) << 2) | (expand21(q[1]) << 1) | expand21(q[0]);
  ids[i] = i;
}
#pragma clang fp contract(on)

__global__ void k_gather_points(const float* xyzn, const uint32_t* ids, uint32_t n, float4* pts) {
  uint32_t j = gid();
  if (j >= n) return;
  uint32_t o = ids[j];
  float w;
  __builtin_memcpy(&w, &o, 4);
  pts[j] = make_float4(xyzn[3 * (size_t)o], xyzn[3 * (size_t)o + 1], xyzn[3 * (size_t)o + 2], w);
}

void launch_minmax(const float* xyz, uint32_t n, uint32_t* mm6, hipStream_t s) {
  k_minmax<<<nblocks(n), kBlock, 0, s>>>(xyz, n, mm6);
}
void launch_normalize(const float* xyz, uint32_t n, const uint32_t* mm6, float* xyzn,
                      uint64_t* keys, uint32_t* ids, hipStream_t s) {
  k_normalize<<<nblocks(n), kBlock, 0, s>>>(xyz, n, mm6, xyzn, keys, ids);
}
void launch_gather_points(const float* xyzn, const uint32_t* ids, uint32_t n, float4* pts,
                          hipStream_t s) {
  k_gather_points<<<nblocks(n), kBlock, 0, s>>>(xyzn, ids, n, pts);
}

size_t sort_temp_bytes(uint32_t n) {
  size_t bytes = 0;
  (void)rocprim::radix_sort_pairs((void*)nullptr, bytes, (uint64_t*)nullptr, (uint64_t*)nullptr,
                                  (uint32_t*)nullptr, (uint32_t*)nullptr, n, 0, 63);
  return bytes;
}

void sort_pairs(uint64_t* keys, uint32_t* ids, uint64_t* keys_tmp, uint32_t* ids_tmp, uint32_t n,
                void* temp, size_t temp_bytes, hipStream_t s) {
  (void)rocprim::radix_sort_pairs(temp, temp_bytes, keys, keys_tmp, ids, ids_tmp, n, 0, 63, s);
}

// Exact duplicates (equal normalized coordinates) share a key, so they sit
// within a run of equal keys after the sort.  All but the highest input id
// are redundant under the lifted SoS and are not inserted; the uninserted
// list is built from the rest (stable, one scan).
__device__ __forceinline__ bool redundant_dup(const float4* pts, const uint64_t* keys, uint32_t n,
                                              uint32_t j, uint32_t* capped) {
  float4 p = pts[j];
  uint32_t id = pt_rank(p);
  uint64_t kj = keys[j];
  const int kCap = 64;
  for (int d = -1; d <= 1; d += 2) {
    int steps = 0;
    for (int64_t k = (int64_t)j + d; k >= 0 && k < (int64_t)n && keys[k] == kj; k += d) {
      if (++steps > kCap) {
        *capped = 1u;
        break;
      }
      float4 q = pts[k];
      if (q.x == p.x && q.y == p.y && q.z == p.z && pt_rank(q) > id) return true;
    }
  }
  return false;
}

__global__ void k_uninit_count(const float4* pts, const uint64_t* keys, uint32_t n,
                               uint32_t* blk, uint32_t* ctr) {
  uint32_t j = gid();
  uint32_t capped = 0;
  bool keep = j < n && !redundant_dup(pts, keys, n, j, &capped);
  if (capped) atomicAdd(&ctr[C_DUPSCAN], 1u);
  uint32_t tot;
  block_rank(keep, &tot);
  if (threadIdx.x == 0) blk[blockIdx.x] = tot;
}

__global__ void k_uninit_scatter(const float4* pts, const uint64_t* keys, uint32_t n,
                                 const uint32_t* blk, uint32_t* uvert, uint32_t* utet) {
  uint32_t j = gid();
  uint32_t capped = 0;
  bool keep = j < n && !redundant_dup(pts, keys, n, j, &capped);
  uint32_t tot;
  uint32_t r = blk[blockIdx.x] + block_rank(keep, &tot);
  if (keep) {
    uvert[r] = j;
    utet[r] = 0u;  // every point starts in the super tet (slot 0)
  }
}

void launch_init_uninserted(const float4* pts, const uint64_t* keys, uint32_t n, uint32_t* blk,
                            uint32_t* uvert, uint32_t* utet, uint32_t* ctr, hipStream_t s) {
  uint32_t nb = nblocks(n);
  k_uninit_count<<<nb, kBlock, 0, s>>>(pts, keys, n, blk, ctr);
  k_scan_blocks<<<1, 1024, 0, s>>>(blk, nb, ctr, 2);
  k_uninit_scatter<<<nb, kBlock, 0, s>>>(pts, keys, n, blk, uvert, utet);
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
__global__ void k_vote(Mesh M, const uint32_t* uvert, const uint32_t* utet, uint32_t U) {
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

__global__ void k_pick_count(Mesh M, const uint32_t* utet, uint32_t U, uint32_t* blk, int ring) {
  uint32_t i = gid();
  bool w = i < U && is_winner(M, utet, i, ring);
  uint32_t tot;
  block_rank(w, &tot);
  if (threadIdx.x == 0) blk[blockIdx.x] = tot;
}

__global__ void k_pick_scatter(Mesh M, const uint32_t* uvert, const uint32_t* utet, uint32_t U,
                               uint32_t round, uint2* wlist, uint32_t* uvert2, uint32_t* utet2,
                               const uint32_t* blk, int ring) {
  uint32_t i = gid();
  bool w = i < U && is_winner(M, utet, i, ring);
  uint32_t tot;
  uint32_t r = blk[blockIdx.x] + block_rank(w, &tot);
  if (i >= U) return;
  uint32_t t = utet[i], v = uvert[i];
  if (w) {
    wlist[r] = make_uint2(t, v);
    M.tsplit[t] = make_uint2(round, M.ctr[C_BASE] + 3u * r);
  } else {
    uvert2[i - r] = v;
    utet2[i - r] = t;
  }
}

void launch_vote(const Mesh& M, const uint32_t* uvert, const uint32_t* utet, uint32_t U,
                 hipStream_t s) {
  k_vote<<<nblocks(U), kBlock, 0, s>>>(M, uvert, utet, U);
}

// votes of tets whose winner was deferred must not survive into the next
// round (list positions change); split resets the votes of winning tets
__global__ void k_clear_votes(Mesh M, const uint32_t* utet2, const uint32_t* ctr) {
  uint32_t i = gid();
  if (i < ctr[C_SURV]) M.tvote[utet2[i]] = kVoteNone;
}

void launch_pick(const Mesh& M, const uint32_t* uvert, const uint32_t* utet, uint32_t U,
                 uint32_t round, uint2* wlist, uint32_t* uvert2, uint32_t* utet2, uint32_t* blk,
                 int ring, hipStream_t s) {
  uint32_t nb = nblocks(U);
  k_pick_count<<<nb, kBlock, 0, s>>>(M, utet, U, blk, ring);
  k_scan_blocks<<<1, 1024, 0, s>>>(blk, nb, M.ctr, 0);
  k_pick_scatter<<<nb, kBlock, 0, s>>>(M, uvert, utet, U, round, wlist, uvert2, utet2, blk, ring);
  if (ring > 0) k_clear_votes<<<nb, kBlock, 0, s>>>(M, utet2, M.ctr);
}

__global__ void k_split(Mesh M, const uint2* wlist, uint32_t W, uint32_t round,
                        uint32_t* act_next) {
  uint32_t k = gid();
  if (k < W) split_body(M, wlist, k, round, act_next);
}

void launch_split(const Mesh& M, const uint2* wlist, uint32_t W, uint32_t round,
                  uint32_t* act_next, hipStream_t s) {
  if (W) k_split<<<nblocks(W), kBlock, 0, s>>>(M, wlist, W, round, act_next);
}

__global__ void k_split_points(Mesh M, const uint32_t* uvert, uint32_t* utet, uint32_t U,
                               uint32_t round, uint64_t* st) {
  uint32_t i = gid();
  PredStats ls = {};
  if (i < U) split_points_body(M, uvert, utet, i, round, &ls);
  flush_stats(ls, st);
}

void launch_split_points(const Mesh& M, const uint32_t* uvert, uint32_t* utet, uint32_t U,
                         uint32_t round, uint64_t* st, hipStream_t s) {
  if (U) k_split_points<<<nblocks(U), kBlock, 0, s>>>(M, uvert, utet, U, round, st);
}

// ---------------------------------------------------------------------------
// flips
// ---------------------------------------------------------------------------
__global__ void k_flip_check(Mesh M, const uint32_t* act, uint4* prop, uint32_t A, uint32_t it,
                             uint64_t* st) {
  uint32_t j = gid();
  PredStats ls = {};
  if (j < A) flip_check_body(M, act, prop, j, it, &ls);
  flush_stats(ls, st);
}

__global__ void k_flip_decide_exec(Mesh M, const uint4* prop, uint32_t A, uint32_t it,
                                   uint32_t round, uint4* out, uint32_t* act_next) {
  uint32_t j = gid();
  if (j < A) flip_decide_exec_body(M, prop, j, it, round, out, act_next);
}

// F is an upper bound (the active-list length); the flip count itself is
// read from the device counter, so the host does not wait for it.
__global__ void k_flip_update(Mesh M, const uint4* out, uint32_t F, uint32_t it) {
  uint32_t k = gid();
  if (k < F && k < M.ctr[C_FLIPS]) flip_update_body(M, out, k, it);
}

// ctr[idx] = v, as a kernel so that a captured graph never records a memset
__global__ void k_set_ctr(uint32_t* ctr, int idx, uint32_t v, int idx2, uint32_t v2) {
  if (threadIdx.x == 0) {
    ctr[idx] = v;
    if (idx2 >= 0) ctr[idx2] = v2;
  }
}

void launch_set_ctr(uint32_t* ctr, int idx, uint32_t v, int idx2, uint32_t v2, hipStream_t s) {
  k_set_ctr<<<1, 64, 0, s>>>(ctr, idx, v, idx2, v2);
}

void launch_flip_check(const Mesh& M, const uint32_t* act, uint4* prop, uint32_t A, uint32_t it,
                       uint64_t* st, hipStream_t s) {
  if (A) k_flip_check<<<nblocks(A), kBlock, 0, s>>>(M, act, prop, A, it, st);
}
void launch_flip_decide_exec(const Mesh& M, const uint4* prop, uint32_t A, uint32_t it,
                             uint32_t round, uint4* out, uint32_t* act_next, hipStream_t s) {
  if (A) k_flip_decide_exec<<<nblocks(A), kBlock, 0, s>>>(M, prop, A, it, round, out, act_next);
}
void launch_flip_update(const Mesh& M, const uint4* out, uint32_t F, uint32_t it,
                        hipStream_t s) {
  if (F) k_flip_update<<<nblocks(F), kBlock, 0, s>>>(M, out, F, it);
}

__global__ void k_relocate(Mesh M, const uint32_t* uvert, uint32_t* utet, uint32_t U,
                           uint32_t round, uint64_t* st) {
  uint32_t i = gid();
  PredStats ls = {};
  if (i < U) relocate_body(M, uvert, utet, i, round, &ls);
  flush_stats(ls, st);
}

void launch_relocate(const Mesh& M, const uint32_t* uvert, uint32_t* utet, uint32_t U,
                     uint32_t round, uint64_t* st, hipStream_t s) {
  if (U) k_relocate<<<nblocks(U), kBlock, 0, s>>>(M, uvert, utet, U, round, st);
}

__global__ void k_audit(Mesh M, uint32_t T, uint32_t* rec, uint32_t cap, uint64_t* st) {
  uint32_t t = gid();
  PredStats ls = {};
  if (t < T) audit_body(M, t, rec, cap, &ls);
  flush_stats(ls, st);
}

void launch_audit(const Mesh& M, uint32_t T, uint32_t* rec, uint32_t cap, uint64_t* st,
                  hipStream_t s) {
  if (T) k_audit<<<nblocks(T), kBlock, 0, s>>>(M, T, rec, cap, st);
}

// ---------------------------------------------------------------------------
// output
// ---------------------------------------------------------------------------
__global__ void k_alive_count(Mesh M, uint32_t T, uint32_t* blk) {
  uint32_t t = gid();
  bool a = t < T && (M.tinfo[t] & kAlive);
  uint32_t tot;
  block_rank(a, &tot);
  if (threadIdx.x == 0) blk[blockIdx.x] = tot;
}

__global__ void k_alive_scatter(Mesh M, uint32_t T, const uint32_t* blk, uint32_t* slot2idx) {
  uint32_t t = gid();
  bool a = t < T && (M.tinfo[t] & kAlive);
  uint32_t tot;
  uint32_t r = blk[blockIdx.x] + block_rank(a, &tot);
  if (t < T) slot2idx[t] = a ? r : kInvalid;
}

__global__ void k_output_write(Mesh M, uint32_t T, uint32_t n, const uint32_t* slot2idx,
                               uint32_t* out_tets, uint32_t* out_adj) {
  uint32_t t = gid();
  if (t >= T) return;
  uint32_t r = slot2idx[t];
  if (r == kInvalid) return;
  uint4 tv = M.tet[t], to = M.opp[t];
  uint4 ov, oa;
  for (int k = 0; k < 4; ++k) {
    uint32_t v = get4(tv, k);
    set4(ov, k, v < n ? pt_rank(M.pts[v]) : v);
    uint32_t e = get4(to, k);
    uint32_t nr = slot2idx[opp_tet(e)];
    set4(oa, k, nr == kInvalid ? kInvalid : (nr << 5) | opp_face(e));
  }
  reinterpret_cast<uint4*>(out_tets)[r] = ov;
  reinterpret_cast<uint4*>(out_adj)[r] = oa;
}

void launch_output(const Mesh& M, uint32_t T, uint32_t n, uint32_t* slot2idx, uint32_t* blk,
                   uint32_t* out_tets, uint32_t* out_adj, hipStream_t s) {
  uint32_t nb = nblocks(T);
  k_alive_count<<<nb, kBlock, 0, s>>>(M, T, blk);
  k_scan_blocks<<<1, 1024, 0, s>>>(blk, nb, M.ctr, 1);
  k_alive_scatter<<<nb, kBlock, 0, s>>>(M, T, blk, slot2idx);
  if (out_tets) k_output_write<<<nb, kBlock, 0, s>>>(M, T, n, slot2idx, out_tets, out_adj);
}

__global__ void k_map_records(uint32_t* rec, uint32_t n, const uint32_t* slot2idx) {
  uint32_t i = gid();
  if (i < n) rec[i] = (slot2idx[rec[i] >> 2] << 2) | (rec[i] & 3u);
}

void launch_map_records(uint32_t* rec, uint32_t n, const uint32_t* slot2idx, hipStream_t s) {
  if (n) k_map_records<<<nblocks(n), kBlock, 0, s>>>(rec, n, slot2idx);
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

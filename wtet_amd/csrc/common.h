// Shared layout, encodings and small helpers for the triangulation kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "predicates.h"

namespace wtet {

constexpr uint32_t kInvalid = 0xFFFFFFFFu;
constexpr uint64_t kVoteNone = 0xFFFFFFFFFFFFFFFFull;

// tinfo bits
constexpr uint32_t kAlive = 1u;   // slot holds a live tet
constexpr uint32_t kQueued = 2u;  // tet id is in the next active list
constexpr uint32_t kDead = 4u;    // slot freed by a 3-2 flip; tfwd holds a live neighbour

// Face f of a tet is the face opposite vertex f.  kFace[f] lists its three
// vertex positions ordered so that (face..., v_f) is an even permutation of
// (v0, v1, v2, v3): for a tet with raw orient3d < 0, orient3d(face, v_f) < 0.
WT_HD int face_vi(int f, int k) {
  // {1,3,2}, {0,2,3}, {0,3,1}, {0,1,2}
  const uint32_t packed = (1u | 3u << 2 | 2u << 4) | (0u | 2u << 2 | 3u << 4) << 6 |
                          (0u | 3u << 2 | 1u << 4) << 12 | (0u | 1u << 2 | 2u << 4) << 18;
  return (int)((packed >> (6 * f + 2 * k)) & 3u);
}

// adjacency: (tet << 2) | face, kInvalid when none
WT_HD uint32_t enc_opp(uint32_t t, uint32_t f) { return (t << 2) | f; }
WT_HD uint32_t opp_tet(uint32_t e) { return e >> 2; }
WT_HD uint32_t opp_face(uint32_t e) { return e & 3u; }

WT_HD uint32_t get4(const uint4& v, int i) {
  return i == 0 ? v.x : (i == 1 ? v.y : (i == 2 ? v.z : v.w));
}
WT_HD void set4(uint4& v, int i, uint32_t x) {
  if (i == 0) v.x = x;
  else if (i == 1) v.y = x;
  else if (i == 2) v.z = x;
  else v.w = x;
}
WT_HD int find4(const uint4& v, uint32_t x) {
  return v.x == x ? 0 : (v.y == x ? 1 : (v.z == x ? 2 : 3));
}

// Points: xyz in f32, w holds the SoS rank (original input id) as bits.
WT_HD uint32_t pt_rank(const float4& p) { return f32_bits(p.w); }

// ---------------------------------------------------------------------------
// atomics (device instructions; plain host builtins for host-side callers)
// ---------------------------------------------------------------------------
WT_HD uint64_t at_min_u64(uint64_t* p, uint64_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return atomicMin((unsigned long long*)p, (unsigned long long)v);
#else
  return __atomic_fetch_min(p, v, __ATOMIC_RELAXED);
#endif
}
WT_HD uint32_t at_add_u32(uint32_t* p, uint32_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return atomicAdd(p, v);
#else
  return __atomic_fetch_add(p, v, __ATOMIC_RELAXED);
#endif
}
WT_HD uint32_t at_or_u32(uint32_t* p, uint32_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return atomicOr(p, v);
#else
  return __atomic_fetch_or(p, v, __ATOMIC_RELAXED);
#endif
}
WT_HD uint32_t at_and_u32(uint32_t* p, uint32_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
  return atomicAnd(p, v);
#else
  return __atomic_fetch_and(p, v, __ATOMIC_RELAXED);
#endif
}

// Append to a list with one atomic per wavefront: the active lanes that call
// this together get consecutive slots.
WT_HD uint32_t append_slot(uint32_t* counter) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint64_t mask = __ballot(1);
  int lane = __lane_id();
  int leader = __ffsll((unsigned long long)mask) - 1;
  uint32_t rank = __popcll(mask & ((1ull << lane) - 1ull));
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(counter, (uint32_t)__popcll(mask));
  base = __shfl(base, leader);
  return base + rank;
#else
  return __atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED);
#endif
}

// 32-bit mixing hash (lowbias32) used to vary walk face order
WT_HD uint32_t hash32(uint32_t x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

}  // namespace wtet

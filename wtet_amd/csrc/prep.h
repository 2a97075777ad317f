// Point preparation, start tets and output compaction for one or many point
// clouds (host + device bodies; the kernels in kernels.hip wrap them).
//
// Layout.  Cloud c owns the input points [off[c], off[c+1]).  Each cloud is
// normalized on its own (src/lib.rs:313-337, per call in the reference) and
// sorted by (cloud, Morton key).  After the sort cloud c's points are the
// vertices [off[c] + 4c, off[c+1] + 4c), followed by its four super vertices
// (src/gpu/mod.rs:152-162); one infinity vertex V = off[C] + 4C closes every
// cloud's mesh.  Tet slots 5c..5c+4 hold cloud c's five start tets
// (src/shaders/init.wgsl:87-148).  A point's SoS rank (float4.w bits) is its
// id within its cloud, and a super vertex's rank is n_c + k, so the rank is
// also the vertex id the result reports.  One cloud (C = 1) is the plain
// layout: vertices 0..n-1, super n..n+3, infinity n+4.
#pragma once

#include "ops.h"

namespace wtet {

struct Clouds {
  const uint32_t* off;  // C + 1 input offsets (device or host memory)
  uint32_t C;
  uint32_t V;           // infinity vertex id
  int kbits;            // Morton bits per axis
  int cshift;           // key bit where the cloud index starts
  int hilbert;          // 1: keys lead with the reference's 8-bit Hilbert index
};

// largest c < C with off[c] <= i (input index)
WT_HD uint32_t cloud_of_point(const uint32_t* off, uint32_t C, uint32_t i) {
  uint32_t lo = 0, hi = C;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (off[mid] <= i) lo = mid;
    else hi = mid;
  }
  return lo;
}

// largest c < C with off[c] + 4c <= v (vertex id, v < V)
WT_HD uint32_t cloud_of_vertex(const uint32_t* off, uint32_t C, uint32_t v) {
  uint32_t lo = 0, hi = C;
  while (hi - lo > 1) {
    uint32_t mid = (lo + hi) >> 1;
    if (off[mid] + 4u * mid <= v) lo = mid;
    else hi = mid;
  }
  return lo;
}

WT_HD uint32_t super_base(const uint32_t* off, uint32_t c) { return off[c + 1] + 4u * c; }

// order-preserving map of f32 to u32 (for atomic min/max)
WT_HD uint32_t ord_f32(float f) {
  uint32_t u = f32_bits(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
WT_HD float unord_f32(uint32_t o) {
  uint32_t u = (o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o;
  float f;
  __builtin_memcpy(&f, &u, 4);
  return f;
}

// spread the low 21 bits of v to every third bit of a 64-bit word
WT_HD uint64_t expand21(uint64_t v) {
  v &= 0x1FFFFFull;
  v = (v | (v << 32)) & 0x001F00000000FFFFull;
  v = (v | (v << 16)) & 0x001F0000FF0000FFull;
  v = (v | (v << 8)) & 0x100F00F00F00F00Full;
  v = (v | (v << 4)) & 0x10C30C30C30C30C3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}

// 3D Hilbert index of an 8-bit-per-axis cell (J. Skilling, "Programming the
// Hilbert curve", AIP Conf. Proc. 707, 2004: axes to transposed index, then
// the transposed bits interleaved, x first).  Consecutive indices are
// face-neighbour cells.
WT_HD uint32_t hilbert3_8(uint32_t x, uint32_t y, uint32_t z) {
  uint32_t X[3] = {x & 255u, y & 255u, z & 255u};
  // undo excess work
  for (uint32_t Q = 128u; Q > 1u; Q >>= 1) {
    uint32_t P = Q - 1u;
    for (int i = 0; i < 3; ++i) {
      if (X[i] & Q) {
        X[0] ^= P;
      } else {
        uint32_t t = (X[0] ^ X[i]) & P;
        X[0] ^= t;
        X[i] ^= t;
      }
    }
  }
  // Gray encode
  X[1] ^= X[0];
  X[2] ^= X[1];
  uint32_t t = 0u;
  for (uint32_t Q = 128u; Q > 1u; Q >>= 1)
    if (X[2] & Q) t ^= Q - 1u;
  X[0] ^= t;
  X[1] ^= t;
  X[2] ^= t;
  uint32_t h = 0u;
  for (int b = 7; b >= 0; --b)
    for (int i = 0; i < 3; ++i) h = (h << 1) | ((X[i] >> b) & 1u);
  return h;
}

// The reference's first insertion stride for n points (power of two in
// [64, 16384] from sqrt(n), doubled above 1e5 points;
// src/gpu/mod.rs:533-548), from integer arithmetic: the smallest power of
// two >= the clamped scale.
WT_HD uint32_t initial_stride(uint32_t n) {
  // scale = sqrt(n) (n <= 1e5) or 2 sqrt(n); stride = 2^ceil(log2(clamp(scale, 64, 16384)))
  // 2^k >= c*sqrt(n)  <=>  4^k >= c^2 n
  uint64_t c2n = (n > 100000u ? 4ull : 1ull) * (uint64_t)n;
  uint32_t s = 64u;
  while (s < 16384u && (uint64_t)s * s < c2n) s <<= 1;
  return s;
}

// Position of sorted element i among n in the reference's stratified order
// with first stride S (src/gpu/mod.rs:573-600): passes of strides S, S/2,
// .., 1 take the not yet taken multiples of the stride, in index order.
WT_HD uint32_t stratified_rank(uint32_t i, uint32_t n, uint32_t S) {
  if (i % S == 0u) return i / S;
  uint32_t before = (n + S - 1u) / S;  // pass 0
  for (uint32_t s = S >> 1; s >= 1u; s >>= 1) {
    // pass of stride s: multiples of s that are not multiples of 2s
    if (i % s == 0u) return before + (i / s - 1u) / 2u;
    before += (n + s - 1u) / s - (n + 2u * s - 1u) / (2u * s);
  }
  return before;  // unreachable: every i is a multiple of 1
}

// Normalize input point i of its cloud (f32 sub then div, scale floor
// 1e-10: src/lib.rs:313-337) and build its sort key (cloud, Morton code).
// mm holds per cloud the ordered-u32 min (3) and max (3) of each axis.
#pragma clang fp contract(off)
WT_HD void normalize_body(const float* xyz, uint32_t i, const Clouds& K, const uint32_t* mm,
                          float* xyzn, uint64_t* keys, uint32_t* ids) {
  uint32_t c = K.C > 1 ? cloud_of_point(K.off, K.C, i) : 0u;
  const uint32_t* m = mm + 6 * (size_t)c;
  float mn[3], ext[3];
  for (int k = 0; k < 3; ++k) {
    mn[k] = unord_f32(m[k]);
    ext[k] = unord_f32(m[3 + k]) - mn[k];
  }
  float s = ext[0];
  if (ext[1] > s) s = ext[1];
  if (ext[2] > s) s = ext[2];
  if (!(s >= 1e-10f)) s = 1e-10f;
  const float qmax = (float)((1u << K.kbits) - 1u);
  uint32_t q[3];
  for (int k = 0; k < 3; ++k) {
    float v = (xyz[3 * (size_t)i + k] - mn[k]) / s;
    xyzn[3 * (size_t)i + k] = v;
    float cl = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
    q[k] = (uint32_t)(cl * qmax);
  }
  uint64_t key;
  if (K.hilbert) {
    // the reference's Hilbert keys: each axis over its own extent, rounded
    // to 0..255 (src/hilbert.rs:20-38)
    uint32_t h[3];
    for (int k = 0; k < 3; ++k) {
      float u = ext[k] > 0.0f ? (xyz[3 * (size_t)i + k] - mn[k]) / ext[k] : 0.0f;
      u = u < 0.0f ? 0.0f : (u > 1.0f ? 1.0f : u);
      h[k] = (uint32_t)rintf(u * 255.0f);
    }
    // a fine Morton code below the 24-bit index keeps runs of equal keys
    // down to (nearly) equal coordinates, as the duplicate scan needs
    key = ((uint64_t)hilbert3_8(h[0], h[1], h[2]) << (3 * K.kbits)) |
          (expand21(q[2]) << 2) | (expand21(q[1]) << 1) | expand21(q[0]);
  } else {
    // the Morton order is for memory locality only (the insertion order
    // cannot change the triangulation)
    key = (expand21(q[2]) << 2) | (expand21(q[1]) << 1) | expand21(q[0]);
  }
  if (K.C > 1) key |= (uint64_t)c << K.cshift;
  keys[i] = key;
  ids[i] = i;
}
#pragma clang fp contract(on)

WT_HD uint32_t cloud_of_key(const Clouds& K, uint64_t key) {
  return K.C > 1 ? (uint32_t)(key >> K.cshift) : 0u;
}

// sorted position j -> vertex j + 4c with its normalized coordinates and rank
WT_HD void gather_body(const float* xyzn, const uint32_t* ids, const uint64_t* keys, uint32_t j,
                       const Clouds& K, float4* pts) {
  uint32_t o = ids[j];
  uint32_t c = cloud_of_key(K, keys[j]);
  uint32_t rank = o - K.off[c];
  float w;
  __builtin_memcpy(&w, &rank, 4);
  pts[j + 4u * c] =
      make_float4(xyzn[3 * (size_t)o], xyzn[3 * (size_t)o + 1], xyzn[3 * (size_t)o + 2], w);
}

// g in [0, 4C]: super vertex g % 4 of cloud g / 4, and (g == 4C) infinity
WT_HD void tail_body(uint32_t g, const Clouds& K, float4* pts) {
  const float sx[5] = {-100.f, 400.f, -100.f, -100.f, 1000.f};
  const float sy[5] = {-100.f, -100.f, 400.f, -100.f, 1000.f};
  const float sz[5] = {-100.f, -100.f, -100.f, 400.f, 1000.f};
  uint32_t slot, rank;
  int k;
  if (g == 4u * K.C) {
    k = 4;
    slot = K.V;
    rank = K.V;
  } else {
    uint32_t c = g >> 2;
    k = (int)(g & 3u);
    slot = super_base(K.off, c) + (uint32_t)k;
    rank = K.off[c + 1] - K.off[c] + (uint32_t)k;
  }
  float w;
  __builtin_memcpy(&w, &rank, 4);
  pts[slot] = make_float4(sx[k], sy[k], sz[k], w);
}

// The five start tets of one cloud with vertices given as codes (0..3:
// super vertex k, 4: infinity) and adjacency as (start tet << 2) | face:
// the super tet (raw orient3d < 0) and one tet joining each of its faces,
// reversed, to the infinity vertex, so every face has a neighbour.
struct StartTets {
  uint32_t tv[20];
  uint32_t to[20];
};

inline StartTets make_start_tets() {
  StartTets S;
  for (int k = 0; k < 4; ++k) S.tv[k] = (uint32_t)k;
  for (int f = 0; f < 4; ++f) {
    uint32_t x = (uint32_t)face_vi(f, 0), y = (uint32_t)face_vi(f, 1), z = (uint32_t)face_vi(f, 2);
    uint32_t* t = &S.tv[4 * (1 + f)];
    t[0] = y;
    t[1] = x;
    t[2] = z;
    t[3] = 4u;
  }
  // adjacency by face matching over the five tets
  for (int t = 0; t < 5; ++t)
    for (int f = 0; f < 4; ++f) {
      uint32_t a[3];
      int j = 0;
      for (int v = 0; v < 4; ++v)
        if (v != f) a[j++] = S.tv[4 * t + v];
      S.to[4 * t + f] = kInvalid;
      for (int u = 0; u < 5 && S.to[4 * t + f] == kInvalid; ++u) {
        if (u == t) continue;
        for (int g = 0; g < 4; ++g) {
          uint32_t b[3];
          int i = 0;
          for (int v = 0; v < 4; ++v)
            if (v != g) b[i++] = S.tv[4 * u + v];
          bool same = true;
          for (int p = 0; p < 3; ++p) {
            bool in = false;
            for (int q = 0; q < 3; ++q) in = in || a[p] == b[q];
            same = same && in;
          }
          if (same) {
            S.to[4 * t + f] = enc_opp((uint32_t)u, (uint32_t)g);
            break;
          }
        }
      }
    }
  return S;
}

// g in [0, 5C): start tet g % 5 of cloud g / 5, in slot g
WT_HD void start_tet_body(const Mesh& M, const StartTets& S, uint32_t g, const Clouds& K) {
  uint32_t c = g / 5u, t = g - 5u * c;
  uint32_t sb = super_base(K.off, c);
  uint4 tv, to, tov;
  for (int k = 0; k < 4; ++k) {
    uint32_t code = S.tv[4 * t + k];
    set4(tv, k, code == 4u ? K.V : sb + code);
    uint32_t e = S.to[4 * t + k];
    set4(to, k, enc_opp(5u * c + opp_tet(e), opp_face(e)));
    uint32_t ocode = S.tv[4 * opp_tet(e) + opp_face(e)];  // the neighbour's vertex opposite
    set4(tov, k, ocode == 4u ? K.V : sb + ocode);
  }
  TV(M, g) = tv;
  set_adj(M, g, to, tov);
  INFO(M, g) = kAlive;
  M.tvote[g] = kVoteNone;
  M.fvote[g] = kVoteNone;
  M.tsplit[g] = make_uint2(0u, 0u);
  FSTAMP(M, g) = 0ull;
}

// Exact duplicates (equal normalized coordinates in one cloud) share a key,
// so they sit in one run of equal keys after the sort.  All but the highest
// input id are redundant under the lifted SoS and are not inserted.
// cpts = pts + 4c (the run lies in one cloud); *capped is set when a run is
// longer than the scan cap (those points are then inserted as usual).
WT_HD bool redundant_dup(const float4* cpts, const uint64_t* keys, uint32_t n, uint32_t j,
                         uint32_t* capped) {
  float4 p = cpts[j];
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
      float4 q = cpts[k];
      if (q.x == p.x && q.y == p.y && q.z == p.z && pt_rank(q) > id) return true;
    }
  }
  return false;
}

// output: cloud of a live tet = cloud of its smallest vertex (never V)
WT_HD uint32_t cloud_of_tet(const Mesh& M, uint32_t t, const Clouds& K) {
  if (K.C <= 1) return 0u;
  uint4 v = TV(M, t);
  uint32_t m = v.x < v.y ? v.x : v.y;
  m = m < v.z ? m : v.z;
  m = m < v.w ? m : v.w;
  return cloud_of_vertex(K.off, K.C, m);
}

// Output row j (cloud order): slot = olist[j] of cloud c = ocloud[j], rows of
// cloud c start at cstart[c].  Vertex ids are cloud-local (rank; infinity
// n_c + 4), adjacency (local row << 5) | face.  slot2pos: slot -> row.
// vrank[v]: vertex v's rank (its input id within its cloud), one word per
// vertex so that the four lookups of a row stay in a small array
WT_HD void output_body(const Mesh& M, uint32_t j, const uint32_t* olist, const uint32_t* ocloud,
                       const uint32_t* cstart, const uint32_t* slot2pos, const uint32_t* vrank,
                       const Clouds& K, uint32_t* out_tets, uint32_t* out_adj) {
  uint32_t t = olist[j];
  uint32_t c = K.C > 1 ? ocloud[j] : 0u;
  uint32_t base = cstart[c];
  uint32_t inf_id = K.off[c + 1] - K.off[c] + 4u;
  uint4 tv = TV(M, t), to = TN(M, t);
  uint4 ov, oa;
  for (int k = 0; k < 4; ++k) {
    uint32_t v = get4(tv, k);
    set4(ov, k, v == K.V ? inf_id : vrank[v]);
    uint32_t e = get4(to, k);
    uint32_t nr = slot2pos[opp_tet(e)];
    set4(oa, k, nr == kInvalid ? kInvalid : ((nr - base) << 5) | opp_face(e));
  }
  reinterpret_cast<uint4*>(out_tets)[j] = ov;
  reinterpret_cast<uint4*>(out_adj)[j] = oa;
}

}  // namespace wtet

// Per-element bodies of the insertion/flip round kernels.
//
// One call handles one element (an uninserted point, a winning tet, an
// active tet, an executed flip).  The __global__ wrappers live in
// kernels.hip.  The algorithm follows the reference's phase-1 loop
// (src/phase1.rs:16-1431): vote -> 1-to-4 split -> point split -> flip
// iterations (check, vote, flip, adjacency update) -> point relocation,
// re-designed around device-resident lists instead of full-array scans.
#pragma once

#include "common.h"

namespace wtet {

// device-side counters (Mesh::ctr)
enum Ctr : int {
  // read by every thread of the flip kernels, written between kernels: the
  // first line.  Counters that take atomics at volume sit 1 KB apart, away
  // from it and from each other.
  C_A = 0,         // length of the current active list
  C_IT = 1,        // flip iteration tag (fvote keys, flip stamps)
  C_PAR = 2,       // which of the two active lists is current
  C_RUN = 3,       // 1 while this iteration runs
  C_STOP = 4,      // set when an iteration could overflow the tet slots
  C_ROUND = 5,     // insertion round (flip stamps)
  C_CAP = 6,       // tet slot capacity
  C_PEAKA = 7,     // longest active list seen
  C_TETS_IT = 8,   // C_TETS when the current flip iteration began
  C_SLOW_SUM = 9,  // tets deferred by the lean flip check, summed over iterations
  C_WALK_FAIL = 10, // relocation walks that hit the step limit
  C_WIN = 11,      // insertions this round
  C_BASE = 12,     // first slot reserved for this round's splits
  C_ALIVE = 13,    // alive tets (output compaction)
  C_BADTET = 14,   // points found in a tet touching infinity (must stay 0)
  C_NPTS = 15,     // points to insert (input minus exact duplicates)
  C_DUPSCAN = 16,  // duplicate scans that hit their length cap
  C_ULEN = 17,     // uninserted list length at the start of the round
  C_SURV = 18,     // points left after this round's insertions
  C_TETS = 256,      // next free tet slot
  C_ACT_NEXT = 512,  // length of the next active list
  C_FLIPS = 768,     // winning flips this iteration
  C_SLOW_T = 1024,   // tets deferred by the lean flip check this iteration
  C_SLOW_P = 1280,   // points deferred by the lean point kernels
  C_VIOL = 1536,     // non-Delaunay faces found by the audit
  C_FAILREC = 1537,  // recorded violation entries
  C_COUNT = 1792
};

// per-call sums over flip iterations (u64, device)
enum IterSum : int {
  S_CHECKED = 0,        // tets checked
  S_FLIPS = 1,          // flips executed
  S_ITERS = 2,          // iterations
  S_FLIPS23 = 3,        // 2-3 flips (the rest are 3-2)
  S_SMALL_CHECKED = 4,  // of S_CHECKED: in k_flip_small
  S_SMALL_FLIPS = 5,    // of S_FLIPS: in k_flip_small
  S_COUNT = 8
};

// Adjacency of a tet slot in one 32-byte record: per face f, w[2f] = the
// neighbour across it ((tet << 2) | face) and w[2f + 1] = that neighbour's
// vertex opposite the face.  The flip check reads its own slot only, and
// the adjacency update rewrites a face with one 8-byte store.
struct __attribute__((aligned(32))) AdjRec {
  uint4 lo;  // faces 0, 1
  uint4 hi;  // faces 2, 3
};

// Where the faces of a tet destroyed by a flip went: for each external face
// f, w[2f] = (new tet << 2) | face and w[2f + 1] = the new tet's vertex
// opposite it (one 8-byte load per face in flip_update_body).
struct __attribute__((aligned(64))) FlipMsg {
  uint4 lo;        // faces 0, 1
  uint4 hi;        // faces 2, 3
  uint64_t stamp;  // the last flip that wrote the slot: round << 32 | flip iteration
  uint64_t pad[3];
};

// message of a destroyed tet: new neighbour pointers e (kInvalid where
// unused), all opposite the vertex v
WT_HD void set_msg(FlipMsg* m, uint4 e, uint32_t v) {
  m->lo = make_uint4(e.x, v, e.y, v);
  m->hi = make_uint4(e.z, v, e.w, v);
}

static_assert(sizeof(AdjRec) == 32, "adjacency record layout");
static_assert(sizeof(FlipMsg) == 64, "flip message layout: one 64-byte line per slot");

struct Mesh {
  const float4* pts;  // n + 5 points (n real, 4 super, 1 infinity)
  uint4* tet;         // vertex ids
  AdjRec* adj;        // neighbours and the vertices across the faces
  uint32_t* tinfo;    // kAlive | kQueued | kDead
  uint32_t* tfwd;     // live neighbour of a slot freed by a 3-2 flip
  uint64_t* tvote;    // insertion vote (min key)
  uint64_t* fvote;    // flip vote (min key, iteration-tagged)
  uint2* tsplit;      // (round, first new slot) for tets split this round
  uint32_t* tpoint;   // the point inserted into a tet split this round
  FlipMsg* tmsg;      // per slot destroyed by a flip this iteration
  uint32_t* ctr;      // Ctr counters
  uint32_t inf;       // id of the infinity vertex
  int mode;           // PredMode
};

WT_HD uint4& TV(const Mesh& M, uint32_t t) { return M.tet[t]; }
WT_HD uint4 TN(const Mesh& M, uint32_t t) {
  uint4 a = M.adj[t].lo, b = M.adj[t].hi;
  return make_uint4(a.x, a.z, b.x, b.z);
}
WT_HD uint4 TOV(const Mesh& M, uint32_t t) {
  uint4 a = M.adj[t].lo, b = M.adj[t].hi;
  return make_uint4(a.y, a.w, b.y, b.w);
}
// both halves of slot t's record: neighbours n and opposite vertices ov
WT_HD void set_adj(const Mesh& M, uint32_t t, uint4 n, uint4 ov) {
  M.adj[t].lo = make_uint4(n.x, ov.x, n.y, ov.y);
  M.adj[t].hi = make_uint4(n.z, ov.z, n.w, ov.w);
}
WT_HD uint32_t& INFO(const Mesh& M, uint32_t t) { return M.tinfo[t]; }
// the stamp shares its slot's message line: the adjacency update reads a
// flipped neighbour's stamp and message together
WT_HD uint64_t& FSTAMP(const Mesh& M, uint32_t t) { return M.tmsg[t].stamp; }
WT_HD uint64_t flip_stamp(uint32_t round, uint32_t it) { return (uint64_t)round << 32 | it; }
WT_HD uint32_t& TFWD(const Mesh& M, uint32_t t) { return M.tfwd[t]; }

WT_HD float3 P3(const Mesh& M, uint32_t v) {
  float4 p = M.pts[v];
  return make_float3(p.x, p.y, p.z);
}

WT_HD void load_pt(const Mesh& M, uint32_t v, float* out, uint32_t* rank) {
  float4 p = M.pts[v];
  out[0] = p.x;
  out[1] = p.y;
  out[2] = p.z;
  *rank = pt_rank(p);
}

// orient3d_sos on four vertex ids
WT_HD int o3(const Mesh& M, uint32_t a, uint32_t b, uint32_t c, uint32_t d, PredStats* st) {
  float A[3], B[3], C[3], D[3];
  uint32_t ra, rb, rc, rd;
  load_pt(M, a, A, &ra);
  load_pt(M, b, B, &rb);
  load_pt(M, c, C, &rc);
  load_pt(M, d, D, &rd);
  return orient3d_sos(A, B, C, D, ra, rb, rc, rd, M.mode, st);
}

// Predicates of the lean (FAST) and full bodies.  FAST runs the FP32 filter
// alone and returns 2 when it cannot decide; the body then stops with no
// side effect and reports the element as deferred, and a second kernel
// runs the full staged predicates on the deferred elements only.
template <bool FAST>
WT_HD int o3s(const Mesh& M, const float* a, const float* b, const float* c, const float* d,
              uint32_t ra, uint32_t rb, uint32_t rc, uint32_t rd, PredStats* st) {
  if (FAST) return orient3d_sos_fast(a, b, c, d, st);
  return orient3d_sos(a, b, c, d, ra, rb, rc, rd, M.mode, st);
}

template <bool FAST>
WT_HD int o4s(const Mesh& M, const float* a, const float* b, const float* c, const float* d,
              const float* e, uint32_t ra, uint32_t rb, uint32_t rc, uint32_t rd, uint32_t re,
              PredStats* st) {
  if (FAST) return orient4d_fast(a, b, c, d, e, st);
  return orient4d(a, b, c, d, e, ra, rb, rc, rd, re, M.mode, st);
}

// The flip check's insphere.  The lean form adds the FP64 filter, which
// settles nearly every face the FP32 filter leaves on non-degenerate input:
// few tets then wait for the deferred full check.  Measured at 1M uniform
// points: the lean check 6.46 -> 6.79 ms (158 VGPRs, occupancy 5 -> 3), the
// deferred check 2.76 -> 1.92 ms, device phase 27.35 -> 26.70 ms; calling
// the FP64 stage out of line (128 VGPRs) made the lean check 7.40 ms.
template <bool FAST>
WT_HD int o4s_check(const Mesh& M, const float* a, const float* b, const float* c, const float* d,
                    const float* e, uint32_t ra, uint32_t rb, uint32_t rc, uint32_t rd, uint32_t re,
                    PredStats* st) {
  if (FAST) {
    int s = orient4d_fast(a, b, c, d, e, st);
    if (s != 2) return s;
    s = insphere_f64(a, b, c, d, e);
    if (s != 2) stat_inc(st, 5);
    return s;
  }
  return orient4d(a, b, c, d, e, ra, rb, rc, rd, re, M.mode, st);
}

template <bool FAST>
WT_HD int o3v(const Mesh& M, uint32_t a, uint32_t b, uint32_t c, uint32_t d, PredStats* st) {
  float A[3], B[3], C[3], D[3];
  uint32_t ra, rb, rc, rd;
  load_pt(M, a, A, &ra);
  load_pt(M, b, B, &rb);
  load_pt(M, c, C, &rc);
  load_pt(M, d, D, &rd);
  return o3s<FAST>(M, A, B, C, D, ra, rb, rc, rd, st);
}

WT_HD uint32_t* opp_word(const Mesh& M, uint32_t t, uint32_t f) {
  return reinterpret_cast<uint32_t*>(&M.adj[t]) + 2 * f;
}

WT_HD uint32_t* oppv_word(const Mesh& M, uint32_t t, uint32_t f) {
  return reinterpret_cast<uint32_t*>(&M.adj[t]) + 2 * f + 1;
}

// face f of slot t: neighbour pointer and opposite vertex in one store
WT_HD void set_face(const Mesh& M, uint32_t t, uint32_t f, uint32_t n, uint32_t ov) {
  reinterpret_cast<uint2*>(&M.adj[t])[f] = make_uint2(n, ov);
}

// ---------------------------------------------------------------------------
// vote: each uninserted point bids for its tet; the key prefers the point
// deepest inside the circumsphere (equivalently nearest the circumcentre:
// -insphere = 6|vol|(r^2 - d^2) for a raw-negative tet), ties to the
// lowest list position.  src/shaders/vote.wgsl:89-281 (two passes there).
// ---------------------------------------------------------------------------
// returns kVoteNone when the point must not vote
WT_HD uint64_t vote_key(const Mesh& M, const uint32_t* uvert, const uint32_t* utet, uint32_t i) {
  // The key is computed before the infinity test (every id, infinity
  // included, is a loadable point), so that the point's own id and
  // coordinates load beside its tet's vertices: two round trips.
  uint32_t t = utet[i], q = uvert[i];
  uint4 tv = TV(M, t);
  float a[3], b[3], c[3], d[3], p[3];
  uint32_t r;
  load_pt(M, q, p, &r);
  load_pt(M, tv.x, a, &r);
  load_pt(M, tv.y, b, &r);
  load_pt(M, tv.z, c, &r);
  load_pt(M, tv.w, d, &r);
  double det = insphere_eval<double>(a, b, c, d, p).det;
  double val = -det;
  if (!(val > 0.0)) val = 0.0;
  float fv = (float)val;
  uint64_t key = ((uint64_t)(0xFFFFFFFFu - f32_bits(fv)) << 32) | i;
  bool bad = tv.x == M.inf || tv.y == M.inf || tv.z == M.inf || tv.w == M.inf;
  if (bad) at_add_u32(&M.ctr[C_BADTET], 1u);
  return bad ? kVoteNone : key;
}

WT_HD void vote_body(const Mesh& M, const uint32_t* uvert, const uint32_t* utet, uint32_t i) {
  uint64_t key = vote_key(M, uvert, utet, i);
  if (key != kVoteNone) at_min_u64(&M.tvote[utet[i]], key);
}

// Point i is inserted this round when it holds its tet's vote and no tet
// within `ring` face-steps holds a smaller key.  Inserting only into
// non-adjacent tets keeps the stars of simultaneous insertions apart, which
// is what leaves concurrent flipping unable to finish (faces that are
// neither 2-3 nor 3-2 flippable); the deferred points win a later round.
// Reads votes only, so every thread sees the same decision.
WT_HD bool is_winner(const Mesh& M, const uint32_t* utet, uint32_t i, int ring) {
  uint32_t t = utet[i];
  uint64_t kt = M.tvote[t];
  if ((uint32_t)(kt & 0xFFFFFFFFull) != i) return false;
  if (ring <= 0) return true;
  uint4 to = TN(M, t);
  for (int f = 0; f < 4; ++f) {
    uint32_t nt = opp_tet(get4(to, f));
    if (M.tvote[nt] < kt) return false;
    if (ring > 1) {
      uint4 no = TN(M, nt);
      for (int g = 0; g < 4; ++g) {
        uint32_t n2 = opp_tet(get4(no, g));
        if (n2 != t && M.tvote[n2] < kt) return false;
      }
    }
  }
  return true;
}

// ---------------------------------------------------------------------------
// 1-to-4 split of tet t at point p (src/shaders/split.wgsl:104-327).
// Child j replaces vertex j by p, so it keeps t's orientation and holds
// t's face j.  Child 0 reuses slot t, children 1..3 use base..base+2.
// ---------------------------------------------------------------------------
WT_HD uint32_t child_slot(uint32_t t, uint32_t base, uint32_t j) {
  return j == 0 ? t : base + j - 1;
}

// c[0..3] receives the children's slots; returns which of them join the
// active list (bit j for child j; slot t only if no earlier list holds it).
// The neighbours' split records are loaded before any store: this kernel
// writes those of fresh slots only, never a neighbour's.
WT_HD uint32_t split_body(const Mesh& M, const uint2* wlist, uint32_t k, uint32_t round,
                          uint32_t c[4]) {
  uint32_t t = wlist[k].x, p = wlist[k].y;
  uint32_t base = M.tsplit[t].y;
  uint4 tv = TV(M, t);
  uint4 to = TN(M, t);
  uint4 tov = TOV(M, t);
  c[0] = t, c[1] = base, c[2] = base + 1, c[3] = base + 2;
  uint2 ns[4];
  uint32_t np[4];
  for (int j = 0; j < 4; ++j) {
    uint32_t nt = opp_tet(get4(to, j));
    ns[j] = M.tsplit[nt];
    np[j] = M.tpoint[nt];  // read for every neighbour, used where it splits
  }
  for (int j = 0; j < 4; ++j) {
    uint4 nv = tv;
    set4(nv, j, p);
    // child i (i != j) keeps t's vertex j opposite their shared face
    uint4 no, nov;
    for (int i = 0; i < 4; ++i)
      if (i != j) {
        set4(no, i, enc_opp(c[i], (uint32_t)j));
        set4(nov, i, get4(tv, j));
      }
    uint32_t e = get4(to, j);
    uint32_t nt = opp_tet(e), nf = opp_face(e);
    if (ns[j].x == round) {
      // the neighbour splits too: its child nf holds its point opposite
      set4(no, j, enc_opp(child_slot(nt, ns[j].y, nf), nf));
      set4(nov, j, np[j]);
    } else {
      set4(no, j, e);
      set4(nov, j, get4(tov, j));
      set_face(M, nt, nf, enc_opp(c[j], (uint32_t)j), p);
    }
    TV(M, c[j]) = nv;
    set_adj(M, c[j], no, nov);
    M.tvote[c[j]] = kVoteNone;
    if (j > 0) {  // fresh slot: initialise every per-slot field
      INFO(M, c[j]) = kAlive | kQueued;
      M.fvote[c[j]] = kVoteNone;
      M.tsplit[c[j]] = make_uint2(0u, 0u);
      FSTAMP(M, c[j]) = 0ull;
    }
  }
  uint32_t old = at_or_u32(&INFO(M, t), kQueued);
  return (old & kQueued) ? 0xEu : 0xFu;
}

// ---------------------------------------------------------------------------
// move an uninserted point of a split tet into its child
// (src/shaders/split_points.wgsl:222-335): three orient3d_sos tests.
// ---------------------------------------------------------------------------
template <bool FAST>
WT_HD bool split_points_body_t(const Mesh& M, const uint32_t* uvert, uint32_t* utet, uint32_t i,
                               uint32_t round, PredStats* st) {
  uint32_t t = utet[i];
  uint32_t q = uvert[i];
  uint2 sp = M.tsplit[t];
  if (sp.x != round) return false;
  uint32_t base = sp.y;
  uint4 c0 = TV(M, t);      // (p, v1, v2, v3)
  uint32_t v0 = TV(M, base).x;  // child 1 = (v0, p, v2, v3)
  uint32_t p = c0.x, v1 = c0.y, v2 = c0.z, v3 = c0.w;
  // the six points in one round trip
  float Pp[3], P0[3], P1[3], P2[3], P3[3], Q[3];
  uint32_t rp, r0, r1, r2, r3, rq;
  load_pt(M, p, Pp, &rp);
  load_pt(M, v0, P0, &r0);
  load_pt(M, v1, P1, &r1);
  load_pt(M, v2, P2, &r2);
  load_pt(M, v3, P3, &r3);
  load_pt(M, q, Q, &rq);
  // children 0 | 1 across plane (p, v2, v3): +1 -> v0's side -> not child 0
  int sa = o3s<FAST>(M, Pp, P3, P2, Q, rp, r3, r2, rq, st);
  if (sa == 2) return true;
  int a = sa > 0 ? 1 : 0;
  // children 2 | 3 across plane (v0, p, v1): +1 -> v2's side -> not child 2
  int sb = o3s<FAST>(M, P0, Pp, P1, Q, r0, rp, r1, rq, st);
  if (sb == 2) return true;
  int b = sb > 0 ? 3 : 2;
  // decide between a and b on their shared face: face b of (v0, v1, v2, v3)
  // with vertex a replaced by p -- (w0, v3, w1) for b = 2, (w0, w1, v2) for
  // b = 3, where w0 = a ? v0 : p and w1 = a ? p : v1 (chosen by value)
  float X[3], Y[3], Z[3];
  for (int k = 0; k < 3; ++k) {
    float w0 = a ? P0[k] : Pp[k], w1 = a ? Pp[k] : P1[k];
    X[k] = w0;
    Y[k] = b == 2 ? P3[k] : w1;
    Z[k] = b == 2 ? w1 : P2[k];
  }
  uint32_t rw0 = a ? r0 : rp, rw1 = a ? rp : r1;
  int s = o3s<FAST>(M, X, Y, Z, Q, rw0, b == 2 ? r3 : rw1, b == 2 ? rw1 : r2, rq, st);
  if (s == 2) return true;
  int child = s > 0 ? a : b;
  utet[i] = child_slot(t, base, (uint32_t)child);
  return false;
}

WT_HD void split_points_body(const Mesh& M, const uint32_t* uvert, uint32_t* utet, uint32_t i,
                             uint32_t round, PredStats* st) {
  (void)split_points_body_t<false>(M, uvert, utet, i, round, st);
}

// ---------------------------------------------------------------------------
// flip proposals (src/shaders/check_delaunay_{fast,exact}.wgsl, flip votes
// src/shaders/mark_rejected_flips.wgsl:61-178)
// proposal: x = t, y = face | kind << 2 | edge << 3, z = neighbour, w = third tet
// kind 0: 2-3 flip, 1: 3-2 flip around edge (edge = rotation of the face triple)
// ---------------------------------------------------------------------------
WT_HD uint64_t flip_key(uint32_t it, uint32_t t) {
  return ((uint64_t)(0xFFFFFFFFu - it) << 32) | t;
}

template <bool FAST>
WT_HD bool flip_check_body_t(const Mesh& M, const uint32_t* act, uint4* prop, uint32_t j,
                             uint32_t it, PredStats* st) {
  uint32_t t = act[j];
  // Two memory round trips: the slot's info, vertices and opposite vertices
  // together (all in the slot's own records), then all eight points (the
  // point array includes the infinity vertex, so every id is loadable).
  uint32_t info = INFO(M, t);
  uint4 tv = TV(M, t);
  uint4 tov = TOV(M, t);
  // each tet is on the list once and no other thread of this kernel writes
  // its info word: a plain read-modify-write clears the queued bit
  INFO(M, t) = info & ~kQueued;
  prop[j].x = kInvalid;
  if (!(info & kAlive)) return false;
  if (tv.x == M.inf || tv.y == M.inf || tv.z == M.inf || tv.w == M.inf) return false;
  float P[4][3], Q[4][3];
  uint32_t R[4], RQ[4];
  for (int k = 0; k < 4; ++k) load_pt(M, get4(tv, k), P[k], &R[k]);
  for (int k = 0; k < 4; ++k) load_pt(M, get4(tov, k), Q[k], &RQ[k]);
  for (int f = 0; f < 4; ++f) {
    uint32_t o = get4(tov, f);
    if (o == M.inf) continue;
    const float* O = Q[f];
    uint32_t ro = RQ[f];
    int ins = o4s_check<FAST>(M, P[0], P[1], P[2], P[3], O, R[0], R[1], R[2], R[3], ro, st);
    if (ins == 2) return true;
    if (ins > 0) continue;
    // o lies inside the circumsphere: classify the flip
    int px = face_vi(f, 0), py = face_vi(f, 1), pz = face_vi(f, 2);
    int pos[3] = {px, py, pz};
    int bad = -1, nbad = 0;
    for (int k = 0; k < 3; ++k) {
      int u = pos[k], w = pos[(k + 1) % 3];
      // edge (u, w): the line a-o should pass on the face side -> raw > 0
      int s = o3s<FAST>(M, P[u], P[w], P[f], O, R[u], R[w], R[f], ro, st);
      if (s == 2) return true;
      if (s > 0) {
        bad = k;
        ++nbad;
      }
    }
    uint32_t kind, third = kInvalid;
    if (nbad == 0) {
      kind = 0;
    } else if (nbad == 1) {
      // 3-2 needs the edge to have degree 3: the tet across t's face
      // opposite the third vertex must have o as its apex
      int pthird = pos[(bad + 2) % 3];
      if (get4(tov, pthird) != o) continue;
      kind = 1;
      third = opp_tet(*opp_word(M, t, (uint32_t)pthird));
    } else {
      continue;
    }
    uint32_t nt = opp_tet(*opp_word(M, t, (uint32_t)f));
    uint64_t key = flip_key(it, t);
    at_min_u64(&M.fvote[t], key);
    at_min_u64(&M.fvote[nt], key);
    if (kind == 1) at_min_u64(&M.fvote[third], key);
    prop[j] = make_uint4(t, (uint32_t)f | kind << 2 | (uint32_t)(kind ? bad : 0) << 3, nt, third);
    return false;
  }
  return false;
}

// point k (0..3, chosen at run time) of a tet's four: selects instead of an
// indexed local array, which would live in scratch memory
WT_HD void pick_pt(const float (*P)[3], const uint32_t* R, int k, float* out, uint32_t* r) {
  for (int c = 0; c < 3; ++c) out[c] = k == 0 ? P[0][c] : (k == 1 ? P[1][c] : (k == 2 ? P[2][c] : P[3][c]));
  *r = k == 0 ? R[0] : (k == 1 ? R[1] : (k == 2 ? R[2] : R[3]));
}

// One face f (chosen at run time) of flip_check_body_t's loop with the full
// predicates: true when the face proposes a flip, with its kind, the
// offending edge (3-2) and the third tet (3-2) as that loop sets them.  P, R:
// the tet's points and ranks; Q, RQ: the vertices across its faces.
WT_HD bool flip_face_full(const Mesh& M, uint32_t t, const uint4& tov, const float (*P)[3],
                          const uint32_t* R, const float (*Q)[3], const uint32_t* RQ, int f,
                          PredStats* st, uint32_t* kind, uint32_t* bad, uint32_t* third) {
  uint32_t o = get4(tov, f);
  if (o == M.inf) return false;
  float O[3];
  uint32_t ro;
  pick_pt(Q, RQ, f, O, &ro);
  int ins = o4s_check<false>(M, P[0], P[1], P[2], P[3], O, R[0], R[1], R[2], R[3], ro, st);
  if (ins > 0) return false;
  // o lies inside the circumsphere: classify the flip
  float A[3], V[3][3];
  uint32_t ra, RV[3];
  pick_pt(P, R, f, A, &ra);
  int pos[3] = {face_vi(f, 0), face_vi(f, 1), face_vi(f, 2)};
  for (int k = 0; k < 3; ++k) pick_pt(P, R, pos[k], V[k], &RV[k]);
  int b = -1, nbad = 0;
  for (int k = 0; k < 3; ++k) {
    int w = (k + 1) % 3;
    if (o3s<false>(M, V[k], V[w], A, O, RV[k], RV[w], ra, ro, st) > 0) {
      b = k;
      ++nbad;
    }
  }
  if (nbad == 0) {
    *kind = 0u;
    *bad = 0u;
    *third = kInvalid;
    return true;
  }
  if (nbad > 1) return false;
  // 3-2 needs the edge to have degree 3 (see flip_check_body_t)
  int pthird = b == 0 ? pos[2] : (b == 1 ? pos[0] : pos[1]);
  if (get4(tov, pthird) != o) return false;
  *kind = 1u;
  *bad = (uint32_t)b;
  *third = opp_tet(*opp_word(M, t, (uint32_t)pthird));
  return true;
}


// ---------------------------------------------------------------------------
// execute a winning flip (src/shaders/flip.wgsl:531-845).  New tets reuse
// the old slots; a 2-3 flip takes one fresh slot, a 3-2 flip frees one.
// External faces keep the old neighbour pointer; flip_update_body fixes
// them once every flip of the iteration is written.
// s: the fresh slot of a 2-3 flip.  The caller queues the new tets (the
// reused slots were claimed by flip_decide_body).
// outk[0]: x,y,z = slots of the new tets (z unused for 3-2), w = kind;
// outk[1]: the flip's vertices o, a and (3-2) z, for the adjacency update
// ---------------------------------------------------------------------------
// The records a flip reads: vertices, neighbours and opposite vertices of
// its two (2-3) or three (3-2) tets.  Loading them is split from the flip so
// that a caller can issue the loads early and overlap them with other work.
struct FlipTets {
  uint4 tv, to, tov, nv, no, nov, mv, mo, mov;
};

WT_HD void flip_load(const Mesh& M, const uint4& pr, FlipTets& F) {
  // a 2-3 flip has no third tet: its slot is replaced by t's, so every load
  // is of a valid slot and none depends on a branch
  uint32_t m = ((pr.y >> 2) & 1u) ? pr.w : pr.x;
  F.tv = TV(M, pr.x), F.to = TN(M, pr.x), F.tov = TOV(M, pr.x);
  F.nv = TV(M, pr.z), F.no = TN(M, pr.z), F.nov = TOV(M, pr.z);
  F.mv = TV(M, m), F.mo = TN(M, m), F.mov = TOV(M, m);
}

WT_HD void flip_exec_loaded(const Mesh& M, const uint4 pr, const FlipTets& F, uint32_t it,
                            uint32_t round, uint32_t s, uint4* outk) {
  uint32_t t = pr.x, f = pr.y & 3u, kind = (pr.y >> 2) & 1u, edge = (pr.y >> 3) & 3u;
  uint32_t n = pr.z;
  const uint4 tv = F.tv, to = F.to, tov = F.tov;
  const uint4 nv = F.nv, no = F.no, nov = F.nov;
  uint32_t nf = opp_face(get4(to, (int)f));
  uint32_t a = get4(tv, (int)f), o = get4(nv, (int)nf);
  int pos[3] = {face_vi((int)f, 0), face_vi((int)f, 1), face_vi((int)f, 2)};
  if (kind == 0) {
    uint32_t x = get4(tv, pos[0]), y = get4(tv, pos[1]), z = get4(tv, pos[2]);
    int qx = find4(nv, x), qy = find4(nv, y), qz = find4(nv, z);
    uint32_t Txy = t, Tyz = n, Tzx = s;
    TV(M, Txy) = make_uint4(x, y, o, a);
    TV(M, Tyz) = make_uint4(y, z, o, a);
    TV(M, Tzx) = make_uint4(z, x, o, a);
    // external faces keep the old tets' opposite vertices until
    // flip_update_body sees a neighbour that flipped too
    set_adj(M, Txy, make_uint4(enc_opp(Tyz, 1), enc_opp(Tzx, 0), get4(to, pos[2]), get4(no, qz)),
            make_uint4(z, z, get4(tov, pos[2]), get4(nov, qz)));
    set_adj(M, Tyz, make_uint4(enc_opp(Tzx, 1), enc_opp(Txy, 0), get4(to, pos[0]), get4(no, qx)),
            make_uint4(x, x, get4(tov, pos[0]), get4(nov, qx)));
    set_adj(M, Tzx, make_uint4(enc_opp(Txy, 1), enc_opp(Tyz, 0), get4(to, pos[1]), get4(no, qy)),
            make_uint4(y, y, get4(tov, pos[1]), get4(nov, qy)));
    uint4 mt = make_uint4(kInvalid, kInvalid, kInvalid, kInvalid);
    set4(mt, pos[2], enc_opp(Txy, 2));
    set4(mt, pos[0], enc_opp(Tyz, 2));
    set4(mt, pos[1], enc_opp(Tzx, 2));
    uint4 mn = make_uint4(kInvalid, kInvalid, kInvalid, kInvalid);
    set4(mn, qz, enc_opp(Txy, 3));
    set4(mn, qx, enc_opp(Tyz, 3));
    set4(mn, qy, enc_opp(Tzx, 3));
    set_msg(&M.tmsg[t], mt, o);
    set_msg(&M.tmsg[n], mn, a);
    FSTAMP(M, t) = flip_stamp(round, it);
    FSTAMP(M, n) = flip_stamp(round, it);
    M.tvote[s] = kVoteNone;
    M.fvote[s] = kVoteNone;
    M.tsplit[s] = make_uint2(0u, 0u);
    FSTAMP(M, s) = 0ull;
    INFO(M, s) = kAlive | kQueued;  // fresh slot: queued by the caller
    outk[0] = make_uint4(Txy, Tyz, Tzx, 0u);
    outk[1] = make_uint4(o, a, 0u, 0u);
  } else {
    // rotate the face triple so the offending edge is (x, y), third z
    uint32_t x = get4(tv, pos[edge]), y = get4(tv, pos[(edge + 1) % 3]),
             z = get4(tv, pos[(edge + 2) % 3]);
    int px = pos[edge], py = pos[(edge + 1) % 3];
    uint32_t m = pr.w;
    const uint4 mv = F.mv, mo = F.mo, mov = F.mov;
    int qx = find4(nv, x), qy = find4(nv, y);
    int mx = find4(mv, x), my = find4(mv, y);
    uint32_t A = t, B = n;
    TV(M, A) = make_uint4(a, z, o, x);
    TV(M, B) = make_uint4(z, a, o, y);
    set_adj(M, A, make_uint4(get4(no, qy), get4(mo, my), get4(to, py), enc_opp(B, 3)),
            make_uint4(get4(nov, qy), get4(mov, my), get4(tov, py), y));
    set_adj(M, B, make_uint4(get4(mo, mx), get4(no, qx), get4(to, px), enc_opp(A, 3)),
            make_uint4(get4(mov, mx), get4(nov, qx), get4(tov, px), x));
    uint4 mt = make_uint4(kInvalid, kInvalid, kInvalid, kInvalid);
    set4(mt, py, enc_opp(A, 2));
    set4(mt, px, enc_opp(B, 2));
    uint4 mn = make_uint4(kInvalid, kInvalid, kInvalid, kInvalid);
    set4(mn, qy, enc_opp(A, 0));
    set4(mn, qx, enc_opp(B, 1));
    uint4 mm = make_uint4(kInvalid, kInvalid, kInvalid, kInvalid);
    set4(mm, my, enc_opp(A, 1));
    set4(mm, mx, enc_opp(B, 0));
    set_msg(&M.tmsg[t], mt, o);
    set_msg(&M.tmsg[n], mn, a);
    set_msg(&M.tmsg[m], mm, z);
    FSTAMP(M, t) = flip_stamp(round, it);
    FSTAMP(M, n) = flip_stamp(round, it);
    FSTAMP(M, m) = flip_stamp(round, it);
    INFO(M, m) = kDead;
    TFWD(M, m) = A;
    outk[0] = make_uint4(A, B, kInvalid, 1u);
    outk[1] = make_uint4(o, a, z, 0u);
  }
}

WT_HD void flip_exec_body(const Mesh& M, const uint4 pr, uint32_t it, uint32_t round,
                          uint32_t s, uint4* outk) {
  FlipTets F;
  flip_load(M, pr, F);
  flip_exec_loaded(M, pr, F, it, round, s, outk);
}

// A proposal wins when it holds the vote of every tet it involves; the
// winner executes at once (its tets belong to no other winner), losers go
// back to the active list (src/shaders/mark_rejected_flips.wgsl:61-178 and
// src/shaders/flip.wgsl:531-845, one pass here).
// Decide proposal j and claim the queue entries it needs: a winner's two
// reused slots (and, for a 2-3 flip, the fresh slot, counted in *nslot),
// or a loser's own tet.  qid[0..*nq) are the tets to append to the next
// active list (kQueued is set here, so each tet is appended once); for a
// 2-3 winner the fresh slot is appended by the caller after qid.
WT_HD bool flip_decide_body(const Mesh& M, const uint4& pr, uint32_t it, uint32_t qid[2],
                            uint32_t* nq, uint32_t* nslot) {
  *nq = 0;
  *nslot = 0;
  if (pr.x == kInvalid) return false;
  uint64_t key = flip_key(it, pr.x);
  bool win = M.fvote[pr.x] == key && M.fvote[pr.z] == key;
  if (((pr.y >> 2) & 1u) && M.fvote[pr.w] != key) win = false;
  if (win) {
    uint32_t reuse[2] = {pr.x, pr.z};
    for (int k = 0; k < 2; ++k) {
      uint32_t old = at_or_u32(&INFO(M, reuse[k]), kAlive | kQueued);
      if (!(old & kQueued)) qid[(*nq)++] = reuse[k];
    }
    if (((pr.y >> 2) & 1u) == 0u) *nslot = 1;
  } else {
    uint32_t old = at_or_u32(&INFO(M, pr.x), kQueued);
    if (!(old & kQueued)) qid[(*nq)++] = pr.x;
  }
  return win;
}

// flip_decide_body for the list kernel: a winner also loads its flip's
// tets (into F), issued after its queue claims so that the loads and the
// claims' round trips overlap.  Same decisions and queue entries.
WT_HD bool flip_decide_load(const Mesh& M, const uint4& pr, uint32_t it, uint32_t qid[2],
                            uint32_t* nq, uint32_t* nslot, FlipTets& F) {
  *nq = 0;
  *nslot = 0;
  if (pr.x == kInvalid) return false;
  uint64_t key = flip_key(it, pr.x);
  bool win = M.fvote[pr.x] == key && M.fvote[pr.z] == key;
  if (((pr.y >> 2) & 1u) && M.fvote[pr.w] != key) win = false;
  if (win) {
    uint32_t o0 = at_or_u32(&INFO(M, pr.x), kAlive | kQueued);
    uint32_t o1 = at_or_u32(&INFO(M, pr.z), kAlive | kQueued);
    flip_load(M, pr, F);
    if (!(o0 & kQueued)) qid[(*nq)++] = pr.x;
    if (!(o1 & kQueued)) qid[(*nq)++] = pr.z;
    if (((pr.y >> 2) & 1u) == 0u) *nslot = 1;
  } else {
    uint32_t old = at_or_u32(&INFO(M, pr.x), kQueued);
    if (!(old & kQueued)) qid[(*nq)++] = pr.x;
  }
  return win;
}

// Sequential form (one proposal at a time, host callers): counters advance
// with one atomic each.
WT_HD void flip_decide_exec_body(const Mesh& M, const uint4* prop, uint32_t j, uint32_t it,
                                 uint32_t round, uint4* out, uint32_t* act_next) {
  uint4 pr = prop[j];
  uint32_t qid[2], nq, nslot;
  bool win = flip_decide_body(M, pr, it, qid, &nq, &nslot);
  for (uint32_t k = 0; k < nq; ++k) act_next[at_add_u32(&M.ctr[C_ACT_NEXT], 1u)] = qid[k];
  if (!win) return;
  uint32_t s = nslot ? at_add_u32(&M.ctr[C_TETS], 1u) : kInvalid;
  uint32_t f = at_add_u32(&M.ctr[C_FLIPS], 1u);
  flip_exec_body(M, pr, it, round, s, &out[2 * (size_t)f]);
  if (nslot) act_next[at_add_u32(&M.ctr[C_ACT_NEXT], 1u)] = s;
}

// src/shaders/update_opp.wgsl:73-220.  The flip's six external faces
// (tet T, face e, T's vertex vT opposite e): a face whose neighbour q
// flipped too takes the tet and vertex that q's flip left in its message;
// otherwise q's face points back at T.  Every load comes before the first
// store (the flip record, the pointers, then tags and messages: three round
// trips): no thread of this kernel
// stores a word that one of these loads reads (a flipped tet's external face
// is rewritten only by its own flip, after the read; an unflipped
// neighbour's face is never read; tags and messages are read-only here).
WT_HD void flip_update_body(const Mesh& M, const uint4* out, uint32_t k, uint32_t it) {
  uint4 r = out[2 * (size_t)k];
  uint4 w = out[2 * (size_t)k + 1];  // o, a, z
  // 2-3: new tets (x, y, o, a), external faces 2 (apex o) and 3 (apex a);
  // 3-2: A = (a, z, o, x), B = (z, a, o, y), external faces 0, 1, 2.
  // Selected without a branch, so the record is read in one round trip.
  bool k32 = r.w != 0u;
  uint32_t T[6] = {r.x, r.x, k32 ? r.x : r.y, r.y, k32 ? r.y : r.z, k32 ? r.y : r.z};
  uint32_t E[6] = {k32 ? 0u : 2u, k32 ? 1u : 3u, 2u, k32 ? 0u : 3u, k32 ? 1u : 2u, k32 ? 2u : 3u};
  uint32_t V[6] = {k32 ? w.y : w.x, k32 ? w.z : w.y, w.x, k32 ? w.z : w.y, k32 ? w.y : w.x,
                   k32 ? w.x : w.y};
  uint32_t ptr[6], fq[6];
  uint2 msg[6];
  for (int i = 0; i < 6; ++i) ptr[i] = *opp_word(M, T[i], E[i]);
  for (int i = 0; i < 6; ++i) fq[i] = (uint32_t)FSTAMP(M, opp_tet(ptr[i]));
  // every neighbour is a valid slot: its message is loaded unconditionally
  // (stale where it did not flip, then unused), so no branch splits the loads
  for (int i = 0; i < 6; ++i)
    msg[i] = reinterpret_cast<const uint2*>(&M.tmsg[opp_tet(ptr[i])])[opp_face(ptr[i])];
  // one 8-byte store per face, its slot and value chosen without a branch
  for (int i = 0; i < 6; ++i) {
    bool fl = fq[i] == it;
    uint32_t dt = fl ? T[i] : opp_tet(ptr[i]), df = fl ? E[i] : opp_face(ptr[i]);
    set_face(M, dt, df, fl ? msg[i].x : enc_opp(T[i], E[i]), fl ? msg[i].y : V[i]);
  }
}

// ---------------------------------------------------------------------------
// relocate points whose tet a flip replaced: a stochastic visibility walk
// (Devillers, Pion, Teillaud 2002) from the old slot.  Replaces the
// reference's flip-trace replay (src/shaders/relocate_points_fast.wgsl).
// ---------------------------------------------------------------------------
constexpr uint32_t kWalkLimit = 1u << 16;

// *state: 0 arrived, 1 step limit, 2 deferred (FAST filter undecided)
template <bool FAST>
WT_HD uint32_t walk_to_t(const Mesh& M, uint32_t t, uint32_t q, uint32_t seed, PredStats* st,
                         int* state) {
  float Q[3];
  uint32_t rq;
  load_pt(M, q, Q, &rq);
  for (uint32_t step = 0; step < kWalkLimit; ++step) {
    uint4 tv = TV(M, t);
    uint4 to = TN(M, t);
    uint32_t r0 = hash32(seed + step * 0x9E3779B9u) & 3u;
    // the tet's four points in one round trip, all four face tests, then
    // the step takes the first face in its random order (r0, r0 + 1, ...)
    // that q lies beyond, or defers if an undecided face comes first
    float P[4][3];
    uint32_t R[4];
    for (int k = 0; k < 4; ++k) load_pt(M, get4(tv, k), P[k], &R[k]);
    uint32_t beyond = 0u, undecided = 0u;
#pragma unroll
    for (int f = 0; f < 4; ++f) {
      int a = face_vi(f, 0), b = face_vi(f, 1), c = face_vi(f, 2);
      int o = o3s<FAST>(M, P[a], P[b], P[c], Q, R[a], R[b], R[c], rq, st);
      if (o == 2) undecided |= 1u << f;
      else if (o < 0) beyond |= 1u << f;
    }
    uint32_t m = beyond | undecided;
    if (!m) {
      *state = 0;
      return t;
    }
    uint32_t rot = ((m >> r0) | (m << (4u - r0))) & 0xFu;
    uint32_t f = ((uint32_t)ctz32(rot) + r0) & 3u;
    if (undecided & (1u << f)) {
      *state = 2;
      return t;
    }
    t = opp_tet(get4(to, (int)f));
  }
  *state = 1;
  return t;
}

template <bool FAST>
WT_HD bool relocate_body_t(const Mesh& M, const uint32_t* uvert, uint32_t* utet, uint32_t i,
                           uint32_t round, PredStats* st) {
  uint32_t t = utet[i];
  if ((uint32_t)(FSTAMP(M, t) >> 32) != round) return false;
  for (int guard = 0; guard < 64 && (INFO(M, t) & kDead); ++guard) t = TFWD(M, t);
  int state = 0;
  t = walk_to_t<FAST>(M, t, uvert[i], i * 0x85EBCA6Bu, st, &state);
  if (state == 2) return true;  // the full walk restarts from utet[i]
  if (state == 1) at_add_u32(&M.ctr[C_WALK_FAIL], 1u);
  utet[i] = t;
  return false;
}

WT_HD void relocate_body(const Mesh& M, const uint32_t* uvert, uint32_t* utet, uint32_t i,
                         uint32_t round, PredStats* st) {
  (void)relocate_body_t<false>(M, uvert, utet, i, round, st);
}

// ---------------------------------------------------------------------------
// audit: every finite face, counted once from its lower slot
// (src/lib.rs:25-78 checks real-only pairs; this covers all finite faces)
// rec: (t << 2 | f) of violations, capacity cap.  FAST: FP32 filter only;
// returns true (nothing recorded) when a face is undecided, for the full
// pass to redo the tet.
// ---------------------------------------------------------------------------
template <bool FAST>
WT_HD bool audit_body_t(const Mesh& M, uint32_t t, uint32_t* rec, uint32_t cap, PredStats* st) {
  if (!(INFO(M, t) & kAlive)) return false;
  uint4 tv = TV(M, t);
  if (tv.x == M.inf || tv.y == M.inf || tv.z == M.inf || tv.w == M.inf) return false;
  uint4 to = TN(M, t);
  uint4 tov = TOV(M, t);
  float P[4][3];
  uint32_t R[4];
  for (int k = 0; k < 4; ++k) load_pt(M, get4(tv, k), P[k], &R[k]);
  int bad[4] = {0, 0, 0, 0};
  for (int f = 0; f < 4; ++f) {
    uint32_t nt = opp_tet(get4(to, f));
    if (nt < t) continue;
    uint32_t o = get4(tov, f);
    if (o == M.inf) continue;
    float O[3];
    uint32_t ro;
    load_pt(M, o, O, &ro);
    int s = o4s<FAST>(M, P[0], P[1], P[2], P[3], O, R[0], R[1], R[2], R[3], ro, st);
    if (s == 2) return true;
    bad[f] = s < 0;
  }
  for (int f = 0; f < 4; ++f)
    if (bad[f]) {
      at_add_u32(&M.ctr[C_VIOL], 1u);
      uint32_t k = at_add_u32(&M.ctr[C_FAILREC], 1u);
      if (k < cap) rec[k] = (t << 2) | (uint32_t)f;
    }
  return false;
}

WT_HD void audit_body(const Mesh& M, uint32_t t, uint32_t* rec, uint32_t cap, PredStats* st) {
  (void)audit_body_t<false>(M, t, rec, cap, st);
}

}  // namespace wtet

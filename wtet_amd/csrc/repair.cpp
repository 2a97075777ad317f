// Host phase 2 (see repair.h) and the host quality check.
//
// Repair by region re-triangulation: around a non-Delaunay face take the
// region R of tets near it, compute the Delaunay triangulation D of R's
// vertices together with the four super vertices (Bowyer-Watson with the
// same SoS predicates as the kernels), and when every boundary face of R is
// a face of D, replace R by the tets of D that lie inside R.  Every Delaunay
// tet of the whole point set with vertices in R is a tet of D, so once the
// new boundary faces pass the insphere test the result is the triangulation
// the Delaunay lemma certifies; otherwise R grows and the step repeats.  In
// the limit R is the whole mesh and D the full triangulation, which is what
// the reference's full_bw_reconstruction computes (src/phase2.rs:383-606).
#include "repair.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <deque>
#include <unordered_map>
#include <unordered_set>
#include <atomic>
#include <new>
#include <thread>

#include "common.h"

namespace {

using namespace wtet;

struct PointSet {
  const float* xyzn;
  uint32_t n;
  float sup[4][3];
  const float* p(uint32_t v) const { return v < n ? xyzn + 3 * (size_t)v : sup[v - n]; }
};

int o3sos(const PointSet& S, uint32_t a, uint32_t b, uint32_t c, uint32_t d, int mode) {
  return orient3d_sos(S.p(a), S.p(b), S.p(c), S.p(d), a, b, c, d, mode, nullptr);
}

int o4(const PointSet& S, const uint32_t* t, uint32_t q, int mode) {
  return orient4d(S.p(t[0]), S.p(t[1]), S.p(t[2]), S.p(t[3]), S.p(q), t[0], t[1], t[2], t[3], q,
                  mode, nullptr);
}

struct FaceKey {
  uint32_t a, b, c;
  bool operator==(const FaceKey& o) const { return a == o.a && b == o.b && c == o.c; }
};
struct FaceHash {
  size_t operator()(const FaceKey& k) const {
    uint64_t h = k.a * 0x9E3779B97F4A7C15ull;
    h ^= (k.b + 0x632BE59BD9B4E019ull + (h << 6) + (h >> 2));
    h ^= (k.c * 0xC2B2AE3D27D4EB4Full + (h << 6) + (h >> 2));
    return (size_t)h;
  }
};
// The faces of a local triangulation in one open-addressing table (a
// region's attempts reuse its storage): sorted vertex triple -> the two
// sides, 4 * tet + face (-1: none yet).
struct FaceTable {
  std::vector<FaceKey> key;
  std::vector<std::array<int64_t, 2>> side;
  std::vector<uint8_t> used;
  size_t mask = 0;
  void reset(size_t faces) {
    size_t cap = 64;
    while (cap < 2 * faces) cap <<= 1;
    if (key.size() < cap) {
      key.resize(cap);
      side.resize(cap);
    }
    used.assign(cap, 0);
    mask = cap - 1;
  }
  size_t slot(const FaceKey& k) const {
    uint64_t h = FaceHash()(k);
    h ^= h >> 33;
    h *= 0xFF51AFD7ED558CCDull;
    h ^= h >> 33;
    size_t i = (size_t)h & mask;
    while (used[i] && !(key[i] == k)) i = (i + 1) & mask;
    return i;
  }
  void add(const FaceKey& k, int64_t s) {
    size_t i = slot(k);
    if (!used[i]) {
      used[i] = 1;
      key[i] = k;
      side[i] = {s, -1};
    } else {
      side[i][1] = s;
    }
  }
  const std::array<int64_t, 2>* get(const FaceKey& k) const {
    size_t i = slot(k);
    return used[i] ? &side[i] : nullptr;
  }
};

FaceKey face_key(uint32_t x, uint32_t y, uint32_t z) {
  uint32_t a[3] = {x, y, z};
  std::sort(a, a + 3);
  return FaceKey{a[0], a[1], a[2]};
}

// ---------------------------------------------------------------------------
// local Bowyer-Watson over a vertex subset plus the super vertices
// ---------------------------------------------------------------------------
struct Local {
  std::vector<uint32_t> v;   // 4 per tet
  std::vector<int64_t> a;    // 4 per tet: neighbour * 4 + face, -1 none
  std::vector<uint8_t> alive;
  std::vector<int64_t> mark;
  std::vector<int64_t> freel;
  int64_t last = 0;          // a tet of the last insertion (the next walk starts there)
  int64_t stamp = 0;         // insertions so far (cavity marks)
  uint32_t rng = 0x2545F491u;

  int64_t make() {
    if (!freel.empty()) {
      int64_t t = freel.back();
      freel.pop_back();
      return t;
    }
    int64_t t = (int64_t)alive.size();
    v.resize(v.size() + 4);
    a.resize(a.size() + 4, -1);
    alive.push_back(0);
    mark.push_back(-1);
    return t;
  }
};

uint32_t morton10(const float* p) {
  auto q = [](float x) {
    float c = x < 0.f ? 0.f : (x > 1.f ? 1.f : x);
    uint32_t v = (uint32_t)(c * 1023.0f);
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
  };
  return (q(p[2]) << 2) | (q(p[1]) << 1) | q(p[0]);
}

// The Delaunay triangulation of the four super vertices alone.
void local_init(const PointSet& S, int mode, Local& L) {
  uint32_t n = S.n;
  L = Local();
  int64_t t0 = L.make();
  uint32_t s[4] = {n, n + 1, n + 2, n + 3};
  if (o3sos(S, s[0], s[1], s[2], s[3], mode) <= 0) std::swap(s[0], s[1]);
  std::memcpy(&L.v[0], s, 16);
  L.alive[t0] = 1;
  L.last = t0;
}

// Bowyer-Watson insertion of `ids` into L (initialised by local_init, maybe
// holding earlier insertions: a region that grows inserts only its new
// vertices).  Morton order: each visibility walk starts at the previous
// point's new tets, a few steps away.
bool local_insert(const PointSet& S, std::vector<uint32_t> ids, int mode, Local& L, std::string& err) {
  std::vector<std::pair<uint32_t, uint32_t>> ord;
  ord.reserve(ids.size());
  for (uint32_t q : ids) ord.emplace_back(morton10(S.p(q)), q);
  std::sort(ord.begin(), ord.end());
  for (size_t i = 0; i < ord.size(); ++i) ids[i] = ord[i].second;
  int64_t last = L.last;
  uint32_t rng = L.rng;
  std::vector<int64_t> cav, stack;
  struct BF {
    uint32_t f[3];
    int64_t out;
  };
  std::vector<BF> bf;
  std::unordered_map<uint64_t, std::pair<int64_t, int>> edges;
  auto conflict = [&](int64_t t, uint32_t q) { return o4(S, &L.v[4 * t], q, mode) < 0; };
  for (size_t it = 0; it < ids.size(); ++it) {
    uint32_t q = ids[it];
    // visibility walk
    int64_t t = last;
    bool found = false;
    for (size_t step = 0; step < 4 * L.alive.size() + 64; ++step) {
      rng = rng * 1664525u + 1013904223u;
      int off = (int)(rng >> 30);
      int moved = 0;
      for (int k = 0; k < 4; ++k) {
        int f = (k + off) & 3;
        const uint32_t* tv = &L.v[4 * t];
        if (o3sos(S, tv[face_vi(f, 0)], tv[face_vi(f, 1)], tv[face_vi(f, 2)], q, mode) < 0) {
          int64_t nb = L.a[4 * t + f];
          if (nb < 0) {
            err = "walk left the super tetrahedron";
            return false;
          }
          t = nb >> 2;
          moved = 1;
          break;
        }
      }
      if (!moved) {
        found = true;
        break;
      }
    }
    if (!found || !conflict(t, q)) {
      t = -1;
      for (int64_t u = 0; u < (int64_t)L.alive.size(); ++u)
        if (L.alive[u] && conflict(u, q)) {
          t = u;
          break;
        }
      if (t < 0) continue;  // redundant under the lifted SoS (an exact duplicate):
                            // the reference's Bowyer-Watson skips it too
                            // (src/phase2.rs:440-443)
    }
    // cavity
    int64_t stamp = L.stamp++;
    cav.clear();
    stack.clear();
    L.mark[t] = stamp;
    cav.push_back(t);
    stack.push_back(t);
    while (!stack.empty()) {
      int64_t u = stack.back();
      stack.pop_back();
      for (int f = 0; f < 4; ++f) {
        int64_t nb = L.a[4 * u + f];
        if (nb < 0) continue;
        int64_t w = nb >> 2;
        if (L.mark[w] == stamp || L.mark[w] == -2 - stamp) continue;
        if (conflict(w, q)) {
          L.mark[w] = stamp;
          cav.push_back(w);
          stack.push_back(w);
        } else {
          L.mark[w] = -2 - stamp;
        }
      }
    }
    bf.clear();
    for (int64_t u : cav)
      for (int f = 0; f < 4; ++f) {
        int64_t nb = L.a[4 * u + f];
        if (nb >= 0 && L.mark[nb >> 2] == stamp) continue;
        BF b;
        for (int k = 0; k < 3; ++k) b.f[k] = L.v[4 * u + face_vi(f, k)];
        b.out = nb;
        bf.push_back(b);
      }
    for (int64_t u : cav) {
      L.alive[u] = 0;
      L.freel.push_back(u);
    }
    edges.clear();
    for (const BF& b : bf) {
      uint32_t x = b.f[0], y = b.f[1], z = b.f[2];
      if (o3sos(S, x, y, z, q, mode) <= 0) std::swap(x, y);
      int64_t nt = L.make();
      uint32_t* tv = &L.v[4 * nt];
      tv[0] = x;
      tv[1] = y;
      tv[2] = z;
      tv[3] = q;
      L.alive[nt] = 1;
      L.mark[nt] = -1;
      L.a[4 * nt + 3] = b.out;
      if (b.out >= 0) L.a[b.out] = 4 * nt + 3;
      for (int f = 0; f < 3; ++f) {
        uint32_t e0 = tv[(f + 1) % 3], e1 = tv[(f + 2) % 3];
        uint64_t key = e0 < e1 ? ((uint64_t)e0 << 32 | e1) : ((uint64_t)e1 << 32 | e0);
        auto itx = edges.find(key);
        if (itx == edges.end()) {
          edges.emplace(key, std::make_pair(nt, f));
        } else {
          L.a[4 * nt + f] = 4 * itx->second.first + itx->second.second;
          L.a[4 * itx->second.first + itx->second.second] = 4 * nt + f;
          edges.erase(itx);
        }
      }
      last = nt;
    }
    if (!edges.empty()) {
      err = "cavity boundary is not closed";
      return false;
    }
  }
  L.last = last;
  L.rng = rng;
  return true;
}

// ---------------------------------------------------------------------------
// in-place repair on the result arrays (ids, adjacency (t << 5) | f)
// ---------------------------------------------------------------------------
struct NotPresent {
  uint32_t row;
};

struct Mesh32 {
  MeshBuf& m;
  uint32_t inf;
  void need(uint32_t t) const {
    if (m.present && t < m.known && m.present[t] != m.present_mark) throw NotPresent{t};
  }
  size_t count() const { return m.count; }
  bool dead(uint32_t t) const {
    need(t);
    return m.tets[4 * (size_t)t] == kInvalid;
  }
  bool has_inf(uint32_t t) const {
    need(t);
    const uint32_t* v = &m.tets[4 * (size_t)t];
    return v[0] == inf || v[1] == inf || v[2] == inf || v[3] == inf;
  }
  uint32_t* tv(uint32_t t) {
    need(t);
    return &m.tets[4 * (size_t)t];
  }
  const uint32_t* tv(uint32_t t) const {
    need(t);
    return &m.tets[4 * (size_t)t];
  }
  uint32_t& ta(uint32_t t, int f) {
    need(t);
    return m.adj[4 * (size_t)t + f];
  }
  uint32_t ta(uint32_t t, int f) const {
    need(t);
    return m.adj[4 * (size_t)t + f];
  }
};

// Fills the dead slots `holes` with live tets moved from the end of the
// rows (their neighbours' back-pointers follow), then drops dead tail rows,
// so that rows [0, count) hold live tets only.
void fill_holes(Mesh32& W, std::vector<uint32_t>& holes, std::vector<uint32_t>& changed) {
  MeshBuf& mesh = W.m;
  std::sort(holes.begin(), holes.end());
  auto trim = [&]() {
    while (mesh.count > 0 && W.dead((uint32_t)(mesh.count - 1))) --mesh.count;
  };
  for (uint32_t h : holes) {
    trim();
    if (h >= mesh.count) continue;
    uint32_t last = (uint32_t)(mesh.count - 1);
    std::memcpy(W.tv(h), W.tv(last), 16);
    for (int f = 0; f < 4; ++f) {
      uint32_t e = W.ta(last, f);
      W.ta(h, f) = e;
      if (e != kInvalid) {
        W.ta(e >> 5, (int)(e & 3u)) = (h << 5) | (uint32_t)f;
        changed.push_back(e >> 5);
      }
    }
    changed.push_back(h);
    --mesh.count;
  }
  trim();
}

bool face_bad(const PointSet& S, const Mesh32& W, uint32_t t, int f, int mode) {
  uint32_t e = W.ta(t, f);
  if (e == kInvalid) return false;
  uint32_t u = e >> 5;
  if (W.has_inf(t) || W.has_inf(u)) return false;
  uint32_t o = W.tv(u)[e & 3u];
  return o4(S, W.tv(t), o, mode) < 0;
}

}  // namespace

static wtet_status repair_impl(const float* xyzn, uint32_t n, MeshBuf& mesh,
                               const std::vector<uint32_t>& viol, int mode, RepairReport* rep,
                               uint32_t region_cap);

wtet_status repair_delaunay(const float* xyzn, uint32_t n, MeshBuf& mesh,
                            const std::vector<uint32_t>& viol, int mode, RepairReport* rep,
                            uint32_t region_cap) {
  try {
    return repair_impl(xyzn, n, mesh, viol, mode, rep, region_cap);
  } catch (const NotPresent&) {
    if (rep) {
      rep->needs_full = true;
      rep->message = "the repair reached rows outside the downloaded neighbourhood";
    }
    return WTET_ERR_NOT_CONVERGED;
  }
}

namespace {

// A region's replacement, computed from the mesh without writing to it, so
// that the regions of distinct faces can be computed on several threads.
enum : int { kPlanOk = 0, kPlanClean = 1, kPlanFetch = 2, kPlanFail = 3 };
struct Plan {
  uint32_t t = 0;
  int f = 0;
  int status = kPlanClean;
  std::vector<uint32_t> R;      // the region's slots
  Local L;                      // Delaunay triangulation of R's vertices and the super vertices
  std::vector<int64_t> order;   // the tets of L inside R
  std::unordered_map<FaceKey, uint32_t, FaceHash> bface;  // R's boundary faces -> outside neighbour
  uint32_t missing_row = 0;     // kPlanFetch: a row the host lacks
  wtet_status err = WTET_OK;    // kPlanFail
  bool too_large = false;
  std::string msg;
};

// The region of the face (t, f): R starts as t, its neighbour across f and
// their neighbours, and grows by one ring of face-neighbours (and by the
// tets across boundary faces that L lacks) until every boundary face of R
// is a face of L; the plan is then the tets of L inside R.
void plan_region(const PointSet& S, const Mesh32& W, uint32_t t, int f, int mode, uint32_t region_cap,
                 Plan& P) {
  P = Plan();
  P.t = t;
  P.f = f;
  const uint32_t n = S.n;
  try {
    if (t >= W.count() || W.dead(t) || !face_bad(S, W, t, f, mode)) return;
    std::unordered_set<uint32_t> inR;
    std::vector<uint32_t>& R = P.R;
    std::vector<uint32_t> ring;
    auto add = [&](uint32_t x) {
      if (x >= W.count() || W.dead(x) || W.has_inf(x) || inR.count(x)) return;
      inR.insert(x);
      R.push_back(x);
    };
    add(t);
    add(W.ta(t, f) >> 5);
    // the local triangulation grows with R: each attempt inserts only the
    // vertices the previous attempts did not
    Local& L = P.L;
    local_init(S, mode, L);
    std::vector<uint32_t> have, fresh, real;
    FaceTable fmap;
    for (;;) {
      ring.clear();
      for (uint32_t x : R)
        for (int h = 0; h < 4; ++h)
          if (W.ta(x, h) != kInvalid) ring.push_back(W.ta(x, h) >> 5);
      for (uint32_t x : ring) add(x);
      if (region_cap && R.size() > region_cap) {
        P.status = kPlanFail;
        P.err = WTET_ERR_NOT_CONVERGED;
        P.too_large = true;
        P.msg = "a repair region outgrew its cap";
        return;
      }
      if (R.size() * 2 > W.count())
        for (uint32_t x = 0; x < W.count(); ++x) add(x);
      real.clear();
      for (uint32_t x : R)
        for (int k = 0; k < 4; ++k)
          if (W.tv(x)[k] < n) real.push_back(W.tv(x)[k]);
      std::sort(real.begin(), real.end());
      real.erase(std::unique(real.begin(), real.end()), real.end());
      fresh.clear();
      std::set_difference(real.begin(), real.end(), have.begin(), have.end(), std::back_inserter(fresh));
      have.swap(real);
      std::string err;
      if (!local_insert(S, fresh, mode, L, err)) {
        P.status = kPlanFail;
        P.err = WTET_ERR_INTERNAL;
        P.msg = err;
        return;
      }
      fmap.reset(2 * L.alive.size() + 16);
      for (int64_t d = 0; d < (int64_t)L.alive.size(); ++d) {
        if (!L.alive[d]) continue;
        const uint32_t* dv = &L.v[4 * d];
        for (int h = 0; h < 4; ++h)
          fmap.add(face_key(dv[face_vi(h, 0)], dv[face_vi(h, 1)], dv[face_vi(h, 2)]), 4 * d + h);
      }
      // boundary faces of R -> outside neighbour (t << 5 | f, or kInvalid)
      auto& bface = P.bface;
      bface.clear();
      std::vector<int64_t> seeds;
      std::vector<uint32_t> grow_with;
      bool missing = false;
      for (uint32_t x : R)
        for (int h = 0; h < 4; ++h) {
          uint32_t nb = W.ta(x, h);
          if (nb != kInvalid && inR.count(nb >> 5)) continue;
          if (nb != kInvalid) (void)W.dead(nb >> 5);  // written by the commit: must be present
          const uint32_t* xv = W.tv(x);
          uint32_t fa = xv[face_vi(h, 0)], fb = xv[face_vi(h, 1)], fc = xv[face_vi(h, 2)];
          FaceKey k = face_key(fa, fb, fc);
          const std::array<int64_t, 2>* it = fmap.get(k);
          if (!it) {
            missing = true;
            if (nb != kInvalid) grow_with.push_back(nb >> 5);
            continue;
          }
          bface.emplace(k, nb);
          for (int s2 = 0; s2 < 2; ++s2) {
            int64_t df = (*it)[s2];
            if (df < 0) continue;
            uint32_t apex = L.v[4 * (df >> 2) + (df & 3)];
            if (o3sos(S, fa, fb, fc, apex, mode) > 0) seeds.push_back(df >> 2);
          }
        }
      if (missing) {
        for (uint32_t x : grow_with) add(x);
        continue;
      }
      // inner tets of L: flood from the seeds, never across a boundary face
      std::unordered_set<int64_t> inner;
      std::vector<int64_t> st = seeds;
      for (int64_t d : seeds)
        if (inner.insert(d).second) P.order.push_back(d);
      while (!st.empty()) {
        int64_t d = st.back();
        st.pop_back();
        const uint32_t* dv = &L.v[4 * d];
        for (int h = 0; h < 4; ++h) {
          FaceKey k = face_key(dv[face_vi(h, 0)], dv[face_vi(h, 1)], dv[face_vi(h, 2)]);
          if (bface.count(k)) continue;
          int64_t nb = L.a[4 * d + h];
          if (nb < 0) {
            P.status = kPlanFail;
            P.err = WTET_ERR_INTERNAL;
            P.msg = "region flood reached the hull";
            return;
          }
          if (inner.insert(nb >> 2).second) {
            st.push_back(nb >> 2);
            P.order.push_back(nb >> 2);
          }
        }
      }
      P.status = kPlanOk;
      return;
    }
  } catch (const NotPresent& np) {
    P.status = kPlanFetch;
    P.missing_row = np.row;
  }
}

// Writes a plan into the mesh: the tets of L inside R take R's slots, then
// appended ones; R's leftover slots become holes.  Returns false when the
// rows run out of room.  The boundary faces go to `recheck`.
bool commit_plan(Mesh32& W, const Plan& P, std::vector<uint32_t>& holes, std::vector<uint32_t>& changed,
                 std::vector<uint32_t>* touched, std::deque<std::pair<uint32_t, int>>& recheck,
                 RepairReport* rep) {
  MeshBuf& mesh = W.m;
  const Local& L = P.L;
  std::unordered_map<int64_t, uint32_t> slot;
  slot.reserve(P.order.size());
  size_t k = 0;
  for (int64_t d : P.order) {
    uint32_t sl;
    if (k < P.R.size()) {
      sl = P.R[k++];
    } else {
      if (mesh.count >= mesh.cap) return false;
      sl = (uint32_t)mesh.count++;
      for (int h = 0; h < 4; ++h) W.ta(sl, h) = kInvalid;
    }
    slot[d] = sl;
    changed.push_back(sl);
    if (touched) touched->push_back(sl);
  }
  for (; k < P.R.size(); ++k) {
    W.tv(P.R[k])[0] = kInvalid;
    holes.push_back(P.R[k]);
    if (touched) touched->push_back(P.R[k]);
  }
  for (int64_t d : P.order) std::memcpy(W.tv(slot[d]), &L.v[4 * d], 16);
  for (int64_t d : P.order) {
    uint32_t sl = slot[d];
    const uint32_t* dv = &L.v[4 * d];
    for (int h = 0; h < 4; ++h) {
      FaceKey fk = face_key(dv[face_vi(h, 0)], dv[face_vi(h, 1)], dv[face_vi(h, 2)]);
      auto bf = P.bface.find(fk);
      if (bf != P.bface.end()) {
        uint32_t out = bf->second;
        W.ta(sl, h) = out;
        if (out != kInvalid) {
          W.ta(out >> 5, (int)(out & 3u)) = (sl << 5) | (uint32_t)h;
          changed.push_back(out >> 5);
          if (touched) touched->push_back(out >> 5);
        }
        recheck.emplace_back(sl, h);
      } else {
        int64_t nb = L.a[4 * d + h];
        W.ta(sl, h) = (slot.at(nb >> 2) << 5) | (uint32_t)(nb & 3);
      }
    }
  }
  if (rep) {
    rep->regions++;
    rep->max_region = std::max<uint32_t>(rep->max_region, (uint32_t)P.R.size());
  }
  return true;
}

// Centroids of the faces planned in a round, hashed on a grid of cells
// `cell` wide (fixed per round); near(c, r): a planned face's centroid lies
// within max(r, its own reach) of c, searched over the cells within two
// cells of c (a face farther out than that is not found: it is planned
// too, which costs a duplicate plan, not a wrong result).
struct KnotGrid {
  float cell = 0.f;
  std::unordered_map<uint64_t, std::vector<std::array<float, 4>>> cells;
  void clear() {
    cells.clear();
    cell = 0.f;
  }
  static int64_t ci(float x, float cell) { return (int64_t)std::floor(x / cell); }
  static uint64_t key(int64_t x, int64_t y, int64_t z) {
    return ((uint64_t)(x & 0x1FFFFF) << 42) | ((uint64_t)(y & 0x1FFFFF) << 21) | (uint64_t)(z & 0x1FFFFF);
  }
  bool near(const float* c, float r) const {
    if (cells.empty() || !(cell > 0.f)) return false;
    int64_t x = ci(c[0], cell), y = ci(c[1], cell), z = ci(c[2], cell);
    for (int64_t dx = -2; dx <= 2; ++dx)
      for (int64_t dy = -2; dy <= 2; ++dy)
        for (int64_t dz = -2; dz <= 2; ++dz) {
          auto it = cells.find(key(x + dx, y + dy, z + dz));
          if (it == cells.end()) continue;
          for (const auto& e : it->second) {
            float d2 = (e[0] - c[0]) * (e[0] - c[0]) + (e[1] - c[1]) * (e[1] - c[1]) + (e[2] - c[2]) * (e[2] - c[2]);
            float rr = std::max(r, e[3]);
            if (d2 <= rr * rr) return true;
          }
        }
    return false;
  }
  void insert(const float* c, float r) {
    if (!(cell > 0.f)) cell = std::max(r, 1e-6f);
    cells[key(ci(c[0], cell), ci(c[1], cell), ci(c[2], cell))].push_back({c[0], c[1], c[2], r});
  }
};

// Host threads for the region plans of one round (the box gives a GPU's
// process 16 cores)
unsigned plan_threads(size_t jobs) {
  unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  return (unsigned)std::min<size_t>(jobs, std::min(hw, 16u));
}

}  // namespace

static wtet_status repair_impl(const float* xyzn, uint32_t n, MeshBuf& mesh,
                               const std::vector<uint32_t>& viol, int mode, RepairReport* rep,
                               uint32_t region_cap) {
  PointSet S;
  S.xyzn = xyzn;
  S.n = n;
  const float sv[4][3] = {{-100.f, -100.f, -100.f}, {400.f, -100.f, -100.f},
                          {-100.f, 400.f, -100.f}, {-100.f, -100.f, 400.f}};
  std::memcpy(S.sup, sv, sizeof(sv));
  mode = mode == 1 ? PRED_EXACT_ONLY : PRED_FILTERED;
  Mesh32 W{mesh, n + 4};
  std::vector<uint32_t>& changed = mesh.changed;
  std::vector<uint32_t> holes;

  auto failed = [&](const Plan& P) {
    if (rep) {
      rep->message = P.msg;
      rep->too_large = rep->too_large || P.too_large;
    }
    return P.err;
  };
  // The serial path of one face: a plan, fetching the rows it lacks, then
  // its commit.  Every row a plan reads is read before the commit writes.
  std::deque<std::pair<uint32_t, int>> recheck;
  Plan P1;
  auto serial = [&](uint32_t t, int f, std::vector<uint32_t>* touched_out) -> wtet_status {
    for (;;) {
      plan_region(S, W, t, f, mode, region_cap, P1);
      if (P1.status != kPlanFetch) break;
      if (!mesh.fetch || !mesh.fetch(P1.missing_row)) throw NotPresent{P1.missing_row};
      if (rep) rep->fetches++;
    }
    if (P1.status == kPlanFail) return failed(P1);
    if (P1.status == kPlanOk && !commit_plan(W, P1, holes, changed, touched_out, recheck, rep)) {
      if (rep) rep->message = "repair needs more room than the staging buffer has";
      return WTET_ERR_OUT_OF_MEMORY;
    }
    return WTET_OK;
  };

  // Rounds: the queued faces that are still non-Delaunay, one per knot (a
  // face within kKnotReach of its longest edge of a face already planned in
  // the round, centroid to centroid, waits for the next round), are planned
  // on several threads, then
  // committed one by one in queue order.  A plan whose region or outer
  // neighbours an earlier commit of the round changed goes the serial way.
  // A face whose rows the host lacks, and a plan that reached a missing
  // row, wait for the next round; the rows they reached are fetched
  // together after the round's commits (with fetch_many; without it they go
  // the serial way, which fetches one row at a time).  Each commit's
  // boundary faces join the next round.
  std::vector<std::pair<uint32_t, int>> cur;
  for (uint32_t r : viol) cur.emplace_back(r >> 2, (int)(r & 3u));
  std::vector<std::pair<uint32_t, int>> next, reps;
  std::vector<Plan> plans;
  std::unordered_set<uint32_t> touched;
  std::vector<uint32_t> tl, want;
  constexpr float kKnotReach = 6.f;
  KnotGrid grid;
  while (!cur.empty()) {
    next.clear();
    reps.clear();
    grid.clear();
    want.clear();
    for (auto [t, f] : cur) {
      bool bad = false, lacks = false;
      uint32_t row = 0;
      try {
        bad = t < W.count() && !W.dead(t) && face_bad(S, W, t, f, mode);
      } catch (const NotPresent& np) {
        lacks = true;
        row = np.row;
      }
      if (lacks && mesh.fetch_many) {
        want.push_back(row);
        next.emplace_back(t, f);
        continue;
      }
      if (lacks) {
        // before this round's plans are computed, so they see its commit
        wtet_status st = serial(t, f, nullptr);
        if (st != WTET_OK) return st;
        for (auto& q : recheck) next.push_back(q);
        recheck.clear();
        continue;
      }
      if (!bad) continue;
      // stuck faces come in knots a few tets across: a face near one
      // already planned waits
      const uint32_t* tv = W.tv(t);
      const float* q[3] = {S.p(tv[face_vi(f, 0)]), S.p(tv[face_vi(f, 1)]), S.p(tv[face_vi(f, 2)])};
      float c[3], len2 = 0.f;
      for (int k = 0; k < 3; ++k) {
        c[k] = (q[0][k] + q[1][k] + q[2][k]) / 3.f;
        const float* a = q[k];
        const float* b = q[(k + 1) % 3];
        float d2 = (a[0] - b[0]) * (a[0] - b[0]) + (a[1] - b[1]) * (a[1] - b[1]) + (a[2] - b[2]) * (a[2] - b[2]);
        len2 = std::max(len2, d2);
      }
      const float reach = kKnotReach * std::sqrt(len2);
      if (grid.near(c, reach)) {
        next.emplace_back(t, f);
        continue;
      }
      grid.insert(c, reach);
      reps.emplace_back(t, f);

    }
    plans.assign(reps.size(), Plan());
    unsigned nt = plan_threads(reps.size());
    if (nt <= 1) {
      for (size_t i = 0; i < reps.size(); ++i)
        plan_region(S, W, reps[i].first, reps[i].second, mode, region_cap, plans[i]);
    } else {
      std::atomic<size_t> nextjob{0};
      std::vector<std::thread> th;
      for (unsigned k = 0; k < nt; ++k)
        th.emplace_back([&] {
          for (size_t i; (i = nextjob.fetch_add(1)) < reps.size();) {
            // an exception must not leave a worker thread (it would end the
            // process): it becomes the plan's failure
            try {
              plan_region(S, W, reps[i].first, reps[i].second, mode, region_cap, plans[i]);
            } catch (const std::bad_alloc&) {
              plans[i].status = kPlanFail;
              plans[i].err = WTET_ERR_OUT_OF_MEMORY;
              plans[i].msg = "host memory ran out while planning a repair region";
            } catch (...) {
              plans[i].status = kPlanFail;
              plans[i].err = WTET_ERR_INTERNAL;
              plans[i].msg = "a repair region's plan failed";
            }
          }
        });
      for (auto& x : th) x.join();
    }
    touched.clear();
    for (size_t i = 0; i < plans.size(); ++i) {
      const Plan& P = plans[i];
      if (P.status == kPlanClean) continue;
      if (P.status == kPlanFail) return failed(P);
      if (P.status == kPlanFetch && mesh.fetch_many) {
        want.push_back(P.missing_row);
        next.emplace_back(P.t, P.f);
        continue;
      }
      bool stale = P.status == kPlanFetch;
      for (size_t k = 0; !stale && k < P.R.size(); ++k) stale = touched.count(P.R[k]) > 0;
      for (auto it = P.bface.begin(); !stale && it != P.bface.end(); ++it)
        stale = it->second != kInvalid && touched.count(it->second >> 5) > 0;
      if (stale) {
        tl.clear();
        wtet_status st = serial(P.t, P.f, &tl);
        if (st != WTET_OK) return st;
        touched.insert(tl.begin(), tl.end());
        for (auto& [t2, f2] : recheck) next.emplace_back(t2, f2);
        recheck.clear();
        continue;
      }
      tl.clear();
      if (!commit_plan(W, P, holes, changed, &tl, recheck, rep)) {
        if (rep) rep->message = "repair needs more room than the staging buffer has";
        return WTET_ERR_OUT_OF_MEMORY;
      }
      touched.insert(tl.begin(), tl.end());
      for (auto& [t2, f2] : recheck) next.emplace_back(t2, f2);
      recheck.clear();
    }
    if (!want.empty()) {
      std::sort(want.begin(), want.end());
      want.erase(std::unique(want.begin(), want.end()), want.end());
      if (!mesh.fetch_many(want)) throw NotPresent{want[0]};
      if (rep) rep->fetches++;
    }
    cur.swap(next);
  }
  fill_holes(W, holes, changed);
  std::sort(changed.begin(), changed.end());
  changed.erase(std::unique(changed.begin(), changed.end()), changed.end());
  while (!changed.empty() && changed.back() >= mesh.count) changed.pop_back();
  return WTET_OK;
}

// ---------------------------------------------------------------------------
// Bowyer-Watson insertion of the points the insertion rounds left out
// ---------------------------------------------------------------------------
wtet_status insert_points(const float* xyzn, uint32_t n, MeshBuf& mesh, std::vector<uint32_t> ids,
                          int mode, RepairReport* rep) {
  if (mesh.present) {
    if (rep) rep->message = "point insertion needs every row on the host";
    return WTET_ERR_INVALID_ARGUMENT;
  }
  PointSet S;
  S.xyzn = xyzn;
  S.n = n;
  const float sv[4][3] = {{-100.f, -100.f, -100.f}, {400.f, -100.f, -100.f},
                          {-100.f, 400.f, -100.f}, {-100.f, -100.f, 400.f}};
  std::memcpy(S.sup, sv, sizeof(sv));
  mode = mode == 1 ? PRED_EXACT_ONLY : PRED_FILTERED;
  Mesh32 W{mesh, n + 4};
  std::vector<uint32_t>& changed = mesh.changed;
  // Morton order: each walk starts from the previous point's new tets
  std::vector<std::pair<uint32_t, uint32_t>> order;
  order.reserve(ids.size());
  for (uint32_t q : ids) {
    if (q >= n) return WTET_ERR_INVALID_ARGUMENT;
    order.emplace_back(morton10(S.p(q)), q);
  }
  std::sort(order.begin(), order.end());
  uint32_t last = kInvalid;
  for (uint32_t t = 0; t < W.count() && last == kInvalid; ++t)
    if (!W.dead(t) && !W.has_inf(t)) last = t;
  if (last == kInvalid) {
    if (rep) rep->message = "no finite tet to start from";
    return WTET_ERR_INTERNAL;
  }
  auto conflict = [&](uint32_t t, uint32_t q) {
    return !W.dead(t) && !W.has_inf(t) && o4(S, W.tv(t), q, mode) < 0;
  };
  std::vector<uint32_t> cav, stack;
  std::unordered_set<uint32_t> incav;
  struct BF {
    uint32_t f[3];
    uint32_t out;
  };
  std::vector<BF> bf;
  std::unordered_map<uint64_t, uint32_t> edges;  // edge -> (slot << 2) | face
  std::vector<uint32_t> holes;
  uint32_t rng = 0x9E3779B9u;
  for (auto& pr : order) {
    const uint32_t q = pr.second;
    // visibility walk through finite tets: q lies inside the super tet, so
    // a face it lies beyond is never a face of the super tet
    uint32_t t = last;
    bool found = false;
    for (size_t step = 0; step < 4 * W.count() + 64 && !found; ++step) {
      rng = rng * 1664525u + 1013904223u;
      int off = (int)(rng >> 30);
      bool moved = false;
      for (int k = 0; k < 4 && !moved; ++k) {
        int f = (k + off) & 3;
        const uint32_t* tv = W.tv(t);
        if (o3sos(S, tv[face_vi(f, 0)], tv[face_vi(f, 1)], tv[face_vi(f, 2)], q, mode) < 0) {
          uint32_t e = W.ta(t, f);
          if (e == kInvalid || W.has_inf(e >> 5)) {
            if (rep) rep->message = "point walk left the super tetrahedron";
            return WTET_ERR_INTERNAL;
          }
          t = e >> 5;
          moved = true;
        }
      }
      found = !moved;
    }
    if (!found || !conflict(t, q)) {
      t = kInvalid;
      for (uint32_t u = 0; u < W.count(); ++u)
        if (conflict(u, q)) {
          t = u;
          break;
        }
      // no tet in conflict: an exact duplicate, redundant under the lifted
      // SoS (the reference's Bowyer-Watson skips it, src/phase2.rs:440-443)
      if (t == kInvalid) continue;
    }
    // cavity: the connected tets whose circumsphere holds q
    cav.assign(1, t);
    stack.assign(1, t);
    incav.clear();
    incav.insert(t);
    std::unordered_set<uint32_t> outside;
    while (!stack.empty()) {
      uint32_t u = stack.back();
      stack.pop_back();
      for (int f = 0; f < 4; ++f) {
        uint32_t e = W.ta(u, f);
        if (e == kInvalid) continue;
        uint32_t w = e >> 5;
        if (incav.count(w) || outside.count(w)) continue;
        if (conflict(w, q)) {
          incav.insert(w);
          cav.push_back(w);
          stack.push_back(w);
        } else {
          outside.insert(w);
        }
      }
    }
    bf.clear();
    for (uint32_t u : cav)
      for (int f = 0; f < 4; ++f) {
        uint32_t e = W.ta(u, f);
        if (e != kInvalid && incav.count(e >> 5)) continue;
        BF b;
        const uint32_t* tv = W.tv(u);
        for (int k = 0; k < 3; ++k) b.f[k] = tv[face_vi(f, k)];
        b.out = e;
        bf.push_back(b);
      }
    // the new tets take the cavity's slots, then appended ones; in 3D a
    // cavity may have fewer boundary faces than tets, and its leftover
    // slots become holes
    if (bf.size() > cav.size() && mesh.count + (bf.size() - cav.size()) > mesh.cap) {
      if (rep) rep->message = "point insertion needs more room than the staging buffer has";
      return WTET_ERR_OUT_OF_MEMORY;
    }
    edges.clear();
    for (size_t i = 0; i < bf.size(); ++i) {
      uint32_t slot = i < cav.size() ? cav[i] : (uint32_t)mesh.count++;
      const BF& b = bf[i];
      uint32_t x = b.f[0], y = b.f[1], z = b.f[2];
      if (o3sos(S, x, y, z, q, mode) <= 0) std::swap(x, y);
      uint32_t* tv = W.tv(slot);
      tv[0] = x;
      tv[1] = y;
      tv[2] = z;
      tv[3] = q;
      W.ta(slot, 3) = b.out;
      if (b.out != kInvalid) {
        W.ta(b.out >> 5, (int)(b.out & 3u)) = (slot << 5) | 3u;
        changed.push_back(b.out >> 5);
      }
      for (int f = 0; f < 3; ++f) {
        uint32_t e0 = tv[(f + 1) % 3], e1 = tv[(f + 2) % 3];
        uint64_t key = e0 < e1 ? ((uint64_t)e0 << 32 | e1) : ((uint64_t)e1 << 32 | e0);
        auto it = edges.find(key);
        if (it == edges.end()) {
          edges.emplace(key, (slot << 2) | (uint32_t)f);
        } else {
          uint32_t o = it->second;
          W.ta(slot, f) = ((o >> 2) << 5) | (o & 3u);
          W.ta(o >> 2, (int)(o & 3u)) = (slot << 5) | (uint32_t)f;
          edges.erase(it);
        }
      }
      changed.push_back(slot);
      last = slot;
    }
    for (size_t k = bf.size(); k < cav.size(); ++k) {
      W.tv(cav[k])[0] = kInvalid;
      holes.push_back(cav[k]);
    }
    if (!edges.empty()) {
      if (rep) rep->message = "point cavity boundary is not closed";
      return WTET_ERR_INTERNAL;
    }
    if (rep) rep->inserted++;
  }
  fill_holes(W, holes, changed);
  std::sort(changed.begin(), changed.end());
  changed.erase(std::unique(changed.begin(), changed.end()), changed.end());
  while (!changed.empty() && changed.back() >= mesh.count) changed.pop_back();
  return WTET_OK;
}

// input ids 0..n-1 that no tet of rows [0, count) uses
std::vector<uint32_t> missing_points(const uint32_t* tets, size_t count, uint32_t n) {
  std::vector<uint8_t> seen(n, 0);
  for (size_t i = 0; i < 4 * count; ++i)
    if (tets[i] < n) seen[tets[i]] = 1;
  std::vector<uint32_t> out;
  for (uint32_t v = 0; v < n; ++v)
    if (!seen[v]) out.push_back(v);
  return out;
}

// ---------------------------------------------------------------------------
// check_quality (src/python.rs:282-339 -> src/lib.rs:25-78)
// ---------------------------------------------------------------------------
extern "C" wtet_status wtet_check_quality(const float* xyz, uint32_t n, const uint32_t* tets,
                                          const uint32_t* adjacency, uint32_t num_tets,
                                          uint64_t* violations, uint64_t* interior_tets) {
  if (!xyz || !tets || !adjacency || !violations || !interior_tets) return WTET_ERR_INVALID_ARGUMENT;
  if (n < 1) return WTET_ERR_INVALID_ARGUMENT;
  // normalize exactly as the triangulation does (src/lib.rs:313-337)
  std::vector<float> pn((size_t)n * 3);
  float mn[3] = {3.40282347e+38f, 3.40282347e+38f, 3.40282347e+38f};
  float mx[3] = {-3.40282347e+38f, -3.40282347e+38f, -3.40282347e+38f};
  for (uint32_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      float v = xyz[3 * (size_t)i + k];
      if (v < mn[k]) mn[k] = v;
      if (v > mx[k]) mx[k] = v;
    }
  float s = mx[0] - mn[0];
  if (mx[1] - mn[1] > s) s = mx[1] - mn[1];
  if (mx[2] - mn[2] > s) s = mx[2] - mn[2];
  if (!(s >= 1e-10f)) s = 1e-10f;
  for (uint32_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) pn[3 * (size_t)i + k] = (xyz[3 * (size_t)i + k] - mn[k]) / s;
  PointSet S;
  S.xyzn = pn.data();
  S.n = n;
  const float sv[4][3] = {{-100.f, -100.f, -100.f}, {400.f, -100.f, -100.f},
                          {-100.f, 400.f, -100.f}, {-100.f, -100.f, 400.f}};
  std::memcpy(S.sup, sv, sizeof(sv));
  uint64_t viol = 0, interior = 0;
  for (uint32_t t = 0; t < num_tets; ++t) {
    const uint32_t* tv = tets + 4 * (size_t)t;
    if (tv[0] >= n || tv[1] >= n || tv[2] >= n || tv[3] >= n) continue;
    ++interior;
    for (int f = 0; f < 4; ++f) {
      uint32_t e = adjacency[4 * (size_t)t + f];
      if (e == kInvalid) continue;
      uint32_t u = e >> 5, g = e & 3u;
      if (t >= u) continue;
      if (u >= num_tets) return WTET_ERR_INVALID_ARGUMENT;
      const uint32_t* uv = tets + 4 * (size_t)u;
      if (uv[0] >= n || uv[1] >= n || uv[2] >= n || uv[3] >= n) continue;
      if (o4(S, tv, uv[g], PRED_FILTERED) < 0) ++viol;
    }
  }
  *violations = viol;
  *interior_tets = interior;
  return WTET_OK;
}

// ---------------------------------------------------------------------------
// phase 2 alone
// ---------------------------------------------------------------------------
extern "C" wtet_status wtet_repair(const float* xyzn, uint32_t n, uint32_t** tets,
                                   uint32_t** adjacency, uint32_t* num_tets, int mode,
                                   uint32_t* regions) {
  if (!xyzn || !tets || !*tets || !adjacency || !num_tets || n < 1) return WTET_ERR_INVALID_ARGUMENT;
  try {
    size_t M = *num_tets;
    std::vector<uint32_t> tv(*tets, *tets + 4 * M), ta(4 * M, kInvalid);
    for (size_t i = 0; i < 4 * M; ++i)
      if (tv[i] > n + 4) return WTET_ERR_INVALID_ARGUMENT;
    if (*adjacency) {
      ta.assign(*adjacency, *adjacency + 4 * M);
    } else {
      std::unordered_map<FaceKey, int64_t, FaceHash> seen;
      for (size_t t = 0; t < M; ++t)
        for (int f = 0; f < 4; ++f) {
          uint32_t a[3];
          int j = 0;
          for (int v = 0; v < 4; ++v)
            if (v != f) a[j++] = tv[4 * t + v];
          FaceKey k = face_key(a[0], a[1], a[2]);
          auto it = seen.find(k);
          if (it == seen.end()) {
            seen.emplace(k, (int64_t)(4 * t + f));
          } else {
            int64_t o = it->second;
            ta[4 * t + f] = ((uint32_t)(o >> 2) << 5) | (uint32_t)(o & 3);
            ta[o] = ((uint32_t)t << 5) | (uint32_t)f;
            seen.erase(it);
          }
        }
    }
    // every face that fails the insphere test seeds the repair
    PointSet S;
    S.xyzn = xyzn;
    S.n = n;
    const float sv[4][3] = {{-100.f, -100.f, -100.f}, {400.f, -100.f, -100.f},
                            {-100.f, 400.f, -100.f}, {-100.f, -100.f, 400.f}};
    std::memcpy(S.sup, sv, sizeof(sv));
    int pm = mode == 1 ? PRED_EXACT_ONLY : PRED_FILTERED;
    std::vector<uint32_t> viol;
    for (size_t t = 0; t < M; ++t) {
      const uint32_t* v = &tv[4 * t];
      bool inf = v[0] == n + 4 || v[1] == n + 4 || v[2] == n + 4 || v[3] == n + 4;
      if (inf) continue;
      for (int f = 0; f < 4; ++f) {
        uint32_t e = ta[4 * t + f];
        if (e == kInvalid) continue;
        uint32_t u = e >> 5, g = e & 3u;
        if (u <= t || u >= M) continue;
        const uint32_t* w = &tv[4 * (size_t)u];
        if (w[0] == n + 4 || w[1] == n + 4 || w[2] == n + 4 || w[3] == n + 4) continue;
        if (o4(S, v, w[g], pm) < 0) viol.push_back(((uint32_t)t << 2) | (uint32_t)f);
      }
    }
    // input ids no tet uses are inserted after the repair (phase 2
    // rebuilds with every point, src/phase2.rs:383-606)
    std::vector<uint32_t> missing = missing_points(tv.data(), M, n);
    size_t room = 65536 + 8 * missing.size();
    RepairReport rr;
    wtet_status s = WTET_OK;
    const std::vector<uint32_t> tv0 = tv, ta0 = ta;
    MeshBuf mb{nullptr, nullptr, 0, 0, {}};
    // out of room (the Delaunay mesh may hold far more tets than the input):
    // start over from the input with twice the room
    for (int attempt = 0;; ++attempt) {
      tv = tv0;
      ta = ta0;
      tv.resize(tv.size() + 4 * room);
      ta.resize(ta.size() + 4 * room);
      mb = MeshBuf{tv.data(), ta.data(), M, M + room, {}};
      rr = RepairReport{};
      s = WTET_OK;
      if (!viol.empty()) s = repair_delaunay(xyzn, n, mb, viol, mode, &rr);
      if (s == WTET_OK && !missing.empty()) s = insert_points(xyzn, n, mb, missing, mode, &rr);
      if (s == WTET_ERR_OUT_OF_MEMORY && attempt < 6) {
        room = 2 * (M + room);
        continue;
      }
      break;
    }
    if (s != WTET_OK) return s;
    tv.resize(4 * mb.count);
    ta.resize(4 * mb.count);
    uint32_t* ot = (uint32_t*)std::malloc(std::max<size_t>(tv.size(), 1) * 4);
    uint32_t* oa = (uint32_t*)std::malloc(std::max<size_t>(ta.size(), 1) * 4);
    if (!ot || !oa) {
      std::free(ot);
      std::free(oa);
      return WTET_ERR_OUT_OF_MEMORY;
    }
    std::memcpy(ot, tv.data(), tv.size() * 4);
    std::memcpy(oa, ta.data(), ta.size() * 4);
    *tets = ot;
    *adjacency = oa;
    *num_tets = (uint32_t)(tv.size() / 4);
    if (regions) *regions = rr.regions;
    return WTET_OK;
  } catch (...) {
    return WTET_ERR_OUT_OF_MEMORY;
  }
}

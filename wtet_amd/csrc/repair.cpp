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
#include <cstring>
#include <deque>
#include <unordered_map>
#include <unordered_set>

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

bool local_delaunay(const PointSet& S, std::vector<uint32_t> ids, int mode, Local& L,
                    std::string& err) {
  uint32_t n = S.n;
  L = Local();
  int64_t t0 = L.make();
  uint32_t s[4] = {n, n + 1, n + 2, n + 3};
  if (o3sos(S, s[0], s[1], s[2], s[3], mode) <= 0) std::swap(s[0], s[1]);
  std::memcpy(&L.v[0], s, 16);
  L.alive[t0] = 1;
  // random insertion order (deterministic): a new point then mostly lies
  // inside the hull of those already inserted, so its cavity stays small
  std::sort(ids.begin(), ids.end());
  uint64_t sh = 0x9E3779B97F4A7C15ull ^ ids.size();
  for (size_t i = ids.size(); i > 1; --i) {
    sh = sh * 6364136223846793005ull + 1442695040888963407ull;
    std::swap(ids[i - 1], ids[(size_t)((sh >> 33) % i)]);
  }
  int64_t last = t0;
  uint32_t rng = 0x2545F491u;
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
    int64_t stamp = (int64_t)it;
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
    if (m.present && t < m.known && !m.present[t]) throw NotPresent{t};
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

  std::deque<std::pair<uint32_t, int>> queue;
  for (uint32_t r : viol) queue.emplace_back(r >> 2, (int)(r & 3u));
  std::vector<uint32_t> holes;
  std::unordered_set<uint32_t> inR;
  std::vector<uint32_t> R, ring;

  wtet_status fail = WTET_OK;
  // one queued face: every row it reads is read before the first write, so
  // a missing row can be fetched and the face processed again
  auto process = [&](uint32_t t, int f) -> void {
    if (t >= W.count() || W.dead(t) || !face_bad(S, W, t, f, mode)) return;
    R.clear();
    inR.clear();
    auto add = [&](uint32_t x) {
      if (x >= W.count() || W.dead(x) || W.has_inf(x) || inR.count(x)) return;
      inR.insert(x);
      R.push_back(x);
    };
    add(t);
    add(W.ta(t, f) >> 5);
    bool done = false;
    while (!done) {
      ring.clear();
      for (uint32_t x : R)
        for (int h = 0; h < 4; ++h)
          if (W.ta(x, h) != kInvalid) ring.push_back(W.ta(x, h) >> 5);
      for (uint32_t x : ring) add(x);
      if (region_cap && R.size() > region_cap) {
        if (rep) {
          rep->too_large = true;
          rep->message = "a repair region outgrew its cap";
        }
        fail = WTET_ERR_NOT_CONVERGED;
        return;
      }
      if (R.size() * 2 > W.count())
        for (uint32_t x = 0; x < W.count(); ++x) add(x);
      std::vector<uint32_t> real;
      for (uint32_t x : R)
        for (int k = 0; k < 4; ++k)
          if (W.tv(x)[k] < n) real.push_back(W.tv(x)[k]);
      std::sort(real.begin(), real.end());
      real.erase(std::unique(real.begin(), real.end()), real.end());
      Local L;
      std::string err;
      if (!local_delaunay(S, real, mode, L, err)) {
        if (rep) rep->message = err;
        fail = WTET_ERR_INTERNAL;
        return;
      }
      std::unordered_map<FaceKey, std::array<int64_t, 2>, FaceHash> fmap;
      fmap.reserve(4 * L.alive.size());
      for (int64_t d = 0; d < (int64_t)L.alive.size(); ++d) {
        if (!L.alive[d]) continue;
        const uint32_t* dv = &L.v[4 * d];
        for (int h = 0; h < 4; ++h) {
          FaceKey k = face_key(dv[face_vi(h, 0)], dv[face_vi(h, 1)], dv[face_vi(h, 2)]);
          auto it = fmap.find(k);
          if (it == fmap.end()) fmap.emplace(k, std::array<int64_t, 2>{4 * d + h, -1});
          else it->second[1] = 4 * d + h;
        }
      }
      // boundary faces of R -> outside neighbour (t << 5 | f, or kInvalid)
      std::unordered_map<FaceKey, uint32_t, FaceHash> bface;
      std::vector<int64_t> seeds;
      std::vector<uint32_t> grow_with;
      bool missing = false;
      for (uint32_t x : R)
        for (int h = 0; h < 4; ++h) {
          uint32_t nb = W.ta(x, h);
          if (nb != kInvalid && inR.count(nb >> 5)) continue;
          if (nb != kInvalid) (void)W.dead(nb >> 5);  // written below: must be present
          const uint32_t* xv = W.tv(x);
          uint32_t fa = xv[face_vi(h, 0)], fb = xv[face_vi(h, 1)], fc = xv[face_vi(h, 2)];
          FaceKey k = face_key(fa, fb, fc);
          auto it = fmap.find(k);
          if (it == fmap.end()) {
            missing = true;
            if (nb != kInvalid) grow_with.push_back(nb >> 5);
            continue;
          }
          bface.emplace(k, nb);
          for (int s2 = 0; s2 < 2; ++s2) {
            int64_t df = it->second[s2];
            if (df < 0) continue;
            uint32_t apex = L.v[4 * (df >> 2) + (df & 3)];
            if (o3sos(S, fa, fb, fc, apex, mode) > 0) seeds.push_back(df >> 2);
          }
        }
      if (missing) {
        for (uint32_t x : grow_with) add(x);
        continue;
      }
      // inner tets of D
      std::unordered_map<int64_t, uint32_t> inner;
      std::vector<int64_t> order;
      std::vector<int64_t> st = seeds;
      for (int64_t d : seeds)
        if (inner.emplace(d, kInvalid).second) order.push_back(d);
      while (!st.empty()) {
        int64_t d = st.back();
        st.pop_back();
        const uint32_t* dv = &L.v[4 * d];
        for (int h = 0; h < 4; ++h) {
          FaceKey k = face_key(dv[face_vi(h, 0)], dv[face_vi(h, 1)], dv[face_vi(h, 2)]);
          if (bface.count(k)) continue;
          int64_t nb = L.a[4 * d + h];
          if (nb < 0) {
            if (rep) rep->message = "region flood reached the hull";
            fail = WTET_ERR_INTERNAL;
            return;
          }
          if (inner.emplace(nb >> 2, kInvalid).second) {
            st.push_back(nb >> 2);
            order.push_back(nb >> 2);
          }
        }
      }
      // slots: R's slots first, then appended ones; leftovers become holes
      size_t k = 0;
      for (int64_t d : order) {
        uint32_t slot;
        if (k < R.size()) {
          slot = R[k++];
        } else {
          if (mesh.count >= mesh.cap) {
            if (rep) rep->message = "repair needs more room than the staging buffer has";
            fail = WTET_ERR_OUT_OF_MEMORY;
            return;
          }
          slot = (uint32_t)mesh.count++;
          for (int h = 0; h < 4; ++h) W.ta(slot, h) = kInvalid;
        }
        inner[d] = slot;
        changed.push_back(slot);
      }
      for (; k < R.size(); ++k) {
        W.tv(R[k])[0] = kInvalid;
        holes.push_back(R[k]);
      }
      for (int64_t d : order) {
        uint32_t slot = inner[d];
        std::memcpy(W.tv(slot), &L.v[4 * d], 16);
      }
      for (int64_t d : order) {
        uint32_t slot = inner[d];
        const uint32_t* dv = &L.v[4 * d];
        for (int h = 0; h < 4; ++h) {
          FaceKey fk = face_key(dv[face_vi(h, 0)], dv[face_vi(h, 1)], dv[face_vi(h, 2)]);
          auto bf = bface.find(fk);
          if (bf != bface.end()) {
            uint32_t out = bf->second;
            W.ta(slot, h) = out;
            if (out != kInvalid) {
              W.ta(out >> 5, (int)(out & 3u)) = (slot << 5) | (uint32_t)h;
              changed.push_back(out >> 5);
            }
            queue.emplace_back(slot, h);
          } else {
            int64_t nb = L.a[4 * d + h];
            W.ta(slot, h) = (inner.at(nb >> 2) << 5) | (uint32_t)(nb & 3);
          }
        }
      }
      if (rep) {
        rep->regions++;
        rep->max_region = std::max<uint32_t>(rep->max_region, (uint32_t)R.size());
      }
      done = true;
    }
  };
  while (!queue.empty()) {
    auto [t, f] = queue.front();
    for (;;) {
      try {
        process(t, f);
        break;
      } catch (const NotPresent& np) {
        if (!mesh.fetch || !mesh.fetch(np.row)) throw;
        if (rep) rep->fetches++;
      }
    }
    if (fail != WTET_OK) return fail;
    queue.pop_front();
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

/*
 * Oracle triangulation -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Serial Bowyer-Watson over the normalized points plus the reference's four
 * finite super vertices, with the reference's SoS predicates and point ids
 * as perturbation ranks.  Follows the semantics of full_bw_reconstruction
 * (src/phase2.rs:383-606): start from the super tetrahedron oriented so that
 * orient3d_sos = +1, carve every tet whose circumsphere holds the new point
 * (orient4d < 0, BFS from a seed), cone the cavity boundary to the point and
 * orient every new tet with orient3d_sos.  The reference seeds the cavity by
 * a linear scan; here a visibility walk finds the seed (the linear scan is
 * kept as the fallback), and points go in Morton order.  Neither choice can
 * change the result: the SoS Delaunay triangulation is unique.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* face f of a tet (opposite vertex f), ordered so that
 * orient3d(face..., v_f) < 0 for a tet with raw orient3d < 0 */
static const int kFace[4][3] = {{1, 3, 2}, {0, 2, 3}, {0, 3, 1}, {0, 1, 2}};

void o_free(void *p) { free(p); }

/* ------------------------------------------------------------------------ */
/* point-set helpers                                                         */
/* ------------------------------------------------------------------------ */

/* src/lib.rs:313-337 */
void o_normalize_points(const float *xyz, int64_t n, float *out,
                        float *bbox_min3, float *scale_out) {
  float mn[3] = {3.40282347e+38f, 3.40282347e+38f, 3.40282347e+38f};
  float mx[3] = {-3.40282347e+38f, -3.40282347e+38f, -3.40282347e+38f};
  for (int64_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      float v = xyz[i * 3 + k];
      if (v < mn[k]) mn[k] = v; /* NaN never replaces (f32::min semantics) */
      if (v > mx[k]) mx[k] = v;
    }
  float ext[3];
  for (int k = 0; k < 3; ++k) ext[k] = mx[k] - mn[k];
  float s = ext[0];
  if (ext[1] > s) s = ext[1];
  if (ext[2] > s) s = ext[2];
  if (!(s >= 1e-10f)) s = 1e-10f;
  for (int64_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) out[i * 3 + k] = (xyz[i * 3 + k] - mn[k]) / s;
  if (bbox_min3) memcpy(bbox_min3, mn, sizeof(mn));
  if (scale_out) *scale_out = s;
}

/* src/morton.rs:73-81 */
uint32_t o_expand_bits(uint32_t v) {
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

static uint32_t quant10(float p, float mn, float mx) {
  float t = (p - mn) / (mx - mn);
  if (isnan(t)) return 0; /* Rust: NaN as u32 == 0 */
  if (t < 0.0f) t = 0.0f;
  if (t > 1.0f) t = 1.0f;
  return (uint32_t)(t * 1023.0f);
}

/* src/morton.rs:22-35 */
uint32_t o_morton_code(const float *p, const float *mn, const float *mx) {
  uint32_t x = quant10(p[0], mn[0], mx[0]);
  uint32_t y = quant10(p[1], mn[1], mx[1]);
  uint32_t z = quant10(p[2], mn[2], mx[2]);
  return (o_expand_bits(z) << 2) | (o_expand_bits(y) << 1) | o_expand_bits(x);
}

/* src/gpu/mod.rs:533-548 */
int64_t o_initial_stride(int64_t n) {
  double s = sqrt((double)n);
  if (n > 100000) s *= 2.0;
  if (s < 64.0) s = 64.0;
  if (s > 16384.0) s = 16384.0;
  double l = ceil(log2(s));
  return (int64_t)pow(2.0, l);
}

/* src/gpu/mod.rs:573-600 */
void o_stratified_order(int64_t n, int64_t stride, int64_t *order) {
  unsigned char *done = (unsigned char *)calloc((size_t)n + 1, 1);
  int64_t k = 0;
  for (int64_t s = stride; s >= 1; s /= 2)
    for (int64_t i = 0; i < n; i += s)
      if (!done[i]) {
        done[i] = 1;
        order[k++] = i;
      }
  free(done);
}

/* ------------------------------------------------------------------------ */
/* Bowyer-Watson                                                             */
/* ------------------------------------------------------------------------ */

typedef struct {
  uint32_t *v;   /* 4 per tet */
  int64_t *adj;  /* 4 per tet: neighbor*4 + face, or -1 */
  unsigned char *alive;
  int64_t *mark; /* cavity stamp */
  int64_t cap, count;
  int64_t *freel;
  int64_t nfree;
} Mesh;

static void mesh_reserve(Mesh *m, int64_t need) {
  if (need <= m->cap) return;
  int64_t cap = m->cap ? m->cap : 1024;
  while (cap < need) cap *= 2;
  m->v = (uint32_t *)realloc(m->v, sizeof(uint32_t) * 4 * cap);
  m->adj = (int64_t *)realloc(m->adj, sizeof(int64_t) * 4 * cap);
  m->alive = (unsigned char *)realloc(m->alive, cap);
  m->mark = (int64_t *)realloc(m->mark, sizeof(int64_t) * cap);
  m->freel = (int64_t *)realloc(m->freel, sizeof(int64_t) * cap);
  for (int64_t i = m->cap; i < cap; ++i) m->mark[i] = -1;
  m->cap = cap;
}

static int64_t mesh_new(Mesh *m) {
  if (m->nfree > 0) return m->freel[--m->nfree];
  mesh_reserve(m, m->count + 1);
  m->mark[m->count] = -1;
  return m->count++;
}

typedef struct {
  const double *P; /* (n+4) x 3 */
  int64_t n;
} Pts;

static inline const double *pt(const Pts *p, uint32_t i) { return p->P + 3 * (size_t)i; }

static int conflict(const Pts *P, const Mesh *m, int64_t t, uint32_t q) {
  const uint32_t *v = m->v + 4 * t;
  return o_orient4d(pt(P, v[0]), pt(P, v[1]), pt(P, v[2]), pt(P, v[3]),
                    pt(P, q), v[0], v[1], v[2], v[3], q) < 0;
}

/* 1 when q lies on the inner side of face f of t */
static int inside_face(const Pts *P, const Mesh *m, int64_t t, int f,
                       uint32_t q) {
  const uint32_t *v = m->v + 4 * t;
  uint32_t a = v[kFace[f][0]], b = v[kFace[f][1]], c = v[kFace[f][2]];
  return o_orient3d_sos(pt(P, a), pt(P, b), pt(P, c), pt(P, q), a, b, c, q) > 0;
}

static uint64_t rng_next(uint64_t *s) {
  *s = *s * 6364136223846793005ull + 1442695040888963407ull;
  return *s >> 33;
}

static int64_t locate(const Pts *P, const Mesh *m, int64_t start, uint32_t q,
                      uint64_t *rng, int64_t max_steps) {
  int64_t t = start;
  for (int64_t step = 0; step < max_steps; ++step) {
    int off = (int)(rng_next(rng) & 3);
    int moved = 0;
    for (int k = 0; k < 4; ++k) {
      int f = (k + off) & 3;
      if (!inside_face(P, m, t, f, q)) {
        int64_t nb = m->adj[4 * t + f];
        if (nb < 0) return -1; /* outside the hull: cannot happen */
        t = nb >> 2;
        moved = 1;
        break;
      }
    }
    if (!moved) return t;
  }
  return -1;
}

typedef struct {
  uint32_t a, b; /* sorted edge */
  int64_t tet;
  int face;
} EdgeRec;

static int cmp_edge(const void *x, const void *y) {
  const EdgeRec *p = (const EdgeRec *)x, *q = (const EdgeRec *)y;
  if (p->a != q->a) return p->a < q->a ? -1 : 1;
  if (p->b != q->b) return p->b < q->b ? -1 : 1;
  return 0;
}

typedef struct {
  uint32_t code;
  uint32_t id;
} KeyId;

/* Morton order; equal codes by descending id, so that of exact duplicates
 * the highest id goes in first and the others -- redundant under the lifted
 * SoS, which perturbs |p|^2 independently -- find no conflicting tet and are
 * skipped, as in the reference's Bowyer-Watson (src/phase2.rs:440-443). */
static int cmp_keyid(const void *x, const void *y) {
  const KeyId *p = (const KeyId *)x, *q = (const KeyId *)y;
  if (p->code != q->code) return p->code < q->code ? -1 : 1;
  return p->id > q->id ? -1 : (p->id < q->id);
}

int64_t o_delaunay_bw(const float *pts, int64_t n, uint32_t **tets_out,
                      int64_t *stats) {
  int64_t st[4] = {0, 0, 0, 0};
  double *P = (double *)malloc(sizeof(double) * 3 * (size_t)(n + 4));
  for (int64_t i = 0; i < 3 * n; ++i) P[i] = (double)pts[i];
  /* src/gpu/mod.rs:152-162 (big = 100) */
  const double sup[4][3] = {{-100, -100, -100}, {400, -100, -100},
                            {-100, 400, -100}, {-100, -100, 400}};
  memcpy(P + 3 * n, sup, sizeof(sup));
  Pts PP = {P, n};

  Mesh m;
  memset(&m, 0, sizeof(m));
  mesh_reserve(&m, 8 * n + 64);
  uint32_t s0 = (uint32_t)n, s1 = s0 + 1, s2 = s0 + 2, s3 = s0 + 3;
  int64_t t0 = mesh_new(&m);
  int o = o_orient3d_sos(pt(&PP, s0), pt(&PP, s1), pt(&PP, s2), pt(&PP, s3),
                         s0, s1, s2, s3);
  uint32_t init[4] = {s0, s1, s2, s3};
  if (o <= 0) { init[0] = s1; init[1] = s0; }
  memcpy(m.v, init, sizeof(init));
  for (int f = 0; f < 4; ++f) m.adj[f] = -1;
  m.alive[t0] = 1;

  /* insertion order: Morton code of the point, ties by id */
  KeyId *ord = (KeyId *)malloc(sizeof(KeyId) * (size_t)(n ? n : 1));
  {
    float mn[3] = {0, 0, 0}, mx[3] = {1, 1, 1};
    for (int64_t i = 0; i < n; ++i) {
      ord[i].code = o_morton_code(pts + 3 * i, mn, mx);
      ord[i].id = (uint32_t)i;
    }
    qsort(ord, (size_t)n, sizeof(KeyId), cmp_keyid);
  }

  int64_t *cav = NULL, *stack = NULL;
  int64_t cav_cap = 0;
  typedef struct { uint32_t f[3]; int64_t out; } BFace;
  BFace *bf = NULL;
  EdgeRec *er = NULL;
  int64_t *newt = NULL;
  int64_t last = t0;
  uint64_t rng = 0x9E3779B97F4A7C15ull;
  int ok = 1;

  for (int64_t it = 0; it < n && ok; ++it) {
    uint32_t q = ord[it].id;
    int64_t seed = locate(&PP, &m, last, q, &rng, 4 * (m.count + 16));
    if (seed < 0 || !conflict(&PP, &m, seed, q)) {
      /* reference behaviour: linear scan for any conflicting tet */
      st[0]++;
      if (seed >= 0) st[1]++;
      seed = -1;
      for (int64_t t = 0; t < m.count; ++t)
        if (m.alive[t] && conflict(&PP, &m, t, q)) { seed = t; break; }
      if (seed < 0) continue; /* point outside all spheres: skipped */
    }
    /* BFS cavity */
    int64_t nc = 0, ns = 0;
    if (cav_cap < m.count + 8) {
      cav_cap = 2 * (m.count + 8);
      cav = (int64_t *)realloc(cav, sizeof(int64_t) * cav_cap);
      stack = (int64_t *)realloc(stack, sizeof(int64_t) * cav_cap);
    }
    m.mark[seed] = it;
    cav[nc++] = seed;
    stack[ns++] = seed;
    while (ns > 0) {
      int64_t t = stack[--ns];
      for (int f = 0; f < 4; ++f) {
        int64_t nb = m.adj[4 * t + f];
        if (nb < 0) continue;
        int64_t u = nb >> 2;
        if (m.mark[u] == it || m.mark[u] == -2 - it) continue;
        if (conflict(&PP, &m, u, q)) {
          m.mark[u] = it;
          cav[nc++] = u;
          stack[ns++] = u;
        } else {
          m.mark[u] = -2 - it; /* tested, outside */
        }
      }
    }
    /* boundary faces */
    bf = (BFace *)realloc(bf, sizeof(BFace) * (size_t)(4 * nc));
    int64_t nb_faces = 0;
    for (int64_t i = 0; i < nc; ++i) {
      int64_t t = cav[i];
      for (int f = 0; f < 4; ++f) {
        int64_t nb = m.adj[4 * t + f];
        if (nb >= 0 && m.mark[nb >> 2] == it) continue;
        BFace *b = &bf[nb_faces++];
        for (int k = 0; k < 3; ++k) b->f[k] = m.v[4 * t + kFace[f][k]];
        b->out = nb;
      }
    }
    /* free cavity tets */
    for (int64_t i = 0; i < nc; ++i) {
      m.alive[cav[i]] = 0;
      m.freel[m.nfree++] = cav[i];
    }
    newt = (int64_t *)realloc(newt, sizeof(int64_t) * (size_t)nb_faces);
    er = (EdgeRec *)realloc(er, sizeof(EdgeRec) * (size_t)(3 * nb_faces));
    int64_t ne = 0;
    for (int64_t i = 0; i < nb_faces; ++i) {
      BFace *b = &bf[i];
      uint32_t a = b->f[0], bb = b->f[1], c = b->f[2];
      int s = o_orient3d_sos(pt(&PP, a), pt(&PP, bb), pt(&PP, c), pt(&PP, q),
                             a, bb, c, q);
      if (s <= 0) { st[1]++; uint32_t t = a; a = bb; bb = t; }
      int64_t t = mesh_new(&m);
      newt[i] = t;
      uint32_t *v = m.v + 4 * t;
      v[0] = a; v[1] = bb; v[2] = c; v[3] = q;
      m.alive[t] = 1;
      m.mark[t] = -1;
      m.adj[4 * t + 3] = b->out;
      if (b->out >= 0) m.adj[b->out] = 4 * t + 3;
      for (int f = 0; f < 3; ++f) {
        /* face f (contains q) is the triangle q + edge of {a,b,c} \ v[f] */
        uint32_t e0 = v[(f + 1) % 3], e1 = v[(f + 2) % 3];
        EdgeRec *r = &er[ne++];
        r->a = e0 < e1 ? e0 : e1;
        r->b = e0 < e1 ? e1 : e0;
        r->tet = t;
        r->face = f;
      }
    }
    qsort(er, (size_t)ne, sizeof(EdgeRec), cmp_edge);
    for (int64_t i = 0; i + 1 < ne; i += 2) {
      EdgeRec *x = &er[i], *y = &er[i + 1];
      if (x->a != y->a || x->b != y->b) { ok = 0; break; }
      m.adj[4 * x->tet + x->face] = 4 * y->tet + y->face;
      m.adj[4 * y->tet + y->face] = 4 * x->tet + x->face;
    }
    if (nb_faces > 0) last = newt[0];
  }

  int64_t cnt = 0;
  for (int64_t t = 0; t < m.count; ++t) cnt += m.alive[t];
  uint32_t *out = (uint32_t *)malloc(sizeof(uint32_t) * 4 * (size_t)(cnt ? cnt : 1));
  int64_t k = 0;
  for (int64_t t = 0; t < m.count; ++t)
    if (m.alive[t]) memcpy(out + 4 * (k++), m.v + 4 * t, 16);
  *tets_out = out;
  if (stats) memcpy(stats, st, sizeof(st));
  free(P); free(ord); free(cav); free(stack); free(bf); free(er); free(newt);
  free(m.v); free(m.adj); free(m.alive); free(m.mark); free(m.freel);
  return ok ? cnt : -1;
}

/* ------------------------------------------------------------------------ */
/* canonical form, adjacency, audit                                          */
/* ------------------------------------------------------------------------ */

static int cmp_quad(const void *x, const void *y) {
  const uint32_t *p = (const uint32_t *)x, *q = (const uint32_t *)y;
  for (int k = 0; k < 4; ++k)
    if (p[k] != q[k]) return p[k] < q[k] ? -1 : 1;
  return 0;
}

int64_t o_canonical_real(const uint32_t *tets, int64_t m, uint32_t n,
                         uint32_t *out) {
  int64_t k = 0;
  for (int64_t i = 0; i < m; ++i) {
    const uint32_t *t = tets + 4 * i;
    if (t[0] >= n || t[1] >= n || t[2] >= n || t[3] >= n) continue;
    uint32_t q[4] = {t[0], t[1], t[2], t[3]};
    for (int a = 1; a < 4; ++a)
      for (int b = a; b > 0 && q[b] < q[b - 1]; --b) {
        uint32_t s = q[b]; q[b] = q[b - 1]; q[b - 1] = s;
      }
    memcpy(out + 4 * k, q, 16);
    ++k;
  }
  qsort(out, (size_t)k, 16, cmp_quad);
  return k;
}

typedef struct {
  uint32_t k[3];
  uint32_t tet;
  uint32_t face;
} FaceRec;

static int cmp_face(const void *x, const void *y) {
  const FaceRec *p = (const FaceRec *)x, *q = (const FaceRec *)y;
  for (int i = 0; i < 3; ++i)
    if (p->k[i] != q->k[i]) return p->k[i] < q->k[i] ? -1 : 1;
  return 0;
}

/* src/phase2.rs:1460-1499 semantics: faces matched by vertex set */
void o_rebuild_adjacency(const uint32_t *tets, int64_t m, uint32_t *adj) {
  FaceRec *r = (FaceRec *)malloc(sizeof(FaceRec) * 4 * (size_t)(m ? m : 1));
  for (int64_t t = 0; t < m; ++t)
    for (int f = 0; f < 4; ++f) {
      FaceRec *x = &r[4 * t + f];
      uint32_t a[3], j = 0;
      for (int v = 0; v < 4; ++v)
        if (v != f) a[j++] = tets[4 * t + v];
      if (a[0] > a[1]) { uint32_t s = a[0]; a[0] = a[1]; a[1] = s; }
      if (a[1] > a[2]) { uint32_t s = a[1]; a[1] = a[2]; a[2] = s; }
      if (a[0] > a[1]) { uint32_t s = a[0]; a[0] = a[1]; a[1] = s; }
      memcpy(x->k, a, sizeof(a));
      x->tet = (uint32_t)t;
      x->face = (uint32_t)f;
      adj[4 * t + f] = 0xFFFFFFFFu;
    }
  qsort(r, 4 * (size_t)m, sizeof(FaceRec), cmp_face);
  for (int64_t i = 0; i + 1 < 4 * m; ++i) {
    if (cmp_face(&r[i], &r[i + 1]) == 0) {
      adj[4 * r[i].tet + r[i].face] = (r[i + 1].tet << 5) | r[i + 1].face;
      adj[4 * r[i + 1].tet + r[i + 1].face] = (r[i].tet << 5) | r[i].face;
      ++i;
    }
  }
  free(r);
}

/* src/lib.rs:25-78 (only_real=1); only_real=0 audits every finite face */
int64_t o_count_violations(const float *pts, int64_t n,
                           const uint32_t *tets, const uint32_t *adj,
                           int64_t m, int only_real) {
  double *P = (double *)malloc(sizeof(double) * 3 * (size_t)(n + 5));
  for (int64_t i = 0; i < 3 * n; ++i) P[i] = (double)pts[i];
  const double sup[5][3] = {{-100, -100, -100}, {400, -100, -100},
                            {-100, 400, -100}, {-100, -100, 400},
                            {1000, 1000, 1000}};
  memcpy(P + 3 * n, sup, sizeof(sup));
  uint32_t lim = only_real ? (uint32_t)n : (uint32_t)n + 4;
  int64_t viol = 0;
#pragma omp parallel for reduction(+ : viol) schedule(dynamic, 1024)
  for (int64_t t = 0; t < m; ++t) {
    const uint32_t *v = tets + 4 * t;
    if (v[0] >= lim || v[1] >= lim || v[2] >= lim || v[3] >= lim) continue;
    for (int f = 0; f < 4; ++f) {
      uint32_t a = adj[4 * t + f];
      if (a == 0xFFFFFFFFu) continue;
      int64_t u = a >> 5;
      uint32_t uf = a & 3;
      if (t >= u || u >= m) continue;
      const uint32_t *w = tets + 4 * u;
      if (w[0] >= lim || w[1] >= lim || w[2] >= lim || w[3] >= lim) continue;
      uint32_t vb = w[uf];
      int s = o_orient4d(P + 3 * v[0], P + 3 * v[1], P + 3 * v[2],
                         P + 3 * v[3], P + 3 * vb, v[0], v[1], v[2], v[3], vb);
      if (s < 0) ++viol;
    }
  }
  free(P);
  return viol;
}

/* ------------------------------------------------------------------------ */
/* full_bw_reconstruction as written (src/phase2.rs:383-606)                 */
/* ------------------------------------------------------------------------ */

static int cmp_i64(const void *x, const void *y) {
  int64_t a = *(const int64_t *)x, b = *(const int64_t *)y;
  return a < b ? -1 : (a > b);
}

static void sort3(uint32_t *f) {
  for (int a = 1; a < 3; ++a)
    for (int b = a; b > 0 && f[b] < f[b - 1]; --b) {
      uint32_t s = f[b]; f[b] = f[b - 1]; f[b - 1] = s;
    }
}

int64_t o_delaunay_bw_reference(const float *pts, int64_t n, uint32_t **tets_out) {
  const uint32_t INV = 0xFFFFFFFFu;
  /* face f = the vertices other than f, in ascending position
   * (src/phase2.rs:395-397) */
  static const int fi[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
  double *P = (double *)malloc(sizeof(double) * 3 * (size_t)(n + 4));
  for (int64_t i = 0; i < 3 * n; ++i) P[i] = (double)pts[i];
  const double sup[4][3] = {{-100, -100, -100}, {400, -100, -100},
                            {-100, 400, -100}, {-100, -100, 400}};
  memcpy(P + 3 * n, sup, sizeof(sup));
  Pts PP = {P, n};
  int64_t cap = 1024, len = 1;
  uint32_t *T = (uint32_t *)malloc(16 * (size_t)cap);  /* tets, 4 each */
  uint32_t *A = (uint32_t *)malloc(16 * (size_t)cap);  /* (t << 5) | f */
  int64_t *stamp = (int64_t *)malloc(8 * (size_t)cap);
  uint32_t s0 = (uint32_t)n, s1 = s0 + 1, s2 = s0 + 2, s3 = s0 + 3;
  int o = o_orient3d_sos(pt(&PP, s0), pt(&PP, s1), pt(&PP, s2), pt(&PP, s3), s0, s1, s2, s3);
  uint32_t init[4] = {s0, s1, s2, s3};
  if (!(o > 0)) { init[0] = s1; init[1] = s0; }
  memcpy(T, init, 16);
  for (int f = 0; f < 4; ++f) A[f] = INV;
  stamp[0] = -1;
  int64_t *cav = NULL, *queue = NULL;
  int64_t cav_cap = 0;
  typedef struct { uint32_t f[3]; int64_t nt; uint32_t nf; } BFace;
  BFace *bd = NULL;
  uint32_t *NT = NULL, *NA = NULL;
  int ok = 1;
  for (int64_t pi = 0; pi < n && ok; ++pi) {
    uint32_t v = (uint32_t)pi;
    /* step 1: the first live tet (in creation order) whose circumsphere holds v */
    int64_t seed = -1;
    for (int64_t ti = 0; ti < len; ++ti) {
      const uint32_t *t = T + 4 * ti;
      if (t[0] == INV || t[1] == INV || t[2] == INV || t[3] == INV) continue;
      if (o_orient4d(pt(&PP, t[0]), pt(&PP, t[1]), pt(&PP, t[2]), pt(&PP, t[3]), pt(&PP, v),
                     t[0], t[1], t[2], t[3], v) < 0) { seed = ti; break; }
    }
    if (seed < 0) continue; /* outside every circumsphere: skipped */
    /* step 2: BFS cavity */
    if (cav_cap < len + 1) {
      cav_cap = 2 * (len + 1);
      cav = (int64_t *)realloc(cav, 8 * (size_t)cav_cap);
      queue = (int64_t *)realloc(queue, 8 * (size_t)cav_cap);
    }
    int64_t nc = 0, nq = 0;
    cav[nc++] = seed;
    queue[nq++] = seed;
    stamp[seed] = pi;
    while (nq > 0) {
      int64_t ti = queue[--nq];
      for (int f = 0; f < 4; ++f) {
        uint32_t pk = A[4 * ti + f];
        if (pk == INV) continue;
        int64_t nb = pk >> 5;
        if (nb >= len || stamp[nb] == pi) continue;
        const uint32_t *t = T + 4 * nb;
        if (t[0] == INV || t[1] == INV || t[2] == INV || t[3] == INV) continue;
        if (o_orient4d(pt(&PP, t[0]), pt(&PP, t[1]), pt(&PP, t[2]), pt(&PP, t[3]), pt(&PP, v),
                       t[0], t[1], t[2], t[3], v) < 0) {
          stamp[nb] = pi;
          cav[nc++] = nb;
          queue[nq++] = nb;
        }
      }
    }
    /* step 3: boundary faces, cavity in ascending order (a BTreeSet) */
    qsort(cav, (size_t)nc, 8, cmp_i64);
    bd = (BFace *)realloc(bd, sizeof(BFace) * 4 * (size_t)nc);
    int64_t nb_faces = 0;
    for (int64_t i = 0; i < nc; ++i) {
      int64_t ti = cav[i];
      const uint32_t *t = T + 4 * ti;
      for (int f = 0; f < 4; ++f) {
        uint32_t pk = A[4 * ti + f];
        BFace b;
        for (int k = 0; k < 3; ++k) b.f[k] = t[fi[f][k]];
        if (pk == INV) {
          b.nt = -1;
          b.nf = 0;
        } else {
          if (stamp[pk >> 5] == pi) continue;
          b.nt = pk >> 5;
          b.nf = pk & 31u;
        }
        bd[nb_faces++] = b;
      }
    }
    /* step 4: new tets joining each boundary face to v */
    NT = (uint32_t *)realloc(NT, 16 * (size_t)nb_faces);
    NA = (uint32_t *)realloc(NA, 16 * (size_t)nb_faces);
    for (int64_t i = 0; i < nb_faces; ++i) {
      uint32_t a = bd[i].f[0], b = bd[i].f[1], c = bd[i].f[2];
      int s = o_orient3d_sos(pt(&PP, a), pt(&PP, b), pt(&PP, c), pt(&PP, v), a, b, c, v);
      uint32_t *q = NT + 4 * i;
      if (s > 0) { q[0] = a; q[1] = b; } else { q[0] = b; q[1] = a; }
      q[2] = c;
      q[3] = v;
      for (int k = 0; k < 4; ++k) NA[4 * i + k] = INV;
    }
    /* step 5: external links, then new tets linked pairwise */
    int64_t base = len;
    for (int64_t i = 0; i < nb_faces; ++i) {
      uint32_t *q = NT + 4 * i;
      int vp = 0;
      while (q[vp] != v) ++vp;
      if (bd[i].nt >= 0) {
        NA[4 * i + vp] = ((uint32_t)bd[i].nt << 5) | bd[i].nf;
        A[4 * bd[i].nt + bd[i].nf] = ((uint32_t)(base + i) << 5) | (uint32_t)vp;
      }
      for (int f = 0; f < 4; ++f) {
        if (f == vp || NA[4 * i + f] != INV) continue;
        uint32_t mf[3] = {q[fi[f][0]], q[fi[f][1]], q[fi[f][2]]};
        sort3(mf);
        for (int64_t j = i + 1; j < nb_faces; ++j) {
          const uint32_t *r = NT + 4 * j;
          for (int f2 = 0; f2 < 4; ++f2) {
            uint32_t of[3] = {r[fi[f2][0]], r[fi[f2][1]], r[fi[f2][2]]};
            sort3(of);
            if (mf[0] == of[0] && mf[1] == of[1] && mf[2] == of[2]) {
              NA[4 * i + f] = ((uint32_t)(base + j) << 5) | (uint32_t)f2;
              NA[4 * j + f2] = ((uint32_t)(base + i) << 5) | (uint32_t)f;
            }
          }
        }
      }
    }
    /* step 6: cavity tets die, new tets are appended */
    for (int64_t i = 0; i < nc; ++i)
      for (int k = 0; k < 4; ++k) {
        T[4 * cav[i] + k] = INV;
        A[4 * cav[i] + k] = INV;
      }
    if (len + nb_faces > cap) {
      while (len + nb_faces > cap) cap *= 2;
      T = (uint32_t *)realloc(T, 16 * (size_t)cap);
      A = (uint32_t *)realloc(A, 16 * (size_t)cap);
      stamp = (int64_t *)realloc(stamp, 8 * (size_t)cap);
    }
    memcpy(T + 4 * len, NT, 16 * (size_t)nb_faces);
    memcpy(A + 4 * len, NA, 16 * (size_t)nb_faces);
    for (int64_t i = 0; i < nb_faces; ++i) stamp[len + i] = -1;
    len += nb_faces;
  }
  int64_t cnt = 0;
  for (int64_t t = 0; t < len; ++t) cnt += T[4 * t] != INV;
  uint32_t *out = (uint32_t *)malloc(16 * (size_t)(cnt ? cnt : 1));
  int64_t k = 0;
  for (int64_t t = 0; t < len; ++t)
    if (T[4 * t] != INV) memcpy(out + 4 * (k++), T + 4 * t, 16);
  *tets_out = out;
  free(P); free(T); free(A); free(stamp); free(cav); free(queue); free(bd); free(NT); free(NA);
  return ok ? cnt : -1;
}

/* Many independent clouds (normalized points, cloud c = rows
 * [off[c], off[c+1])), each triangulated by o_delaunay_bw on its own, the
 * clouds spread over `threads` OpenMP threads.  counts[c] receives cloud
 * c's tet count (-1 on failure).  Returns the total, or -1.  The CPU
 * baseline of the batched workload: the reference triangulates one cloud
 * per delaunay_3d call (src/lib.rs:160-310). */
int64_t o_delaunay_bw_many(const float *pts, const uint64_t *off, int64_t C, int threads,
                           int64_t *counts) {
  int64_t total = 0, bad = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads) reduction(+ : total, bad)
  for (int64_t c = 0; c < C; ++c) {
    uint32_t *t = NULL;
    int64_t m = o_delaunay_bw(pts + 3 * off[c], (int64_t)(off[c + 1] - off[c]), &t, NULL);
    counts[c] = m;
    if (m < 0) bad++;
    else total += m;
    free(t);
  }
  return bad ? -1 : total;
}

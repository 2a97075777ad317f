/*
 * Oracle predicates -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Exact signs are computed with floating-point expansion arithmetic
 * (J. R. Shewchuk, "Adaptive Precision Floating-Point Arithmetic and Fast
 * Robust Geometric Predicates", DCG 18, 1997): Two-Sum, Dekker split
 * Two-Product, grow/scale expansions with zero elimination.  Determinants
 * are expanded by Laplace's rule directly on the raw coordinates, so every
 * value is exact; a conservative f64 filter answers first.
 *
 * Must be compiled with -ffp-contract=off (no fused multiply-add) on an
 * SSE2 target, so that every double operation rounds once to nearest.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* expansion arithmetic                                                      */
/* ------------------------------------------------------------------------ */

#define SPLITTER 134217729.0 /* 2^27 + 1 */

static inline void two_sum(double a, double b, double *x, double *y) {
  double s = a + b;
  double bv = s - a;
  double av = s - bv;
  double br = b - bv;
  double ar = a - av;
  *x = s;
  *y = ar + br;
}

static inline void fast_two_sum(double a, double b, double *x, double *y) {
  double s = a + b;
  double bv = s - a;
  *x = s;
  *y = b - bv;
}

static inline void split(double a, double *hi, double *lo) {
  double c = SPLITTER * a;
  double abig = c - a;
  *hi = c - abig;
  *lo = a - *hi;
}

static inline void two_prod(double a, double b, double *x, double *y) {
  double p = a * b;
  double ah, al, bh, bl;
  split(a, &ah, &al);
  split(b, &bh, &bl);
  double e1 = p - ah * bh;
  double e2 = e1 - al * bh;
  double e3 = e2 - ah * bl;
  *x = p;
  *y = al * bl - e3;
}

/* thread-local bump arena for expansion storage */
#define ARENA_DOUBLES (1u << 20)
static _Thread_local double *g_arena = NULL;
static _Thread_local size_t g_top = 0;

static double *aalloc(size_t k) {
  if (!g_arena) {
    g_arena = (double *)malloc(sizeof(double) * ARENA_DOUBLES);
    if (!g_arena) abort();
  }
  if (g_top + k > ARENA_DOUBLES) abort(); /* sizes here are bounded */
  double *p = g_arena + g_top;
  g_top += k;
  return p;
}

typedef struct {
  int n;
  double *v; /* increasing magnitude, nonoverlapping, no zeros (n>=1) */
} X;

static X x_double(double a) {
  X r;
  r.v = aalloc(1);
  r.v[0] = a;
  r.n = 1;
  return r;
}

/* h = e + b */
static X x_grow(X e, double b) {
  X h;
  h.v = aalloc((size_t)e.n + 1);
  int k = 0;
  double q = b;
  for (int i = 0; i < e.n; ++i) {
    double s, t;
    two_sum(q, e.v[i], &s, &t);
    q = s;
    if (t != 0.0) h.v[k++] = t;
  }
  if (q != 0.0 || k == 0) h.v[k++] = q;
  h.n = k;
  return h;
}

static X x_add(X e, X f) {
  X h = e;
  for (int i = 0; i < f.n; ++i) h = x_grow(h, f.v[i]);
  return h;
}

static X x_neg(X e) {
  X h;
  h.v = aalloc((size_t)e.n);
  h.n = e.n;
  for (int i = 0; i < e.n; ++i) h.v[i] = -e.v[i];
  return h;
}

static X x_scale(X e, double b) {
  X h;
  h.v = aalloc(2 * (size_t)e.n);
  int k = 0;
  double q, t;
  two_prod(e.v[0], b, &q, &t);
  if (t != 0.0) h.v[k++] = t;
  for (int i = 1; i < e.n; ++i) {
    double p1, p0, s;
    two_prod(e.v[i], b, &p1, &p0);
    two_sum(q, p0, &s, &t);
    if (t != 0.0) h.v[k++] = t;
    fast_two_sum(p1, s, &q, &t);
    if (t != 0.0) h.v[k++] = t;
  }
  if (q != 0.0 || k == 0) h.v[k++] = q;
  h.n = k;
  return h;
}

static X x_mul(X e, X f) {
  X acc = x_scale(e, f.v[0]);
  for (int i = 1; i < f.n; ++i) acc = x_add(acc, x_scale(e, f.v[i]));
  return acc;
}

static int x_sign(X e) {
  double top = e.v[e.n - 1];
  return (top > 0.0) - (top < 0.0);
}

/* |p|^2 as an exact expansion */
static X x_lift(const double *p) {
  double a1, a0, b1, b0, c1, c0;
  two_prod(p[0], p[0], &a1, &a0);
  two_prod(p[1], p[1], &b1, &b0);
  two_prod(p[2], p[2], &c1, &c0);
  X h = x_double(a0);
  h = x_grow(h, a1);
  h = x_grow(h, b0);
  h = x_grow(h, b1);
  h = x_grow(h, c0);
  h = x_grow(h, c1);
  return h;
}

/*
 * Exact determinant of a k x k matrix (k <= 5) whose entries are expansions,
 * by Laplace expansion along successive columns with all row-subset minors
 * memoized.  m[r*k + c].
 */
static X x_det(const X *m, int k) {
  X minor[32];
  int have[32];
  memset(have, 0, sizeof(have));
  /* level 1: single rows, column 0 */
  for (int r = 0; r < k; ++r) {
    minor[1 << r] = m[r * k + 0];
    have[1 << r] = 1;
  }
  for (int level = 2; level <= k; ++level) {
    int col = level - 1;
    for (int mask = 0; mask < (1 << k); ++mask) {
      if (__builtin_popcount((unsigned)mask) != level) continue;
      X acc;
      int first = 1;
      int pos = 0;
      for (int r = 0; r < k; ++r) {
        if (!(mask & (1 << r))) continue;
        int sub = mask & ~(1 << r);
        X term = x_mul(m[r * k + col], minor[sub]);
        /* cofactor sign (-1)^(pos + col) with pos the row's rank in mask */
        if ((pos + col) & 1) term = x_neg(term);
        acc = first ? term : x_add(acc, term);
        first = 0;
        ++pos;
      }
      minor[mask] = acc;
      have[mask] = 1;
    }
  }
  return minor[(1 << k) - 1];
}

/* ------------------------------------------------------------------------ */
/* exact predicates                                                          */
/* ------------------------------------------------------------------------ */

static const double kEps = 1.1102230246251565e-16; /* 2^-53 */

static int exact_orient3d(const double *a, const double *b, const double *c,
                          const double *d) {
  g_top = 0;
  const double *p[4] = {a, b, c, d};
  X m[16];
  for (int r = 0; r < 4; ++r) {
    m[r * 4 + 0] = x_double(p[r][0]);
    m[r * 4 + 1] = x_double(p[r][1]);
    m[r * 4 + 2] = x_double(p[r][2]);
    m[r * 4 + 3] = x_double(1.0);
  }
  return x_sign(x_det(m, 4));
}

int o_orient3d_sign(const double *a, const double *b, const double *c,
                    const double *d) {
  double adx = a[0] - d[0], bdx = b[0] - d[0], cdx = c[0] - d[0];
  double ady = a[1] - d[1], bdy = b[1] - d[1], cdy = c[1] - d[1];
  double adz = a[2] - d[2], bdz = b[2] - d[2], cdz = c[2] - d[2];
  double bc = bdx * cdy - cdx * bdy;
  double ca = cdx * ady - adx * cdy;
  double ab = adx * bdy - bdx * ady;
  double det = adz * bc + bdz * ca + cdz * ab;
  double perm = (fabs(bdx * cdy) + fabs(cdx * bdy)) * fabs(adz) +
                (fabs(cdx * ady) + fabs(adx * cdy)) * fabs(bdz) +
                (fabs(adx * bdy) + fabs(bdx * ady)) * fabs(cdz);
  /* deliberately loose bound (the oracle favours certainty over speed) */
  double bound = 64.0 * kEps * perm;
  if (det > bound) return 1;
  if (-det > bound) return -1;
  return exact_orient3d(a, b, c, d);
}

static int exact_insphere(const double *a, const double *b, const double *c,
                          const double *d, const double *e) {
  g_top = 0;
  const double *p[5] = {a, b, c, d, e};
  X m[25];
  for (int r = 0; r < 5; ++r) {
    m[r * 5 + 0] = x_double(p[r][0]);
    m[r * 5 + 1] = x_double(p[r][1]);
    m[r * 5 + 2] = x_double(p[r][2]);
    m[r * 5 + 3] = x_lift(p[r]);
    m[r * 5 + 4] = x_double(1.0);
  }
  /* det[x y z w 1] over rows a..e equals Shewchuk's insphere */
  return x_sign(x_det(m, 5));
}

int o_insphere_sign(const double *a, const double *b, const double *c,
                    const double *d, const double *e) {
  double ax = a[0] - e[0], ay = a[1] - e[1], az = a[2] - e[2];
  double bx = b[0] - e[0], by = b[1] - e[1], bz = b[2] - e[2];
  double cx = c[0] - e[0], cy = c[1] - e[1], cz = c[2] - e[2];
  double dx = d[0] - e[0], dy = d[1] - e[1], dz = d[2] - e[2];
  double ab = ax * by - bx * ay, bc = bx * cy - cx * by;
  double cd = cx * dy - dx * cy, da = dx * ay - ax * dy;
  double ac = ax * cy - cx * ay, bd = bx * dy - dx * by;
  double abc = az * bc - bz * ac + cz * ab;
  double bcd = bz * cd - cz * bd + dz * bc;
  double cda = cz * da + dz * ac + az * cd;
  double dab = dz * ab + az * bd + bz * da;
  double al = ax * ax + ay * ay + az * az;
  double bl = bx * bx + by * by + bz * bz;
  double cl = cx * cx + cy * cy + cz * cz;
  double dl = dx * dx + dy * dy + dz * dz;
  double det = (dl * abc - cl * dab) + (bl * cda - al * bcd);
  /* permanent: every product magnitude that enters det */
  double pab = fabs(ax * by) + fabs(bx * ay), pbc = fabs(bx * cy) + fabs(cx * by);
  double pcd = fabs(cx * dy) + fabs(dx * cy), pda = fabs(dx * ay) + fabs(ax * dy);
  double pac = fabs(ax * cy) + fabs(cx * ay), pbd = fabs(bx * dy) + fabs(dx * by);
  double pabc = fabs(az) * pbc + fabs(bz) * pac + fabs(cz) * pab;
  double pbcd = fabs(bz) * pcd + fabs(cz) * pbd + fabs(dz) * pbc;
  double pcda = fabs(cz) * pda + fabs(dz) * pac + fabs(az) * pcd;
  double pdab = fabs(dz) * pab + fabs(az) * pbd + fabs(bz) * pda;
  double perm = dl * pabc + cl * pdab + bl * pcda + al * pbcd;
  double bound = 128.0 * kEps * perm;
  if (det > bound) return 1;
  if (-det > bound) return -1;
  return exact_insphere(a, b, c, d, e);
}

int o_orient3d_lifted_sign(const double *a, const double *b, const double *c,
                           const double *d) {
  g_top = 0;
  const double *p[4] = {a, b, c, d};
  X m[16];
  for (int r = 0; r < 4; ++r) {
    m[r * 4 + 0] = x_double(p[r][0]);
    m[r * 4 + 1] = x_double(p[r][1]);
    m[r * 4 + 2] = x_lift(p[r]);
    m[r * 4 + 3] = x_double(1.0);
  }
  return x_sign(x_det(m, 4));
}

int o_orient2d_lifted_sign(const double *a, const double *b, const double *c) {
  g_top = 0;
  const double *p[3] = {a, b, c};
  X m[9];
  for (int r = 0; r < 3; ++r) {
    m[r * 3 + 0] = x_double(p[r][0]);
    m[r * 3 + 1] = x_lift(p[r]);
    m[r * 3 + 2] = x_double(1.0);
  }
  return x_sign(x_det(m, 3));
}

int o_orient1d_lifted_sign(const double *a, const double *b) {
  g_top = 0;
  X d = x_add(x_lift(a), x_neg(x_lift(b)));
  return x_sign(d);
}

double o_orient2d_plain(const double *a, const double *b, const double *c) {
  volatile double t1 = a[0] - c[0];
  volatile double t2 = b[1] - c[1];
  volatile double t3 = a[1] - c[1];
  volatile double t4 = b[0] - c[0];
  volatile double p1 = t1 * t2;
  volatile double p2 = t3 * t4;
  return p1 - p2;
}

/* ------------------------------------------------------------------------ */
/* SoS cascades (reference semantics)                                        */
/* ------------------------------------------------------------------------ */

static inline int sgn(double v) { return (v > 0.0) - (v < 0.0); }

/* orient3d with 14-level SoS, src/predicates.rs:106-208 */
int o_orient3d_sos_depth(const double *a, const double *b, const double *c,
                         const double *d, uint32_t ia, uint32_t ib,
                         uint32_t ic, uint32_t id, int *depth_out) {
  int s = o_orient3d_sign(a, b, c, d);
  if (s != 0) {
    if (depth_out) *depth_out = -1;
    return s < 0 ? 1 : -1;
  }
  uint32_t v[4] = {ia, ib, ic, id};
  const double *p[4] = {a, b, c, d};
  int parity = 1;
  /* sort ascending by index; parity of the permutation */
  for (int i = 0; i < 3; ++i) {
    int mi = i;
    for (int j = i + 1; j < 4; ++j)
      if (v[j] < v[mi]) mi = j;
    if (mi != i) {
      uint32_t tv = v[i]; v[i] = v[mi]; v[mi] = tv;
      const double *tp = p[i]; p[i] = p[mi]; p[mi] = tp;
      parity = -parity;
    }
  }
  int depth = 0, r = 0;
  for (; depth < 14; ++depth) {
    double q0[2], q1[2], q2[2];
    switch (depth) {
#define O2(P, Q, R, U, W)                              \
  q0[0] = p[P][U]; q0[1] = p[P][W];                    \
  q1[0] = p[Q][U]; q1[1] = p[Q][W];                    \
  q2[0] = p[R][U]; q2[1] = p[R][W];                    \
  r = sgn(o_orient2d_plain(q0, q1, q2));               \
  break;
      case 0: O2(1, 2, 3, 0, 1)
      case 1: O2(1, 2, 3, 0, 2)
      case 2: O2(1, 2, 3, 1, 2)
      case 3: O2(0, 2, 3, 0, 1)
      case 4: r = sgn(p[2][0] - p[3][0]); break;
      case 5: r = sgn(p[2][1] - p[3][1]); break;
      case 6: O2(0, 2, 3, 0, 2)
      case 7: r = sgn(p[2][2] - p[3][2]); break;
      case 8: O2(0, 2, 3, 1, 2)
      case 9: O2(0, 1, 3, 0, 1)
      case 10: r = sgn(p[1][0] - p[3][0]); break;
      case 11: r = sgn(p[1][1] - p[3][1]); break;
      case 12: r = sgn(p[0][0] - p[3][0]); break;
      default: r = 1; break;
#undef O2
    }
    if (r != 0) break;
  }
  if (depth == 1 || depth == 3 || depth == 5 || depth == 8 || depth == 10)
    r = -r;
  r *= parity;
  if (depth_out) *depth_out = depth;
  return r < 0 ? 1 : -1;
}

int o_orient3d_sos(const double *a, const double *b, const double *c,
                   const double *d, uint32_t ia, uint32_t ib, uint32_t ic,
                   uint32_t id) {
  return o_orient3d_sos_depth(a, b, c, d, ia, ib, ic, id, NULL);
}

/*
 * The 49-level insphere cascade (src/predicates.rs:478-675), as data.
 * kind: 3 = orient3d-type (4 ranks), 2 = orient2d-type (3 ranks),
 *       1 = orient1d-type (2 ranks), 0 = constant +1.
 * perm: coordinate order applied to each point before the sub-determinant:
 *       0 xyz, 1 xzy, 2 yzx, 3 yxz, 4 zxy.
 */
typedef struct {
  unsigned char kind, lifted, perm, r[4];
} Step;

static const Step kCascade[50] = {
    {0, 0, 0, {0, 0, 0, 0}}, /* unused depth 0 */
    {3, 0, 0, {1, 2, 3, 4}}, {3, 1, 0, {1, 2, 3, 4}},
    {3, 1, 1, {1, 2, 3, 4}}, {3, 1, 2, {1, 2, 3, 4}},
    {3, 0, 0, {0, 2, 3, 4}}, {2, 0, 0, {2, 3, 4, 0}},
    {2, 0, 1, {2, 3, 4, 0}}, {2, 0, 2, {2, 3, 4, 0}},
    {3, 1, 0, {0, 2, 3, 4}}, {2, 1, 0, {2, 3, 4, 0}},
    {2, 1, 3, {2, 3, 4, 0}}, {3, 1, 1, {0, 2, 3, 4}},
    {2, 1, 4, {2, 3, 4, 0}}, {3, 1, 2, {0, 2, 3, 4}},
    {3, 0, 0, {0, 1, 3, 4}}, {2, 0, 0, {1, 3, 4, 0}},
    {2, 0, 1, {1, 3, 4, 0}}, {2, 0, 2, {1, 3, 4, 0}},
    {2, 0, 0, {0, 3, 4, 0}}, {1, 0, 0, {3, 4, 0, 0}},
    {1, 0, 3, {3, 4, 0, 0}}, {2, 0, 1, {0, 3, 4, 0}},
    {1, 0, 4, {3, 4, 0, 0}}, {2, 0, 2, {0, 3, 4, 0}},
    {3, 1, 0, {0, 1, 3, 4}}, {2, 1, 0, {1, 3, 4, 0}},
    {2, 1, 3, {1, 3, 4, 0}}, {2, 1, 0, {0, 3, 4, 0}},
    {1, 1, 0, {3, 4, 0, 0}}, {2, 1, 3, {0, 3, 4, 0}},
    {3, 1, 1, {0, 1, 3, 4}}, {2, 1, 4, {1, 3, 4, 0}},
    {2, 1, 4, {0, 3, 4, 0}}, {3, 1, 2, {0, 1, 3, 4}},
    {3, 0, 0, {0, 1, 2, 4}}, {2, 0, 0, {1, 2, 4, 0}},
    {2, 0, 1, {1, 2, 4, 0}}, {2, 0, 2, {1, 2, 4, 0}},
    {2, 0, 0, {0, 2, 4, 0}}, {1, 0, 0, {2, 4, 0, 0}},
    {1, 0, 3, {2, 4, 0, 0}}, {2, 0, 1, {0, 2, 4, 0}},
    {1, 0, 4, {2, 4, 0, 0}}, {2, 0, 2, {0, 2, 4, 0}},
    {2, 0, 0, {0, 1, 4, 0}}, {1, 0, 0, {1, 4, 0, 0}},
    {1, 0, 3, {1, 4, 0, 0}}, {1, 0, 0, {0, 4, 0, 0}},
    {0, 0, 0, {0, 0, 0, 0}}, /* depth 49: constant */
};

static const unsigned char kPerm[5][3] = {
    {0, 1, 2}, {0, 2, 1}, {1, 2, 0}, {1, 0, 2}, {2, 0, 1}};

static int flips4(int depth) {
  switch (depth) {
    case 1: case 3: case 7: case 9: case 11: case 14: case 15: case 16:
    case 18: case 20: case 22: case 23: case 26: case 30: case 31: case 32:
    case 37: case 39: case 41: case 44: case 46:
      return 1;
    default:
      return 0;
  }
}

int o_orient4d_sos_only(const double *a, const double *b, const double *c,
                        const double *d, const double *e, uint32_t ia,
                        uint32_t ib, uint32_t ic, uint32_t id, uint32_t ie,
                        int *depth_out) {
  uint32_t idx[5] = {ia, ib, ic, id, ie};
  const double *pt[5] = {a, b, c, d, e};
  int swaps = 0;
  /* ascending by index; only the permutation parity matters */
  for (int i = 0; i < 4; ++i)
    for (int j = 4; j > i; --j)
      if (idx[j] < idx[j - 1]) {
        uint32_t t = idx[j]; idx[j] = idx[j - 1]; idx[j - 1] = t;
        const double *tp = pt[j]; pt[j] = pt[j - 1]; pt[j - 1] = tp;
        ++swaps;
      }
  int depth, r = 0;
  for (depth = 1; depth <= 49; ++depth) {
    const Step *s = &kCascade[depth];
    if (s->kind == 0) { r = 1; break; }
    double q[4][3];
    int np = s->kind + 1;
    for (int k = 0; k < np; ++k)
      for (int j = 0; j < 3; ++j) q[k][j] = pt[s->r[k]][kPerm[s->perm][j]];
    if (s->kind == 3) {
      r = s->lifted ? o_orient3d_lifted_sign(q[0], q[1], q[2], q[3])
                    : o_orient3d_sign(q[0], q[1], q[2], q[3]);
    } else if (s->kind == 2) {
      r = s->lifted ? o_orient2d_lifted_sign(q[0], q[1], q[2])
                    : sgn(o_orient2d_plain(q[0], q[1], q[2]));
    } else {
      r = s->lifted ? o_orient1d_lifted_sign(q[0], q[1]) : sgn(q[0][0] - q[1][0]);
    }
    if (r != 0) break;
  }
  if (depth > 49) depth = 49;
  if (flips4(depth)) r = -r;
  if (swaps & 1) r = -r;
  if (depth_out) *depth_out = depth;
  return r > 0 ? 1 : -1;
}

int o_orient4d(const double *a, const double *b, const double *c,
               const double *d, const double *e, uint32_t ia, uint32_t ib,
               uint32_t ic, uint32_t id, uint32_t ie) {
  int s = o_insphere_sign(a, b, c, d, e);
  if (s != 0) return s;
  return o_orient4d_sos_only(a, b, c, d, e, ia, ib, ic, id, ie, NULL);
}

/* ------------------------------------------------------------------------ */
/* batch helpers                                                             */
/* ------------------------------------------------------------------------ */

void o_orient3d_sign_batch(const double *pts, int64_t n, int32_t *out) {
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < n; ++i) {
    const double *q = pts + i * 12;
    out[i] = o_orient3d_sign(q, q + 3, q + 6, q + 9);
  }
}

void o_insphere_sign_batch(const double *pts, int64_t n, int32_t *out) {
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < n; ++i) {
    const double *q = pts + i * 15;
    out[i] = o_insphere_sign(q, q + 3, q + 6, q + 9, q + 12);
  }
}

void o_orient3d_sos_batch(const double *pts, const uint32_t *idx, int64_t n,
                          int32_t *out, int32_t *depth) {
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < n; ++i) {
    const double *q = pts + i * 12;
    const uint32_t *k = idx + i * 4;
    int dd = 0;
    out[i] = o_orient3d_sos_depth(q, q + 3, q + 6, q + 9, k[0], k[1], k[2],
                                  k[3], &dd);
    if (depth) depth[i] = dd;
  }
}

void o_orient4d_batch(const double *pts, const uint32_t *idx, int64_t n,
                      int32_t *out, int32_t *depth) {
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < n; ++i) {
    const double *q = pts + i * 15;
    const uint32_t *k = idx + i * 5;
    int s = o_insphere_sign(q, q + 3, q + 6, q + 9, q + 12);
    int dd = 0;
    if (s == 0)
      s = o_orient4d_sos_only(q, q + 3, q + 6, q + 9, q + 12, k[0], k[1],
                              k[2], k[3], k[4], &dd);
    out[i] = s;
    if (depth) depth[i] = dd;
  }
}

// Staged geometric predicates for the triangulation kernels (host + device).
//
// Inputs are f32 coordinates (the normalized points and the four finite
// super vertices).  Each sign is decided by the first stage that can certify
// it:
//   1. FP32 semi-static filter: the determinant in f32 against Shewchuk's
//      stage-A bound (J. R. Shewchuk, DCG 18, 1997) with eps = 2^-24, plus an
//      absolute slack and a magnitude floor so f32 underflow cannot fool it;
//   2. FP64 filter: the same formula in f64 with eps = 2^-53 (f32 inputs are
//      exact in f64);
//   3. exact stage: every coordinate is an integer multiple of 2^E, where E
//      is the lowest set-bit exponent over the coordinates involved, so the
//      determinant is evaluated exactly in fixed-width two's-complement
//      integers (32-bit limbs): one limb per coordinate difference when the
//      differences to the last point are exact in f64 and span <= 30 bits,
//      else 64-bit limb pairs per coordinate when the coordinates span
//      <= 60 bits, 192-bit limbs otherwise.
// When the exact determinant is zero, Simulation of Simplicity decides, with
// the cascades of the reference (src/predicates.rs:106-208 for orient3d,
// :478-675 for insphere): Edelsbrunner-Muecke perturbations ranked by point
// id, whose non-lifted 2D steps are the plain f64 expression of
// src/predicates.rs:73-75 (reproduced bit for bit, no fused multiply-add).
//
// Conventions (as in the reference):
//   orient3d_raw(a,b,c,d) = sign det[a-d; b-d; c-d]
//   insphere_raw          = sign of Shewchuk's insphere
//   orient3d_sos -> +1 when the perturbed raw determinant is < 0, else -1
//   orient4d     -> +1 (outside) / -1 (inside) for a tet with raw orient < 0
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define WT_HD __host__ __device__ __forceinline__
#define WT_HDN static __host__ __device__ __noinline__

namespace wtet {

enum PredMode : int { PRED_FILTERED = 0, PRED_EXACT_ONLY = 1 };

// Per-stage resolution counters: [orient3d f32, f64, exact!=0, exact==0,
//                                 insphere f32, f64, exact!=0, exact==0]
struct PredStats {
  uint32_t c[8];
};

WT_HD void stat_inc(PredStats* s, int k) {
  if (s) s->c[k] += 1;
}

// ---------------------------------------------------------------------------
// fixed-width signed integers, little-endian 32-bit limbs, two's complement
// ---------------------------------------------------------------------------
template <int N>
struct BI {
  uint32_t w[N];
};

template <int N>
WT_HD BI<N> bi_zero() {
  BI<N> r;
#pragma unroll
  for (int i = 0; i < N; ++i) r.w[i] = 0u;
  return r;
}

template <int N>
WT_HD bool bi_isneg(const BI<N>& a) {
  return (a.w[N - 1] >> 31) != 0u;
}

template <int N>
WT_HD int bi_sign(const BI<N>& a) {
  if (bi_isneg(a)) return -1;
  uint32_t o = 0u;
#pragma unroll
  for (int i = 0; i < N; ++i) o |= a.w[i];
  return o ? 1 : 0;
}

template <int N>
WT_HD BI<N> bi_add(const BI<N>& a, const BI<N>& b) {
  BI<N> r;
  uint64_t c = 0;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    uint64_t s = (uint64_t)a.w[i] + b.w[i] + c;
    r.w[i] = (uint32_t)s;
    c = s >> 32;
  }
  return r;
}

template <int N>
WT_HD BI<N> bi_sub(const BI<N>& a, const BI<N>& b) {
  BI<N> r;
  uint64_t c = 1;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    uint64_t s = (uint64_t)a.w[i] + (uint32_t)~b.w[i] + c;
    r.w[i] = (uint32_t)s;
    c = s >> 32;
  }
  return r;
}

template <int N>
WT_HD BI<N> bi_neg(const BI<N>& a) {
  return bi_sub(bi_zero<N>(), a);
}

// sign-extend (or truncate) to M limbs
template <int M, int N>
WT_HD BI<M> bi_ext(const BI<N>& a) {
  BI<M> r;
  uint32_t e = bi_isneg(a) ? 0xFFFFFFFFu : 0u;
#pragma unroll
  for (int i = 0; i < M; ++i) r.w[i] = i < N ? a.w[i] : e;
  return r;
}

template <int A, int B>
WT_HD BI<A + B> bi_mul(const BI<A>& a, const BI<B>& b) {
  bool na = bi_isneg(a), nb = bi_isneg(b);
  BI<A> ua = na ? bi_neg(a) : a;
  BI<B> ub = nb ? bi_neg(b) : b;
  BI<A + B> r = bi_zero<A + B>();
#pragma unroll
  for (int i = 0; i < A; ++i) {
    uint64_t c = 0;
#pragma unroll
    for (int j = 0; j < B; ++j) {
      uint64_t t = (uint64_t)ua.w[i] * ub.w[j] + r.w[i + j] + c;
      r.w[i + j] = (uint32_t)t;
      c = t >> 32;
    }
    r.w[i + B] = (uint32_t)c;
  }
  return (na != nb) ? bi_neg(r) : r;
}

// ---------------------------------------------------------------------------
// f32 -> integer multiple of 2^E
// ---------------------------------------------------------------------------
struct F32Parts {
  uint32_t m;  // significand (0 for zero)
  int e;       // value = +-m * 2^e
  bool neg;
};

WT_HD uint32_t f32_bits(float v) {
  uint32_t u;
  __builtin_memcpy(&u, &v, 4);
  return u;
}

WT_HD F32Parts f32_parts(float v) {
  uint32_t u = f32_bits(v);
  F32Parts p;
  p.neg = (u >> 31) != 0u;
  uint32_t be = (u >> 23) & 0xFFu;
  uint32_t fr = u & 0x7FFFFFu;
  if (be == 0u) {
    p.m = fr;
    p.e = -149;
  } else {
    p.m = fr | 0x800000u;
    p.e = (int)be - 150;
  }
  return p;
}

WT_HD int ctz32(uint32_t x) { return __builtin_ctz(x); }
WT_HD int clz32(uint32_t x) { return __builtin_clz(x); }

// lowest / highest set-bit exponents over k coordinates; returns false when
// every coordinate is zero
WT_HD bool span_of(const float* v, int k, int* lo, int* hi) {
  int L = 100000, H = -100000;
  for (int i = 0; i < k; ++i) {
    F32Parts p = f32_parts(v[i]);
    if (p.m == 0u) continue;
    int l = p.e + ctz32(p.m);
    int h = p.e + 31 - clz32(p.m);
    L = l < L ? l : L;
    H = h > H ? h : H;
  }
  *lo = L;
  *hi = H;
  return L != 100000;
}

// coordinate as a D-limb integer in units of 2^E (requires shift + 24 bits
// to fit below the sign bit)
template <int D>
WT_HD BI<D> to_fixed(float v, int E) {
  F32Parts p = f32_parts(v);
  BI<D> r = bi_zero<D>();
  if (p.m == 0u) return r;
  // drop the mantissa's trailing zeros first: E is the lowest set bit over
  // all coordinates, so the shift below is >= 0 only from that bit
  int tz = ctz32(p.m);
  int s = p.e + tz - E;  // >= 0
  int limb = s >> 5, bit = s & 31;
  uint64_t x = (uint64_t)(p.m >> tz) << bit;
  // unrolled selects instead of r.w[limb]: a run-time index into a local
  // array would put it in scratch memory
#pragma unroll
  for (int i = 0; i < D; ++i)
    r.w[i] = i == limb ? (uint32_t)x : (i == limb + 1 ? (uint32_t)(x >> 32) : 0u);
  return p.neg ? bi_neg(r) : r;
}

// ---------------------------------------------------------------------------
// Short differences.  orient3d and insphere depend only on the differences
// of their points to the last one.  Each difference of two f32 coordinates
// is computed in f64 and kept when exact (checked with Knuth's two-sum
// error term); when every one is exact and their set bits span at most
// kRelSpan bits, they are integers below 2^30 in units of their lowest set
// bit, and the determinant runs on single 32-bit limbs.  Nearby points have
// short differences even where the coordinates themselves span many bits
// (a 128^3 lattice: coordinates span up to 31 bits, neighbour differences
// 24), and the limb products fall about fourfold.
// ---------------------------------------------------------------------------
constexpr int kRelSpan = 30;

struct F64Parts {
  uint64_t m;  // significand (0 for zero)
  int e;       // value = +-m * 2^e
  bool neg;
};

WT_HD F64Parts f64_parts(double v) {
  uint64_t u;
  __builtin_memcpy(&u, &v, 8);
  F64Parts p;
  p.neg = (u >> 63) != 0u;
  uint32_t be = (uint32_t)((u >> 52) & 0x7FFu);
  uint64_t fr = u & ((1ull << 52) - 1ull);
  if (be == 0u) {
    p.m = fr;
    p.e = -1074;
  } else {
    p.m = fr | (1ull << 52);
    p.e = (int)be - 1075;
  }
  return p;
}

// a - b in f64; false when the rounded difference is not the exact one
WT_HD bool exact_sub(double a, double b, double* d) {
  double s = a - b;
  double bv = s - a;
  double err = (a - (s - bv)) + (-b - bv);
  *d = s;
  return err == 0.0;
}

// the R points before the last as differences to it, 1-limb integers in
// units of 2^E (E the lowest set bit over them); false when the tier does
// not apply (a difference not exact in f64, or a span over kRelSpan)
template <int R>
WT_HD bool rel_fixed1(const float* P, BI<1> (*d)[3]) {
  double v[R][3];
  int L = 100000, H = -100000;
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (!exact_sub((double)P[r * 3 + k], (double)P[R * 3 + k], &v[r][k])) return false;
      F64Parts p = f64_parts(v[r][k]);
      if (p.m == 0u) continue;
      int l = p.e + __builtin_ctzll(p.m);
      int h = p.e + 63 - __builtin_clzll(p.m);
      L = l < L ? l : L;
      H = h > H ? h : H;
    }
  if (L != 100000 && H - L + 1 > kRelSpan) return false;
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      F64Parts p = f64_parts(v[r][k]);
      BI<1> x;
      x.w[0] = 0u;
      if (p.m != 0u) {
        int tz = __builtin_ctzll(p.m);
        x.w[0] = (uint32_t)((p.m >> tz) << (p.e + tz - L));  // < 2^30
        if (p.neg) x = bi_neg(x);
      }
      d[r][k] = x;
    }
  return true;
}

// ---------------------------------------------------------------------------
// exact determinants (templated on the coordinate width D in limbs)
// ---------------------------------------------------------------------------
// det of the 3x3 differences d (rows a-d, b-d, c-d)
template <int D>
WT_HD int orient3d_core(const BI<D> (*d)[3]) {
  BI<2 * D> bc = bi_sub(bi_mul(d[1][0], d[2][1]), bi_mul(d[2][0], d[1][1]));
  BI<2 * D> ca = bi_sub(bi_mul(d[2][0], d[0][1]), bi_mul(d[0][0], d[2][1]));
  BI<2 * D> ab = bi_sub(bi_mul(d[0][0], d[1][1]), bi_mul(d[1][0], d[0][1]));
  BI<3 * D> det = bi_add(bi_add(bi_mul(d[0][2], bc), bi_mul(d[1][2], ca)),
                         bi_mul(d[2][2], ab));
  return bi_sign(det);
}

template <int D>
WT_HD int exact_orient3d_T(const float* P, int E) {
  // P: 4 points x 3 coords; det[a-d; b-d; c-d]
  BI<D> d[3][3];
  for (int r = 0; r < 3; ++r)
    for (int k = 0; k < 3; ++k)
      d[r][k] = bi_sub(to_fixed<D>(P[r * 3 + k], E), to_fixed<D>(P[9 + k], E));
  return orient3d_core<D>(d);
}

// Shewchuk's insphere on the differences d (rows a-e .. d-e)
template <int D>
WT_HD int insphere_core(const BI<D> (*d)[3]) {
  BI<D> x[4], y[4], z[4];
  BI<2 * D> l[4];
  for (int r = 0; r < 4; ++r) {
    x[r] = d[r][0];
    y[r] = d[r][1];
    z[r] = d[r][2];
    l[r] = bi_add(bi_add(bi_mul(x[r], x[r]), bi_mul(y[r], y[r])), bi_mul(z[r], z[r]));
  }
  // 2x2 minors on (x, y)
  BI<2 * D> ab = bi_sub(bi_mul(x[0], y[1]), bi_mul(x[1], y[0]));
  BI<2 * D> bc = bi_sub(bi_mul(x[1], y[2]), bi_mul(x[2], y[1]));
  BI<2 * D> cd = bi_sub(bi_mul(x[2], y[3]), bi_mul(x[3], y[2]));
  BI<2 * D> da = bi_sub(bi_mul(x[3], y[0]), bi_mul(x[0], y[3]));
  BI<2 * D> ac = bi_sub(bi_mul(x[0], y[2]), bi_mul(x[2], y[0]));
  BI<2 * D> bd = bi_sub(bi_mul(x[1], y[3]), bi_mul(x[3], y[1]));
  // 3x3 minors with z
  BI<3 * D> abc = bi_add(bi_sub(bi_mul(z[0], bc), bi_mul(z[1], ac)), bi_mul(z[2], ab));
  BI<3 * D> bcd = bi_add(bi_sub(bi_mul(z[1], cd), bi_mul(z[2], bd)), bi_mul(z[3], bc));
  BI<3 * D> cda = bi_add(bi_add(bi_mul(z[2], da), bi_mul(z[3], ac)), bi_mul(z[0], cd));
  BI<3 * D> dab = bi_add(bi_add(bi_mul(z[3], ab), bi_mul(z[0], bd)), bi_mul(z[1], da));
  BI<5 * D> det = bi_add(bi_sub(bi_mul(l[3], abc), bi_mul(l[2], dab)),
                         bi_sub(bi_mul(l[1], cda), bi_mul(l[0], bcd)));
  return bi_sign(det);
}

template <int D>
WT_HD int exact_insphere_T(const float* P, int E) {
  // P: 5 points; Shewchuk's insphere with a..d relative to e.  The same
  // expression as insphere_core, written out on its own: built through the
  // core's array argument, the wide forms spilled more.
  BI<D> x[4], y[4], z[4];
  BI<2 * D> l[4];
  for (int r = 0; r < 4; ++r) {
    x[r] = bi_sub(to_fixed<D>(P[r * 3 + 0], E), to_fixed<D>(P[12], E));
    y[r] = bi_sub(to_fixed<D>(P[r * 3 + 1], E), to_fixed<D>(P[13], E));
    z[r] = bi_sub(to_fixed<D>(P[r * 3 + 2], E), to_fixed<D>(P[14], E));
    l[r] = bi_add(bi_add(bi_mul(x[r], x[r]), bi_mul(y[r], y[r])), bi_mul(z[r], z[r]));
  }
  BI<2 * D> ab = bi_sub(bi_mul(x[0], y[1]), bi_mul(x[1], y[0]));
  BI<2 * D> bc = bi_sub(bi_mul(x[1], y[2]), bi_mul(x[2], y[1]));
  BI<2 * D> cd = bi_sub(bi_mul(x[2], y[3]), bi_mul(x[3], y[2]));
  BI<2 * D> da = bi_sub(bi_mul(x[3], y[0]), bi_mul(x[0], y[3]));
  BI<2 * D> ac = bi_sub(bi_mul(x[0], y[2]), bi_mul(x[2], y[0]));
  BI<2 * D> bd = bi_sub(bi_mul(x[1], y[3]), bi_mul(x[3], y[1]));
  BI<3 * D> abc = bi_add(bi_sub(bi_mul(z[0], bc), bi_mul(z[1], ac)), bi_mul(z[2], ab));
  BI<3 * D> bcd = bi_add(bi_sub(bi_mul(z[1], cd), bi_mul(z[2], bd)), bi_mul(z[3], bc));
  BI<3 * D> cda = bi_add(bi_add(bi_mul(z[2], da), bi_mul(z[3], ac)), bi_mul(z[0], cd));
  BI<3 * D> dab = bi_add(bi_add(bi_mul(z[3], ab), bi_mul(z[0], bd)), bi_mul(z[1], da));
  BI<5 * D> det = bi_add(bi_sub(bi_mul(l[3], abc), bi_mul(l[2], dab)),
                         bi_sub(bi_mul(l[1], cda), bi_mul(l[0], bcd)));
  return bi_sign(det);
}

// det [[p0 p1 |p|^2 1]] over 4 points (coordinates already permuted)
template <int D>
WT_HD int exact_orient3d_lifted_T(const float* P, int E) {
  BI<D> X[4], Y[4], Z[4];
  BI<2 * D> W[4];
  for (int r = 0; r < 4; ++r) {
    X[r] = to_fixed<D>(P[r * 3 + 0], E);
    Y[r] = to_fixed<D>(P[r * 3 + 1], E);
    Z[r] = to_fixed<D>(P[r * 3 + 2], E);
    W[r] = bi_add(bi_add(bi_mul(X[r], X[r]), bi_mul(Y[r], Y[r])), bi_mul(Z[r], Z[r]));
  }
  BI<D> dx[3], dy[3];
  BI<2 * D> dw[3];
  for (int r = 0; r < 3; ++r) {
    dx[r] = bi_sub(X[r], X[3]);
    dy[r] = bi_sub(Y[r], Y[3]);
    dw[r] = bi_sub(W[r], W[3]);
  }
  BI<2 * D> bc = bi_sub(bi_mul(dx[1], dy[2]), bi_mul(dx[2], dy[1]));
  BI<2 * D> ca = bi_sub(bi_mul(dx[2], dy[0]), bi_mul(dx[0], dy[2]));
  BI<2 * D> ab = bi_sub(bi_mul(dx[0], dy[1]), bi_mul(dx[1], dy[0]));
  BI<4 * D> det = bi_add(bi_add(bi_mul(dw[0], bc), bi_mul(dw[1], ca)), bi_mul(dw[2], ab));
  return bi_sign(det);
}

// det [[p0 p1 |p|^2 1]] over 4 COPLANAR points from their differences to
// the last one: det[[u_i - u_3, v_i - v_3, |p_i - p_3|^2]].  Translating
// the points by t changes the lifted determinant by -2 t_w det[[p0 p1 p2 1]]
// (t_w the coordinate left out of the first two columns), which vanishes
// when the four points are coplanar; the sign is then the absolute form's.
WT_HD int orient3d_lifted_rel1(const BI<1> (*d)[3]) {
  BI<2> w[3];
  for (int r = 0; r < 3; ++r)
    w[r] = bi_add(bi_add(bi_mul(d[r][0], d[r][0]), bi_mul(d[r][1], d[r][1])), bi_mul(d[r][2], d[r][2]));
  BI<2> bc = bi_sub(bi_mul(d[1][0], d[2][1]), bi_mul(d[2][0], d[1][1]));
  BI<2> ca = bi_sub(bi_mul(d[2][0], d[0][1]), bi_mul(d[0][0], d[2][1]));
  BI<2> ab = bi_sub(bi_mul(d[0][0], d[1][1]), bi_mul(d[1][0], d[0][1]));
  BI<4> det = bi_add(bi_add(bi_mul(w[0], bc), bi_mul(w[1], ca)), bi_mul(w[2], ab));
  return bi_sign(det);
}

// det [[p0 |p|^2 1]] over 3 points
template <int D>
WT_HD int exact_orient2d_lifted_T(const float* P, int E) {
  BI<D> X[3], Y[3], Z[3];
  BI<2 * D> W[3];
  for (int r = 0; r < 3; ++r) {
    X[r] = to_fixed<D>(P[r * 3 + 0], E);
    Y[r] = to_fixed<D>(P[r * 3 + 1], E);
    Z[r] = to_fixed<D>(P[r * 3 + 2], E);
    W[r] = bi_add(bi_add(bi_mul(X[r], X[r]), bi_mul(Y[r], Y[r])), bi_mul(Z[r], Z[r]));
  }
  BI<3 * D> det = bi_sub(bi_mul(bi_sub(X[0], X[2]), bi_sub(W[1], W[2])),
                         bi_mul(bi_sub(W[0], W[2]), bi_sub(X[1], X[2])));
  return bi_sign(det);
}

template <int D>
WT_HD int exact_orient1d_lifted_T(const float* P, int E) {
  BI<2 * D> W[2];
  for (int r = 0; r < 2; ++r) {
    BI<D> X = to_fixed<D>(P[r * 3 + 0], E), Y = to_fixed<D>(P[r * 3 + 1], E),
          Z = to_fixed<D>(P[r * 3 + 2], E);
    W[r] = bi_add(bi_add(bi_mul(X, X), bi_mul(Y, Y)), bi_mul(Z, Z));
  }
  return bi_sign(bi_sub(W[0], W[1]));
}

// Narrow (64-bit) coordinates when the span fits, else 192-bit.  Inputs are
// normalized points and super vertices, so the span is below 190 bits.
constexpr int kNarrowSpan = 60;

#define WT_EXACT_DISPATCH(fn, P, k)                 \
  int lo_, hi_;                                     \
  if (!span_of(P, k, &lo_, &hi_)) return 0;         \
  if (hi_ - lo_ + 1 <= kNarrowSpan) return fn<2>(P, lo_); \
  return fn<6>(P, lo_);

// The short-difference tier as small leaf functions (few registers, so a
// call saves none: no scratch traffic on the path degenerate inputs take
// millions of times), returning 2 when the tier does not apply; the wide
// forms are separate calls for the rest.
WT_HDN int exact_orient3d_short(const float* P) {
  BI<1> d[3][3];
  return rel_fixed1<3>(P, d) ? orient3d_core<1>(d) : 2;
}
WT_HDN int exact_insphere_short(const float* P) {
  BI<1> d[4][3];
  return rel_fixed1<4>(P, d) ? insphere_core<1>(d) : 2;
}
WT_HDN int exact_orient3d_wide(const float* P) { WT_EXACT_DISPATCH(exact_orient3d_T, P, 12) }
WT_HDN int exact_insphere_wide(const float* P) { WT_EXACT_DISPATCH(exact_insphere_T, P, 15) }
// The short tier cannot apply with a super vertex among the points (its
// coordinates are -100 or 400, the differences to it span more than 30
// bits): such calls go to the wide forms at once, without a failed try.
WT_HD bool near_unit(const float* P, int k) {
  bool ok = true;
  for (int i = 0; i < k; ++i) ok = ok && P[i] > -4.0f && P[i] < 4.0f;
  return ok;
}
WT_HD int exact_orient3d(const float* P) {
  int s = near_unit(P, 12) ? exact_orient3d_short(P) : 2;
  return s != 2 ? s : exact_orient3d_wide(P);
}
WT_HD int exact_insphere(const float* P) {
  int s = near_unit(P, 15) ? exact_insphere_short(P) : 2;
  return s != 2 ? s : exact_insphere_wide(P);
}
// the lifted orient3d of four points known to be coplanar (the SoS cascade
// reaches a lifted 4-point step only after the plain orient3d of the same
// four points vanished): short differences when they apply
WT_HDN int exact_orient3d_lifted_coplanar_short(const float* P) {
  BI<1> d[3][3];
  return rel_fixed1<3>(P, d) ? orient3d_lifted_rel1(d) : 2;
}
WT_HDN int exact_orient3d_lifted_wide(const float* P) { WT_EXACT_DISPATCH(exact_orient3d_lifted_T, P, 12) }
WT_HD int exact_orient3d_lifted_coplanar(const float* P) {
  int s = near_unit(P, 12) ? exact_orient3d_lifted_coplanar_short(P) : 2;
  return s != 2 ? s : exact_orient3d_lifted_wide(P);
}
WT_HDN int exact_orient2d_lifted(const float* P) { WT_EXACT_DISPATCH(exact_orient2d_lifted_T, P, 9) }
WT_HDN int exact_orient1d_lifted(const float* P) { WT_EXACT_DISPATCH(exact_orient1d_lifted_T, P, 6) }

#undef WT_EXACT_DISPATCH

// ---------------------------------------------------------------------------
// floating-point filters (Shewchuk stage A)
// ---------------------------------------------------------------------------
#pragma clang fp contract(off)

WT_HD int fsign(float v) { return (v > 0.0f) - (v < 0.0f); }
WT_HD int dsign(double v) { return (v > 0.0) - (v < 0.0); }

// returns 2 when undecided
WT_HD int orient3d_f32(const float* a, const float* b, const float* c, const float* d) {
  float adx = a[0] - d[0], bdx = b[0] - d[0], cdx = c[0] - d[0];
  float ady = a[1] - d[1], bdy = b[1] - d[1], cdy = c[1] - d[1];
  float adz = a[2] - d[2], bdz = b[2] - d[2], cdz = c[2] - d[2];
  float bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
  float cdxady = cdx * ady, adxcdy = adx * cdy;
  float adxbdy = adx * bdy, bdxady = bdx * ady;
  float det = adz * (bdxcdy - cdxbdy) + bdz * (cdxady - adxcdy) + cdz * (adxbdy - bdxady);
  float perm = (fabsf(bdxcdy) + fabsf(cdxbdy)) * fabsf(adz) +
               (fabsf(cdxady) + fabsf(adxcdy)) * fabsf(bdz) +
               (fabsf(adxbdy) + fabsf(bdxady)) * fabsf(cdz);
  const float kB = 4.1724e-07f;  // (7 + 56*2^-24) * 2^-24, rounded up
  if (!(perm >= 8.0e-28f)) return 2;  // 2^-90: keep f32 underflow out
  float eb = kB * perm + 1.0e-36f;
  if (det > eb) return 1;
  if (-det > eb) return -1;
  return 2;
}

WT_HD int orient3d_f64(const float* af, const float* bf, const float* cf, const float* df) {
  double adx = (double)af[0] - df[0], bdx = (double)bf[0] - df[0], cdx = (double)cf[0] - df[0];
  double ady = (double)af[1] - df[1], bdy = (double)bf[1] - df[1], cdy = (double)cf[1] - df[1];
  double adz = (double)af[2] - df[2], bdz = (double)bf[2] - df[2], cdz = (double)cf[2] - df[2];
  double bdxcdy = bdx * cdy, cdxbdy = cdx * bdy;
  double cdxady = cdx * ady, adxcdy = adx * cdy;
  double adxbdy = adx * bdy, bdxady = bdx * ady;
  double det = adz * (bdxcdy - cdxbdy) + bdz * (cdxady - adxcdy) + cdz * (adxbdy - bdxady);
  double perm = (fabs(bdxcdy) + fabs(cdxbdy)) * fabs(adz) +
                (fabs(cdxady) + fabs(adxcdy)) * fabs(bdz) +
                (fabs(adxbdy) + fabs(bdxady)) * fabs(cdz);
  const double kB = 7.7715611723762e-16;  // (7 + 56*2^-53) * 2^-53, rounded up
  double eb = kB * perm;
  if (det > eb) return 1;
  if (-det > eb) return -1;
  return 2;
}

template <typename T>
struct InsphereParts {
  T det, perm;
};

template <typename T>
WT_HD InsphereParts<T> insphere_eval(const float* a, const float* b, const float* c,
                                     const float* d, const float* e) {
  T aex = (T)a[0] - (T)e[0], bex = (T)b[0] - (T)e[0], cex = (T)c[0] - (T)e[0], dex = (T)d[0] - (T)e[0];
  T aey = (T)a[1] - (T)e[1], bey = (T)b[1] - (T)e[1], cey = (T)c[1] - (T)e[1], dey = (T)d[1] - (T)e[1];
  T aez = (T)a[2] - (T)e[2], bez = (T)b[2] - (T)e[2], cez = (T)c[2] - (T)e[2], dez = (T)d[2] - (T)e[2];
  T aexbey = aex * bey, bexaey = bex * aey, ab = aexbey - bexaey;
  T bexcey = bex * cey, cexbey = cex * bey, bc = bexcey - cexbey;
  T cexdey = cex * dey, dexcey = dex * cey, cd = cexdey - dexcey;
  T dexaey = dex * aey, aexdey = aex * dey, da = dexaey - aexdey;
  T aexcey = aex * cey, cexaey = cex * aey, ac = aexcey - cexaey;
  T bexdey = bex * dey, dexbey = dex * bey, bd = bexdey - dexbey;
  T abc = aez * bc - bez * ac + cez * ab;
  T bcd = bez * cd - cez * bd + dez * bc;
  T cda = cez * da + dez * ac + aez * cd;
  T dab = dez * ab + aez * bd + bez * da;
  T alift = aex * aex + aey * aey + aez * aez;
  T blift = bex * bex + bey * bey + bez * bez;
  T clift = cex * cex + cey * cey + cez * cez;
  T dlift = dex * dex + dey * dey + dez * dez;
  InsphereParts<T> r;
  r.det = (dlift * abc - clift * dab) + (blift * cda - alift * bcd);
  T ez = aez < 0 ? -aez : aez, fz = bez < 0 ? -bez : bez, gz = cez < 0 ? -cez : cez,
    hz = dez < 0 ? -dez : dez;
#define WT_ABS(v) ((v) < 0 ? -(v) : (v))
  T aexbeyp = WT_ABS(aexbey), bexaeyp = WT_ABS(bexaey);
  T bexceyp = WT_ABS(bexcey), cexbeyp = WT_ABS(cexbey);
  T cexdeyp = WT_ABS(cexdey), dexceyp = WT_ABS(dexcey);
  T dexaeyp = WT_ABS(dexaey), aexdeyp = WT_ABS(aexdey);
  T aexceyp = WT_ABS(aexcey), cexaeyp = WT_ABS(cexaey);
  T bexdeyp = WT_ABS(bexdey), dexbeyp = WT_ABS(dexbey);
#undef WT_ABS
  r.perm = ((cexdeyp + dexceyp) * fz + (dexbeyp + bexdeyp) * gz + (bexceyp + cexbeyp) * hz) * alift +
           ((dexaeyp + aexdeyp) * gz + (aexceyp + cexaeyp) * hz + (cexdeyp + dexceyp) * ez) * blift +
           ((aexbeyp + bexaeyp) * hz + (bexdeyp + dexbeyp) * ez + (dexaeyp + aexdeyp) * fz) * clift +
           ((bexceyp + cexbeyp) * ez + (cexaeyp + aexceyp) * fz + (aexbeyp + bexaeyp) * gz) * dlift;
  return r;
}

WT_HD int insphere_f32(const float* a, const float* b, const float* c, const float* d,
                       const float* e) {
  InsphereParts<float> r = insphere_eval<float>(a, b, c, d, e);
  const float kB = 9.5368e-07f;  // (16 + 224*2^-24) * 2^-24, rounded up
  if (!(r.perm >= 8.271806e-25f)) return 2;  // 2^-80
  float eb = kB * r.perm + 7.7e-34f;           // + 2^-110
  if (r.det > eb) return 1;
  if (-r.det > eb) return -1;
  return 2;
}

// perm (optional): receives the permanent, for the adaptive stage
WT_HD int insphere_f64(const float* a, const float* b, const float* c, const float* d,
                       const float* e, double* perm = nullptr) {
  InsphereParts<double> r = insphere_eval<double>(a, b, c, d, e);
  const double kB = 1.7763568394003e-15;  // (16 + 224*2^-53) * 2^-53, rounded up
  double eb = kB * r.perm;
  if (perm) *perm = r.perm;
  if (r.det > eb) return 1;
  if (-r.det > eb) return -1;
  return 2;
}

// ---------------------------------------------------------------------------
// FP64 adaptive stage for insphere: the same determinant in double-double
// arithmetic (error-free products by fused multiply-add, Knuth's two-sum;
// T. J. Dekker, Numer. Math. 18, 1971; Y. Hida, X. S. Li, D. H. Bailey,
// ARITH-15, 2001).  Each double-double operation errs by at most about
// 4 * 2^-106 times the magnitudes it combines, so the whole evaluation errs
// by less than 2^-100 times the permanent (the same expression over
// absolute values); a sign is certified when it clears 2^-90 times the
// permanent.  Requires every coordinate difference to be exact in f64
// (checked), which also keeps the low words far above the subnormal range
// (differences of f32 inputs are >= 2^-149 in magnitude, products of five
// of them >= 2^-745).
// ---------------------------------------------------------------------------
struct DD {
  double hi, lo;
};

WT_HD DD dd_two_sum(double a, double b) {
  double s = a + b;
  double bb = s - a;
  return {s, (a - (s - bb)) + (b - bb)};
}

WT_HD DD dd_quick_two_sum(double a, double b) {
  double s = a + b;
  return {s, b - (s - a)};
}

WT_HD DD dd_two_prod(double a, double b) {
  double p = a * b;
  return {p, __builtin_fma(a, b, -p)};
}

WT_HD DD dd_add(DD a, DD b) {
  DD s = dd_two_sum(a.hi, b.hi);
  DD t = dd_two_sum(a.lo, b.lo);
  s.lo += t.hi;
  s = dd_quick_two_sum(s.hi, s.lo);
  s.lo += t.lo;
  return dd_quick_two_sum(s.hi, s.lo);
}

WT_HD DD dd_neg(DD a) { return {-a.hi, -a.lo}; }

WT_HD DD dd_mul(DD a, DD b) {
  DD p = dd_two_prod(a.hi, b.hi);
  p.lo += a.hi * b.lo + a.lo * b.hi;
  return dd_quick_two_sum(p.hi, p.lo);
}

WT_HD DD dd_mul_d(DD a, double b) {
  DD p = dd_two_prod(a.hi, b);
  p.lo += a.lo * b;
  return dd_quick_two_sum(p.hi, p.lo);
}

// x_i y_j - x_j y_i, exact products
WT_HD DD dd_minor(double xi, double yj, double xj, double yi) {
  return dd_add(dd_two_prod(xi, yj), dd_neg(dd_two_prod(xj, yi)));
}

// perm: the FP64 filter's permanent of the same points
WT_HD int insphere_dd(const float* a, const float* b, const float* c, const float* d,
                      const float* e, double perm) {
  const float* P[4] = {a, b, c, d};
  double X[4], Y[4], Z[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    DD dx = dd_two_sum((double)P[r][0], -(double)e[0]);
    DD dy = dd_two_sum((double)P[r][1], -(double)e[1]);
    DD dz = dd_two_sum((double)P[r][2], -(double)e[2]);
    if (dx.lo != 0.0 || dy.lo != 0.0 || dz.lo != 0.0) return 2;  // difference not exact
    X[r] = dx.hi;
    Y[r] = dy.hi;
    Z[r] = dz.hi;
  }
  DD ab = dd_minor(X[0], Y[1], X[1], Y[0]);
  DD bc = dd_minor(X[1], Y[2], X[2], Y[1]);
  DD cd = dd_minor(X[2], Y[3], X[3], Y[2]);
  DD da = dd_minor(X[3], Y[0], X[0], Y[3]);
  DD ac = dd_minor(X[0], Y[2], X[2], Y[0]);
  DD bd = dd_minor(X[1], Y[3], X[3], Y[1]);
  DD abc = dd_add(dd_add(dd_mul_d(bc, Z[0]), dd_neg(dd_mul_d(ac, Z[1]))), dd_mul_d(ab, Z[2]));
  DD bcd = dd_add(dd_add(dd_mul_d(cd, Z[1]), dd_neg(dd_mul_d(bd, Z[2]))), dd_mul_d(bc, Z[3]));
  DD cda = dd_add(dd_add(dd_mul_d(da, Z[2]), dd_mul_d(ac, Z[3])), dd_mul_d(cd, Z[0]));
  DD dab = dd_add(dd_add(dd_mul_d(ab, Z[3]), dd_mul_d(bd, Z[0])), dd_mul_d(da, Z[1]));
  DD L[4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
    L[r] = dd_add(dd_add(dd_two_prod(X[r], X[r]), dd_two_prod(Y[r], Y[r])), dd_two_prod(Z[r], Z[r]));
  DD det = dd_add(dd_add(dd_mul(L[3], abc), dd_neg(dd_mul(L[2], dab))),
                  dd_add(dd_mul(L[1], cda), dd_neg(dd_mul(L[0], bcd))));
  const double kBound = 8.077935669463161e-28;  // 2^-90
  double eb = kBound * perm;
  double lo = det.lo < 0 ? -det.lo : det.lo;
  if (det.hi > 2.0 * eb + lo) return 1;
  if (-det.hi > 2.0 * eb + lo) return -1;
  return 2;
}

// src/predicates.rs:73-75, bit for bit: four f64 subtractions, two products,
// one subtraction, each rounded to nearest (contraction is off in this file)
WT_HD double orient2d_plain(double a0, double a1, double b0, double b1, double c0, double c1) {
  double t1 = a0 - c0;
  double t2 = b1 - c1;
  double t3 = a1 - c1;
  double t4 = b0 - c0;
  double p1 = t1 * t2;
  double p2 = t3 * t4;
  return p1 - p2;
}

#pragma clang fp contract(on)

// ---------------------------------------------------------------------------
// staged raw signs
// ---------------------------------------------------------------------------
WT_HD int orient3d_raw(const float* a, const float* b, const float* c, const float* d, int mode,
                       PredStats* st) {
  if (mode == PRED_FILTERED) {
    int s = orient3d_f32(a, b, c, d);
    if (s != 2) {
      stat_inc(st, 0);
      return s;
    }
    s = orient3d_f64(a, b, c, d);
    if (s != 2) {
      stat_inc(st, 1);
      return s;
    }
  }
  float P[12] = {a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2], d[0], d[1], d[2]};
  // the short tier pays where the filters have left the near-degenerate
  // cases (a lattice's co-planar and co-spherical sets); on every call of
  // the exact-only mode the wide form alone is faster
  int s = mode == PRED_FILTERED ? exact_orient3d(P) : exact_orient3d_wide(P);
  stat_inc(st, s != 0 ? 2 : 3);
  return s;
}

WT_HD int insphere_raw(const float* a, const float* b, const float* c, const float* d,
                       const float* e, int mode, PredStats* st) {
  if (mode == PRED_FILTERED) {
    int s = insphere_f32(a, b, c, d, e);
    if (s != 2) {
      stat_inc(st, 4);
      return s;
    }
    double perm = 0.0;
    s = insphere_f64(a, b, c, d, e, &perm);
    if (s == 2) s = insphere_dd(a, b, c, d, e, perm);
    if (s != 2) {
      stat_inc(st, 5);
      return s;
    }
  }
  float P[15] = {a[0], a[1], a[2], b[0], b[1], b[2], c[0], c[1], c[2],
                 d[0], d[1], d[2], e[0], e[1], e[2]};
  int s = mode == PRED_FILTERED ? exact_insphere(P) : exact_insphere_wide(P);
  stat_inc(st, s != 0 ? 6 : 7);
  return s;
}

// ---------------------------------------------------------------------------
// SoS cascades (reference semantics)
// ---------------------------------------------------------------------------
WT_HD int sgn_diff(float a, float b) { return (a > b) - (a < b); }

// src/predicates.rs:106-208 after the exact determinant vanished
WT_HDN int orient3d_sos_only(const float* a, const float* b, const float* c, const float* d,
                             uint32_t ia, uint32_t ib, uint32_t ic, uint32_t id) {
  uint32_t v[4] = {ia, ib, ic, id};
  const float* p[4] = {a, b, c, d};
  int parity = 1;
  for (int i = 0; i < 3; ++i) {
    int mi = i;
    for (int j = i + 1; j < 4; ++j)
      if (v[j] < v[mi]) mi = j;
    if (mi != i) {
      uint32_t tv = v[i];
      v[i] = v[mi];
      v[mi] = tv;
      const float* tp = p[i];
      p[i] = p[mi];
      p[mi] = tp;
      parity = -parity;
    }
  }
  // (point, point, point, axis u, axis w) for the orient2d depths
  int depth = 0, r = 0;
  for (; depth < 13; ++depth) {
    int P0 = -1, P1 = 0, P2 = 0, U = 0, W = 0;
    switch (depth) {
      case 0: P0 = 1; P1 = 2; P2 = 3; U = 0; W = 1; break;
      case 1: P0 = 1; P1 = 2; P2 = 3; U = 0; W = 2; break;
      case 2: P0 = 1; P1 = 2; P2 = 3; U = 1; W = 2; break;
      case 3: P0 = 0; P1 = 2; P2 = 3; U = 0; W = 1; break;
      case 4: r = sgn_diff(p[2][0], p[3][0]); break;
      case 5: r = sgn_diff(p[2][1], p[3][1]); break;
      case 6: P0 = 0; P1 = 2; P2 = 3; U = 0; W = 2; break;
      case 7: r = sgn_diff(p[2][2], p[3][2]); break;
      case 8: P0 = 0; P1 = 2; P2 = 3; U = 1; W = 2; break;
      case 9: P0 = 0; P1 = 1; P2 = 3; U = 0; W = 1; break;
      case 10: r = sgn_diff(p[1][0], p[3][0]); break;
      case 11: r = sgn_diff(p[1][1], p[3][1]); break;
      default: r = sgn_diff(p[0][0], p[3][0]); break;
    }
    if (P0 >= 0)
      r = dsign(orient2d_plain(p[P0][U], p[P0][W], p[P1][U], p[P1][W], p[P2][U], p[P2][W]));
    if (r != 0) break;
  }
  if (depth == 13) r = 1;
  if (depth == 1 || depth == 3 || depth == 5 || depth == 8 || depth == 10) r = -r;
  r *= parity;
  return r < 0 ? 1 : -1;
}

WT_HD int orient3d_sos(const float* a, const float* b, const float* c, const float* d,
                       uint32_t ia, uint32_t ib, uint32_t ic, uint32_t id, int mode,
                       PredStats* st) {
  int s = orient3d_raw(a, b, c, d, mode, st);
  if (s != 0) return s < 0 ? 1 : -1;
  return orient3d_sos_only(a, b, c, d, ia, ib, ic, id);
}

// FP32 filter stage alone, for the lean kernels: the staged functions
// sign conventions, or 2 when the filter cannot decide (the caller then
// defers the whole element to a kernel that runs the full stages).
WT_HD int orient3d_sos_fast(const float* a, const float* b, const float* c, const float* d,
                            PredStats* st) {
  int s = orient3d_f32(a, b, c, d);
  if (s == 2) return 2;
  stat_inc(st, 0);
  return s < 0 ? 1 : -1;
}

// The insphere cascade as data: kind (3: 4-point, 2: 3-point, 1: 2-point,
// 0: constant +1), lifted, coordinate order (0 xyz, 1 xzy, 2 yzx, 3 yxz,
// 4 zxy), ranks of the points used.  Depths 1..49 of src/predicates.rs:513-652.
struct SosStep {
  uint8_t kind, lifted, perm, r0, r1, r2, r3;
};

// packed: kind<<24 | lifted<<20 | perm<<16 | r0<<12 | r1<<8 | r2<<4 | r3
// (a table in constant memory: as a local array it was built on the stack
// at every call)
constexpr uint32_t kSos4Steps[50] = {
    0u,
    0x3001234u, 0x3101234u, 0x3111234u, 0x3121234u, 0x3000234u, 0x2002340u,
    0x2012340u, 0x2022340u, 0x3100234u, 0x2102340u, 0x2132340u, 0x3110234u,
    0x2142340u, 0x3120234u, 0x3000134u, 0x2001340u, 0x2011340u, 0x2021340u,
    0x2000340u, 0x1003400u, 0x1033400u, 0x2010340u, 0x1043400u, 0x2020340u,
    0x3100134u, 0x2101340u, 0x2131340u, 0x2100340u, 0x1103400u, 0x2130340u,
    0x3110134u, 0x2141340u, 0x2140340u, 0x3120134u, 0x3000124u, 0x2001240u,
    0x2011240u, 0x2021240u, 0x2000240u, 0x1002400u, 0x1032400u, 0x2010240u,
    0x1042400u, 0x2020240u, 0x2000140u, 0x1001400u, 0x1031400u, 0x1000400u,
    0u};

WT_HD SosStep sos4_step(int depth) {
  uint32_t x = kSos4Steps[depth];
  SosStep s;
  s.kind = (uint8_t)(x >> 24);
  s.lifted = (uint8_t)((x >> 20) & 0xF);
  s.perm = (uint8_t)((x >> 16) & 0xF);
  s.r0 = (uint8_t)((x >> 12) & 0xF);
  s.r1 = (uint8_t)((x >> 8) & 0xF);
  s.r2 = (uint8_t)((x >> 4) & 0xF);
  s.r3 = (uint8_t)(x & 0xF);
  return s;
}

WT_HD bool sos4_flip(int depth) {
  // depths 1,3,7,9,11,14,15,16,18,20,22,23,26,30,31,32,37,39,41,44,46
  const uint64_t mask = (1ull << 1) | (1ull << 3) | (1ull << 7) | (1ull << 9) | (1ull << 11) |
                        (1ull << 14) | (1ull << 15) | (1ull << 16) | (1ull << 18) |
                        (1ull << 20) | (1ull << 22) | (1ull << 23) | (1ull << 26) |
                        (1ull << 30) | (1ull << 31) | (1ull << 32) | (1ull << 37) |
                        (1ull << 39) | (1ull << 41) | (1ull << 44) | (1ull << 46);
  return (mask >> depth) & 1ull;
}

WT_HD void perm_coords(const float* p, int perm, float* q) {
  int i0 = 0, i1 = 1, i2 = 2;
  switch (perm) {
    case 1: i1 = 2; i2 = 1; break;
    case 2: i0 = 1; i1 = 2; i2 = 0; break;
    case 3: i0 = 1; i1 = 0; i2 = 2; break;
    case 4: i0 = 2; i1 = 0; i2 = 1; break;
    default: break;
  }
  q[0] = p[i0];
  q[1] = p[i1];
  q[2] = p[i2];
}

// src/predicates.rs:478-675 after the exact insphere vanished
WT_HDN int orient4d_sos_only(const float* a, const float* b, const float* c, const float* d,
                             const float* e, uint32_t ia, uint32_t ib, uint32_t ic,
                             uint32_t id, uint32_t ie) {
  uint32_t idx[5] = {ia, ib, ic, id, ie};
  const float* pt[5] = {a, b, c, d, e};
  int swaps = 0;
  for (int i = 0; i < 4; ++i)
    for (int j = 4; j > i; --j)
      if (idx[j] < idx[j - 1]) {
        uint32_t t = idx[j];
        idx[j] = idx[j - 1];
        idx[j - 1] = t;
        const float* tp = pt[j];
        pt[j] = pt[j - 1];
        pt[j - 1] = tp;
        ++swaps;
      }
  int depth = 1, r = 1;
  for (; depth < 49; ++depth) {
    SosStep s = sos4_step(depth);
    float Q[12];
    int rk[4] = {s.r0, s.r1, s.r2, s.r3};
    for (int k = 0; k <= (int)s.kind; ++k) perm_coords(pt[rk[k]], s.perm, Q + 3 * k);
    if (s.kind == 3) {
      // a plain orient3d sign goes through the filters first: the same
      // exact sign, almost always from the FP32 stage
      // a lifted step's four points had their plain orient3d vanish at an
      // earlier depth (depths 1, 5, 15 for the lifted 2-4, 9/12/14, 25/31/34)
      r = s.lifted ? exact_orient3d_lifted_coplanar(Q)
                   : orient3d_raw(Q, Q + 3, Q + 6, Q + 9, PRED_FILTERED, nullptr);
    } else if (s.kind == 2) {
      r = s.lifted ? exact_orient2d_lifted(Q)
                   : dsign(orient2d_plain(Q[0], Q[1], Q[3], Q[4], Q[6], Q[7]));
    } else {
      r = s.lifted ? exact_orient1d_lifted(Q) : sgn_diff(Q[0], Q[3]);
    }
    if (r != 0) break;
  }
  if (depth == 49) r = 1;
  if (sos4_flip(depth)) r = -r;
  if (swaps & 1) r = -r;
  return r > 0 ? 1 : -1;
}

WT_HD int orient4d(const float* a, const float* b, const float* c, const float* d,
                   const float* e, uint32_t ia, uint32_t ib, uint32_t ic, uint32_t id,
                   uint32_t ie, int mode, PredStats* st) {
  int s = insphere_raw(a, b, c, d, e, mode, st);
  if (s != 0) return s;
  return orient4d_sos_only(a, b, c, d, e, ia, ib, ic, id, ie);
}

WT_HD int orient4d_fast(const float* a, const float* b, const float* c, const float* d,
                        const float* e, PredStats* st) {
  int s = insphere_f32(a, b, c, d, e);
  if (s == 2) return 2;
  stat_inc(st, 4);
  return s;
}

}  // namespace wtet

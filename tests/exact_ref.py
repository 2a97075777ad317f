"""Independent exact-rational references used to pin the oracle (tests only).

* Determinants are evaluated with ``fractions.Fraction`` (exact), straight
  from the definition -- a method unrelated to the oracle's floating-point
  expansions and to the product's fixed-point integers.
* ``sos_sign`` derives Simulation of Simplicity from first principles
  (Edelsbrunner & Muecke, ACM TOG 9(1), 1990): the matrix entry (row of rank
  i, perturbable column j) receives +eps**(2**(d*i - j)), and the sign of the
  perturbed determinant is the sign of the first non-vanishing coefficient
  in increasing powers of eps.  The coefficient of a set S of perturbed
  entries is the determinant with each row of S replaced by the unit vector
  of its column (multilinearity of the determinant).
"""
from __future__ import annotations

import itertools
from fractions import Fraction


def det(m):
    """Exact determinant (Fractions) by fraction-free elimination."""
    a = [[Fraction(x) for x in row] for row in m]
    n = len(a)
    sign = 1
    d = Fraction(1)
    for c in range(n):
        p = next((r for r in range(c, n) if a[r][c] != 0), None)
        if p is None:
            return Fraction(0)
        if p != c:
            a[c], a[p] = a[p], a[c]
            sign = -sign
        d *= a[c][c]
        for r in range(c + 1, n):
            if a[r][c] != 0:
                f = a[r][c] / a[c][c]
                for k in range(c, n):
                    a[r][k] -= f * a[c][k]
    return sign * d


def sgn(x) -> int:
    return (x > 0) - (x < 0)


def orient3d(a, b, c, d) -> int:
    """sign det[a-d; b-d; c-d] (Shewchuk's orient3d convention)."""
    F = Fraction
    m = [[F(p[k]) - F(d[k]) for k in range(3)] for p in (a, b, c)]
    return sgn(det(m))


def insphere(a, b, c, d, e) -> int:
    """sign of Shewchuk's insphere = det[[p-e, |p-e|^2]] over a,b,c,d."""
    F = Fraction
    rows = []
    for p in (a, b, c, d):
        q = [F(p[k]) - F(e[k]) for k in range(3)]
        rows.append(q + [q[0] * q[0] + q[1] * q[1] + q[2] * q[2]])
    return sgn(det(rows))


def _lift(p):
    F = Fraction
    return F(p[0]) ** 2 + F(p[1]) ** 2 + F(p[2]) ** 2


def sos_sign(points, idx, lifted: bool) -> int:
    """Sign of the E-M perturbed determinant.

    lifted=False: rows (x, y, z, 1)    for 4 points (orient3d), d = 3.
    lifted=True : rows (x, y, z, w, 1) for 5 points (insphere), d = 4.
    Rows stay in the given order; ranks come from idx.
    """
    F = Fraction
    k = len(points)
    ncol = 4 if lifted else 3  # perturbable columns
    dd = ncol
    rank = {r: sorted(range(k), key=lambda i: idx[i]).index(r) for r in range(k)}
    base = []
    for p in points:
        row = [F(p[0]), F(p[1]), F(p[2])]
        if lifted:
            row.append(_lift(p))
        row.append(F(1))
        base.append(row)
    # all partial matchings (row -> perturbable column), sorted by exponent sum
    terms = []
    rows = list(range(k))
    cols = list(range(ncol))
    for s in range(0, min(k, ncol) + 1):
        for rsub in itertools.combinations(rows, s):
            for csub in itertools.permutations(cols, s):
                # exponent of entry (rank i, column j with j = col+1): 2**(d*i - j)
                e = sum(Fraction(2) ** (dd * rank[r] - (c + 1)) for r, c in zip(rsub, csub))
                terms.append((e, rsub, csub))
    terms.sort(key=lambda t: t[0])
    for _, rsub, csub in terms:
        m = [list(r) for r in base]
        for r, c in zip(rsub, csub):
            m[r] = [F(0)] * len(m[r])
            m[r][c] = F(1)
        v = det(m)
        if v != 0:
            return sgn(v)
    raise AssertionError("perturbed determinant vanished identically")


def orient3d_sos_ref(points, idx) -> int:
    """Reference convention: +1 when the perturbed raw determinant is < 0."""
    return 1 if sos_sign(points, idx, lifted=False) < 0 else -1


def orient4d_ref(points, idx) -> int:
    """Reference convention: sign of the perturbed insphere determinant."""
    return sos_sign(points, idx, lifted=True)

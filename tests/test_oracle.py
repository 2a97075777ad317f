"""Pins the CPU oracle before anything is compared against it.

Checks: the reference's own known-answer tests (tests/golden), exact
determinants in rational arithmetic, SoS derived from first principles,
the plain-f64 orient2d rounding, brute-force Delaunay on tiny sets, and
closed-form invariants of the triangulation (exact volume, Euler, adjacency).
"""
import itertools
import json
import os
import random
from fractions import Fraction

import numpy as np
import pytest

import exact_ref as R

GOLD = os.path.join(os.path.dirname(__file__), "golden", "reference_known_answers.json")


@pytest.fixture(scope="module")
def gold():
    with open(GOLD) as f:
        return json.load(f)


# --------------------------------------------------------------------------
# reference known answers
# --------------------------------------------------------------------------

def test_known_orient3d(oracle, gold):
    for case in gold["orient3d"]:
        s = int(oracle.orient3d_sign(np.array([case["pts"]]))[0])
        if "sign" in case:
            assert s == case["sign"], case["src"]
        else:
            assert s != 0, case["src"]


def test_known_insphere(oracle, gold):
    for case in gold["insphere_inside_outside"]:
        e = [float(np.sqrt(3.0)) - 1e-8 if v == "sqrt3_minus_1e-8" else v for v in case["e"]]
        q = np.array([case["tet"] + [e]], dtype=np.float64)
        o = int(oracle.orient3d_sign(q[:, :4])[0])
        s = int(oracle.insphere_sign(q)[0])
        inside = s > 0 if o > 0 else s < 0
        assert inside == case["inside"], case["src"]
    for case in gold["insphere_zero"]:
        q = np.array([case["pts"]], dtype=np.float64)
        assert int(oracle.insphere_sign(q)[0]) == 0, case["src"]
        assert int(oracle.orient4d(q, np.array([case["idx"]]))[0]) != 0
    for case in gold["insphere_permutation_flips_sign"]:
        q = np.array(case["abcde"], dtype=np.float64)
        i, j = case["swap"]
        q2 = q.copy()
        q2[[i, j]] = q2[[j, i]]
        s1 = int(oracle.insphere_sign(q[None])[0])
        s2 = int(oracle.insphere_sign(q2[None])[0])
        assert s1 * s2 < 0, case["src"]


def test_known_sos(oracle, gold):
    for case in gold["orient3d_sos"]:
        r = int(oracle.orient3d_sos(np.array([case["pts"]]), np.array([case["idx"]]))[0])
        assert r != 0
        if case.get("opposite"):
            r2 = int(oracle.orient3d_sos(np.array([case["swapped_pts"]]),
                                         np.array([case["swapped_idx"]]))[0])
            assert r == -r2, case["src"]
        if "expected" in case:
            assert r == case["expected"], case["src"]
    for case in gold["orient4d_nondegenerate"]:
        q = np.array([case["pts"]], dtype=np.float64)
        raw = int(oracle.insphere_sign(q)[0])
        assert raw != 0
        assert int(oracle.orient4d(q, np.array([case["idx"]]))[0]) == raw, case["src"]


def test_known_pointset_helpers(oracle, gold):
    for c in gold["expand_bits"]:
        assert oracle.expand_bits(c["in"]) == c["out"], c["src"]
    for c in gold["interleave"]:
        x, y, z = c["xyz"]
        v = (oracle.expand_bits(z) << 2) | (oracle.expand_bits(y) << 1) | oracle.expand_bits(x)
        assert v == c["out"], c["src"]
    for c in gold["morton_code"]:
        assert oracle.morton_code(c["p"], c["min"], c["max"]) == c["out"], c["src"]
    for c in gold["initial_stride"]:
        assert oracle.initial_stride(c["n"]) == c["stride"], c["src"]
    for c in gold["stratified_order"]:
        order = oracle.stratified_order(c["n"], c["stride"])
        if "order" in c:
            assert order.tolist() == c["order"], c["src"]
        else:
            assert order[: len(c["prefix"])].tolist() == c["prefix"], c["src"]
        assert sorted(order.tolist()) == list(range(c["n"]))
    for c in gold["normalize"]:
        out, mn, sc = oracle.normalize_points(np.array(c["pts"], np.float32))
        assert out.min() == c["min"] and out.max() == c["max"], c["src"]


def test_normalize_closed_form(oracle):
    rng = np.random.default_rng(7)
    p = (rng.standard_normal((5000, 3)) * 37.5 + 11).astype(np.float32)
    out, mn, sc = oracle.normalize_points(p)
    mn_ref = p.min(axis=0)
    ext = (p.max(axis=0) - mn_ref).astype(np.float32)
    sc_ref = np.float32(max(ext.max(), np.float32(1e-10)))
    ref = ((p - mn_ref).astype(np.float32) / sc_ref).astype(np.float32)
    assert sc == sc_ref
    assert np.array_equal(out, ref)


# --------------------------------------------------------------------------
# exact predicates vs rational arithmetic
# --------------------------------------------------------------------------

def _cases_f32(rng, k, npts):
    """Random f32 points: plain, clustered (near-degenerate) and on a grid."""
    out = []
    for mode in range(3):
        if mode == 0:
            q = rng.random((k, npts, 3), dtype=np.float32)
        elif mode == 1:
            base = rng.random((k, 1, 3), dtype=np.float32)
            q = (base + rng.standard_normal((k, npts, 3)).astype(np.float32) * np.float32(1e-6))
        else:
            q = (rng.integers(0, 4, (k, npts, 3)) / np.float32(3.0)).astype(np.float32)
        out.append(q.astype(np.float32))
    return np.concatenate(out)


def test_orient3d_exact_vs_fraction(oracle):
    rng = np.random.default_rng(1)
    q = _cases_f32(rng, 300, 4)
    s = oracle.orient3d_sign(q)
    for i in range(len(q)):
        pts = [[Fraction(float(v)) for v in p] for p in q[i]]
        assert s[i] == R.orient3d(*pts), i


def test_insphere_exact_vs_fraction(oracle):
    rng = np.random.default_rng(2)
    q = _cases_f32(rng, 200, 5)
    # also exact cosphericity: box corners of random boxes
    lo = rng.random((100, 3), dtype=np.float32)
    hi = lo + rng.random((100, 3), dtype=np.float32)
    box = []
    for a, b in zip(lo, hi):
        c = [[(a, b)[(m >> k) & 1][k] for k in range(3)] for m in range(8)]
        box.append(np.array(c[:5], np.float32))
    q = np.concatenate([q, np.array(box)])
    s = oracle.insphere_sign(q)
    zero = 0
    for i in range(len(q)):
        ref = R.insphere(*[[Fraction(float(v)) for v in p] for p in q[i]])
        assert s[i] == ref, i
        zero += ref == 0
    assert zero >= 100  # the box corners are exactly cospherical


def test_lifted_subdeterminants_vs_fraction(oracle):
    rng = np.random.default_rng(3)
    F = Fraction
    for _ in range(200):
        p = (rng.integers(0, 3, (4, 3)) / np.float32(2.0)).astype(np.float64)
        w = [F(float(a)) ** 2 + F(float(b)) ** 2 + F(float(c)) ** 2 for a, b, c in p]
        m3 = [[F(float(p[r][0])), F(float(p[r][1])), w[r], 1] for r in range(4)]
        assert oracle.orient3d_lifted_sign(*p) == R.sgn(R.det(m3))
        m2 = [[F(float(p[r][0])), w[r], 1] for r in range(3)]
        assert oracle.orient2d_lifted_sign(*p[:3]) == R.sgn(R.det(m2))
        assert oracle.orient1d_lifted_sign(p[0], p[1]) == R.sgn(w[0] - w[1])


def test_orient2d_plain_matches_ieee_double():
    """src/predicates.rs:73-75 is a plain f64 expression; Python floats are
    IEEE doubles with no fused multiply-add, so they fix every rounding."""
    from oracle import oracle as O
    rng = np.random.default_rng(4)
    for _ in range(2000):
        a, b, c = (rng.random(2).astype(np.float32).astype(np.float64) for _ in range(3))
        if rng.random() < 0.5:  # near-collinear
            c = (a + (b - a) * np.float64(np.float32(rng.random()))).astype(np.float32).astype(np.float64)
        ref = (a[0] - c[0]) * (b[1] - c[1]) - (a[1] - c[1]) * (b[0] - c[0])
        got = O.orient2d_plain(a, b, c)
        assert got == ref or (np.isnan(got) and np.isnan(ref))


# --------------------------------------------------------------------------
# SoS: restated cascades vs Edelsbrunner-Muecke from first principles
# --------------------------------------------------------------------------

@pytest.mark.parametrize("vals", [[0.0, 0.5, 1.0], [0.0, 1.0], [0.25, 0.5, 0.75, 1.0]])
def test_orient4d_sos_matches_em(oracle, vals):
    rnd = random.Random(11)
    n = 700
    pts = np.array([[[rnd.choice(vals) for _ in range(3)] for _ in range(5)] for _ in range(n)])
    idx = np.array([rnd.sample(range(50), 5) for _ in range(n)], np.uint32)
    got, depth = oracle.orient4d(pts, idx, with_depth=True)
    for i in range(n):
        assert got[i] == R.orient4d_ref(pts[i].tolist(), idx[i].tolist()), (pts[i], idx[i])
    assert len(set(depth.tolist())) >= 12  # many cascade depths exercised


@pytest.mark.parametrize("vals", [[0.0, 0.5, 1.0], [0.0, 1.0]])
def test_orient3d_sos_matches_em(oracle, vals):
    rnd = random.Random(12)
    n = 1500
    pts = np.array([[[rnd.choice(vals) for _ in range(3)] for _ in range(4)] for _ in range(n)])
    idx = np.array([rnd.sample(range(50), 4) for _ in range(n)], np.uint32)
    got, depth = oracle.orient3d_sos(pts, idx, with_depth=True)
    for i in range(n):
        assert got[i] == R.orient3d_sos_ref(pts[i].tolist(), idx[i].tolist())
    assert len(set(depth.tolist())) >= 10


def test_sos_antisymmetry(oracle):
    rnd = random.Random(13)
    for _ in range(300):
        pts = [[rnd.choice([0.0, 1.0]) for _ in range(3)] for _ in range(5)]
        idx = rnd.sample(range(30), 5)
        s = int(oracle.orient4d(np.array([pts]), np.array([idx]))[0])
        i, j = rnd.sample(range(5), 2)
        p2, i2 = list(pts), list(idx)
        p2[i], p2[j] = p2[j], p2[i]
        i2[i], i2[j] = i2[j], i2[i]
        assert int(oracle.orient4d(np.array([p2]), np.array([i2]))[0]) == -s


# --------------------------------------------------------------------------
# the triangulation
# --------------------------------------------------------------------------

def _grid(nx, ny, nz):
    """src/tests.rs:1103-1117 gen_grid: ix-major, iz fastest, i/(n-1) in f32."""
    pts = []
    for ix in range(nx):
        for iy in range(ny):
            for iz in range(nz):
                pts.append([np.float32(ix) / np.float32(max(nx, 2) - 1),
                            np.float32(iy) / np.float32(max(ny, 2) - 1),
                            np.float32(iz) / np.float32(max(nz, 2) - 1)])
    return np.array(pts, np.float32)


def _brute_force_dt(oracle, pn):
    """All 4-subsets of points+super vertices with an empty SoS circumsphere."""
    n = len(pn)
    P = oracle.extended_points(pn)[: n + 4]
    ids = list(range(n + 4))
    quads = list(itertools.combinations(ids, 4))
    q = P[np.array(quads)]
    o = oracle.orient3d_sos(q, np.array(quads, np.uint32))
    oriented = [t if s > 0 else (t[1], t[0], t[2], t[3]) for t, s in zip(quads, o)]
    out = []
    for t in oriented:
        others = [v for v in ids if v not in t]
        qq = np.array([np.vstack([P[list(t)], P[v]]) for v in others])
        ii = np.array([list(t) + [v] for v in others], np.uint32)
        if np.all(oracle.orient4d(qq, ii) > 0):
            out.append(tuple(sorted(t)))
    return sorted(out)


@pytest.mark.parametrize("kind", ["uniform", "grid", "cospherical", "duplicates", "coplanar"])
def test_bw_matches_brute_force(oracle, kind):
    rng = np.random.default_rng(5)
    if kind == "uniform":
        p = rng.random((9, 3), dtype=np.float32)
    elif kind == "grid":
        p = _grid(3, 2, 2)
    elif kind == "cospherical":
        p = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
                      [0.6, 0.8, 0], [0, 0.6, 0.8]], np.float32)
    elif kind == "duplicates":
        p = rng.random((6, 3), dtype=np.float32)
        p = np.vstack([p, p[:3]])
    else:
        p = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0], [0.5, 0.5, 0],
                      [0.25, 0.75, 0], [0.5, 0.5, 1]], np.float32)
    pn, _, _ = oracle.normalize_points(p)
    tets = oracle.delaunay(pn)
    got = sorted(tuple(sorted(int(v) for v in t)) for t in tets)
    assert got == _brute_force_dt(oracle, pn)


def _exact_volume_ok(pts_ext, tets):
    """Sum of raw orient3d determinants of all tets == the super tet's, exactly;
    and no tet has positive raw orientation."""
    F = Fraction
    total = F(0)
    for t in tets:
        a, b, c, d = (pts_ext[int(v)] for v in t)
        m = [[F(float(x[k])) - F(float(d[k])) for k in range(3)] for x in (a, b, c)]
        v = R.det(m)
        if v > 0:
            return False
        total += v
    s = pts_ext[-5:-1]
    m = [[F(float(x[k])) - F(float(s[3][k])) for k in range(3)] for x in s[:3]]
    return total == R.det(m)


@pytest.mark.parametrize("kind", ["uniform", "grid"])
def test_bw_invariants(oracle, kind):
    rng = np.random.default_rng(6)
    p = rng.random((400, 3), dtype=np.float32) if kind == "uniform" else _grid(5, 5, 5)
    pn, _, _ = oracle.normalize_points(p)
    n = len(pn)
    tets, stats = oracle.delaunay(pn, return_stats=True)
    assert stats[0] == 0 and stats[1] == 0  # walk always found a conflicting seed
    ext = oracle.extended_points(pn)
    assert _exact_volume_ok(ext, tets)
    adj = oracle.rebuild_adjacency(tets)
    # closed 3-ball: hull faces are exactly the 4 faces of the super tet
    assert int((adj == 0xFFFFFFFF).sum()) == 4
    # symmetric adjacency
    for t in range(len(tets)):
        for f in range(4):
            a = int(adj[t, f])
            if a != 0xFFFFFFFF:
                assert int(adj[a >> 5, a & 3]) == (t << 5) | f
    # Euler characteristic of a triangulated 3-ball: V - E + F - T = 1
    V = n + 4
    E = len({tuple(sorted((int(t[i]), int(t[j])))) for t in tets for i in range(4) for j in range(i + 1, 4)})
    Fc = len({tuple(sorted(int(v) for k, v in enumerate(t) if k != f)) for t in tets for f in range(4)})
    assert V - E + Fc - len(tets) == 1
    assert set(np.unique(tets).tolist()) == set(range(n + 4))
    assert oracle.count_violations(pn, tets, adj, only_real=False) == 0


def test_bw_agrees_with_qhull_on_uniform(oracle):
    """Informational cross-check against scipy's Qhull (general position):
    every real-only oracle tet is a Qhull tet, and every Qhull tet the
    oracle lacks has a super vertex inside its circumsphere."""
    scipy_spatial = pytest.importorskip("scipy.spatial")
    rng = np.random.default_rng(8)
    p = rng.random((3000, 3), dtype=np.float32)
    pn, _, _ = oracle.normalize_points(p)
    n = len(pn)
    tets = oracle.delaunay(pn)
    mine = {tuple(t) for t in oracle.canonical_real(tets, n).tolist()}
    qh = {tuple(sorted(s)) for s in scipy_spatial.Delaunay(pn.astype(np.float64)).simplices.tolist()}
    assert mine <= qh
    missing = sorted(qh - mine)
    assert len(missing) < 0.05 * len(qh)
    ext = oracle.extended_points(pn)
    for t in missing:
        q = []
        idx = []
        o = oracle.orient3d_sos(ext[list(t)][None], np.array([t], np.uint32))[0]
        tt = t if o > 0 else (t[1], t[0], t[2], t[3])
        for s in range(n, n + 4):
            q.append(np.vstack([ext[list(tt)], ext[s]]))
            idx.append(list(tt) + [s])
        assert np.any(oracle.orient4d(np.array(q), np.array(idx, np.uint32)) < 0)


@pytest.mark.parametrize("kind", ["uniform", "grid", "cospherical", "duplicates", "coplanar"])
def test_reference_bw_restatement_matches_brute_force(oracle, kind):
    """full_bw_reconstruction as written (input order, linear seed scan,
    src/phase2.rs:383-606) gives the brute-force SoS Delaunay set on tiny
    inputs, degenerate ones included."""
    rng = np.random.default_rng(9)
    if kind == "uniform":
        p = rng.random((9, 3), dtype=np.float32)
    elif kind == "grid":
        p = _grid(3, 2, 2)
    elif kind == "cospherical":
        p = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
                      [0.6, 0.8, 0], [0, 0.6, 0.8]], np.float32)
    elif kind == "duplicates":
        p = rng.random((6, 3), dtype=np.float32)
        p = np.vstack([p, p[:3]])
    else:
        p = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0], [0.5, 0.5, 0],
                      [0.25, 0.75, 0], [0.5, 0.5, 1]], np.float32)
    pn, _, _ = oracle.normalize_points(p)
    got = sorted(tuple(sorted(int(v) for v in t)) for t in oracle.delaunay_reference(pn))
    assert got == _brute_force_dt(oracle, pn)


def test_reference_bw_restatement_matches_oracle_10k(oracle):
    """The restated reference CPU path on the 10,000-point config-1 input
    (seed 0) gives the oracle's committed canonical set, and on a 6^3
    lattice (SoS everywhere) the oracle's set."""
    import hashlib
    import sys
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
    from wtet_amd import workloads as W
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "inputs_sha256.json")))
    p = W.uniform(10000, 0)
    assert W.sha256(p) == gold["uniform_10000_seed0"]
    pn, _, _ = oracle.normalize_points(p)
    can = oracle.canonical_real(oracle.delaunay_reference(pn), len(p))
    g = gold["oracle_canonical_uniform_10000_seed0"]
    assert len(can) == g["count"]
    assert hashlib.sha256(can.astype("<u4").tobytes()).hexdigest() == g["sha256"]
    q = _grid(6, 6, 6)
    qn, _, _ = oracle.normalize_points(q)
    assert np.array_equal(oracle.canonical_real(oracle.delaunay_reference(qn), len(q)),
                          oracle.canonical_real(oracle.delaunay(qn), len(q)))

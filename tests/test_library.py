"""CPU-side checks of the product library: it loads, exports every symbol
include/wtet.h declares, and its host code (predicates used by the repair
stage, the repair itself, check_quality, the Python surface) agrees with
the oracle.  No device calls."""
import os
import re

import numpy as np
import pytest

import wtet_amd
from wtet_amd import _core

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    txt = open(os.path.join(ROOT, "include", "wtet.h")).read()
    return sorted(set(re.findall(r"\b(wtet_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = _core.lib()
    syms = _declared_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(L, s), s


def test_adjacency_encoding_matches_reference_known_answers():
    import json
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_known_answers.json")))
    L = _core.lib()
    import ctypes
    for c in gold["encode_opp"]:
        p = wtet_amd.encode_adjacency(c["tet"], c["face"])
        assert wtet_amd.decode_adjacency(p) == (c["tet"], c["face"])
        assert L.wtet_encode_adjacency(c["tet"], c["face"]) == p
        t, f = ctypes.c_uint32(), ctypes.c_uint32()
        L.wtet_decode_adjacency(p, ctypes.byref(t), ctypes.byref(f))
        assert (t.value, f.value) == (c["tet"], c["face"])
    assert wtet_amd.INVALID == 2 ** 32 - 1


def _pred_cases(rng, npts, k):
    A = rng.random((k, npts, 3), dtype=np.float32)
    base = rng.random((k, 1, 3), dtype=np.float32)
    B = (base + rng.standard_normal((k, npts, 3)).astype(np.float32) * np.float32(1e-6)).astype(np.float32)
    C = (rng.integers(0, 4, (k, npts, 3)) / np.float32(3)).astype(np.float32)
    D = (rng.integers(0, 128, (k, npts, 3)).astype(np.float32) / np.float32(127)).astype(np.float32)
    S = rng.random((k, npts, 3), dtype=np.float32)
    sup = np.array([[-100, -100, -100], [400, -100, -100], [-100, 400, -100], [-100, -100, 400]], np.float32)
    S[:, 0, :] = sup[rng.integers(0, 4, k)]
    T = (rng.random((k, npts, 3), dtype=np.float32) * np.float32(1e-30)).astype(np.float32)
    T[:, :2] = rng.random((k, 2, 3), dtype=np.float32)
    return np.concatenate([A, B, C, D, S, T])


@pytest.mark.parametrize("mode", [0, 1])
def test_host_predicates_match_oracle(oracle, mode):
    rng = np.random.default_rng(21)
    for kind, npts, ofn in [(0, 4, oracle.orient3d_sos), (1, 5, oracle.orient4d)]:
        q = _pred_cases(rng, npts, 3000)
        idx = np.array([rng.permutation(40)[:npts] for _ in range(len(q))], np.uint32)
        ref = ofn(q.astype(np.float64), idx)
        got = _core.predicates_host(kind, q, idx, mode)
        assert np.array_equal(got, ref)
    for kind, npts, ofn in [(2, 4, oracle.orient3d_sign), (3, 5, oracle.insphere_sign)]:
        q = _pred_cases(rng, npts, 2000)
        assert np.array_equal(_core.predicates_host(kind, q, None, mode), ofn(q.astype(np.float64)))


def _random_valid_flips(tets, rounds, rng):
    """Apply random 2-3 flips to a triangulation (valid, no longer Delaunay)."""
    from collections import defaultdict
    tets = [list(t) for t in tets]
    for _ in range(rounds):
        faces = defaultdict(list)
        for i, t in enumerate(tets):
            for f in range(4):
                faces[tuple(sorted(t[:f] + t[f + 1:]))].append(i)
        shared = [v for v in faces.values() if len(v) == 2]
        i, j = shared[rng.integers(len(shared))]
        yield i, j, tets


@pytest.mark.parametrize("npts,nflips", [(400, 40), (3000, 300)])
def test_repair_restores_delaunay_from_perturbed_mesh(oracle, npts, nflips):
    """Flip random interior faces of the exact triangulation with 2-3 flips
    wherever geometrically valid (a non-Delaunay but valid mesh), then the
    host repair must return exactly the oracle's tet set.  With 300 flips
    spread over 3,000 points the repair plans many separate regions in one
    round on several threads, and some plans meet an earlier commit."""
    rng = np.random.default_rng(3)
    p = rng.random((npts, 3), dtype=np.float32)
    pn, _, _ = oracle.normalize_points(p)
    n = len(pn)
    ref = oracle.delaunay(pn)
    ext = oracle.extended_points(pn)
    tets = [list(map(int, t)) for t in ref]
    flipped = 0
    for _ in range(10 * nflips):
        if flipped >= nflips:
            break
        adj = oracle.rebuild_adjacency(np.array(tets, np.uint32))
        t = int(rng.integers(len(tets)))
        f = int(rng.integers(4))
        a = int(adj[t, f])
        if a == 0xFFFFFFFF:
            continue
        u, g = a >> 5, a & 3
        T, U = tets[t], tets[u]
        apex, o = T[f], U[g]
        face = [T[k] for k in ((1, 3, 2), (0, 2, 3), (0, 3, 1), (0, 1, 2))[f]]
        ok = True
        new = []
        for k in range(3):
            x, y = face[k], face[(k + 1) % 3]
            s = oracle.orient3d_sos(ext[[x, y, apex, o]][None], np.array([[x, y, apex, o]], np.uint32))[0]
            if s != -1:
                ok = False
                break
            new.append([x, y, o, apex])
        if not ok:
            continue
        for idx in sorted([t, u], reverse=True):
            tets.pop(idx)
        tets.extend(new)
        flipped += 1
    assert flipped >= nflips // 4
    bad = np.array(tets, np.uint32)
    adjb = oracle.rebuild_adjacency(bad)
    assert oracle.count_violations(pn, bad, adjb, only_real=False) > 0
    out_t, out_a, regions = _core.repair(pn, bad, None)
    assert regions >= (1 if nflips < 100 else 20), regions
    got = oracle.canonical_real(out_t, n)
    assert np.array_equal(got, oracle.canonical_real(ref, n))
    assert oracle.count_violations(pn, out_t, oracle.rebuild_adjacency(out_t), only_real=False) == 0


@pytest.mark.parametrize("mode", [0, 1])
def test_repair_inserts_points_no_tet_uses(oracle, mode):
    """Points absent from the mesh (the insertion rounds stopped at
    max_insert_iterations, src/phase1.rs:26) are inserted by the host phase:
    from the super tet alone, and from the exact triangulation of the first
    half of the points.  In 3D a cavity may have fewer boundary faces than
    tets; the slots it leaves must not survive as stale tets, so the tet set
    and the adjacency are checked against the oracle's."""
    rng = np.random.default_rng(11 + mode)
    p = np.concatenate([rng.random((3000, 3), dtype=np.float32),
                        (0.5 + 0.05 * rng.standard_normal((1000, 3))).astype(np.float32)])
    pn, _, _ = oracle.normalize_points(p)
    n = len(pn)
    ref = oracle.canonical_real(oracle.delaunay(pn), n)
    k = n // 2
    half = oracle.delaunay(pn[:k]).astype(np.int64)
    half[half >= k] += n - k  # super vertices k..k+3 -> n..n+3
    for start in (np.array([[n, n + 1, n + 2, n + 3]], np.uint32), half.astype(np.uint32)):
        out_t, out_a, _ = _core.repair(pn, start, None, mode=mode)
        assert np.array_equal(oracle.canonical_real(out_t, n), ref)
        m = len(out_t)
        assert np.all(out_t < n + 4)
        assert np.array_equal(np.sort(out_a.ravel()), np.sort(oracle.rebuild_adjacency(out_t).ravel()))
        a = out_a.ravel()
        own = (np.repeat(np.arange(m), 4) << 5) | np.tile(np.arange(4), m)
        ok = a != 0xFFFFFFFF  # the super tet's outer faces have no neighbour
        assert ok.sum() == 4 * m - 4
        back = out_a[a[ok] >> 5, a[ok] & 3]
        assert np.array_equal(back, own[ok])


def test_check_quality_counts_like_reference(oracle):
    rng = np.random.default_rng(9)
    p = rng.random((300, 3), dtype=np.float32)
    pn, _, _ = oracle.normalize_points(p)
    tets = oracle.delaunay(pn)
    adj = oracle.rebuild_adjacency(tets)
    res = wtet_amd.DelaunayResult(tets, adj, np.zeros(0, np.uint32))
    v, interior = wtet_amd.check_quality(p, res)
    assert v == 0
    assert interior == len(oracle.canonical_real(tets, len(p)))


def test_python_surface_errors_and_reprs():
    with pytest.raises(ValueError, match=r"Points must have shape \(N, 3\), got \[5, 2\]"):
        wtet_amd.delaunay(np.zeros((5, 2), np.float32))
    with pytest.raises(ValueError, match="Need at least 4 points for 3D Delaunay, got 3"):
        wtet_amd.delaunay(np.zeros((3, 3), np.float32))
    with pytest.raises(ValueError, match="Invalid insertion_rule 'midpoint'"):
        wtet_amd.Config(insertion_rule="midpoint")
    c = wtet_amd.Config()
    assert repr(c) == ("Config(insertion_rule='circumcenter', flipping=true, sorting=false, "
                       "hilbert=false, splaying=true, max_insert_iter=100, max_flip_iter=10)")
    r = wtet_amd.DelaunayResult(np.zeros((7, 4), np.uint32), np.zeros((7, 4), np.uint32),
                                np.zeros(0, np.uint32))
    assert repr(r) == "DelaunayResult(num_tets=7, failed_verts=0)"
    assert wtet_amd.__version__ == "0.1.0"


def test_batch_input_validation():
    from wtet_amd import _core
    with pytest.raises(ValueError, match="at least 4 points"):
        _core._batch_input([np.zeros((3, 3), np.float32)], None)
    with pytest.raises(ValueError, match=r"offsets\[-1\]"):
        _core._batch_input(np.zeros((10, 3), np.float32), np.array([0, 5, 9]))
    with pytest.raises(ValueError, match="nondecreasing"):
        _core._check_offsets([0, 8, 4])
    with pytest.raises(ValueError, match="starting at 0"):
        _core._check_offsets([1, 8])
    with pytest.raises(ValueError, match=r"shape \(N, 3\)"):
        _core._batch_input([np.zeros((5, 2), np.float32)], None)
    pts, off = _core._batch_input([np.zeros((4, 3)), np.ones((5, 3))], None)
    assert off.tolist() == [0, 4, 9] and pts.dtype == np.float32 and pts.shape == (9, 3)


def test_mixed_batch_workload_shape():
    from wtet_amd import workloads as W
    pts, off = W.mixed_batch(64, 9)
    sizes = np.diff(off.astype(np.int64))
    assert len(off) == 65 and off[0] == 0 and int(off[-1]) == len(pts)
    assert sizes.min() >= 64 and sizes.max() <= 50000
    assert pts.dtype == np.float32 and pts.min() >= 0 and pts.max() <= 1


def test_insertion_order_matches_reference_known_answers(oracle):
    """The library's stride and stratified order (the code the device runs)
    against the reference's known answers (src/gpu/mod.rs:606-696, as data
    in tests/golden) and the oracle's restatement of the same functions."""
    import json
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                       "reference_known_answers.json")))
    for c in gold["initial_stride"]:
        assert _core.initial_stride(c["n"]) == c["stride"], c["src"]
    for c in gold["stratified_order"]:
        rank = _core.stratified_ranks(c["n"], c["stride"])
        order = np.empty(c["n"], np.int64)
        order[rank] = np.arange(c["n"])
        if "order" in c:
            assert order.tolist() == c["order"], c["src"]
        else:
            assert order[: len(c["prefix"])].tolist() == c["prefix"], c["src"]
    for n in (1, 2, 3, 17, 100, 1000, 4097, 65536, 100001, 250000):
        s = oracle.initial_stride(n)
        assert _core.initial_stride(n) == s
        rank = _core.stratified_ranks(n, s)
        order = np.empty(n, np.int64)
        order[rank] = np.arange(n)
        assert np.array_equal(order, oracle.stratified_order(n, s)), n


def test_hilbert_index_is_a_hilbert_curve():
    """The defining properties on the whole 256^3 grid: a bijection onto
    0..2^24-1 whose consecutive indices are face-neighbour cells, starting at
    the origin; plus the reference's own checks (unique corners, locality;
    src/hilbert.rs:45-120)."""
    g = np.arange(256, dtype=np.uint32)
    x, y, z = np.meshgrid(g, g, g, indexing="ij")
    cells = np.stack([x.ravel(), y.ravel(), z.ravel()], 1)
    h = _core.hilbert_index(cells)
    inv = np.empty(1 << 24, np.int64)
    inv[h] = np.arange(1 << 24)
    assert np.array_equal(np.sort(h), np.arange(1 << 24, dtype=np.uint32))
    walk = cells[inv].astype(np.int64)
    assert np.all(np.abs(np.diff(walk, axis=0)).sum(1) == 1)
    assert walk[0].tolist() == [0, 0, 0]
    corners = np.array([[a, b, c] for a in (0, 255) for b in (0, 255) for c in (0, 255)])
    assert len(set(_core.hilbert_index(corners).tolist())) == 8
    q = lambda v: np.rint(np.array(v, np.float32) * np.float32(255)).astype(np.uint32)  # noqa: E731
    pts = [q([.5, .5, .5]), q([.51, .5, .5]), q([.5, .51, .5]), q([.5, .5, .51]), q([.1, .1, .1]), q([.9, .9, .9])]
    hh = _core.hilbert_index(np.array(pts)).astype(np.int64)
    near = np.abs(hh[1:4] - hh[0]).mean()
    far = np.abs(hh[4:6] - hh[0]).mean()
    assert len(set(hh.tolist())) == 6 and near < far


@pytest.mark.parametrize("mode", [0, 1])
def test_host_insphere_far_and_clustered(oracle, mode):
    """insphere on sets mixing the far super vertices with points clustered
    within 1e-7..1e-3 of each other: the FP64 filter's bound is dominated by
    the far points, so these reach the adaptive (double-double) stage and
    the exact stage; every sign equals the oracle's."""
    rng = np.random.default_rng(77)
    sup = np.array([[-100, -100, -100], [400, -100, -100], [-100, 400, -100], [-100, -100, 400]], np.float32)
    qs = []
    for k in range(6000):
        base = rng.random(3).astype(np.float32)
        spread = np.float32(10.0 ** rng.uniform(-7, -3))
        q = (base + rng.standard_normal((5, 3)).astype(np.float32) * spread).astype(np.float32)
        nsup = int(rng.integers(0, 4))
        pick = rng.permutation(4)[:nsup]
        q[rng.permutation(5)[:nsup]] = sup[pick]
        qs.append(q)
    q = np.stack(qs)
    idx = np.array([rng.permutation(40)[:5] for _ in range(len(q))], np.uint32)
    assert np.array_equal(_core.predicates_host(1, q, idx, mode), oracle.orient4d(q.astype(np.float64), idx))
    assert np.array_equal(_core.predicates_host(3, q, None, mode), oracle.insphere_sign(q.astype(np.float64)))

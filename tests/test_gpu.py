"""GPU parity tests: the HIP path (through the C ABI) against the oracle.

Bar: the canonical real-only tet set (sorted quads, sorted list) equals the
oracle's exactly; the device predicates equal the oracle's signs exactly.
At sizes the oracle cannot finish quickly the size-independent properties
are checked instead (no insphere violation on any face, exact counts).
"""
import hashlib
import json
import os

import numpy as np
import pytest

import wtet_amd
from wtet_amd import _core
from wtet_amd import workloads as W

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "inputs_sha256.json")))


@pytest.fixture(scope="module")
def ctx():
    c = wtet_amd.Context(0)
    yield c
    c.close()


def _check(oracle, ctx, pts, mode="filtered", **kw):
    n = len(pts)
    res = ctx.delaunay(pts, wtet_amd.Config(predicate_mode=mode, collect_stats=True, **kw))
    pn, _, _ = oracle.normalize_points(pts)
    ref = oracle.canonical_real(oracle.delaunay(pn), n)
    got = oracle.canonical_real(res.tets, n)
    assert got.shape == ref.shape, (got.shape, ref.shape, res.stats)
    assert np.array_equal(got, ref), res.stats
    # adjacency is symmetric and every face has a neighbour (infinity tets close the mesh)
    adj = res.adjacency
    t = np.repeat(np.arange(len(adj)), 4)
    f = np.tile(np.arange(4), len(adj))
    a = adj.ravel()
    assert np.all(a != wtet_amd.INVALID)
    back = adj[a >> 5, a & 3]
    assert np.array_equal(back, (t << 5) | f)
    return res


# -- predicates ------------------------------------------------------------------

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
    # corners of small cubes of a 128^3 lattice: co-spherical and coplanar
    # subsets, so SoS runs deep (the lifted steps on coplanar points) and
    # the short-difference exact tier applies
    base = rng.integers(0, 120, (k, 1, 3)).astype(np.float32)
    step = rng.integers(1, 8, (k, 1, 1)).astype(np.float32)
    E = ((base + rng.integers(0, 2, (k, npts, 3)).astype(np.float32) * step) / np.float32(127)).astype(np.float32)
    # all points in one axis plane: every insphere vanishes and SoS is
    # decided by the lifted steps on coplanar points (depths 2-4)
    F = (rng.integers(0, 128, (k, npts, 3)).astype(np.float32) / np.float32(127)).astype(np.float32)
    ax = rng.integers(0, 3, k)
    F[np.arange(k), :, ax] = F[np.arange(k), :1, ax]
    return np.concatenate([A, B, C, D, S, T, E, F])


@pytest.mark.parametrize("mode", [0, 1])
def test_device_predicates_match_oracle(oracle, ctx, mode):
    rng = np.random.default_rng(31)
    for kind, npts, ofn in [(0, 4, oracle.orient3d_sos), (1, 5, oracle.orient4d)]:
        q = _pred_cases(rng, npts, 20000)
        idx = np.array([rng.permutation(40)[:npts] for _ in range(len(q))], np.uint32)
        ref = ofn(q.astype(np.float64), idx)
        got, st = ctx.predicates(kind, q, idx, mode)
        assert np.array_equal(got, ref), int((got != ref).sum())
        base = 0 if kind == 0 else 4
        assert int(st[base:base + 4].sum()) == len(q)  # every query counted once
        if mode == 1:
            assert st[base] == 0 and st[base + 1] == 0  # exact-only never filters
    for kind, npts, ofn in [(2, 4, oracle.orient3d_sign), (3, 5, oracle.insphere_sign)]:
        q = _pred_cases(rng, npts, 10000)
        got, _ = ctx.predicates(kind, q, None, mode)
        assert np.array_equal(got, ofn(q.astype(np.float64)))


# -- triangulations ----------------------------------------------------------------

def test_config1_uniform_10k_exact_path(oracle, ctx):
    """10,000 U[0,1)^3 f32 seed 0, Hilbert insertion order, exact predicates only."""
    p = W.uniform(10000, 0)
    assert W.sha256(p) == GOLD["uniform_10000_seed0"]
    res = _check(oracle, ctx, p, mode="exact", enable_hilbert_sorting=True)
    can = oracle.canonical_real(res.tets, len(p))
    g = GOLD["oracle_canonical_uniform_10000_seed0"]
    assert len(can) == g["count"]
    assert hashlib.sha256(can.astype("<u4").tobytes()).hexdigest() == g["sha256"]
    st = res.stats
    assert st["pred"][0] == st["pred"][1] == st["pred"][4] == st["pred"][5] == 0


@pytest.mark.parametrize("n,seed", [(4, 1), (5, 2), (20, 3), (100, 4), (1000, 5), (10000, 0), (50000, 7)])
def test_uniform_filtered(oracle, ctx, n, seed):
    _check(oracle, ctx, W.uniform(n, seed))


@pytest.mark.parametrize("k", [2, 3, 4, 8, 16])
def test_lattice_degenerate(oracle, ctx, k):
    res = _check(oracle, ctx, W.lattice(k))
    if k == 16:
        can = oracle.canonical_real(res.tets, k ** 3)
        g = GOLD["oracle_canonical_lattice_16"]
        assert len(can) == g["count"]
        assert hashlib.sha256(can.astype("<u4").tobytes()).hexdigest() == g["sha256"]


@pytest.mark.parametrize("kind", ["sphere", "shell", "clusters", "duplicates", "plane", "line",
                                  "scaled", "cospherical"])
def test_special_sets(oracle, ctx, kind):
    rng = np.random.default_rng(17)
    if kind == "sphere":
        p = W.sphere(4000, 1)
    elif kind == "shell":
        p = W.shell(3000, 2)
    elif kind == "clusters":
        p = W.clusters(5000, 3)
    elif kind == "duplicates":
        q = rng.random((400, 3), dtype=np.float32)
        p = np.vstack([q, q[:200]])
    elif kind == "plane":
        p = np.hstack([rng.random((600, 2), dtype=np.float32), np.full((600, 1), 0.25, np.float32)])
    elif kind == "line":
        p = np.hstack([rng.random((60, 1), dtype=np.float32), np.zeros((60, 2), np.float32)])
    elif kind == "scaled":
        p = (rng.standard_normal((3000, 3)) * 1e4 + 5e5).astype(np.float32)
    else:
        p = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
                      [0.6, 0.8, 0], [0, 0.6, 0.8], [0.8, 0, 0.6]], np.float32)
    _check(oracle, ctx, p)


def test_exact_and_filtered_agree_on_lattice(oracle, ctx):
    p = W.lattice(12)
    a = ctx.delaunay(p, wtet_amd.Config(predicate_mode="exact"))
    b = ctx.delaunay(p, wtet_amd.Config(predicate_mode="filtered"))
    assert np.array_equal(oracle.canonical_real(a.tets, len(p)), oracle.canonical_real(b.tets, len(p)))


def test_result_arrays_outlive_the_result(oracle, ctx):
    """The tets and adjacency arrays of a host result use the library's
    memory without a copy: they stay valid, unchanged and writable after the
    result object is gone and other calls have allocated and freed results,
    and a view keeps the memory alive on its own."""
    import gc
    p = W.uniform(3000, 7)
    r = ctx.delaunay(p)
    tets, adj = r.tets, r.adjacency
    want_t, want_a = tets.copy(), adj.copy()
    view = r.tets[10:20]
    want_v = view.copy()
    del r
    gc.collect()
    for s in range(4):
        ctx.delaunay(W.uniform(3500 + 500 * s, 100 + s))
        gc.collect()
    assert np.array_equal(tets, want_t) and np.array_equal(adj, want_a)
    del tets
    gc.collect()
    ctx.delaunay(W.uniform(4000, 9))
    assert np.array_equal(view, want_v)
    assert adj.flags.writeable and adj.dtype == np.uint32 and adj.flags.c_contiguous
    n = len(p)
    pn, _, _ = oracle.normalize_points(p)
    assert np.array_equal(oracle.canonical_real(want_t, n), oracle.canonical_real(oracle.delaunay(pn), n))


def test_uniform_200k_matches_oracle(oracle, ctx):
    _check(oracle, ctx, W.uniform(200000, 11))


def _assert_gold_set(oracle, tets, n, name):
    """The canonical real-only set equals the oracle's, committed as count +
    SHA-256 by tests/golden/make_fullsize_fixtures.py."""
    g = GOLD["oracle_canonical_" + name]
    can = oracle.canonical_real(tets, n)
    assert len(can) == g["count"], (len(can), g["count"])
    assert hashlib.sha256(can.astype("<u4").tobytes()).hexdigest() == g["sha256"]


def test_uniform_1m_matches_oracle(oracle, ctx):
    """1,000,000 U[0,1)^3 f32 points, seed 1: the canonical set equals the
    oracle's (committed fixture); no face fails the insphere test and the
    counts are consistent."""
    p = W.uniform(1000000, 1)
    assert W.sha256(p) == GOLD["uniform_1000000_seed1"]
    res = ctx.delaunay(p, wtet_amd.Config(collect_stats=True))
    _assert_gold_set(oracle, res.tets, len(p), "uniform_1000000_seed1")
    v, interior = wtet_amd.check_quality(p, res)
    assert v == 0
    assert interior > 6 * len(p)
    ids = np.unique(res.tets)
    assert ids[0] == 0 and len(ids) == len(p) + 5  # every point, 4 super, infinity
    # the exact-only path gives the same set
    ex = ctx.delaunay(p, wtet_amd.Config(predicate_mode="exact"))
    _assert_gold_set(oracle, ex.tets, len(p), "uniform_1000000_seed1")


def test_context_reuse_and_smaller_second_call(oracle, ctx):
    _check(oracle, ctx, W.uniform(20000, 21))
    _check(oracle, ctx, W.uniform(3000, 22))


def test_module_level_api(oracle):
    p = W.uniform(2000, 5)
    r = wtet_amd.delaunay(p)
    assert r.tets.dtype == np.uint32 and r.adjacency.shape == r.tets.shape
    assert r.failed_verts.dtype == np.uint32
    assert "GPU" in wtet_amd.initialize_gpu()
    assert "GPU Device" in wtet_amd.gpu_info()
    v, _ = wtet_amd.check_quality(p, r)
    assert v == 0


def test_resident_matches_host_result(oracle, ctx):
    """Device-resident points and result: same mesh as the host-result call."""
    p = W.uniform(30000, 8)
    d = ctx.upload(p)
    try:
        m, dt, da = ctx.delaunay_resident(d, len(p))
        tets = ctx.download(dt, (m, 4), np.uint32)
        adj = ctx.download(da, (m, 4), np.uint32)
    finally:
        ctx.free(d)
    host = ctx.delaunay(p)
    assert m == host.num_tets
    assert np.array_equal(oracle.canonical_real(tets, len(p)), oracle.canonical_real(host.tets, len(p)))
    back = adj[(adj >> 5).ravel(), (adj & 3).ravel()]
    idx = (np.repeat(np.arange(m), 4) << 5) | np.tile(np.arange(4), m)
    assert np.array_equal(back, idx)


def test_flip_graph_replays_match_direct_launches(oracle, ctx):
    """Each round replays the captured graph of its flip iterations, with
    eager launches and host reads between replays; every call must equal
    the direct-launch result and the oracle."""
    p = W.uniform(50000, 61)
    ref = oracle.canonical_real(_check(oracle, ctx, p).tets, len(p))
    ctx.set_flip_graphs(False)
    try:
        direct = ctx.delaunay(p)
    finally:
        ctx.set_flip_graphs(True)
    assert np.array_equal(oracle.canonical_real(direct.tets, len(p)), ref)
    _, before = ctx.flip_items()
    small = W.uniform(7000, 62)
    small_ref = oracle.canonical_real(oracle.delaunay(oracle.normalize_points(small)[0]), len(small))
    for q, want in ((p, ref), (small, small_ref), (p, ref)):  # the smaller call recaptures
        r = ctx.delaunay(q)
        assert np.array_equal(oracle.canonical_real(r.tets, len(q)), want)
    _, after = ctx.flip_items()
    assert after > before


@pytest.mark.parametrize("items", [1, 8, 64])
def test_flip_grid_sizes(oracle, ctx, items):
    i0, _ = ctx.flip_items()
    ctx.set_flip_items(items)
    try:
        _check(oracle, ctx, W.uniform(20000, 63))
    finally:
        ctx.set_flip_items(i0)


def test_autotune_picks_a_candidate(oracle, ctx):
    p = W.uniform(100000, 64)
    d = ctx.upload(p)
    i0, _ = ctx.flip_items()
    try:
        g, ms = ctx.autotune(d, [0, len(p)])
        assert g in (1, 2, 4, 8) and ms > 0
        assert ctx.flip_items()[0] == g
    finally:
        ctx.free(d)
        ctx.set_flip_items(i0)
    _check(oracle, ctx, p)


@pytest.mark.parametrize("kind", ["sphere", "uniform"])
def test_resident_repair_on_neighbourhood(oracle, ctx, kind):
    """A resident result is repaired on the downloaded neighbourhood of the
    stuck faces only; the mesh equals the host-result call's."""
    p = W.sphere(200000, 71) if kind == "sphere" else W.uniform(300000, 72)
    d = ctx.upload(p)
    try:
        m, dt, da = ctx.delaunay_resident(d, len(p))
        st = ctx.stats()
        tets = ctx.download(dt, (m, 4), np.uint32)
        adj = ctx.download(da, (m, 4), np.uint32)
    finally:
        ctx.free(d)
    assert st["violations_gpu"] > 0, st
    assert 0 < st["repair_rows_staged"] < m or st["repair_full_downloads"] == 1, st
    host = ctx.delaunay(p)
    assert np.array_equal(oracle.canonical_real(tets, len(p)), oracle.canonical_real(host.tets, len(p)))
    back = adj[(adj >> 5).ravel(), (adj & 3).ravel()]
    assert np.array_equal(back, (np.repeat(np.arange(m), 4) << 5) | np.tile(np.arange(4), m))
    v, _ = wtet_amd.check_quality(p, wtet_amd.DelaunayResult(tets, adj, np.zeros(0, np.uint32)))
    assert v == 0


@pytest.mark.parametrize("name", ["lattice_128", "sphere_1000000_seed2"])
def test_degenerate_full_size_matches_oracle(oracle, ctx, name):
    """Degenerate inputs at full size: the 128^3 lattice (2,097,152 points,
    co-spherical everywhere: exact stage and SoS) and 1M points on the unit
    sphere.  The canonical set equals the oracle's (committed fixture); the
    reference's own audit (check_quality, src/lib.rs:25-78) finds no
    violation, every point is a vertex, and the adjacency is closed."""
    p = W.lattice(128) if name == "lattice_128" else W.sphere(1_000_000, 2)
    assert W.sha256(p) == GOLD[name]
    res = ctx.delaunay(p, wtet_amd.Config(collect_stats=True))
    st = res.stats
    _assert_gold_set(oracle, res.tets, len(p), name)
    v, interior = wtet_amd.check_quality(p, res)
    assert v == 0, (v, st)
    assert interior > len(p)
    ids = np.unique(res.tets)
    assert len(ids) == len(p) + 5 and ids[0] == 0
    adj = res.adjacency
    m = len(adj)
    back = adj[(adj >> 5).ravel(), (adj & 3).ravel()]
    assert np.array_equal(back, (np.repeat(np.arange(m), 4) << 5) | np.tile(np.arange(4), m))
    if name == "lattice_128":
        assert st["pred"][7] > 0  # SoS decided some insphere calls


@pytest.mark.parametrize("case", ["uniform", "lattice", "duplicates", "clusters"])
def test_hilbert_insertion_order(oracle, ctx, case):
    """enable_hilbert_sorting: the reference's 8-bit Hilbert keys and
    stratified insertion order (src/gpu/mod.rs:473-600).  The order cannot
    change the triangulation: same set as the oracle."""
    if case == "uniform":
        p = W.uniform(30000, 81)
    elif case == "lattice":
        p = W.lattice(12)
    elif case == "duplicates":
        u = W.uniform(5000, 82)
        p = np.concatenate([u, u[:2500], u[:100]])
    else:
        p = W.clusters(20000, 83)
    r = _check(oracle, ctx, p, enable_hilbert_sorting=True)
    if case == "duplicates":
        assert r.stats["duplicates"] == 2600


def test_hilbert_order_in_a_batch(oracle, ctx):
    clouds = [W.uniform(3000, 84), W.lattice(6), W.clusters(2000, 85), W.uniform(5, 86)]
    res = ctx.delaunay_batch(clouds, config=wtet_amd.Config(enable_hilbert_sorting=True))
    for c, p in enumerate(clouds):
        pn, _, _ = oracle.normalize_points(p)
        ref = oracle.canonical_real(oracle.delaunay(pn), len(p))
        assert np.array_equal(oracle.canonical_real(res[c].tets, len(p)), ref), c


def test_oversized_repair_region_reruns_with_wider_ring(oracle, ctx):
    """A repair region past the cap stops the host repair and reruns the
    device phase with the insertion ring one wider; the rerun's repair has
    no cap.  A cap of one tet forces it whenever the GPU leaves a stuck face
    (200,000 points on a sphere always leave some).  The host-result and the
    resident calls both give the oracle's tets."""
    p = W.sphere(200000, 5)
    n = len(p)
    d = ctx.upload(p)
    ctx.set_repair_region_cap(1)
    try:
        forced = ctx.delaunay(p)
        st = ctx.stats()
        m, dt, da = ctx.delaunay_resident(d, n)
        st_r = ctx.stats()
        tets = ctx.download(dt, (m, 4), np.uint32)
        adj = ctx.download(da, (m, 4), np.uint32)
    finally:
        ctx.set_repair_region_cap(1 << 14)
        ctx.free(d)
    assert st["repair_retries"] == 1 and st_r["repair_retries"] == 1, (st, st_r)
    pn, _, _ = oracle.normalize_points(p)
    want = oracle.canonical_real(oracle.delaunay(pn), n)
    assert np.array_equal(oracle.canonical_real(forced.tets, n), want)
    assert np.array_equal(oracle.canonical_real(tets, n), want)
    back = adj[(adj >> 5).ravel(), (adj & 3).ravel()]
    assert np.array_equal(back, (np.repeat(np.arange(m), 4) << 5) | np.tile(np.arange(4), m))


@pytest.mark.parametrize("rings", [1, 2])
def test_resident_repair_fetches_missing_rows(oracle, ctx, rings):
    """With a thin staged neighbourhood the repair reaches rows the host
    lacks; it fetches the rows around them and goes on (after a few fetches
    it takes every row instead).  The mesh equals the host-result call's."""
    p = W.uniform(300000, 72)
    d = ctx.upload(p)
    ctx.set_repair_rings(rings)
    try:
        m, dt, da = ctx.delaunay_resident(d, len(p))
        st = ctx.stats()
        tets = ctx.download(dt, (m, 4), np.uint32)
        adj = ctx.download(da, (m, 4), np.uint32)
    finally:
        ctx.set_repair_rings(7)
        ctx.free(d)
    assert st["violations_gpu"] > 0, st
    if rings == 1:  # one ring is never enough for this input's regions
        assert st["repair_fetches"] > 0, st
    host = ctx.delaunay(p)
    assert np.array_equal(oracle.canonical_real(tets, len(p)), oracle.canonical_real(host.tets, len(p)))
    back = adj[(adj >> 5).ravel(), (adj & 3).ravel()]
    assert np.array_equal(back, (np.repeat(np.arange(m), 4) << 5) | np.tile(np.arange(4), m))


@pytest.mark.parametrize("cap", [0, 1, 3])
def test_insertion_cap_then_host_insertion(oracle, ctx, cap):
    """max_insert_iterations stops the insertion rounds (src/phase1.rs:26;
    the reference's own test runs max_insert_iterations=1,
    src/tests.rs:1583-1595); the host phase inserts the points they left, so
    the set still equals the oracle's, and failed_verts lists those points.
    Host result, resident result and a batch all agree."""
    p = W.uniform(4000, 81)
    n = len(p)
    res = _check(oracle, ctx, p, max_insert_iterations=cap)
    st = res.stats
    assert st["rounds"] == cap and st["points_left"] > 0, st
    assert len(res.failed_verts) >= st["points_left"] - st["duplicates"]
    assert np.all(np.diff(res.failed_verts.astype(np.int64)) > 0)
    ref = oracle.canonical_real(res.tets, n)
    d = ctx.upload(p)
    try:
        m, dt, da = ctx.delaunay_resident(d, n, wtet_amd.Config(max_insert_iterations=cap))
        tets = ctx.download(dt, (m, 4), np.uint32)
        adj = ctx.download(da, (m, 4), np.uint32)
    finally:
        ctx.free(d)
    assert np.array_equal(oracle.canonical_real(tets, n), ref)
    back = adj[(adj >> 5).ravel(), (adj & 3).ravel()]
    assert np.array_equal(back, (np.repeat(np.arange(m), 4) << 5) | np.tile(np.arange(4), m))
    q = W.clusters(2500, 82)
    b = ctx.delaunay_batch([p, q], config=wtet_amd.Config(max_insert_iterations=cap))
    assert np.array_equal(oracle.canonical_real(b[0].tets, n), ref)
    qn = oracle.normalize_points(q)[0]
    assert np.array_equal(oracle.canonical_real(b[1].tets, len(q)),
                          oracle.canonical_real(oracle.delaunay(qn), len(q)))


def test_insertion_cap_without_phase2_leaves_points_out(oracle, ctx):
    """With enable_splaying off nothing finishes the mesh (src/lib.rs:227
    skips phase 2): the points the rounds left are absent from the tets and
    listed in failed_verts."""
    p = W.uniform(3000, 83)
    res = ctx.delaunay(p, wtet_amd.Config(max_insert_iterations=2, enable_splaying=False))
    used = np.unique(res.tets[res.tets < len(p)])
    absent = np.setdiff1d(np.arange(len(p)), used)
    assert len(absent) == res.stats["points_left"] > 0
    assert np.all(np.isin(absent, res.failed_verts))


def test_no_flipping_overflows_audit_records(oracle, ctx):
    """enable_flipping off: every non-Delaunay face the splits leave goes to
    the host repair, about one per point.  With the first audit's record
    buffer cut to 1,024 entries these overflow it; the audit runs again into
    a larger one, the repair starts from every violated face and reaches the
    oracle's set (a host result, a resident result and a batch)."""
    p = W.uniform(20000, 84)
    n = len(p)
    ctx.set_audit_records(1024)
    try:
        res = _check(oracle, ctx, p, enable_flipping=False)
        assert res.stats["violations_gpu"] > 1024, res.stats
        assert res.stats["flip_iterations"] == 0
        ref = oracle.canonical_real(res.tets, n)
        d = ctx.upload(p)
        try:
            m, dt, _ = ctx.delaunay_resident(d, n, wtet_amd.Config(enable_flipping=False))
            tets = ctx.download(dt, (m, 4), np.uint32)
        finally:
            ctx.free(d)
        assert ctx.stats()["violations_gpu"] > 1024
        assert np.array_equal(oracle.canonical_real(tets, n), ref)
        q = W.clusters(3000, 85)
        b = ctx.delaunay_batch([p, q], config=wtet_amd.Config(enable_flipping=False))
        assert np.array_equal(oracle.canonical_real(b[0].tets, n), ref)
        qn = oracle.normalize_points(q)[0]
        assert np.array_equal(oracle.canonical_real(b[1].tets, len(q)),
                              oracle.canonical_real(oracle.delaunay(qn), len(q)))
    finally:
        ctx.set_audit_records(1 << 20)

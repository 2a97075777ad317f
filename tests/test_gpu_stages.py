"""Stage-level entry points (wtet_stage_*): the device loop driven one
dispatch group at a time, as the reference's tests drive phase1::run
(src/tests.rs:305-306, 383-384, 624-637, 814-816).

Bar: the loop through the stages gives the oracle's canonical tet set, the
same as one wtet_delaunay call; after every stage the current mesh is a
valid closed triangulation (symmetric adjacency, negative orientation of
every finite tet, Euler characteristic 0 of the S^3 closure), the counts
move as the stage says, every uninserted point lies in the tet it is
recorded in (exact SoS orient3d by the oracle), and the device audit counts
the oracle's non-Delaunay faces of the same mesh.
"""
import numpy as np
import pytest

import wtet_amd
from wtet_amd import stages
from wtet_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = wtet_amd.Context(0)
    yield c
    c.close()


def _mesh_checks(oracle, P, n, res):
    """Closed S^3 triangulation: adjacency, orientation, Euler."""
    tets, adj = res.tets.astype(np.int64), res.adjacency
    m = len(tets)
    a = adj.ravel().astype(np.int64)
    assert np.all(a != wtet_amd.INVALID)
    t = np.repeat(np.arange(m), 4)
    f = np.tile(np.arange(4), m)
    assert np.all((a >> 5) < m)
    assert np.array_equal(adj[a >> 5, a & 3].astype(np.int64), (t << 5) | f)
    # neighbours across a face share its three vertices
    for k in range(4):
        nb = tets[a[k::4] >> 5]
        face = np.delete(tets, k, axis=1)
        for j in range(3):
            assert np.all((nb == face[:, j:j + 1]).any(axis=1))
    # finite tets: raw orient3d < 0
    fin = np.all(tets < n + 4, axis=1)
    q = P[tets[fin]]
    assert np.all(oracle.orient3d_sign(q) < 0)
    # Euler characteristic of the closed 3-sphere: V - E + F - T = 0
    verts = np.unique(tets)
    faces = {tuple(sorted(x)) for tt in tets for x in (tt[[1, 2, 3]], tt[[0, 2, 3]], tt[[0, 1, 3]], tt[[0, 1, 2]])}
    edges = {(x, y) for fa in faces for x, y in ((fa[0], fa[1]), (fa[0], fa[2]), (fa[1], fa[2]))}
    assert len(verts) - len(edges) + len(faces) - m == 0
    assert len(faces) * 2 == 4 * m
    return set(verts.tolist())


def _located(oracle, P, n, ctx, res):
    """Every uninserted point lies in (SoS) the tet it is recorded in."""
    ids, rows = ctx.stage_locations()
    if len(ids) == 0:
        return ids
    tv = res.tets[rows].astype(np.int64)
    assert np.all(tv < n + 4), "a point is recorded in an infinity tet"
    for k in range(4):
        q = tv.copy()
        q[:, k] = ids
        s = oracle.orient3d_sos(P[q], q.astype(np.uint32))
        assert np.all(s == 1), int((s != 1).sum())  # +1: perturbed raw det < 0, as the tet
    return ids


@pytest.mark.parametrize("mode", ["filtered", "exact"])
def test_stage_loop_equals_oracle_and_one_call(oracle, ctx, mode):
    for pts in (W.uniform(4000, 11), W.clusters(3000, 12), W.lattice(7)):
        n = len(pts)
        cfg = wtet_amd.Config(predicate_mode=mode)
        res = stages.run(ctx, pts, cfg)
        pn, _, _ = oracle.normalize_points(pts)
        ref = oracle.canonical_real(oracle.delaunay(pn), n)
        got = oracle.canonical_real(res.tets, n)
        assert np.array_equal(got, ref)
        one = ctx.delaunay(pts, cfg)
        assert np.array_equal(oracle.canonical_real(one.tets, n), got)
        assert sorted(map(tuple, np.sort(res.tets, 1).tolist())) == \
            sorted(map(tuple, np.sort(one.tets, 1).tolist()))


def test_stage_invariants_after_every_stage(oracle, ctx):
    pts = W.clusters(700, 21)
    n = len(pts)
    pn, _, _ = oracle.normalize_points(pts)
    P = oracle.extended_points(pn)
    distinct = len(np.unique(pn, axis=0))  # lower ids of exact duplicates are never inserted
    seen = {"vote": 0, "split": 0, "split_points": 0, "flip": 0, "relocate": 0}
    prev = {}

    def observe(stage, info):
        res = ctx.stage_readback()
        verts = _mesh_checks(oracle, P, n, res)
        real = {v for v in verts if v < n}
        assert len(res.tets) <= info["tet_slots"]
        if stage == "init":
            assert real == set() and info["uninserted"] == n and len(res.tets) == 5
        if stage == "vote":
            assert 0 < info["winners"] <= prev["uninserted"]
            assert info["uninserted"] == prev["uninserted"] - info["winners"]
            assert len(res.tets) == prev["alive"]  # the vote changes no tet
        if stage == "split":
            # each winner: one tet becomes four, the winner joins the mesh
            assert len(res.tets) == prev["alive"] + 3 * info["winners"]
            assert len(real) == len(prev["real"]) + info["winners"]
            assert info["active"] >= 4 * info["winners"]
        if stage == "flip":
            # 2-3 and 3-2 flips change the tet count by one each; no vertex moves
            assert abs(len(res.tets) - prev["alive"]) <= info["flips"]
            assert real == prev["real"]
        if stage in ("split_points", "relocate"):
            ids = _located(oracle, P, n, ctx, res)
            assert len(ids) == info["uninserted"]
            assert set(ids.tolist()).isdisjoint(real)
            assert len(set(ids.tolist()) | real) == distinct
        if stage in seen:
            seen[stage] += 1
        prev.update(info)
        prev["alive"] = len(res.tets)
        prev["real"] = real

    res = stages.run(ctx, pts, wtet_amd.Config(), observe=observe)
    assert all(v > 0 for v in seen.values()), seen
    ref = oracle.canonical_real(oracle.delaunay(pn), n)
    assert np.array_equal(oracle.canonical_real(res.tets, n), ref)


def test_stage_gather_counts_the_oracle_violations(oracle, ctx):
    pts = W.uniform(2500, 31)
    n = len(pts)
    pn, _, _ = oracle.normalize_points(pts)
    info = ctx.stage_init(pts)
    checked = 0
    for _ in range(6):
        ctx.stage_vote()
        ctx.stage_split()
        ctx.stage_split_points()
        # before any flip the new tets are rarely Delaunay
        g = ctx.stage_gather()
        res = ctx.stage_readback()
        want = oracle.count_violations(pn, res.tets, res.adjacency, only_real=False)
        assert g["violations"] == want
        # failed_verts: vertices opposite the violating real-only faces
        if oracle.count_violations(pn, res.tets, res.adjacency, only_real=True) > 0:
            assert len(res.failed_verts) > 0
        assert np.all(res.failed_verts < n)
        checked += want
        for _ in range(40):
            info = ctx.stage_flip()
            if info["active"] == 0:
                break
        info = ctx.stage_relocate()
        if info["uninserted"] == 0:
            break
    assert checked > 0
    # with the flips done the current mesh has no violation
    while info["active"] > 0:
        info = ctx.stage_flip()
    g = ctx.stage_gather()
    res = ctx.stage_readback()
    assert g["violations"] == oracle.count_violations(pn, res.tets, res.adjacency, only_real=False) == 0
    assert len(res.failed_verts) == 0
    ctx.stage_end()


def test_stage_exact_flip_iterations_mix(oracle, ctx):
    pts = W.lattice(6)
    n = len(pts)
    info = ctx.stage_init(pts)
    k = 0
    while info["uninserted"] > 0:
        ctx.stage_vote()
        ctx.stage_split()
        info = ctx.stage_split_points()
        while info["active"] > 0:
            info = ctx.stage_flip(k % 2)  # alternate staged and exact-only checks
            k += 1
        info = ctx.stage_relocate()
    res = ctx.stage_finish()
    pn, _, _ = oracle.normalize_points(pts)
    assert np.array_equal(oracle.canonical_real(res.tets, n), oracle.canonical_real(oracle.delaunay(pn), n))


def test_stage_order_is_enforced(ctx):
    pts = W.uniform(500, 41)
    with pytest.raises(ValueError, match="no stage session"):
        ctx.stage_vote()
    with pytest.raises(ValueError, match="at least 4 points"):
        ctx.stage_init(pts[:3])
    ctx.stage_init(pts)
    with pytest.raises(ValueError, match="out of order"):
        ctx.stage_split()
    with pytest.raises(ValueError, match="out of order"):
        ctx.stage_split_points()
    with pytest.raises(ValueError, match="not inserted"):
        ctx.stage_finish()
    ctx.stage_init(pts)
    ctx.stage_vote()
    with pytest.raises(ValueError, match="out of order"):
        ctx.stage_vote()
    with pytest.raises(ValueError, match="out of order"):
        ctx.stage_flip()
    ctx.stage_split()
    with pytest.raises(ValueError, match="out of order"):
        ctx.stage_relocate()
    ctx.stage_split_points()
    for _ in range(20):
        info = ctx.stage_flip()
        if info["flips"] > 0:
            break
        ctx.stage_relocate()
        ctx.stage_vote()
        ctx.stage_split()
        ctx.stage_split_points()
    assert info["flips"] > 0
    with pytest.raises(ValueError, match="relocate"):
        ctx.stage_vote()
    with pytest.raises(ValueError, match="relocate"):
        ctx.stage_locations()
    ctx.stage_relocate()
    ctx.stage_vote()
    # another triangulation call on the context ends the session
    ctx.delaunay(pts)
    with pytest.raises(ValueError, match="no stage session"):
        ctx.stage_split()

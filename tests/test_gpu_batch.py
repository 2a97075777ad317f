"""GPU parity tests of the batched path: many independent clouds in one pass.

Bar: every cloud's canonical real-only tet set equals the oracle's for that
cloud alone (each cloud normalized on its own, like one delaunay_3d call,
src/lib.rs:160-310), and its adjacency is symmetric within the cloud.
"""
import json
import os

import numpy as np
import pytest

import wtet_amd
from wtet_amd import workloads as W

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = json.load(open(os.path.join(ROOT, "tests", "golden", "inputs_sha256.json")))


@pytest.fixture(scope="module")
def ctx():
    c = wtet_amd.Context(0)
    yield c
    c.close()


def _adjacency_closed(adj):
    m = len(adj)
    a = adj.ravel()
    assert np.all(a != wtet_amd.INVALID)
    assert np.all((a >> 5) < m)
    back = adj[a >> 5, a & 3]
    return np.array_equal(back, (np.repeat(np.arange(m), 4) << 5) | np.tile(np.arange(4), m))


def _oracle_set(oracle, p):
    pn, _, _ = oracle.normalize_points(p)
    return oracle.canonical_real(oracle.delaunay(pn), len(p))


def _mixed_clouds(seed, count):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(count):
        n = int(rng.integers(4, 1500))
        kind = i % 7
        if kind == 0:
            p = W.uniform(n, seed * 1000 + i)
        elif kind == 1:
            p = W.clusters(max(n, 8), seed * 1000 + i)
        elif kind == 2:
            p = W.lattice(int(rng.integers(2, 7)))
        elif kind == 3:  # exact duplicates
            p = W.uniform(n, seed * 1000 + i)
            p = np.concatenate([p, p[: n // 2]])
        elif kind == 4:
            p = W.sphere(max(n, 8), seed * 1000 + i)
        elif kind == 5:  # far from the unit cube, anisotropic
            p = (W.uniform(n, seed * 1000 + i) * np.float32([100, 1, 0.01]) - 50).astype(np.float32)
        else:  # the smallest clouds
            p = W.uniform(int(rng.integers(4, 9)), seed * 1000 + i)
        out.append(np.ascontiguousarray(p, np.float32))
    return out


@pytest.mark.parametrize("mode", ["filtered", "exact"])
def test_batch_matches_oracle_per_cloud(oracle, ctx, mode):
    clouds = _mixed_clouds(3, 70)
    res = ctx.delaunay_batch(clouds, config=wtet_amd.Config(predicate_mode=mode))
    assert res.num_clouds == len(clouds)
    for c, p in enumerate(clouds):
        r = res[c]
        assert np.array_equal(oracle.canonical_real(r.tets, len(p)), _oracle_set(oracle, p)), c
        assert _adjacency_closed(r.adjacency), c
        assert int(r.tets.max()) == len(p) + 4  # cloud-local ids, infinity n_c + 4


def test_batch_of_one_equals_single_call(oracle, ctx):
    p = W.uniform(20000, 41)
    b = ctx.delaunay_batch([p])
    s = ctx.delaunay(p)
    assert np.array_equal(oracle.canonical_real(b[0].tets, len(p)), oracle.canonical_real(s.tets, len(p)))
    assert b.tets.shape[0] == s.num_tets


def test_batch_result_views_outlive_the_result(oracle, ctx):
    """Batch tets and adjacency use the library's memory without a copy: a
    cloud's views stay valid and unchanged after the batch result is gone
    and other batches have allocated and freed theirs."""
    import gc
    clouds = [W.uniform(800 + 300 * k, 140 + k) for k in range(4)]
    res = ctx.delaunay_batch(clouds)
    c2 = res[2]
    tets, adj = c2.tets, c2.adjacency
    want_t, want_a = tets.copy(), adj.copy()
    del res, c2
    gc.collect()
    for s in range(3):
        ctx.delaunay_batch([W.uniform(900 + 100 * s, 150 + s), W.uniform(700, 160 + s)])
        gc.collect()
    assert np.array_equal(tets, want_t) and np.array_equal(adj, want_a)
    assert np.array_equal(oracle.canonical_real(want_t, len(clouds[2])), _oracle_set(oracle, clouds[2]))


def test_batch_identical_clouds(oracle, ctx):
    """The same coordinates in several clouds never interact."""
    p = W.uniform(500, 42)
    res = ctx.delaunay_batch([p, p, (p * 3).astype(np.float32), p])
    ref = _oracle_set(oracle, p)
    for c in range(4):
        assert np.array_equal(oracle.canonical_real(res[c].tets, len(p)), ref)


def test_batch_flat_input_with_offsets(oracle, ctx):
    clouds = _mixed_clouds(5, 12)
    flat = np.concatenate(clouds)
    off = np.zeros(len(clouds) + 1, np.uint64)
    off[1:] = np.cumsum([len(c) for c in clouds])
    a = ctx.delaunay_batch(flat, off)
    b = ctx.delaunay_batch(clouds)
    # row order belongs to the run (concurrent appends); the sets and the
    # per-cloud counts do not
    assert np.array_equal(a.tet_offsets, b.tet_offsets)
    for c, p in enumerate(clouds):
        assert np.array_equal(oracle.canonical_real(a[c].tets, len(p)), oracle.canonical_real(b[c].tets, len(p)))


def test_batch_resident_matches_host(oracle, ctx):
    clouds = _mixed_clouds(7, 40)
    flat = np.concatenate(clouds)
    off = np.zeros(len(clouds) + 1, np.uint64)
    off[1:] = np.cumsum([len(c) for c in clouds])
    host = ctx.delaunay_batch(flat, off)
    d = ctx.upload(flat)
    try:
        dt, da, start, count = ctx.delaunay_batch_resident(d, off)
        rows = int((start + count).max())
        tets = ctx.download(dt, (rows, 4), np.uint32)
        adj = ctx.download(da, (rows, 4), np.uint32)
    finally:
        ctx.free(d)
    for c, p in enumerate(clouds):
        t = tets[start[c]:start[c] + count[c]]
        assert np.array_equal(oracle.canonical_real(t, len(p)), oracle.canonical_real(host[c].tets, len(p)))
        assert _adjacency_closed(adj[start[c]:start[c] + count[c]])


def test_mixed_batch_4096_clouds(oracle, ctx):
    """4,096 clouds, log-normal sizes (median 2,000), ~10M points: resident
    run, then a sample of clouds checked against the oracle."""
    pts, off = W.mixed_batch(4096, 4)
    assert W.sha256(pts) == GOLD["mixed_batch_4096_seed4"]["sha256"]
    d = ctx.upload(pts)
    try:
        dt, da, start, count = ctx.delaunay_batch_resident(d, off)
        st = ctx.stats()
        rng = np.random.default_rng(0)
        sample = sorted(rng.choice(4096, 24, replace=False).tolist())
        for c in sample:
            a, b = int(off[c]), int(off[c + 1])
            p = pts[a:b]
            t = ctx.download(dt + 16 * int(start[c]), (int(count[c]), 4), np.uint32)
            adj = ctx.download(da + 16 * int(start[c]), (int(count[c]), 4), np.uint32)
            assert np.array_equal(oracle.canonical_real(t, len(p)), _oracle_set(oracle, p)), c
            assert _adjacency_closed(adj), c
    finally:
        ctx.free(d)
    assert st["walk_failures"] == 0
    assert st["num_points"] == len(pts)

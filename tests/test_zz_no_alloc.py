"""wtet_presize_features_noalloc: the presizing features with every buffer
beyond the caller's points and features in the library's static per-GPU
scratch array.  Its output must equal wtet_presize_features' on the same
input (histogram counts bit for bit, moments to float rounding) on two calls
in a row, the second with fewer clouds and points, so that nothing the first
call left in the shared array reaches the second."""
import numpy as np
import pytest

import wtet_amd
from wtet_amd import workloads as W
from oracle import mlp_ref as R

pytestmark = pytest.mark.gpu

MAX_CLOUDS = 2048  # WTET_NOALLOC_MAX_CLOUDS


@pytest.fixture(scope="module")
def ctx():
    c = wtet_amd.Context(0)
    yield c
    c.close()


def _flat(clouds):
    off = np.zeros(len(clouds) + 1, np.uint64)
    off[1:] = np.cumsum([len(c) for c in clouds])
    return np.concatenate(clouds), off


def _compare(ctx, clouds):
    flat, off = _flat(clouds)
    d = ctx.upload(flat)
    try:
        got = ctx.presize_features_noalloc(d, off)
        ref = ctx.presize_features(d, off)
    finally:
        ctx.free(d)
    assert got.shape == ref.shape == (len(clouds), 522)
    assert np.array_equal(got[:, :512], ref[:, :512])  # occupancy counts / n
    assert np.array_equal(got[:, 521], ref[:, 521])  # log(n)
    assert np.allclose(got[:, 512:521], ref[:, 512:521], rtol=1e-6, atol=1e-9)
    for c, p in enumerate(clouds):
        o = R.features(p)
        assert np.array_equal(got[c, :512], o[:512].astype(np.float32)), c
        assert np.allclose(got[c, 512:], o[512:], rtol=1e-5, atol=1e-7), c


def test_noalloc_features_match_allocating_entry_point(ctx):
    print("no-alloc features: call 1, 300 clouds", flush=True)
    rng = np.random.default_rng(21)
    big = [W.uniform(int(rng.integers(64, 3000)), 100 + i) for i in range(296)]
    big += [W.clusters(5000, 2), W.lattice(9), W.sphere(2000, 3), W.anisotropic(4000, 4, k=5)]
    _compare(ctx, big)
    print("no-alloc features: call 2, 3 clouds", flush=True)
    _compare(ctx, [W.uniform(500, 7), W.sphere(300, 8), W.clusters(200, 9)])


def test_noalloc_features_reject_too_many_clouds(ctx):
    pts = W.uniform(4 * (MAX_CLOUDS + 1), 5)
    off = np.arange(0, len(pts) + 1, 4, dtype=np.uint64)
    d = ctx.upload(pts)
    try:
        with pytest.raises(ValueError, match="at most 2048 clouds"):
            ctx.presize_features_noalloc(d, off)
        # the limit itself is accepted
        got = ctx.presize_features_noalloc(d, off[:MAX_CLOUDS + 1])
        ref = ctx.presize_features(d, off[:MAX_CLOUDS + 1])
        assert got.shape == (MAX_CLOUDS, 522)
        assert np.array_equal(got[:, :512], ref[:, :512])
        assert np.allclose(got[:, 512:], ref[:, 512:], rtol=1e-6, atol=1e-9)
    finally:
        ctx.free(d)

"""GPU tests of the library's stable radix sort (the point order's sort;
the reference sorts on the host with a stable sort, src/gpu/mod.rs:422-520)
against numpy's stable argsort: every size class of the tile and scan
loops, partial bit ranges, heavy ties, 32- and 64-bit keys."""
import numpy as np
import pytest

import wtet_amd

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = wtet_amd.Context(0)
    yield c
    c.close()


def _sort(ctx, keys, end_bit):
    n = len(keys)
    vals = np.arange(n, dtype=np.uint32)
    kb = keys.dtype.itemsize * 8
    dk, dv = ctx.upload(keys), ctx.upload(vals)
    ok, ov = ctx.upload(np.zeros_like(keys)), ctx.upload(np.zeros_like(vals))
    try:
        ctx.sort_pairs(dk, dv, ok, ov, n, kb, end_bit)
        gk = ctx.download(ok, (n,), keys.dtype)
        gv = ctx.download(ov, (n,), np.uint32)
        # inputs are left as they were
        assert np.array_equal(ctx.download(dk, (n,), keys.dtype), keys)
    finally:
        for p in (dk, dv, ok, ov):
            ctx.free(p)
    mask = (1 << end_bit) - 1 if end_bit < kb else (1 << kb) - 1
    order = np.argsort(keys & keys.dtype.type(mask), kind="stable")
    assert np.array_equal(gv, order.astype(np.uint32))
    assert np.array_equal(gk, keys[order])


def _keys(rng, n, bits, kind):
    if kind == "random":
        k = rng.integers(0, 1 << 62, n, dtype=np.uint64) * np.uint64(4) + rng.integers(0, 4, n, dtype=np.uint64)
    elif kind == "ties":
        k = rng.integers(0, 5, n, dtype=np.uint64) << np.uint64(max(0, bits - 3))
    else:  # sorted runs with equal high digits
        k = np.sort(rng.integers(0, 1 << 20, n, dtype=np.uint64)) << np.uint64(20)
    return k


@pytest.mark.parametrize("n", [1, 63, 2047, 2048, 2049, 100_003])
@pytest.mark.parametrize("end_bit", [1, 8, 13, 40, 63, 64])
def test_sort_u64_matches_stable_argsort(ctx, n, end_bit):
    rng = np.random.default_rng(n * 131 + end_bit)
    for kind in ("random", "ties", "runs"):
        print(f"sort u64 n={n} end_bit={end_bit} {kind}", flush=True)
        _sort(ctx, _keys(rng, n, end_bit, kind), end_bit)


def test_sort_u64_full_size(ctx):
    """1M keys: the point order of the 1M-point workloads (63-bit Morton
    keys of one cloud)."""
    rng = np.random.default_rng(7)
    _sort(ctx, _keys(rng, 1_000_000, 63, "random"), 63)


@pytest.mark.parametrize("n", [5000, 2_500_000])
@pytest.mark.parametrize("end_bit", [1, 12, 32])
def test_sort_u32_matches_stable_argsort(ctx, n, end_bit):
    """2.5M keys make more than 1,024 tiles, so every digit's scan runs
    over several chunks."""
    rng = np.random.default_rng(n + end_bit)
    keys = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    keys[rng.random(n) < 0.5] &= np.uint32(7)
    print(f"sort u32 n={n} end_bit={end_bit}", flush=True)
    _sort(ctx, keys, end_bit)


def test_sort_zero_bits_and_bad_arguments(ctx):
    rng = np.random.default_rng(3)
    _sort(ctx, rng.integers(0, 1 << 63, 3000, dtype=np.uint64), 0)  # a copy
    with pytest.raises(ValueError):
        ctx.sort_pairs(0, 0, 0, 0, 10, 16, 8)
    with pytest.raises(ValueError):
        ctx.sort_pairs(0, 0, 0, 0, 10, 32, 33)

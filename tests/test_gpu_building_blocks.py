"""Device scan and radix sort against numpy (sizes from 0 to several tiles; ragged ends)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [0, 1, 5, 4095, 4096, 4097, 100000, 3_000_001])
def test_exclusive_scan(ctx, n):
    rng = np.random.default_rng(n)
    a = rng.integers(0, 9, n).astype(np.uint32)
    out = ctx.exclusive_scan(a)
    want = np.concatenate([[0], np.cumsum(a, dtype=np.uint64)]).astype(np.uint32)
    np.testing.assert_array_equal(out, want)


@pytest.mark.parametrize("n,bits", [(1, 64), (2047, 63), (2049, 63), (100003, 40), (1_000_000, 64)])
def test_radix_sort_is_stable(ctx, n, bits):
    rng = np.random.default_rng(7 + n)
    keys = rng.integers(0, 1 << min(bits, 20), n, dtype=np.uint64) << np.uint64(max(0, bits - 20))
    vals = np.arange(n, dtype=np.uint32)
    k, v = ctx.radix_sort(keys, vals, bits)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(k, keys[order])
    np.testing.assert_array_equal(v, vals[order])


@pytest.mark.parametrize("n,bits", [(4095, 8), (4096, 16), (4097, 17), (8193, 24), (65537, 9)])
def test_radix_sort_pass_counts_and_ragged_tiles(ctx, n, bits):
    rng = np.random.default_rng(11 + n)
    keys = rng.integers(0, 1 << bits, n, dtype=np.uint64)
    vals = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    k, v = ctx.radix_sort(keys, vals, bits)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(k, keys[order])
    np.testing.assert_array_equal(v, vals[order])


def test_radix_sort_heavy_duplicates(ctx):
    n = 250_000
    rng = np.random.default_rng(3)
    keys = rng.integers(0, 4, n, dtype=np.uint64)
    vals = np.arange(n, dtype=np.uint32)
    k, v = ctx.radix_sort(keys, vals, 16)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(k, keys[order])
    np.testing.assert_array_equal(v, vals[order])


def test_radix_sort_full_width_random_keys(ctx):
    n = 300_001
    rng = np.random.default_rng(5)
    keys = rng.integers(0, np.iinfo(np.uint64).max, n, dtype=np.uint64, endpoint=True)
    vals = np.arange(n, dtype=np.uint32)
    k, v = ctx.radix_sort(keys, vals, 64)
    order = np.argsort(keys, kind="stable")
    np.testing.assert_array_equal(k, keys[order])
    np.testing.assert_array_equal(v, vals[order])


def test_radix_sort_low_bits_only_with_payload_above(ctx):
    # Keys carry a payload in their upper half; only the low 17 bits are sorted on (stably).
    n = 200_003
    rng = np.random.default_rng(9)
    low = rng.integers(0, 1 << 17, n, dtype=np.uint64)
    high = rng.integers(0, 1 << 32, n, dtype=np.uint64)
    keys = (high << np.uint64(32)) | low
    vals = np.arange(n, dtype=np.uint32)
    k, v = ctx.radix_sort(keys, vals, 17)
    order = np.argsort(low, kind="stable")
    np.testing.assert_array_equal(k, keys[order])
    np.testing.assert_array_equal(v, vals[order])


@pytest.mark.parametrize("n,bit0,bits", [(2, 32, 17), (4096, 32, 8), (300_007, 32, 17), (100_001, 5, 23), (65537, 56, 8)])
def test_radix_sort_keys_on_a_bit_range(ctx, n, bit0, bits):
    # Keys sorted stably on bits [bit0, bit0 + bits) alone: pairs (T << 32 | S) grouped by T keep
    # their input order within a group.
    rng = np.random.default_rng(13 + n)
    keys = rng.integers(0, np.iinfo(np.uint64).max, n, dtype=np.uint64, endpoint=True)
    mask = np.uint64((1 << bits) - 1)
    field = (keys >> np.uint64(bit0)) & mask
    k = ctx.radix_sort_keys(keys, bit0, bits)
    np.testing.assert_array_equal(k, keys[np.argsort(field, kind="stable")])


def test_radix_sort_keys_rejects_bad_ranges(ctx):
    import antimatter_amd as am
    keys = np.arange(10, dtype=np.uint64)
    for bit0, bits in ((60, 8), (-1, 4), (0, 0)):
        with pytest.raises(am.AmError):
            ctx.radix_sort_keys(keys, bit0, bits)

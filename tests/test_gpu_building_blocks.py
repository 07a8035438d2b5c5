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

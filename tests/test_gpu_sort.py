"""The library's radix sorts and exclusive scan (rdk_sort.hip) against numpy.

Every stage that groups data (cell binning, pair canonicalisation, colour
and island order, body CSR, history merge) relies on these being stable
sorts, so the expected results are numpy's stable argsort and cumsum.  Each
case also runs with the count read from device memory (cap > n), the form a
captured step uses.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZES = [0, 1, 2, 63, 100, 4095, 4096, 4097, 100003, (1 << 20) + 17]


def _bits(m):
    b = 1
    while b < 64 and (1 << b) <= m:
        b += 1
    return b


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("key_bits", [7, 17, 32])
def test_sort_pairs_u32_is_stable_numpy_order(gpu_ctx_factory, n, key_bits):
    ctx = gpu_ctx_factory()
    rng = np.random.default_rng(1000 * n + key_bits)
    hi = (1 << key_bits) - 1
    # many equal keys, so stability is exercised
    keys = rng.integers(0, min(hi, max(1, n // 8)) + 1, size=n, dtype=np.uint64).astype(np.uint32)
    if key_bits == 32 and n > 4:
        keys[: n // 4] = rng.integers(0, 1 << 32, size=n // 4, dtype=np.uint64).astype(np.uint32)
    vals = np.arange(n, dtype=np.int32)
    order = np.argsort(keys, kind="stable")
    for cap in (0, n + 5000):
        k, v = ctx.sort_pairs_u32(keys, vals, key_bits, cap=cap)
        assert np.array_equal(k, keys[order]), (n, key_bits, cap)
        assert np.array_equal(v, vals[order]), (n, key_bits, cap)


@pytest.mark.parametrize("n_index", [3, 1000, 100000, 1520000])
def test_sort_pair_keys_u64_matches_numpy(gpu_ctx_factory, n_index):
    ctx = gpu_ctx_factory()
    rng = np.random.default_rng(n_index)
    m = min(300000, 4 * n_index)
    a = rng.integers(0, n_index, size=m, dtype=np.uint64)
    b = rng.integers(0, n_index, size=m, dtype=np.uint64)
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    keys = (lo << np.uint64(32)) | hi
    bits = _bits(n_index)
    for cap in (0, m + 777):
        k = ctx.sort_pair_keys_u64(keys, bits, cap=cap)
        assert np.array_equal(k, np.sort(keys)), (n_index, cap)


@pytest.mark.parametrize("n", [0, 1, 5, 2047, 2048, 2049, 70000, 4 * 400000 + 3])
def test_exclusive_scan_matches_cumsum(gpu_ctx_factory, n):
    ctx = gpu_ctx_factory()
    rng = np.random.default_rng(n)
    x = rng.integers(0, 3, size=n).astype(np.int32)
    want = np.concatenate([[0], np.cumsum(x)[:-1]]).astype(np.int32) if n else np.zeros(0, np.int32)
    for cap in (0, n + 3000):
        got = ctx.exclusive_scan(x, cap=cap)
        assert np.array_equal(got, want), (n, cap)

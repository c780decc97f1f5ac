"""The library's own device primitives (sbp_amd/csrc/primitives.hpp) against numpy: the stable LSD radix sort
(64- and 32-bit keys, sub-ranges of bits, carried values), the int64 prefix sums and the run-length encoding, at
sizes around the tile and chunk boundaries (4,096) and beyond."""
import numpy as np
import pytest

from sbp_amd import native

pytestmark = pytest.mark.gpu

SIZES = [0, 1, 63, 4095, 4096, 4097, 100_003, 1_000_003]


@pytest.fixture(scope="module")
def ctx():
    c = native.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("n", SIZES)
def test_sort_full_u64(ctx, n):
    rng = np.random.default_rng(n)
    k = rng.integers(0, 2 ** 63, n, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, n, dtype=np.uint64)
    ko, _ = ctx.debug_sort(k)
    assert np.array_equal(ko, np.sort(k))


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("begin,end", [(0, 12), (8, 20), (3, 3), (40, 64)])
def test_sort_is_stable_on_a_bit_range_and_carries_values(ctx, n, begin, end):
    rng = np.random.default_rng(7 + n)
    k = rng.integers(0, 2 ** 64 - 1, n, dtype=np.uint64, endpoint=True)
    v = np.arange(n, dtype=np.int32)
    ko, vo = ctx.debug_sort(k, v, begin, end)
    field = (k >> np.uint64(begin)) & np.uint64((1 << (end - begin)) - 1) if end > begin else np.zeros(n, np.uint64)
    order = np.argsort(field, kind="stable")
    assert np.array_equal(vo, v[order])
    assert np.array_equal(ko, k[order])


@pytest.mark.parametrize("n", SIZES)
def test_sort_u32_keys_with_values(ctx, n):
    rng = np.random.default_rng(11 + n)
    k = rng.integers(0, 50, n, dtype=np.uint64)  # many equal keys: stability decides the values' order
    v = rng.integers(-2 ** 31, 2 ** 31 - 1, n, dtype=np.int64).astype(np.int32)
    ko, vo = ctx.debug_sort(k, v, 0, 32, key_bits=32)
    order = np.argsort(k, kind="stable")
    assert np.array_equal(ko, k[order]) and np.array_equal(vo, v[order])


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("inclusive", [False, True])
def test_scan_int64(ctx, n, inclusive):
    rng = np.random.default_rng(3 + n)
    x = rng.integers(-10 ** 9, 10 ** 9, n, dtype=np.int64)
    want = np.cumsum(x) if inclusive else np.concatenate([[0], np.cumsum(x)[:-1]]).astype(np.int64)[:n]
    assert np.array_equal(ctx.debug_scan(x, inclusive), want)


@pytest.mark.parametrize("n", SIZES)
def test_run_length_encode(ctx, n):
    rng = np.random.default_rng(5 + n)
    k = np.sort(rng.integers(0, max(1, n // 3), n, dtype=np.uint64) * np.uint64(2 ** 33))
    u, c = ctx.debug_rle(k)
    wu, wc = np.unique(k, return_counts=True)
    assert np.array_equal(u, wu) and np.array_equal(c, wc.astype(np.int32))

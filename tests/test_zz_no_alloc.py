"""The sort that allocates nothing (rt_sort_i32_noalloc) against rt_sort_i32, through the C ABI: the same
bytes on two calls in a row, the second with a smaller input, plus the edge sizes and the size limit."""
import ctypes

import numpy as np
import pytest

from rayt_amd import _lib, api

pytestmark = pytest.mark.gpu

MAX_N = 1 << 23  # RT_SORT_NOALLOC_MAX_N


def _sorted_both(gpu_ctx, values):
    n = len(values)
    a = api.buffer_from(gpu_ctx, values)
    b = api.buffer_from(gpu_ctx, values)
    temp = api.Buffer(gpu_ctx, max(4, 4 * n))
    lib = _lib.lib()
    assert lib.rt_sort_i32(gpu_ctx.h, a.h, n, temp.h) == 0, lib.rt_last_error()
    assert lib.rt_sort_i32_noalloc(gpu_ctx.h, b.h, n) == 0, lib.rt_last_error()
    gpu_ctx.WaitForAll()
    return a.read_all(np.int32)[:n], b.read_all(np.int32)[:n]


def test_noalloc_sort_equals_sort_twice_in_a_row(gpu_ctx):
    rng = np.random.default_rng(17)
    first = rng.integers(-2**31, 2**31 - 1, 2073600, dtype=np.int64).astype(np.int32)  # a 1080p frame of codes
    second = rng.integers(-5000, 5000, 100003, dtype=np.int64).astype(np.int32)  # smaller, many duplicates
    for values in (first, second):
        ref, got = _sorted_both(gpu_ctx, values)
        np.testing.assert_array_equal(got, ref)
        np.testing.assert_array_equal(got, np.sort(values))


@pytest.mark.parametrize("n", [0, 1, 2, 2013, 2048, 2049, MAX_N])
def test_noalloc_sort_edge_sizes(gpu_ctx, n):
    rng = np.random.default_rng(n)
    values = rng.integers(-2**31, 2**31 - 1, max(n, 1), dtype=np.int64).astype(np.int32)[:n]
    if n == 0:
        values = np.zeros(1, np.int32)
        buf = api.buffer_from(gpu_ctx, values)
        assert _lib.lib().rt_sort_i32_noalloc(gpu_ctx.h, buf.h, 0) == 0
        return
    ref, got = _sorted_both(gpu_ctx, values)
    np.testing.assert_array_equal(got, ref)
    np.testing.assert_array_equal(got, np.sort(values))


def test_noalloc_sort_refuses_beyond_its_maximum(gpu_ctx):
    buf = api.Buffer(gpu_ctx, 4 * (MAX_N + 1))
    lib = _lib.lib()
    assert lib.rt_sort_i32_noalloc(gpu_ctx.h, buf.h, MAX_N + 1) == _lib.RT_ERR_RANGE
    assert b"maximum" in lib.rt_last_error()

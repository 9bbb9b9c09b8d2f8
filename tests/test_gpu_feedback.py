"""GPU parity of the GPU algorithms and the feedback extractor against the oracle
and brute force (std::sort / sequential sum, as the reference's own
TestGPUSort / TestGPUScan in application/test-application.cpp:150-240)."""
import numpy as np
import pytest

import oracle as orc
from rayt_amd import api

pytestmark = pytest.mark.gpu

NO_HIT = 0x7FFFFFFF


def _rand_reference_style(rng, n):
    # rand() % 201 - 100, test-application.cpp:158-161
    return rng.integers(0, 201, n).astype(np.int32) - 100


@pytest.mark.parametrize("n", [1, 2, 2013, 2048, 2049, 100_000, 2_073_600])
def test_gpu_sort(gpu_ctx, n):
    rng = np.random.default_rng(n)
    trials = 10 if n == 2013 else 2
    buf = api.Buffer(gpu_ctx, 4 * n)
    s = api.GPUSort(n, gpu_ctx)
    for k in range(trials):
        src = _rand_reference_style(rng, n) if k % 2 == 0 else rng.integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32)
        buf.Write(0, src)
        s.Sort(buf)
        np.testing.assert_array_equal(buf.read_all(np.int32), np.sort(src))


@pytest.mark.parametrize("n", [1, 2013, 4096, 4097, 1_000_003, 2_073_600])
def test_gpu_scan(gpu_ctx, n):
    rng = np.random.default_rng(n)
    trials = 10 if n == 2013 else 2
    buf = api.Buffer(gpu_ctx, 4 * n)
    s = api.GPUScan(n, gpu_ctx)
    for k in range(trials):
        src = _rand_reference_style(rng, n) if k % 2 == 0 else rng.integers(-2**31, 2**31, n, dtype=np.int64).astype(np.int32)
        buf.Write(0, src)
        s.Scan(buf)
        exp = np.cumsum(src.astype(np.int64)).astype(np.uint64).astype(np.uint32).view(np.int32)  # int32 wrap-around
        np.testing.assert_array_equal(buf.read_all(np.int32), exp)


def test_difference_flag_and_compress(gpu_ctx):
    rng = np.random.default_rng(5)
    n = 50_000
    src = np.sort(rng.integers(-300, 300, n).astype(np.int32))
    a = api.buffer_from(gpu_ctx, src)
    flags = api.Buffer(gpu_ctx, 4 * n)
    api.call("rt_difference_flag", gpu_ctx.h, n, a.h, flags.h)
    f = flags.read_all(np.int32)
    exp = np.concatenate([[0], (src[1:] != src[:-1]).astype(np.int32)])
    np.testing.assert_array_equal(f, exp)
    api.GPUScan(n, gpu_ctx).Scan(flags)
    out = api.Buffer(gpu_ctx, 4 * n)
    api.call("rt_compress", gpu_ctx.h, n, flags.h, a.h, out.h)
    u = np.unique(src)
    np.testing.assert_array_equal(out.Read(0, 4 * len(u), np.int32), u)


def test_node_to_block(gpu_ctx):
    rng = np.random.default_rng(9)
    n = 100_000
    nodes = rng.integers(0, 2**28, n)
    codes = (((nodes + 1) << 1) | rng.integers(0, 2, n)).astype(np.int32)
    codes[::7] = NO_HIT
    codes[::11] = -rng.integers(1, 2**21, len(codes[::11])).astype(np.int32)
    b = api.buffer_from(gpu_ctx, codes)
    api.call("rt_node_to_block", gpu_ctx.h, n, 8192, b.h)
    np.testing.assert_array_equal(b.read_all(np.int32), orc.node_to_block(codes, 8192))


def _synthetic_codes(rng, n, n_faults, n_slots, nohit_frac=0.2, block_level=True, nib=8192):
    kinds = rng.random(n)
    out = np.empty(n, np.int64)
    faults = rng.choice(np.arange(1 << 21), size=max(n_faults, 1), replace=False)
    slots = rng.integers(0, n_slots, n)
    dup = rng.integers(0, 2, n) * (rng.random(n) < 0.1)
    if block_level:
        hits = ((slots + 1) << 1) | dup
    else:
        hits = (((slots * nib + rng.integers(0, nib, n)) + 1) << 1) | dup
    out[:] = hits
    fmask = kinds < (0.3 if n_faults else 0.0)
    out[fmask] = -(faults[rng.integers(0, len(faults), fmask.sum())] + 1)
    out[kinds > 1 - nohit_frac] = NO_HIT
    return out.astype(np.int32)


@pytest.mark.parametrize("case", ["hits_only", "faults", "many_faults", "node_level", "all_nohit", "empty_window"])
def test_feedback_matches_oracle(gpu_ctx, case):
    rng = np.random.default_rng(sum(map(ord, case)))
    n = 640 * 480
    if case == "hits_only":
        codes, bl = _synthetic_codes(rng, n, 0, 5000), True
    elif case == "faults":
        codes, bl = _synthetic_codes(rng, n, 300, 2000), True
    elif case == "many_faults":
        codes, bl = _synthetic_codes(rng, n, 70_000, 500), True
    elif case == "node_level":
        codes, bl = _synthetic_codes(rng, n, 50, 300, block_level=False), False
    elif case == "all_nohit":
        codes, bl = np.full(n, NO_HIT, np.int32), True
    else:
        codes, bl = _synthetic_codes(rng, n, 20, 3, nohit_frac=0.0), True
    fb = api.RayTracingFeedbackExtractor(640, 480, gpu_ctx)
    fb.set_seed(1234)
    buf = api.buffer_from(gpu_ctx, codes)
    oc = codes if bl else orc.node_to_block(codes, 8192)
    ef, rng_state = 1 << 16, 1234
    for it in range(3):  # expected_faults adapts across calls
        buf.Write(0, codes)
        fb.Process(buf, 8192, block_level=bl)
        r = orc.feedback_extract(oc, expected_faults=ef, seed=rng_state)
        assert fb.fault_blocks() == r["faults"], f"faults differ (iteration {it})"
        assert fb.hit_blocks() == r["hits"], f"hits differ (iteration {it})"
        assert fb.duplicate_hit_blocks() == r["dups"], f"dup hits differ (iteration {it})"
        u, e = fb.stats()
        assert (u, e) == (r["unique_count"], r["expected_faults"])
        ef, rng_state = r["expected_faults"], r["rng"]


def test_feedback_out_of_domain_falls_back_to_sort(gpu_ctx):
    rng = np.random.default_rng(3)
    n = 100_000
    codes = rng.integers(-1000, 2**30, n).astype(np.int32)  # hit codes beyond the bitmap domain
    fb = api.RayTracingFeedbackExtractor(n, 1, gpu_ctx)
    fb.set_seed(7)
    buf = api.buffer_from(gpu_ctx, codes)
    fb.Process(buf, 8192, block_level=True)
    r = orc.feedback_extract(codes, seed=7)
    assert fb.fault_blocks() == r["faults"]
    assert fb.hit_blocks() == r["hits"]
    assert fb.duplicate_hit_blocks() == r["dups"]

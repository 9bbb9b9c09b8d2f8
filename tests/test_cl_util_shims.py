"""The CL util shims (include/rayt_cl_util.h): CLContext, CLEvent,
CLEventList, CLBuffer and CLKernel over the C ABI, used the way the
reference's classes use CL util (GPU algorithms/gpu-sort.cpp:16-49,
gpu-scan.cpp:15-43, GPU_renderer/gpu-ray-tracer.cpp:66-227).

CPU: the shim header and the program that drives it compile warning-free.
GPU: the program runs the reference's TestGPUSort / TestGPUScan checks
(application/test-application.cpp:150-240) pass by pass through the shims,
and one frame through the InitTracingFrame / RaytracingPass /
NodeToBlockKernel / FinishTracingFrame shims, whose image and codes equal the
oracle's three-kernel frame.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle as orc
from tests import shim_build
from tests.scenes import POSES, camera_vpi, terrain_tree


def test_shims_compile():
    r = subprocess.run(shim_build.compile_command("", syntax_only=True), capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_shims_sort_scan_and_frame(tmp_path):
    assert os.path.exists(shim_build.BIN), "tests/_bin/cl_util_shim_test is not built: run __graft_entry__.build()"
    tree, path = terrain_tree(8, 8192, 8, 0)
    W, H = 256, 160
    vpi = camera_vpi(*POSES[0], aspect=W / H)
    vpi_file = tmp_path / "vpi.bin"
    np.ascontiguousarray(vpi, np.float32).tofile(vpi_file)
    img_file, codes_file = tmp_path / "image.bin", tmp_path / "codes.bin"
    n = 100003  # merge passes over a partial last pair of runs, scan passes past 2^16
    r = subprocess.run([shim_build.BIN, str(n), path, str(vpi_file), str(W), str(H), str(img_file), str(codes_file)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "sort passed" in r.stdout and "scan passed" in r.stdout and "frame done" in r.stdout
    t = orc.Tree(path)
    c = orc.Cache(t, t.blocks)
    c.initial_fill()
    normals = c.channel_bytes(1, 4).view(np.uint32)
    colors = c.channel_bytes(2, 4).view(np.uint32)
    st = orc.init_tracing_frame(W, H, vpi, 0.0)
    codes = orc.raytracing_pass(st, W, H, c.links(), c.far(), c.root(), normals, colors, 0, 0)
    img = orc.finish_tracing_frame(st, W, H, normals, colors)
    got_codes = np.fromfile(codes_file, np.int32)
    np.testing.assert_array_equal(got_codes, orc.node_to_block(codes, t.nodes_in_block))
    np.testing.assert_array_equal(np.fromfile(img_file, np.uint32), img)
    assert (got_codes != 0x7FFFFFFF).sum() > 1000


@pytest.mark.gpu
@pytest.mark.parametrize("n,length", [(1, 1), (2013, 1), (2013, 256), (2013, 1024), (100003, 4096), (65536, 32768)])
def test_merge_pass_merges_run_pairs(gpu_ctx, n, length):
    """GlobalMergeSortPassKernel (merge_sort.cl:47-73): runs of `length` sorted in, runs of 2 length sorted out."""
    from rayt_amd import api
    rng = np.random.default_rng(n + length)
    x = rng.integers(-50, 50, n).astype(np.int32)  # many equal keys
    for s in range(0, n, length):
        x[s:s + length] = np.sort(x[s:s + length])
    a, b = api.buffer_from(gpu_ctx, x), api.Buffer(gpu_ctx, 4 * n)
    api.call("rt_merge_pass_i32", gpu_ctx.h, n, length, a.h, b.h)
    exp = x.copy()
    for s in range(0, n, 2 * length):
        exp[s:s + 2 * length] = np.sort(exp[s:s + 2 * length])
    np.testing.assert_array_equal(b.read_all(np.int32)[:n], exp)


@pytest.mark.gpu
@pytest.mark.parametrize("n,length", [(1, 1), (2013, 1), (2013, 512), (70000, 65536)])
def test_scan_pass_adds_the_element_length_back(gpu_ctx, n, length):
    """ScanPassKernel (scan.cl:2-12): out[i] = in[i] + in[i - length] for i >= length, with int32 wrap-around."""
    from rayt_amd import api
    rng = np.random.default_rng(n)
    x = rng.integers(-2**31, 2**31 - 1, n, dtype=np.int64).astype(np.int32)
    a, b = api.buffer_from(gpu_ctx, x), api.Buffer(gpu_ctx, 4 * n)
    api.call("rt_scan_pass_i32", gpu_ctx.h, n, length, a.h, b.h)
    exp = x.astype(np.int64)
    exp[length:] += x[:-length].astype(np.int64) if length < n else 0
    np.testing.assert_array_equal(b.read_all(np.int32)[:n], (exp & 0xFFFFFFFF).astype(np.uint32).view(np.int32))

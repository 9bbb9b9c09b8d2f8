"""GPU parity of the batched persistent kernel: every view of a mixed batch is
bit-identical to the plain single-view render of that view, for every launch
shape, and the counting variant adds up to the per-view counts."""
import numpy as np
import pytest

from rayt_amd import api
from tests.scenes import POSES, camera_vpi, terrain_tree

pytestmark = pytest.mark.gpu

SIZES = [(256, 256), (640, 480), (1280, 720), (100, 37), (7, 5), (320, 200)]


def _views(n, seed=0):
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(n):
        w, h = SIZES[i % len(SIZES)]
        yaw, pitch, pos = POSES[i % len(POSES)]
        if i >= len(POSES):  # random poses inside the scene bounds
            yaw, pitch = float(rng.uniform(-180, 180)), float(rng.uniform(-60, 30))
            pos = tuple(float(x) for x in rng.uniform(0, 1, 3))
        specs.append((camera_vpi(yaw, pitch, pos, aspect=w / h), w, h))
    return specs


@pytest.fixture(scope="module")
def scene(gpu_ctx):
    tree, path = terrain_tree(9, 8192, 0, 16)
    m = api.GPUOctreeCacheManager(tree.info().blocks_count, tree, gpu_ctx)
    m.InitialFillCache()
    return m


def _plain(gpu_ctx, m, vpi, w, h, flags):
    out, codes, _ = api.render_plain(gpu_ctx, m, w, h, vpi, flags=flags)
    return out.read_all(np.uint32), codes.read_all(np.int32)


@pytest.mark.parametrize("tile_shape", [0, 1, 2])
@pytest.mark.parametrize("lds_levels", [8, 12, 16])
@pytest.mark.parametrize("wgs_per_cu", [0, 8])  # one wave per tile; persistent
def test_batch_matches_plain(gpu_ctx, scene, tile_shape, lds_levels, wgs_per_cu):
    m = scene
    specs = _views(12, seed=tile_shape * 10 + lds_levels)
    flags = api.L.RT_FLAG_SHADOWS
    b = api.BatchRenderer(m, 64)
    b.config(tile_shape=tile_shape, lds_levels=lds_levels, wgs_per_cu=wgs_per_cu, tree_depth=9)
    views, total, offsets = b.make_views(specs, flags=flags)
    out = api.Buffer(gpu_ctx, 4 * total)
    codes = api.Buffer(gpu_ctx, 4 * total)
    b.render(views, len(specs), out, codes)
    img = out.read_all(np.uint32)
    cod = codes.read_all(np.int32)
    for (vpi, w, h), off in zip(specs, offsets):
        pi, pc = _plain(gpu_ctx, m, vpi, w, h, flags)
        np.testing.assert_array_equal(img[off:off + w * h], pi, err_msg=f"view {w}x{h} image")
        np.testing.assert_array_equal(cod[off:off + w * h], pc, err_msg=f"view {w}x{h} codes")


@pytest.mark.parametrize("lds_levels,tree_depth", [(8, 0), (8, 14), (12, 0)])
def test_batch_spill_stack_matches_plain(gpu_ctx, scene, lds_levels, tree_depth):
    """The spill stack (depth unknown, or declared deeper than the LDS window holds) gives the same bytes."""
    m = scene
    specs = _views(8, seed=lds_levels + tree_depth)
    flags = api.L.RT_FLAG_SHADOWS
    b = api.BatchRenderer(m, 16)
    b.config(tile_shape=2, lds_levels=lds_levels, wgs_per_cu=16, tree_depth=tree_depth)
    views, total, offsets = b.make_views(specs, flags=flags)
    out = api.Buffer(gpu_ctx, 4 * total)
    codes = api.Buffer(gpu_ctx, 4 * total)
    b.render(views, len(specs), out, codes)
    img = out.read_all(np.uint32)
    cod = codes.read_all(np.int32)
    for (vpi, w, h), off in zip(specs, offsets):
        pi, pc = _plain(gpu_ctx, m, vpi, w, h, flags)
        np.testing.assert_array_equal(img[off:off + w * h], pi, err_msg=f"view {w}x{h} image")
        np.testing.assert_array_equal(cod[off:off + w * h], pc, err_msg=f"view {w}x{h} codes")


def test_batch_views_with_lod(gpu_ctx, scene):
    """Views with a LOD voxel size (some 1.5, some 0 in one launch) equal the plain render at their LOD."""
    m = scene
    specs = _views(10, seed=11)
    lods = [1.5 if i % 2 else 0.0 for i in range(len(specs))]
    flags = api.L.RT_FLAG_SHADOWS
    b = api.BatchRenderer(m, 16)
    b.config(tile_shape=0, lds_levels=12, wgs_per_cu=0, tree_depth=9)
    views, total, offsets = b.make_views(specs, flags=flags)
    for i, lod in enumerate(lods):
        views[i].lod_voxel_size = lod
    out = api.Buffer(gpu_ctx, 4 * total)
    codes = api.Buffer(gpu_ctx, 4 * total)
    b.render(views, len(specs), out, codes)
    img, cod = out.read_all(np.uint32), codes.read_all(np.int32)
    for (vpi, w, h), off, lod in zip(specs, offsets, lods):
        po, pc, _ = api.render_plain(gpu_ctx, m, w, h, vpi, lod=lod, flags=flags)
        np.testing.assert_array_equal(img[off:off + w * h], po.read_all(np.uint32), err_msg=f"lod {lod} image")
        np.testing.assert_array_equal(cod[off:off + w * h], pc.read_all(np.int32), err_msg=f"lod {lod} codes")


@pytest.mark.parametrize("wgs_per_cu", [0, 16])
def test_batch_counts_add_up(gpu_ctx, scene, wgs_per_cu):
    m = scene
    specs = _views(9, seed=3)
    b = api.BatchRenderer(m, 16)
    b.config(tile_shape=0, lds_levels=12, wgs_per_cu=wgs_per_cu, tree_depth=9)
    views, total, _ = b.make_views(specs, flags=api.L.RT_FLAG_SHADOWS)
    out = api.Buffer(gpu_ctx, 4 * total)
    st = b.render(views, len(specs), out, None, count=True)
    exp = np.zeros(4, np.int64)
    for vpi, w, h in specs:
        _, _, s = api.render_plain(gpu_ctx, m, w, h, vpi, flags=api.L.RT_FLAG_SHADOWS, count=True)
        exp += np.array(s)
    assert st == [int(x) for x in exp]
    assert st[0] == total


def test_batch_repeatable_and_single_view(gpu_ctx, scene):
    """Launching twice gives the same bytes; a batch of one equals the plain render."""
    m = scene
    vpi = camera_vpi(*POSES[0], aspect=640 / 480)
    b = api.BatchRenderer(m, 4)
    b.config(tile_shape=1, lds_levels=10, wgs_per_cu=32, tree_depth=9)
    views, total, _ = b.make_views([(vpi, 640, 480)])
    out = api.Buffer(gpu_ctx, 4 * total)
    b.render(views, 1, out)
    first = out.read_all(np.uint32)
    b.render(views, 1, out)
    np.testing.assert_array_equal(out.read_all(np.uint32), first)
    pi, _ = _plain(gpu_ctx, m, vpi, 640, 480, 0)
    np.testing.assert_array_equal(first, pi)

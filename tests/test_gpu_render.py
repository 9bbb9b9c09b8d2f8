"""GPU parity: the HIP renderer, cache and feedback pipeline against the oracle.

Hit node ids, feedback codes, pool words and RGBA8 pixels are compared
bit-exactly; hit distances are compared through the pixels and codes they
determine (the renderer does not export t).
"""
import numpy as np
import pytest

import oracle as orc
from rayt_amd import api
from tests.obj_scenes import random_textured_scene
from tests.scenes import POSES, camera_vpi, corrupt_root_links, sphere_tree, terrain_tree

pytestmark = pytest.mark.gpu


def _oracle_pool(path, max_blocks=None):
    t = orc.Tree(path)
    c = orc.Cache(t, max_blocks or t.blocks)
    c.initial_fill()
    return t, c


def _gpu_cache(ctx, tree, max_blocks=None):
    info = tree.info()
    m = api.GPUOctreeCacheManager(max_blocks or info.blocks_count, tree, ctx)
    m.InitialFillCache()
    return m


def _oracle_arrays(t, c):
    links, far = c.links(), c.far()
    normals = c.channel_bytes(1, 4).view(np.uint32)
    colors = c.channel_bytes(2, 4).view(np.uint32)
    return links, far, normals, colors


@pytest.mark.parametrize("level,nib", [(3, 131), (5, 111), (7, 111), (6, 16)])
def test_sphere_pool_layout(gpu_ctx, level, nib):
    """CheckTestOctreeSphereWithGPUData on the device pool after InitialFillCache (whole tree)."""
    center, radius = (0.3, 0.4, 0.65), 0.23
    tree, path = sphere_tree(center, radius, level, nib)
    m = _gpu_cache(gpu_ctx, tree)
    pool = m.pool_arrays()
    err = orc.check_sphere_pool(pool["links"], pool["far"], pool["node index"], pool["node index % 211"],
                                m.root_node_index(), center, radius, level, faultless=True)
    assert err == 0
    t, c = _oracle_pool(path)
    assert m.root_node_index() == c.root()
    np.testing.assert_array_equal(pool["links"], c.links())
    np.testing.assert_array_equal(pool["far"], c.far())
    np.testing.assert_array_equal(pool["node index"], c.channel_bytes(1, 4))
    np.testing.assert_array_equal(pool["node index % 211"], c.channel_bytes(2, 1))


@pytest.mark.parametrize("depth,nib,spheres,boxes", [(8, 8192, 8, 0), (9, 1024, 0, 16), (10, 8192, 0, 64)])
def test_terrain_pool_matches_oracle(gpu_ctx, depth, nib, spheres, boxes):
    tree, path = terrain_tree(depth, nib, spheres, boxes)
    m = _gpu_cache(gpu_ctx, tree)
    t, c = _oracle_pool(path)
    pool = m.pool_arrays()
    assert m.root_node_index() == c.root()
    np.testing.assert_array_equal(pool["links"], c.links())
    np.testing.assert_array_equal(pool["far"], c.far())
    np.testing.assert_array_equal(pool["normals"], c.channel_bytes(1, 4))
    np.testing.assert_array_equal(pool["colors"], c.channel_bytes(2, 4))
    assert m.lru_order() == c.lru_order()


@pytest.mark.parametrize("lod", [0.0, 1.0])
@pytest.mark.parametrize("flags", [0, api.L.RT_FLAG_SHADOWS])
def test_render_plain_matches_oracle(gpu_ctx, lod, flags):
    tree, path = terrain_tree(8, 8192, 8, 0)
    m = _gpu_cache(gpu_ctx, tree)
    t, c = _oracle_pool(path)
    links, far, normals, colors = _oracle_arrays(t, c)
    W, H = 320, 240
    for yaw, pitch, pos in POSES:
        vpi = camera_vpi(yaw, pitch, pos)
        out, codes, _ = api.render_plain(gpu_ctx, m, W, H, vpi, lod=lod, flags=flags)
        img = out.read_all(np.uint32).reshape(H, W)
        cod = codes.read_all(np.int32).reshape(H, W)
        r = orc.render_plain(links, far, c.root(), normals, colors, t.nodes_in_block, W, H, vpi, lod=lod, flags=flags)
        np.testing.assert_array_equal(cod, r["codes"], err_msg=f"codes differ at pose {yaw, pitch, pos}")
        np.testing.assert_array_equal(img, r["image"], err_msg=f"image differs at pose {yaw, pitch, pos}")


@pytest.mark.parametrize("seed,level,nib", [(0, 8, 512), (1, 10, 8192)])
def test_render_imported_obj_matches_oracle(gpu_ctx, tmp_path, seed, level, nib):
    """An OBJ scene (textures, bump map, vertex normals) imported by the product renders as the oracle
    renders the same .tree file, pool words included."""
    obj = random_textured_scene(str(tmp_path / "obj"), seed=seed, n_tris=60)
    tree = api.StoredOctree.obj(obj, level, nib)
    path = str(tmp_path / "scene.tree")
    tree.write(path)
    m = _gpu_cache(gpu_ctx, tree)
    t, c = _oracle_pool(path)
    links, far, normals, colors = _oracle_arrays(t, c)
    pool = m.pool_arrays()
    np.testing.assert_array_equal(pool["links"], c.links())
    np.testing.assert_array_equal(pool["colors"], c.channel_bytes(2, 4))
    W, H = 256, 192
    hits = 0
    for yaw, pitch, pos in POSES:
        vpi = camera_vpi(yaw, pitch, pos)
        out, codes, _ = api.render_plain(gpu_ctx, m, W, H, vpi, flags=api.L.RT_FLAG_SHADOWS)
        r = orc.render_plain(links, far, c.root(), normals, colors, t.nodes_in_block, W, H, vpi,
                             flags=api.L.RT_FLAG_SHADOWS)
        cod = codes.read_all(np.int32).reshape(H, W)
        np.testing.assert_array_equal(cod, r["codes"], err_msg=f"codes differ at pose {yaw, pitch, pos}")
        np.testing.assert_array_equal(out.read_all(np.uint32).reshape(H, W), r["image"])
        hits += int((cod != 0x7FFFFFFF).sum())
    assert hits > 1000
    tree.close()


# (waves, LDS levels, declared depth): window-only stacks (depth - 1 <= levels), spill stacks with the
# window at level 0 (depth unknown) and at level 3 (declared 12 over 8 levels), and the full LDS stack.
@pytest.mark.parametrize("waves,levels,depth", [(1, 8, 8), (1, 10, 0), (4, 8, 8), (1, 22, 8), (4, 14, 8),
                                                (4, 8, 12), (4, 16, 0)])
def test_render_plain_launch_shapes(gpu_ctx, waves, levels, depth):
    """Every launch shape (workgroup size, LDS stack window, spill to private memory) gives the same bytes."""
    tree, path = terrain_tree(8, 8192, 8, 0)
    m = _gpu_cache(gpu_ctx, tree)
    t, c = _oracle_pool(path)
    links, far, normals, colors = _oracle_arrays(t, c)
    W, H = 200, 120
    try:
        api.plain_config(gpu_ctx, levels, waves, depth)
        for yaw, pitch, pos in POSES[:3]:
            vpi = camera_vpi(yaw, pitch, pos)
            out, codes, _ = api.render_plain(gpu_ctx, m, W, H, vpi, flags=api.L.RT_FLAG_SHADOWS)
            r = orc.render_plain(links, far, c.root(), normals, colors, t.nodes_in_block, W, H, vpi,
                                 flags=api.L.RT_FLAG_SHADOWS)
            np.testing.assert_array_equal(out.read_all(np.uint32).reshape(H, W), r["image"])
            np.testing.assert_array_equal(codes.read_all(np.int32).reshape(H, W), r["codes"])
    finally:
        api.plain_config(gpu_ctx, 16, 4, 0)


def test_render_plain_declared_depth_too_small(gpu_ctx):
    """A window-only stack on a tree deeper than declared ends the deeper rays as error pixels (red, no-hit
    code) and leaves every other pixel as the spill stack renders it."""
    tree, _ = terrain_tree(10, 8192, 0, 16)
    m = _gpu_cache(gpu_ctx, tree)
    W, H = 320, 240
    red, no_hit = 0xFF0000FF, 0x7FFFFFFF
    try:
        errors = 0
        for yaw, pitch, pos in POSES[:4]:
            vpi = camera_vpi(yaw, pitch, pos)
            api.plain_config(gpu_ctx, 16, 4, 0)
            out, codes, _ = api.render_plain(gpu_ctx, m, W, H, vpi)
            ref_img, ref_codes = out.read_all(np.uint32), codes.read_all(np.int32)
            assert not np.any(ref_img == red)
            api.plain_config(gpu_ctx, 8, 4, 9)  # window of 8 levels; the tree pushes up to level 8
            out, codes, _ = api.render_plain(gpu_ctx, m, W, H, vpi)
            img, cod = out.read_all(np.uint32), codes.read_all(np.int32)
            bad = img == red
            assert np.all(cod[bad] == no_hit)
            np.testing.assert_array_equal(img[~bad], ref_img[~bad])
            np.testing.assert_array_equal(cod[~bad], ref_codes[~bad])
            errors += int(bad.sum())
        assert errors > 0
    finally:
        api.plain_config(gpu_ctx, 16, 4, 0)


def test_render_counts_match_oracle(gpu_ctx):
    """The counting kernel's totals (rays, shadow rays, fetches F, hits) equal the oracle's."""
    tree, path = terrain_tree(8, 8192, 8, 0)
    m = _gpu_cache(gpu_ctx, tree)
    t, c = _oracle_pool(path)
    links, far, normals, colors = _oracle_arrays(t, c)
    vpi = camera_vpi(*POSES[0])
    _, _, st = api.render_plain(gpu_ctx, m, 640, 480, vpi, flags=api.L.RT_FLAG_SHADOWS, count=True)
    r = orc.render_plain(links, far, c.root(), normals, colors, t.nodes_in_block, 640, 480, vpi,
                         flags=api.L.RT_FLAG_SHADOWS)
    assert st == r["stats"]


def test_bad_links_end_rays_like_the_oracle(gpu_ctx):
    """A pool with bad links (error code 2, a red pixel with no feedback code): the plain kernel, both
    batched kernels and the counting kernel give the oracle's image, codes and counts."""
    tree, path = terrain_tree(8, 8192, 8, 0)
    m = _gpu_cache(gpu_ctx, tree)
    t, c = _oracle_pool(path)
    links, far, normals, colors = _oracle_arrays(t, c)
    words, far_words = corrupt_root_links(links, far, c.root())
    for i, w in words:
        links[i] = w
        m.channel(0).Write(4 * i, np.array([w], np.uint32))
    for i, w in far_words:
        far[i] = w
        m.far_pointers().Write(4 * i, np.array([w], np.uint32))
    W, H = 320, 240
    red = 255 | (255 << 24)
    flags = api.L.RT_FLAG_SHADOWS
    specs = []
    errors = 0
    for yaw, pitch, pos in POSES:
        vpi = camera_vpi(yaw, pitch, pos)
        specs.append((vpi, W, H))
        out, codes, _ = api.render_plain(gpu_ctx, m, W, H, vpi, flags=flags)
        _, _, st = api.render_plain(gpu_ctx, m, W, H, vpi, flags=flags, count=True)
        r = orc.render_plain(links, far, c.root(), normals, colors, t.nodes_in_block, W, H, vpi, flags=flags)
        np.testing.assert_array_equal(codes.read_all(np.int32).reshape(H, W), r["codes"])
        np.testing.assert_array_equal(out.read_all(np.uint32).reshape(H, W), r["image"])
        assert st == r["stats"]
        errors += int(np.count_nonzero(r["image"] == red))
    assert errors > 1000, "the bad links were not reached"
    for wgs in (0, 16):
        b = api.BatchRenderer(m, len(specs))
        b.config(tile_shape=0, lds_levels=10, wgs_per_cu=wgs, tree_depth=8)
        views, total, offsets = b.make_views(specs, flags=flags)
        out = api.Buffer(gpu_ctx, 4 * total)
        cod = api.Buffer(gpu_ctx, 4 * total)
        b.render(views, len(specs), out, cod)
        img, cc = out.read_all(np.uint32), cod.read_all(np.int32)
        for (vpi, w, h), off in zip(specs, offsets):
            r = orc.render_plain(links, far, c.root(), normals, colors, t.nodes_in_block, w, h, vpi, flags=flags)
            np.testing.assert_array_equal(cc[off:off + w * h], r["codes"].reshape(-1), err_msg=f"wgs {wgs}")
            np.testing.assert_array_equal(img[off:off + w * h], r["image"].reshape(-1), err_msg=f"wgs {wgs}")
        b.close()
    m.close()


# (waves per workgroup, LDS levels, declared depth) of the pass kernel: spill from level 0 (the default),
# window-only, spill above a window at level 3, the full LDS stack; one-wave workgroups with a window and
# with spill above it.
@pytest.mark.parametrize("waves,levels,depth", [(4, 16, 0), (4, 8, 8), (4, 8, 12), (4, 22, 0), (1, 8, 8), (1, 8, 12)])
@pytest.mark.parametrize("flags", [0, api.L.RT_FLAG_SHADOWS])
def test_init_pass_finish_match_oracle(gpu_ctx, waves, levels, depth, flags):
    """InitTracingFrame, one RaytracingPass (node-level codes) and FinishTracingFrame as separate kernels."""
    api.plain_config(gpu_ctx, levels, waves, depth)
    try:
        _init_pass_finish(gpu_ctx, flags)
    finally:
        api.plain_config(gpu_ctx, 16, 4, 0)


def _init_pass_finish(gpu_ctx, flags):
    tree, path = terrain_tree(8, 8192, 8, 0)
    m = _gpu_cache(gpu_ctx, tree)
    t, c = _oracle_pool(path)
    links, far, normals, colors = _oracle_arrays(t, c)
    W, H = 200, 150
    P = W * H
    lib = api.lib()
    for yaw, pitch, pos in POSES[:3]:
        vpi = camera_vpi(yaw, pitch, pos)
        states = api.Buffer(gpu_ctx, P * 48)
        api.call("rt_init_tracing_frame", gpu_ctx.h, states.h, W, H, api._fp(vpi), np.float32(1.5))
        raw = states.read_all(np.uint8)
        org = raw[:16 * P].view(np.float32).reshape(P, 4)
        dr = raw[16 * P:32 * P].view(np.float32).reshape(P, 4)
        misc = raw[32 * P:].view(np.int32).reshape(P, 4)
        ost = orc.init_tracing_frame(W, H, vpi, 1.5).reshape(P, 12)
        np.testing.assert_array_equal(org.view(np.uint32), ost[:, 0:4].view(np.uint32))
        np.testing.assert_array_equal(dr.view(np.uint32), ost[:, 4:8].view(np.uint32))
        np.testing.assert_array_equal(misc, ost[:, 8:12].view(np.int32))
        codes = api.Buffer(gpu_ctx, P * 4)
        nrm, col = m.channel_by_name("normals"), m.channel_by_name("colors")
        api.call("rt_raytracing_pass", gpu_ctx.h, codes.h, states.h, W, H, m.channel(0).h, m.far_pointers().h,
                 m.root_node_index(), nrm.h, col.h, 0, flags)
        ocodes = orc.raytracing_pass(ost.reshape(-1), W, H, links, far, c.root(), normals, colors, 0, flags)
        np.testing.assert_array_equal(codes.read_all(np.int32), ocodes)
        out = api.Buffer(gpu_ctx, P * 4)
        api.call("rt_finish_tracing_frame", gpu_ctx.h, out.h, states.h, W, H, api._fp(api.BACKGROUND), nrm.h, col.h)
        np.testing.assert_array_equal(out.read_all(np.uint32),
                                      orc.finish_tracing_frame(ost.reshape(-1), W, H, normals, colors))


def test_depth_kernel_matches_oracle(gpu_ctx):
    tree, path = terrain_tree(8, 8192, 8, 0)
    m = _gpu_cache(gpu_ctx, tree)
    t, c = _oracle_pool(path)
    links, far, _, _ = _oracle_arrays(t, c)
    W, H = 256, 192
    for yaw, pitch, pos in POSES[:4]:
        vpi = camera_vpi(yaw, pitch, pos)
        out = api.Buffer(gpu_ctx, W * H * 4)
        api.call("rt_raytrace_depth", gpu_ctx.h, out.h, W, H, m.channel(0).h, m.far_pointers().h,
                 m.root_node_index(), api._fp(vpi), np.float32(1.0))
        got = out.read_all(np.uint32)
        exp, _ = orc.raytrace_depth(links, far, c.root(), W, H, vpi, 1.0)
        # the grey's logarithm has a fixed operation order on both sides: bit-exact
        np.testing.assert_array_equal(got, exp, err_msg=f"depth image differs at pose {yaw, pitch, pos}")


def test_restart_kernel_matches_oracle(gpu_ctx):
    tree, path = terrain_tree(8, 8192, 8, 0)
    m = _gpu_cache(gpu_ctx, tree)
    t, c = _oracle_pool(path)
    links, far, _, _ = _oracle_arrays(t, c)
    n = len(links)
    legacy = api.Buffer(gpu_ctx, n * 4)
    api.call("rt_legacy_links_from_cache", gpu_ctx.h, legacy.h, m.channel(0).h, m.far_pointers().h, n)
    oleg = orc.legacy_links(links, far)
    np.testing.assert_array_equal(legacy.read_all(np.uint32), oleg)
    W, H = 256, 192
    for yaw, pitch, pos in POSES[:4]:
        vpi = camera_vpi(yaw, pitch, pos)
        out = api.Buffer(gpu_ctx, W * H * 4)
        api.call("rt_raytrace_restart", gpu_ctx.h, out.h, W, H, legacy.h, m.root_node_index(), api._fp(vpi),
                 np.float32(1.0))
        exp, _ = orc.raytrace_restart(oleg, c.root(), W, H, vpi, 1.0)
        np.testing.assert_array_equal(out.read_all(np.uint32), exp)


def test_undersized_channels_are_refused_and_foreign_fault_parents_shade_as_no_node(gpu_ctx):
    """The raw pass and the plain render refuse normals / colour buffers smaller than the node links (the
    shading reads them at hit node indices); the finish kernel shades a fault parent outside the channels
    (a state no pass wrote) as no node instead of reading past them."""
    tree, path = terrain_tree(6, 512, 2, 2)
    m = _gpu_cache(gpu_ctx, tree)
    W, H = 64, 48
    P = W * H
    vpi = camera_vpi(*POSES[0])
    nrm, col = m.channel_by_name("normals"), m.channel_by_name("colors")
    short = api.Buffer(gpu_ctx, nrm.size() - 4)
    states = api.Buffer(gpu_ctx, P * 48)
    api.call("rt_init_tracing_frame", gpu_ctx.h, states.h, W, H, api._fp(vpi), np.float32(0.0))
    codes, out = api.Buffer(gpu_ctx, P * 4), api.Buffer(gpu_ctx, P * 4)
    for n, c in ((short, col), (nrm, short)):
        with pytest.raises(api.RaytError) as e:
            api.call("rt_raytracing_pass", gpu_ctx.h, codes.h, states.h, W, H, m.channel(0).h, m.far_pointers().h,
                     m.root_node_index(), n.h, c.h, 0, 0)
        assert e.value.code == api.L.RT_ERR_RANGE
        with pytest.raises(api.RaytError) as e:
            api.call("rt_render_plain", gpu_ctx.h, out.h, codes.h, W, H, m.channel(0).h, m.far_pointers().h,
                     m.root_node_index(), n.h, c.h, 0, api._fp(vpi), np.float32(0.0), api._fp(api.BACKGROUND), 0,
                     None)
        assert e.value.code == api.L.RT_ERR_RANGE
    # unfinished rays whose fault parent is the last node, one past it, and far beyond
    nodes = nrm.size() // 4
    misc = np.zeros((P, 4), np.int32)
    misc[:, 2] = 1  # colour multiplier: unfinished
    misc[:, 0] = np.resize(np.array([nodes - 1, nodes, 0x7ffffff0], np.int64), P).astype(np.int32)
    states.Write(32 * P, misc.reshape(-1))
    api.call("rt_finish_tracing_frame", gpu_ctx.h, out.h, states.h, W, H, api._fp(api.BACKGROUND), nrm.h, col.h)
    got = out.read_all(np.uint32)
    t, c = _oracle_pool(path)
    _, _, normals, colors = _oracle_arrays(t, c)
    st = np.zeros((P, 12), np.float32)
    st.view(np.int32)[:, 8:12] = misc
    st.view(np.int32)[:, 8][misc[:, 0] >= nodes] = -1  # the oracle's no-node fallback
    np.testing.assert_array_equal(got, orc.finish_tracing_frame(st.reshape(-1), W, H, normals, colors))

"""GPU tests of the continuous-batching server: every request's image and
feedback codes are the same alone, in a mix of scenes and sizes, under a
framebuffer arena too small for the queue, and with a block-pool budget that
holds one scene at a time; and all of them equal the plain single-view render
and the oracle's render of the same view. Out-of-core scenes are replayed
request by request in the oracle's cache and compared bit for bit."""
import numpy as np
import pytest

import oracle as orc
from rayt_amd import api
from tests.scenes import POSES, OraclePartialScene, camera_vpi, terrain_tree

pytestmark = pytest.mark.gpu

SIZES = [(320, 240), (256, 256), (100, 37), (640, 480), (7, 5), (200, 150)]
SHADOWS = api.L.RT_FLAG_SHADOWS


def _pool_bytes(tree):
    i = tree.info()
    return i.blocks_count * (i.bytes_in_block_content + 36)


@pytest.fixture(scope="module")
def scenes(gpu_ctx):
    out = []
    for depth, nib, spheres, boxes in [(8, 8192, 8, 0), (9, 1024, 0, 16)]:
        tree, path = terrain_tree(depth, nib, spheres, boxes)
        m = api.GPUOctreeCacheManager(tree.info().blocks_count, tree, gpu_ctx)
        m.InitialFillCache()
        out.append((tree, path, m, depth))
    return out


def _requests(n, nscenes, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for i in range(n):
        w, h = SIZES[i % len(SIZES)]
        if i < len(POSES):
            yaw, pitch, pos = POSES[i]
        else:
            yaw, pitch = float(rng.uniform(-180, 180)), float(rng.uniform(-60, 30))
            pos = tuple(float(x) for x in rng.uniform(0, 1, 3))
        reqs.append((camera_vpi(yaw, pitch, pos, aspect=w / h), w, h, i % nscenes))
    return reqs


def _plain(gpu_ctx, scenes, req):
    vpi, w, h, s = req
    out, codes, _ = api.render_plain(gpu_ctx, scenes[s][2], w, h, vpi, flags=SHADOWS)
    return out.read_all(np.uint32), codes.read_all(np.int32)


def _serve(server, reqs, one_at_a_time=False):
    results = {}
    if one_at_a_time:
        ids = []
        for vpi, w, h, s in reqs:
            ids.append(server.submit(vpi, w, h, scene=s, flags=SHADOWS, want_codes=True))
            server.drain()
    else:  # one burst: the server sees the whole queue at once
        ids = server.submit_many(reqs, flags=SHADOWS, want_codes=True)
    server.drain()
    got = server.poll(cap=4096)
    for c, img, cod in got:
        assert c.status == 0, f"request {c.id} failed with {c.status}"
        assert c.submit_ms <= c.admit_ms <= c.done_ms
        results[c.id] = (img, cod, c.round_views)
    assert sorted(results) == sorted(ids)
    return [results[i] for i in ids]


def _server(gpu_ctx, scenes, **kw):
    srv = api.RenderServer(gpu_ctx, **kw)
    for tree, _, _, depth in scenes:
        srv.add_scene(tree, depth)
    return srv


def test_server_alone_equals_mix_equals_plain(gpu_ctx, scenes):
    reqs = _requests(24, len(scenes), seed=11)
    srv = _server(gpu_ctx, scenes, max_views_per_round=16, rounds_in_flight=2, framebuffer_bytes=64 << 20)
    try:
        alone = _serve(srv, reqs, one_at_a_time=True)
        assert all(rv == 1 for _, _, rv in alone)
        srv.reset_stats()
        mix = _serve(srv, reqs)
        st = srv.stats()
    finally:
        srv.close()
    assert st.completed == len(reqs) and st.failed == 0
    assert st.rounds == 2, "a burst of 24 requests fills rounds of up to 16 views"
    assert max(rv for _, _, rv in mix) == 16
    for i, req in enumerate(reqs):
        pi, pc = _plain(gpu_ctx, scenes, req)
        np.testing.assert_array_equal(alone[i][0], mix[i][0], err_msg=f"request {i} image alone vs mix")
        np.testing.assert_array_equal(alone[i][1], mix[i][1], err_msg=f"request {i} codes alone vs mix")
        np.testing.assert_array_equal(mix[i][0], pi, err_msg=f"request {i} image vs plain")
        np.testing.assert_array_equal(mix[i][1], pc, err_msg=f"request {i} codes vs plain")
    # every request directly against the oracle's render over the resident pool of its scene
    pools = []
    for _, path, _, _ in scenes:
        t = orc.Tree(path)
        c = orc.Cache(t, t.blocks)
        c.initial_fill()
        pools.append((c.links(), c.far(), c.root(), c.channel_bytes(1, 4).view(np.uint32),
                      c.channel_bytes(2, 4).view(np.uint32), t.nodes_in_block))
    shadow_rays = 0
    for i, (vpi, w, h, s) in enumerate(reqs):
        r = orc.render_plain(*pools[s], w, h, vpi, flags=SHADOWS)
        np.testing.assert_array_equal(mix[i][0], r["image"].reshape(-1), err_msg=f"request {i} image vs oracle")
        np.testing.assert_array_equal(mix[i][1], r["codes"].reshape(-1), err_msg=f"request {i} codes vs oracle")
        shadow_rays += r["stats"][1]
    # the rounds' ray counts (counted without node fetches) are the oracle's
    assert st.primary_rays == st.pixels == sum(w * h for _, w, h, _ in reqs)
    assert st.shadow_rays == shadow_rays > 0
    assert 0 < st.p50_ms <= st.p99_ms <= st.max_ms


def test_server_large_and_small_rounds_equal_plain(gpu_ctx, scenes):
    """A round over two 3840x2160 views' pixels renders into device memory and is copied after the kernel; a
    smaller one has the kernel write its images into the pinned mirror. Both give the plain render."""
    rng = np.random.default_rng(5)
    big = []
    for k in range(3):
        yaw, pitch = float(rng.uniform(-180, 180)), float(rng.uniform(-40, 10))
        pos = tuple(float(x) for x in rng.uniform(0, 1, 3))
        big.append((camera_vpi(yaw, pitch, pos, aspect=3840 / 2160), 3840, 2160, k % len(scenes)))
    small = _requests(4, len(scenes), seed=6)
    srv = _server(gpu_ctx, scenes, max_views_per_round=8, rounds_in_flight=2, framebuffer_bytes=512 << 20)
    try:
        served = _serve(srv, big + small)  # one round: 24.9 M pixels
        st = srv.stats()
        assert st.rounds == 1
        served_small = _serve(srv, small)  # one round: under 1 M pixels
    finally:
        srv.close()
    for i, req in enumerate(big + small):
        pi, pc = _plain(gpu_ctx, scenes, req)
        np.testing.assert_array_equal(served[i][0], pi, err_msg=f"request {i} image (large round) vs plain")
        np.testing.assert_array_equal(served[i][1], pc, err_msg=f"request {i} codes (large round) vs plain")
    for i, req in enumerate(small):
        pi, pc = _plain(gpu_ctx, scenes, req)
        np.testing.assert_array_equal(served_small[i][0], pi, err_msg=f"request {i} image (small round) vs plain")
        np.testing.assert_array_equal(served_small[i][1], pc, err_msg=f"request {i} codes (small round) vs plain")


def test_server_admission_by_framebuffer(gpu_ctx, scenes):
    """An arena that holds two 640x480 outputs still serves a queue of ten; bigger requests are refused."""
    reqs = [(camera_vpi(*POSES[i % len(POSES)], aspect=4 / 3), 640, 480, i % 2) for i in range(10)]
    srv = _server(gpu_ctx, scenes, max_views_per_round=8, rounds_in_flight=2, framebuffer_bytes=8 * 640 * 480 * 2)
    try:
        with pytest.raises(api.RaytError):
            srv.submit(reqs[0][0], 1280, 720, scene=0)
        res = _serve(srv, reqs)
        st = srv.stats()
    finally:
        srv.close()
    assert st.completed == 10 and st.mean_views_per_round == 2.0
    for i, req in enumerate(reqs):
        pi, pc = _plain(gpu_ctx, scenes, req)
        np.testing.assert_array_equal(res[i][0], pi)
        np.testing.assert_array_equal(res[i][1], pc)


def test_server_admission_by_block_pool(gpu_ctx, scenes):
    """A pool budget that holds one scene: scenes are loaded and evicted between rounds, results unchanged."""
    budget = max(_pool_bytes(t) for t, _, _, _ in scenes)
    reqs = _requests(12, len(scenes), seed=5)
    srv = _server(gpu_ctx, scenes, max_views_per_round=4, rounds_in_flight=2, framebuffer_bytes=64 << 20,
                  pool_bytes=budget)
    try:
        res = _serve(srv, reqs)
        st = srv.stats()
    finally:
        srv.close()
    assert st.completed == 12
    assert st.scene_loads >= 2 and st.scene_evictions >= 1
    for i, req in enumerate(reqs):
        pi, pc = _plain(gpu_ctx, scenes, req)
        np.testing.assert_array_equal(res[i][0], pi)
        np.testing.assert_array_equal(res[i][1], pc)


def test_server_refuses_scene_over_budget(gpu_ctx, scenes):
    tree = scenes[0][0]
    srv = api.RenderServer(gpu_ctx, pool_bytes=_pool_bytes(tree) - 1, framebuffer_bytes=1 << 20)
    try:
        with pytest.raises(api.RaytError):
            srv.add_scene(tree, 8)
    finally:
        srv.close()


def test_batch_views_over_several_scenes(gpu_ctx, scenes):
    """rt_batch_add_scene: one launch over views of two resident scenes equals their plain renders."""
    b = api.BatchRenderer(scenes[0][2], 32)
    assert b.add_scene(scenes[1][2]) == 1
    b.config(tile_shape=0, lds_levels=8, wgs_per_cu=8, tree_depth=9)
    reqs = _requests(10, 2, seed=2)
    views, total, offsets = b.make_views(reqs, flags=SHADOWS)
    out = api.Buffer(gpu_ctx, 4 * total)
    codes = api.Buffer(gpu_ctx, 4 * total)
    b.render(views, len(reqs), out, codes)
    img, cod = out.read_all(np.uint32), codes.read_all(np.int32)
    for req, off in zip(reqs, offsets):
        w, h = req[1], req[2]
        pi, pc = _plain(gpu_ctx, scenes, req)
        np.testing.assert_array_equal(img[off:off + w * h], pi)
        np.testing.assert_array_equal(cod[off:off + w * h], pc)
    b.close()


def test_objects_outlive_their_context_handle(gpu_ctx, scenes):
    """Closing a context first leaves its cache, batch and buffers working (each holds a context reference)."""
    ctx2 = api.Context(0)
    tree = scenes[0][0]
    m = api.GPUOctreeCacheManager(tree.info().blocks_count, tree, ctx2)
    m.InitialFillCache()
    b = api.BatchRenderer(m, 4)
    b.config(tile_shape=0, lds_levels=8, wgs_per_cu=8, tree_depth=8)
    vpi = camera_vpi(*POSES[0], aspect=160 / 120)
    views, total, _ = b.make_views([(vpi, 160, 120)], flags=SHADOWS)
    out = api.Buffer(ctx2, 4 * total)
    ctx2.close()
    b.render(views, 1, out)
    img = out.read_all(np.uint32)
    b.close()
    m.close()
    out.close()
    pi, _ = _plain(gpu_ctx, scenes, (vpi, 160, 120, 0))
    np.testing.assert_array_equal(img, pi)


def test_server_pinned_delivery_and_release(gpu_ctx, scenes):
    """Pinned results stay valid until released, hold their arena range, and equal the plain render."""
    reqs = [(camera_vpi(*POSES[i], aspect=4 / 3), 640, 480, i % 2) for i in range(4)]
    srv = _server(gpu_ctx, scenes, max_views_per_round=8, rounds_in_flight=2, framebuffer_bytes=8 * 640 * 480 * 2)
    try:
        ids = srv.submit_many(reqs, flags=SHADOWS, want_codes=True, pinned=True)
        got = []
        while len(got) < 2:
            got += srv.poll(cap=8, timeout_ms=5000, keep_pinned=True)
        assert len(got) == 2
        # the arena is full of held results: nothing else can be admitted
        assert srv.poll(cap=8, timeout_ms=300, keep_pinned=True) == []
        for c, img, cod in got:
            k = ids.index(c.id)
            pi, pc = _plain(gpu_ctx, scenes, reqs[k])
            np.testing.assert_array_equal(img, pi)
            np.testing.assert_array_equal(cod, pc)
            srv.release(c.id)
        with pytest.raises(api.RaytError):
            srv.release(got[0][0].id)  # already released
        srv.drain(10000)
        rest = srv.poll(cap=8)  # copied out and released by the wrapper
        assert sorted(c.id for c, _, _ in rest) == sorted(set(ids) - {c.id for c, _, _ in got})
        for c, img, cod in rest:
            pi, pc = _plain(gpu_ctx, scenes, reqs[ids.index(c.id)])
            np.testing.assert_array_equal(img, pi)
            np.testing.assert_array_equal(cod, pc)
    finally:
        srv.close()


NO_HIT = 0x7FFFFFFF


def test_partial_scene_converges_to_the_resident_render(gpu_ctx, scenes):
    """A scene served from a pool of 40 % of its blocks: the first result misses blocks (their pixels show
    the coarser parent), the feedback between rounds loads them, and repeating the request converges to the
    fully resident render (same pixels; no fault codes, the same pixels hit; hit ids are pool slots).
    Every request's image, codes and fault count equal the oracle's replay of the same requests and loads,
    and the converged image equals the oracle's render over the fully resident pool.
    Primary rays only: a shadow ray that meets a missing block counts as unoccluded and is not fed back
    (the pixel's code is its primary ray's), in the frame path as here."""
    tree, path, _, depth = scenes[1]
    blocks = tree.info().blocks_count
    vpi = camera_vpi(0.0, -89.0, (0.43, 0.66, 0.57), fov=30.0, aspect=1.0)
    t = orc.Tree(path)
    full = orc.Cache(t, t.blocks)
    full.initial_fill()
    r = orc.render_plain(full.links(), full.far(), full.root(), full.channel_bytes(1, 4).view(np.uint32),
                         full.channel_bytes(2, 4).view(np.uint32), t.nodes_in_block, 256, 256, vpi)
    ref_img, ref_cod = r["image"].reshape(-1), r["codes"].reshape(-1)
    oracle = OraclePartialScene(path, int(blocks * 0.4))
    srv = api.RenderServer(gpu_ctx, max_views_per_round=4, rounds_in_flight=2, framebuffer_bytes=8 << 20)
    try:
        s = srv.add_scene(tree, depth, max_blocks=int(blocks * 0.4))
        faults = []
        for k in range(16):
            srv.submit(vpi, 256, 256, scene=s, want_codes=True)
            srv.drain()
            (c, img, cod), = srv.poll(cap=4)
            assert c.status == 0
            oi, oc, of = oracle.request(vpi, 256, 256)
            np.testing.assert_array_equal(cod, oc, err_msg=f"codes of request {k}")
            np.testing.assert_array_equal(img, oi, err_msg=f"image of request {k}")
            assert c.fault_blocks == of, k
            faults.append(c.fault_blocks)
            assert (c.fault_blocks > 0) == bool(np.any(cod < 0))
            if c.fault_blocks == 0:
                break
        assert faults[0] > 0, "the initial fill already held every block of the view"
        assert faults[-1] == 0, f"no convergence: faulted blocks per request {faults}"
        np.testing.assert_array_equal(img, ref_img)
        np.testing.assert_array_equal(cod == NO_HIT, ref_cod == NO_HIT)
        st = srv.stats()
        assert st.block_uploads > 0 and st.load_transactions > 0
    finally:
        srv.close()


def test_partial_scene_in_a_mixed_burst(gpu_ctx, scenes):
    """A burst over a resident scene and a partial one: the resident scene's results still equal the
    oracle's render, every partial-scene result is consistent with its fault count, and loads happen between rounds."""
    tree0, path0, _, d0 = scenes[0]
    tree1, _, _, d1 = scenes[1]
    t0 = orc.Tree(path0)
    c0 = orc.Cache(t0, t0.blocks)
    c0.initial_fill()
    resident0 = (c0.links(), c0.far(), c0.root(), c0.channel_bytes(1, 4).view(np.uint32),
                 c0.channel_bytes(2, 4).view(np.uint32), t0.nodes_in_block)
    srv = api.RenderServer(gpu_ctx, max_views_per_round=6, rounds_in_flight=2, framebuffer_bytes=64 << 20)
    try:
        s0 = srv.add_scene(tree0, d0)
        s1 = srv.add_scene(tree1, d1, max_blocks=int(tree1.info().blocks_count * 0.4))
        reqs = [(vpi, w, h, s0 if s == 0 else s1) for vpi, w, h, s in _requests(24, 2, seed=5)]
        ids = srv.submit_many(reqs, flags=SHADOWS, want_codes=True)
        srv.drain()
        got = {c.id: (c, img, cod) for c, img, cod in srv.poll(cap=64)}
        assert sorted(got) == sorted(ids)
        for rid, req in zip(ids, reqs):
            c, img, cod = got[rid]
            assert c.status == 0
            if req[3] == s0:
                r = orc.render_plain(*resident0, req[1], req[2], req[0], flags=SHADOWS)
                np.testing.assert_array_equal(img, r["image"].reshape(-1))
                np.testing.assert_array_equal(cod, r["codes"].reshape(-1))
                assert c.fault_blocks == 0
            else:
                assert (c.fault_blocks > 0) == bool(np.any(cod < 0))
        assert srv.stats().load_transactions > 0
    finally:
        srv.close()


def test_partial_scene_load_budget(gpu_ctx, scenes):
    """max_block_loads bounds every load between rounds; the scene still converges, in more requests, and
    every request's image and codes equal the oracle's replay under the same budget."""
    tree, path, _, depth = scenes[1]
    vpi = camera_vpi(0.0, -89.0, (0.43, 0.66, 0.57), fov=30.0, aspect=1.0)
    max_blocks = int(tree.info().blocks_count * 0.4)
    oracle = OraclePartialScene(path, max_blocks, max_block_loads=3)
    srv = api.RenderServer(gpu_ctx, max_views_per_round=4, rounds_in_flight=2, framebuffer_bytes=8 << 20,
                           max_block_loads=3)
    try:
        s = srv.add_scene(tree, depth, max_blocks=max_blocks)
        faults = []
        for k in range(64):
            srv.submit(vpi, 256, 256, scene=s, want_codes=True)
            srv.drain()
            (c, img, cod), = srv.poll(cap=4)
            assert c.status == 0
            oi, oc, of = oracle.request(vpi, 256, 256)
            np.testing.assert_array_equal(cod, oc, err_msg=f"codes of request {k}")
            np.testing.assert_array_equal(img, oi, err_msg=f"image of request {k}")
            assert c.fault_blocks == of, k
            faults.append(c.fault_blocks)
            st = srv.stats()
            assert st.block_uploads <= 3 * st.load_transactions
            if c.fault_blocks == 0:
                break
        assert faults[0] > 3 and faults[-1] == 0, f"faulted blocks per request {faults}"
    finally:
        srv.close()

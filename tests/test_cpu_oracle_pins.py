"""CPU tests that pin the oracle functions not covered by test_cpu_oracle.py
against facts they do not compute themselves:

  * the three-kernel frame (InitTracingFrame + RaytracingPass +
    FinishTracingFrame) against the fused single-pass render, which the
    brute-force and closed-form tests pin; node-level codes through
    NodeToBlock against the fused block-level codes;
  * the log-depth kernel and the stackless restart traversal against the
    brute-force nearest leaf of their own rays;
  * the legacy link layout against the cache-format decode of the same pool;
  * the multi-pass out-of-core frame against the resident render: every
    pixel that ends on a hit or a miss shows the resident colour;
  * the bad-link extension (a link outside the pool ends the ray with
    error 2) against the clean pool: a ray either meets a corrupted link and
    ends red with no feedback code, or never does and renders as before.
"""
import numpy as np
import pytest

import oracle as orc
from tests.scenes import POSES, camera_vpi, corrupt_root_links, terrain_tree
from tests.test_cpu_oracle import NO_HIT, _children, _leaves, _oracle_resident, _slab

SHADOWS = 1


def _arrays(c):
    return c.links(), c.far(), c.channel_bytes(1, 4).view(np.uint32), c.channel_bytes(2, 4).view(np.uint32)


@pytest.mark.parametrize("flags", [0, SHADOWS])
def test_three_kernel_frame_equals_fused_render(flags):
    tree, path = terrain_tree(7, 512, 4, 4)
    t, c = _oracle_resident(path)
    links, far, normals, colors = _arrays(c)
    W, H = 160, 120
    for yaw, pitch, pos in POSES[:4]:
        vpi = camera_vpi(yaw, pitch, pos)
        fused = orc.render_plain(links, far, c.root(), normals, colors, t.nodes_in_block, W, H, vpi, lod=1.0,
                                 flags=flags)
        st = orc.init_tracing_frame(W, H, vpi, 1.0)
        codes_block = orc.raytracing_pass(st.copy(), W, H, links, far, c.root(), normals, colors,
                                          t.nodes_in_block, flags)
        codes_node = orc.raytracing_pass(st, W, H, links, far, c.root(), normals, colors, 0, flags)
        img = orc.finish_tracing_frame(st, W, H, normals, colors)
        np.testing.assert_array_equal(img.reshape(H, W), fused["image"])
        np.testing.assert_array_equal(codes_block.reshape(H, W), fused["codes"])
        # node-level codes are ((node + 1) << 1) | dup; NodeToBlock maps them to the block-level ones
        np.testing.assert_array_equal(orc.node_to_block(codes_node, t.nodes_in_block), codes_block)
        hits = codes_node[(codes_node > 0) & (codes_node != NO_HIT)]
        assert len(hits) > 0
        np.testing.assert_array_equal(orc.node_to_block(hits, t.nodes_in_block),
                                      ((((hits >> 1) - 1) // t.nodes_in_block + 1) << 1) | (hits & 1))


def _near_plane_rays(vpi, W, H):
    """Rays of the single-pass kernels (raytracer.cl:29-50), in double precision: origin = the unprojected
    pixel centre at clip z = -1, direction toward its unprojection at z = 0."""
    m = np.asarray(vpi, np.float64).reshape(4, 4)

    def unproject(u, v, z):
        p = m @ np.array([u, v, z, 1.0])
        return p[:3] / p[3]

    out = []
    for y in range(H):
        for x in range(W):
            u, v = (x + 0.5) / W * 2 - 1, (y + 0.5) / H * 2 - 1
            o = unproject(u, v, -1.0)
            d = unproject(u, v, 0.0) - o
            out.append((o, d / np.linalg.norm(d)))
    return out


def _brute_force(vpi, W, H, links, far, root):
    """Per pixel: (nearest leaf id or -1, its entry t) or None where the nearest leaf is ambiguous."""
    leaves = _leaves(links, far, root)
    lo = np.array([[x, y, z] for _, x, y, z, _ in leaves])
    hi = lo + np.array([s for *_, s in leaves])[:, None]
    ids = np.array([n for n, *_ in leaves])
    out = []
    for o, d in _near_plane_rays(vpi, W, H):
        tn, span = _slab(o, d, lo, hi)
        order = np.argsort(tn)
        if not np.isfinite(tn[order[0]]):
            out.append((-1, np.inf))
        elif (len(order) > 1 and tn[order[1]] - tn[order[0]] < 1e-4) or span[order[0]] < 1e-4:
            out.append(None)  # edges and corners: two leaves entered at almost the same t
        else:
            out.append((int(ids[order[0]]), float(tn[order[0]])))
    return out


def _agreement(hit_nodes, bf):
    checked = [(k, e) for k, e in enumerate(bf) if e is not None]
    agree = sum(int(hit_nodes[k] == e[0]) for k, e in checked)
    return agree, len(checked)


def test_ln_fixed_matches_libm():
    """The log-depth grey's fixed-order logarithm against libm's double log rounded to float: equal to the
    bit on every sample (log-uniform over the float range, uniform over the depth range, powers of two and
    the reduction's boundary at sqrt(2)), and the special values ln 0 = -inf, ln 1 = 0, ln inf = inf."""
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(0, 10, 200000), np.exp(rng.uniform(-100, 88, 200000)),
                        2.0 ** np.arange(-149, 128), [np.sqrt(2.0)]]).astype(np.float32)
    x = np.concatenate([x, np.nextafter(x, np.float32(np.inf)), np.nextafter(x, np.float32(0))])
    x = x[(x > 0) & np.isfinite(x)]
    got = orc.ln_fixed(x)
    ref = np.log(x.astype(np.float64)).astype(np.float32)
    np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))
    sp = orc.ln_fixed(np.array([0.0, 1.0, np.inf], np.float32))
    assert sp[0] == -np.inf and sp[1] == 0.0 and sp[2] == np.inf


def test_depth_kernel_against_brute_force():
    """Hit voxels equal the brute-force nearest leaf, and the grey level is the log-depth map of raytracer.cl
    applied to the brute-force distance: a least-squares line of grey against ln t must have the map's slope
    -255 / ln(5 / 0.01) and its value 255 at t = 0.01."""
    tree, path = terrain_tree(5, 64, 2, 2)
    t, c = _oracle_resident(path)
    links, far, _, _ = _arrays(c)
    W, H = 48, 36
    lnt, grey = [], []
    for yaw, pitch, pos in POSES[:3]:
        vpi = camera_vpi(yaw, pitch, pos)
        img, hn = orc.raytrace_depth(links, far, c.root(), W, H, vpi, 0.0)
        bf = _brute_force(vpi, W, H, links, far, c.root())
        agree, checked = _agreement(hn, bf)
        assert checked > 0.8 * W * H
        assert agree >= checked - 2, (agree, checked)
        g = img.view(np.uint8).reshape(-1, 4)[:, 0].astype(float)
        for k, e in enumerate(bf):
            if e is not None and e[0] >= 0 and hn[k] == e[0] and 0 < g[k] < 255:
                lnt.append(np.log(e[1]))
                grey.append(g[k])
    assert len(lnt) > 500
    slope, icept = np.polyfit(np.array(lnt), np.array(grey), 1)
    assert abs(slope - (-255.0 / np.log(500.0))) < 0.02 * 255.0 / np.log(500.0), slope
    assert abs((slope * np.log(0.01) + icept) - 255.0) < 2.0, (slope, icept)


def test_restart_traversal_hits_brute_force_nearest_leaf():
    tree, path = terrain_tree(5, 64, 2, 2)
    t, c = _oracle_resident(path)
    links, far, _, _ = _arrays(c)
    legacy = orc.legacy_links(links, far)
    W, H = 48, 36
    for yaw, pitch, pos in POSES[:3]:
        vpi = camera_vpi(yaw, pitch, pos)
        _, hn = orc.raytrace_restart(legacy, c.root(), W, H, vpi, 0.0)
        agree, checked = _agreement(hn, _brute_force(vpi, W, H, links, far, c.root()))
        assert checked > 0.8 * W * H
        # the restart steps past an empty child by 1e-6 (raytracer_volatile.cl:94-97), which can skip a
        # sliver of a voxel: allow a few of the checked rays
        assert agree >= checked * 0.98, (agree, checked)


def test_legacy_links_decode_like_cache_links():
    tree, path = terrain_tree(6, 512, 2, 3)
    t, c = _oracle_resident(path)
    links, far, _, _ = _arrays(c)
    legacy = orc.legacy_links(links, far)
    todo, seen = [c.root()], 0
    while todo:  # every node reachable from the root
        i = todo.pop()
        seen += 1
        mask, first = _children(links, far, i)
        assert first is not None or not mask, "fault link in a fully resident pool"
        assert int(legacy[i]) & 255 == mask
        if mask:
            assert int(legacy[i]) >> 9 == first, i
            todo += [first + k for k in range(bin(mask).count("1"))]
    assert seen > 1000


def test_out_of_core_frame_matches_resident_colours():
    """An undersized pool, frame loop to the end: every pixel whose ray ends on a hit or a miss has the
    colour of the fully resident render (pool slots differ, the voxel and its channels do not)."""
    tree, path = terrain_tree(8, 256, 0, 8)
    t, full = _oracle_resident(path)
    links, far, normals, colors = _arrays(full)
    small = orc.Cache(t, max(8, int(t.blocks * 0.6)))
    small.initial_fill()
    W, H = 160, 120
    checked, passes = 0, []
    for yaw, pitch, pos in POSES[:4]:
        vpi = camera_vpi(yaw, pitch, pos)
        r = small.render_frame(W, H, vpi, max_passes=50, seed=5)
        res = orc.render_plain(links, far, full.root(), normals, colors, t.nodes_in_block, W, H, vpi)
        done = r["codes"] >= 0  # hits and misses of the last pass (faults are negative)
        same = (r["image"] == res["image"]) | ~done
        assert same.mean() >= 0.999, (1 - same.mean())
        checked += int(done.sum())
        passes.append(r["passes"])
    assert max(passes) >= 2, f"the pool was expected to fault and reload: passes {passes}"
    assert checked > 0.5 * 4 * W * H


@pytest.mark.parametrize("flags", [0, SHADOWS])
def test_bad_links_end_rays_with_error_two(flags):
    """Three of the root's inner children get bad links (a near pointer past the pool, a far index past
    the far table, a far pointer past the pool). Every pixel is then either the clean render's pixel (its
    rays never reach those links) or red with NO_HIT (the traversal stops at the link instead of reading
    outside the arrays), or, with shadow rays, a primary hit whose shadow ray met one and counts as
    unoccluded; the red ones are many, and the counts stay consistent: fewer fetches, fewer hits."""
    tree, path = terrain_tree(8, 8192, 8, 0)
    t, c = _oracle_resident(path)
    links, far, normals, colors = _arrays(c)
    bad_links, bad_far = links.copy(), far.copy()
    words, far_words = corrupt_root_links(links, far, c.root())
    for i, w in words:
        bad_links[i] = w
    for i, w in far_words:
        bad_far[i] = w
    red = 255 | (255 << 24)
    W, H = 200, 150
    total_red = 0
    for yaw, pitch, pos in POSES:
        vpi = camera_vpi(yaw, pitch, pos)
        ok = orc.render_plain(links, far, c.root(), normals, colors, t.nodes_in_block, W, H, vpi, flags=flags)
        bad = orc.render_plain(bad_links, bad_far, c.root(), normals, colors, t.nodes_in_block, W, H, vpi,
                               flags=flags)
        same = (bad["image"] == ok["image"]) & (bad["codes"] == ok["codes"])
        errored = (bad["image"] == red) & (bad["codes"] == NO_HIT)
        # a shadow ray that meets a bad link ends in error and counts as unoccluded: the pixel keeps its
        # code and takes its colour without the shadow (the render without shadow rays)
        unlit = orc.render_plain(links, far, c.root(), normals, colors, t.nodes_in_block, W, H, vpi, flags=0)
        lit = (bad["codes"] == ok["codes"]) & (bad["image"] == unlit["image"]) if flags else errored
        assert np.all(same | errored | lit)
        assert not np.any(ok["image"] == red)
        total_red += int(np.count_nonzero(errored))
        assert bad["stats"][2] <= ok["stats"][2] and bad["stats"][3] <= ok["stats"][3]
    assert total_red > 1000


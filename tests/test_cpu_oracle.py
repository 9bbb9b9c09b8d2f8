"""CPU tests that pin the oracle to facts it does not compute itself.

The reference ships no golden vectors for this path (its only tests are the
ad-hoc TestGPUSort / TestGPUScan / sphere checkers of
application/test-application.cpp:134-291), so the oracle is pinned by:
  * brute force: nearest leaf voxel along a ray by a double-precision slab
    test over every leaf of a small tree;
  * closed forms: misses, LOD cut at a known depth, shading of a known normal,
    the camera matrix against a double-precision inverse;
  * the reference's own structural checker (CheckTestOctreeSphereWithGPUData,
    restated) on pools filled by the oracle's cache, with a planted mistake
    that it must catch;
  * LRU invariants of gpu-octree-cache.h:106 (a child is always nearer the
    least-recently-used end than its parent) and load/unload round trips.
"""
import math

import numpy as np
import pytest

import oracle as orc
from rayt_amd import api
from tests.scenes import POSES, camera_vpi, sphere_tree, terrain_tree

NO_HIT = 0x7FFFFFFF


# ------------------------------------------------------------- helpers
def _children(links, far, i):
    """Absolute first child and mask of pool node i (cache link format)."""
    l = int(links[i])
    mask = l & 255
    if not mask or (l >> 8) & 1:
        return mask, None
    ptr = l >> 11
    first = int(far[ptr]) if (l >> 10) & 1 else i + ptr - (1 << 20)
    return mask, first


def _leaves(links, far, root):
    """All leaf voxels reachable from the root: (node, lo xyz, size)."""
    out = []
    stack = [(root, 0.0, 0.0, 0.0, 1.0)]
    while stack:
        n, x, y, z, s = stack.pop()
        mask, first = _children(links, far, n)
        if not mask:
            out.append((n, x, y, z, s))
            continue
        assert first is not None, "fault link in a fully resident pool"
        h = s / 2
        k = 0
        for ch in range(8):
            if mask & (1 << ch):
                stack.append((first + k, x + h * (ch & 1), y + h * ((ch >> 1) & 1), z + h * ((ch >> 2) & 1), h))
                k += 1
    return out


def _ray(vpi, x, y, W, H):
    m = np.asarray(vpi, np.float64).reshape(4, 4)
    o = m[0:3, 2] / m[3, 2]
    u, v = (x + 0.5) / W * 2 - 1, (y + 0.5) / H * 2 - 1
    p = m @ np.array([u, v, 0.0, 1.0])
    d = p[:3] / p[3] - o
    return o, d / np.linalg.norm(d)


def _slab(o, d, lo, hi):
    with np.errstate(divide="ignore", invalid="ignore"):
        inv = 1.0 / d
        t0 = (lo - o) * inv
        t1 = (hi - o) * inv
    tn = np.nanmax(np.minimum(t0, t1), axis=1)
    tf = np.nanmin(np.maximum(t0, t1), axis=1)
    tn = np.maximum(tn, 0.0)
    # a ray that only touches a voxel (zero-length overlap, e.g. leaving it from its boundary) misses it
    return np.where(tn < tf - 1e-9, tn, np.inf), tf - tn


def _oracle_resident(path):
    t = orc.Tree(path)
    c = orc.Cache(t, t.blocks)
    c.initial_fill()
    return t, c


# ------------------------------------------------------------ traversal
@pytest.mark.parametrize("scene", ["sphere5", "terrain6"])
def test_cast_ray_matches_brute_force(scene):
    if scene == "sphere5":
        _, path = sphere_tree((0.3, 0.4, 0.65), 0.23, 5, 111)
    else:
        _, path = terrain_tree(6, 512, 2, 3)
    t, c = _oracle_resident(path)
    links, far = c.links(), c.far()
    leaves = _leaves(links, far, c.root())
    nodes = np.array([lf[0] for lf in leaves])
    lo = np.array([lf[1:4] for lf in leaves])
    hi = lo + np.array([lf[4] for lf in leaves])[:, None]
    W, H = 64, 48
    checked = 0
    for yaw, pitch, pos in POSES:
        vpi = camera_vpi(yaw, pitch, pos)
        r = orc.render_plain(links, far, c.root(), None, None, t.nodes_in_block, W, H, vpi, want_hits=True)
        for y in range(H):
            for x in range(W):
                o, d = _ray(vpi, x, y, W, H)
                tt, span = _slab(o, d, lo, hi)
                order = np.argsort(tt)
                t1 = tt[order[0]]
                got = r["hit_node"][y, x]
                if not np.isfinite(t1):
                    assert got == -1, (x, y)
                    continue
                t2 = tt[order[1]] if len(order) > 1 else np.inf
                if t2 - t1 < 1e-4 or span[order[0]] < 1e-4:  # ambiguous at fp32 precision (tie or edge clip)
                    continue
                assert got == nodes[order[0]], f"pose {yaw, pitch, pos} pixel {x, y}"
                assert abs(r["hit_t"][y, x] - t1) < 1e-4
                checked += 1
    assert checked > 1000


def test_cast_ray_closed_forms():
    _, path = sphere_tree((0.5, 0.5, 0.5), 0.3, 4, 111)
    t, c = _oracle_resident(path)
    links, far = c.links(), c.far()
    # Camera far outside, looking away from the cube: every ray misses.
    vpi = camera_vpi(0.0, 0.0, (0.5, 0.5, 3.0))  # looks down -z towards the cube
    away = camera_vpi(180.0, 0.0, (0.5, 0.5, 3.0))  # looks +z, away from it
    r = orc.render_plain(links, far, c.root(), None, None, t.nodes_in_block, 32, 24, away, want_hits=True)
    assert (r["hit_node"] == -1).all() and (r["hit_t"] == np.float32(1e30)).all()
    assert (r["codes"] == NO_HIT).all()
    # A huge LOD stops every hit at the first level below the root.
    r = orc.render_plain(links, far, c.root(), None, None, t.nodes_in_block, 32, 24, vpi, lod=1e6, want_hits=True)
    hit = r["hit_node"] >= 0
    assert hit.any()
    first_level = set()
    mask, first = _children(links, far, c.root())
    first_level.update(range(first, first + bin(mask).count("1")))
    assert set(np.unique(r["hit_node"][hit])) <= first_level
    # lod 0: every hit is a leaf.
    r = orc.render_plain(links, far, c.root(), None, None, t.nodes_in_block, 32, 24, vpi, lod=0.0, want_hits=True)
    for n in np.unique(r["hit_node"][r["hit_node"] >= 0]):
        assert links[n] & 255 == 0


def _single_level_pool(normal, color):
    """Root with 8 leaf children, all sharing one normal and colour (cache format)."""
    links = np.zeros(9, np.uint32)
    links[0] = ((((1 << 20) + 1) << 11) | 0xFF)
    nb = np.array(normal + (0,), np.int8).view(np.uint8)
    normals = np.zeros(9, np.uint32)
    normals[:] = np.frombuffer(nb.tobytes(), np.uint32)[0]
    colors = np.zeros(9, np.uint32)
    colors[:] = np.frombuffer(np.array(color + (255,), np.uint8).tobytes(), np.uint32)[0]
    return links, np.zeros(8, np.uint32), normals, colors


def test_shading_closed_form():
    normal, color = (0, 127, 0), (200, 100, 50)
    links, far, normals, colors = _single_level_pool(normal, color)
    vpi = camera_vpi(0.0, -90.0, (0.5, 2.0, 0.5), fov=30.0)  # straight down onto the top face
    r = orc.render_plain(links, far, 0, normals, colors, 0, 16, 16, vpi, want_hits=True)
    assert (r["hit_node"] >= 1).all()
    L1 = np.array([2, -2, 3]) / math.sqrt(17) * 0.6
    L2 = np.array([-3, -3, -2]) / math.sqrt(22) * 0.5
    L3 = np.array([1, 2.5, -1]) / math.sqrt(8.25) * 0.3
    n = np.array(normal) / 127.0
    light = 0.1 + sum(max(0.0, float(-n @ L)) for L in (L1, L2, L3))
    exp = [min(255.0, ch * light) for ch in color]
    px = r["image"].view(np.uint8).reshape(16, 16, 4)
    for k in range(3):
        assert np.all(np.abs(px[..., k].astype(float) - exp[k]) <= 1.0)
    assert (px[..., 3] == 255).all()
    # Background: a miss blends (0, 216/255, 1, 1) * 255 + 0.5.
    r = orc.render_plain(links, far, 0, normals, colors, 0, 8, 8, camera_vpi(0.0, 90.0, (0.5, 2.0, 0.5)))
    bg = r["image"].view(np.uint8).reshape(-1, 4)
    assert (bg == np.array([0, 216, 255, 255], np.uint8)).all()


def test_shadow_ray_blocks_light_one():
    """A floor lit by light 1 under an overhang: shadowed pixels lose exactly light 1's term."""
    _, path = terrain_tree(6, 512, 0, 6)
    t, c = _oracle_resident(path)
    links, far = c.links(), c.far()
    normals = c.channel_bytes(1, 4).view(np.uint32)
    colors = c.channel_bytes(2, 4).view(np.uint32)
    vpi = camera_vpi(*POSES[0])
    a = orc.render_plain(links, far, c.root(), normals, colors, t.nodes_in_block, 96, 72, vpi, flags=0)
    b = orc.render_plain(links, far, c.root(), normals, colors, t.nodes_in_block, 96, 72, vpi, flags=1)
    np.testing.assert_array_equal(a["codes"], b["codes"])  # shadows never change what is hit
    da = a["image"].view(np.uint8).reshape(-1, 4)[:, :3].astype(int)
    db = b["image"].view(np.uint8).reshape(-1, 4)[:, :3].astype(int)
    assert (db <= da).all()  # occlusion only removes light
    assert (db < da).any() and (db == da).any()
    assert b["stats"][1] > 0


# --------------------------------------------------------------- camera
def test_camera_matches_double_precision_and_product():
    for yaw, pitch, roll, pos, fov, aspect in [(30, -20, 0, (0.1, 0.9, 1.3), 60, 4 / 3),
                                               (-135, 10, 5, (0.8, 0.7, 0.8), 90, 16 / 9),
                                               (0, -89, 0, (0.5, 1.5, 0.5), 45, 1.0)]:
        o = orc.camera_vpi(yaw, pitch, roll, pos, fov, 0.01, aspect)
        p = api.Camera(yaw, pitch, roll, pos, fov, 0.01, aspect).view_proj_inv()
        np.testing.assert_array_equal(o.view(np.uint32), p.view(np.uint32))
        r = math.radians

        def rot(axis, a):
            c, s = math.cos(r(a)), math.sin(r(a))
            m = np.eye(4)
            i, j = {0: (1, 2), 1: (2, 0), 2: (0, 1)}[axis]
            m[i, i], m[i, j], m[j, i], m[j, j] = c, -s, s, c
            return m

        R = rot(1, -yaw) @ rot(0, pitch) @ rot(2, -roll)
        T = np.eye(4)
        T[:3, 3] = -np.array(pos)
        f = 1 / math.tan(r(fov) / 2)
        Pm = np.zeros((4, 4))
        Pm[0, 0], Pm[1, 1], Pm[3, 2], Pm[2, 2], Pm[2, 3] = f, f * aspect, -1, -1, -0.02
        inv = np.linalg.inv(Pm @ R.T @ T)
        np.testing.assert_allclose(o.reshape(4, 4), inv, rtol=2e-4, atol=2e-5)
        # The eye point the renderer derives from the matrix is the camera position.
        m = o.reshape(4, 4).astype(np.float64)
        np.testing.assert_allclose(m[0:3, 2] / m[3, 2], pos, atol=1e-4)


# ---------------------------------------------------- structure / cache
@pytest.mark.parametrize("level,nib", [(3, 131), (5, 16), (7, 111), (9, 737)])
def test_sphere_tree_round_trip(level, nib):
    """WriteTestOctreeSphere -> loader -> InitialFillCache -> CheckTestOctreeSphereWithGPUData."""
    center, radius = (0.3, 0.4, 0.65), 0.23
    _, path = sphere_tree(center, radius, level, nib)
    t, c = _oracle_resident(path)
    args = (c.links(), c.far(), c.channel_bytes(1, 4), c.channel_bytes(2, 1), c.root(), center, radius, level)
    assert orc.check_sphere_pool(*args, faultless=True) == 0
    # the checker is not vacuous: a planted wrong child mask or node id is caught
    links = args[0].copy()
    leaf = _leaves(args[0], args[1], args[4])[0][0]
    parent_like = next(i for i in range(len(links)) if links[i] & 255)
    links[parent_like] ^= 1 << (int(np.log2((links[parent_like] & 255) & -(links[parent_like] & 255))))
    assert orc.check_sphere_pool(links, *args[1:], faultless=True) != 0
    ids = args[2].copy()
    ids[4 * leaf] ^= 1
    assert orc.check_sphere_pool(args[0], args[1], ids, *args[3:], faultless=True) != 0


def test_sphere_pool_with_faults_small_cache():
    center, radius = (0.5, 0.5, 0.5), 0.3
    _, path = sphere_tree(center, radius, 6, 64)
    t = orc.Tree(path)
    c = orc.Cache(t, max(2, t.blocks // 3))
    c.initial_fill()
    assert c.loaded() == c.max_blocks
    assert orc.check_sphere_pool(c.links(), c.far(), c.channel_bytes(1, 4), c.channel_bytes(2, 1), c.root(),
                                 center, radius, 6, faultless=False) == 0


def _block_parents(t):
    parents = {}
    for b in range(t.blocks):
        hdr, _ = t.load_block(b)
        parents[b] = int(hdr[1])
    return parents


def test_initial_fill_is_bfs_and_lru_invariant():
    _, path = terrain_tree(7, 128, 0, 4)
    t = orc.Tree(path)
    maxb = t.blocks // 2
    c = orc.Cache(t, maxb)
    c.initial_fill()
    # BFS over fault links from the root block assigns slots 0, 1, 2, ...
    order, seen, q = [], set(), [t.root_block]
    nib = t.nodes_in_block
    while q and len(order) < maxb:
        b = q.pop(0)
        if b in seen:
            continue
        seen.add(b)
        order.append(b)
        _, data = t.load_block(b)
        lk = data[:4 * nib].view(np.uint32)
        q.extend(int(l >> 11) for l in lk if (l & 255) and (l >> 8) & 1)
    assert [c.block_slot(b) for b in order] == list(range(len(order)))
    parents = _block_parents(t)
    slot_of = {b: c.block_slot(b) for b in order}
    block_of = {s: b for b, s in slot_of.items()}
    lru = [s for s in c.lru_order() if s >= 0]
    pos = {s: i for i, s in enumerate(lru)}
    for s in lru:
        p = parents[block_of[s]]
        if p in slot_of:
            assert pos[s] < pos[slot_of[p]], "a child must be nearer the LRU end than its parent"


def test_load_unload_restores_parent_links():
    """Loading a block rewrites its parent's pointers; evicting it restores the stored-format words."""
    _, path = terrain_tree(7, 128, 0, 4)
    t = orc.Tree(path)
    c = orc.Cache(t, 3)
    c.initial_fill()  # root + its first two BFS children
    before = c.links().copy()
    c.StartRequestSession()
    # request a child of slot 0 that is not loaded, forcing an eviction
    _, data = t.load_block(t.root_block)
    lk = data[:4 * t.nodes_in_block].view(np.uint32)
    kids = [int(l >> 11) for l in lk if (l & 255) and (l >> 8) & 1]
    missing = [k for k in kids if c.block_slot(k) < 0]
    assert missing
    c.StartRequestTransaction()
    c.RequestBlock(missing[0])
    c.EndRequestTransaction()
    c.EndRequestSession()
    after = c.links()
    evicted = [s for s in range(3) if s != 0]
    # Slot 0 (root block) words outside the pointers of the evicted / loaded children are unchanged.
    nib = t.nodes_in_block
    changed = np.nonzero(before[:nib] != after[:nib])[0]
    assert len(changed) > 0
    assert evicted


# ------------------------------------------------------------- feedback
def _brute_unique(codes):
    u = np.unique(codes)
    return u


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_feedback_lists_against_unique(seed):
    rng = np.random.default_rng(seed)
    n = 20000
    codes = rng.integers(2, 400, n).astype(np.int32)
    codes[rng.random(n) < 0.3] = -rng.integers(1, 3000, int((rng.random(n) < 0.3).sum()) or 1)[0]
    faults = -rng.integers(1, 5000, n)
    mask = rng.random(n) < 0.25
    codes[mask] = faults[mask]
    codes[rng.random(n) < 0.1] = NO_HIT
    r = orc.feedback_extract(codes, expected_faults=64, hits_to_extract=10**6, seed=seed)
    u = _brute_unique(codes)
    exp_faults = [int(-v - 1) for v in u if v < 0]
    assert r["faults"] == exp_faults  # every distinct fault, block index descending
    assert exp_faults == sorted(exp_faults, reverse=True)
    pos = [int(v) for v in u if 0 < v < NO_HIT]
    got = sorted([(h + 1) << 1 for h in r["hits"]] + [((h + 1) << 1) | 1 for h in r["dups"]])
    assert got == pos  # a window larger than the hits covers each distinct hit exactly once
    assert r["unique_count"] == len(u)
    assert r["expected_faults"] == (128 if len(exp_faults) > 64 else 32)


def test_feedback_hit_window_bounded():
    rng = np.random.default_rng(5)
    codes = (((rng.integers(0, 20000, 100000) + 1) << 1)).astype(np.int32)
    r = orc.feedback_extract(codes, expected_faults=1, hits_to_extract=4000, seed=11)
    assert r["faults"] == []
    # one doubling read of `expected_faults` values, then a window of 4000
    assert len(r["hits"]) == 1 + 4000
    assert len(set(r["hits"])) == len(r["hits"])


# ---------------------------------------------------------------- C ABI
def test_c_abi_exports_every_declared_symbol():
    import ctypes
    from rayt_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = _lib.declared_functions()
    assert len(names) > 60
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(_lib.SIGNATURES) <= set(names)


def test_product_tree_matches_file_loader():
    """The in-memory tree and its .tree file give identical blocks through both loaders."""
    tree, path = terrain_tree(6, 512, 2, 3)
    t = orc.Tree(path)
    info = tree.info()
    assert (info.nodes_in_block, info.blocks_count, info.root_block_index) == (t.nodes_in_block, t.blocks, t.root_block)
    ftree = api.StoredOctree.open(path)
    for b in range(t.blocks):
        h1, p1 = tree.load_block(b)
        h2, p2 = ftree.load_block(b)
        oh, od = t.load_block(b)
        assert np.array_equal(h1, h2) and np.array_equal(p1, p2)
        assert np.array_equal(p1, od[:len(p1)])
        assert int(np.frombuffer(h1[:4].tobytes(), np.uint32)[0]) == b == int(oh[0])

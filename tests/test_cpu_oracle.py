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
    least-recently-used end than its parent), and after every transaction of
    loads and evictions the pool words and far pointers implied by the set of
    loaded blocks, derived from the .tree file alone;
  * a closed-form shadow scene: exactly the floor pixels whose ray towards
    light 1 crosses an occluder lose light 1's term.
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


def _voxel_pool(depth, cells):
    """Cache-format pool (one block, near pointers) holding the leaf cells {(i, j, k): (normal, rgb)} of a
    2^depth grid; inner nodes get a zero normal and a white colour. Nodes are laid out breadth first with
    the children of a node contiguous, root first."""
    nodes, level = [(0, (0, 0, 0))], [(0, 0, 0)]  # (depth, cell)
    for d in range(1, depth + 1):
        present = {(i >> (depth - d), j >> (depth - d), k >> (depth - d)) for i, j, k in cells}
        level = [c for i, j, k in level for c in ((2 * i + (ch & 1), 2 * j + ((ch >> 1) & 1),
                                                   2 * k + ((ch >> 2) & 1)) for ch in range(8)) if c in present]
        nodes += [(d, c) for c in level]
    index = {n: x for x, n in enumerate(nodes)}
    links = np.zeros(len(nodes), np.uint32)
    normals = np.zeros(len(nodes), np.uint32)
    colors = np.full(len(nodes), 0xFFFFFFFF, np.uint32)
    for x, (d, (i, j, k)) in enumerate(nodes):
        if d == depth:
            n, rgb = cells[(i, j, k)]
            normals[x] = np.frombuffer(np.array(n + (0,), np.int8).tobytes(), np.uint32)[0]
            colors[x] = np.frombuffer(np.array(rgb + (255,), np.uint8).tobytes(), np.uint32)[0]
            continue
        kids = [(ch, (2 * i + (ch & 1), 2 * j + ((ch >> 1) & 1), 2 * k + ((ch >> 2) & 1))) for ch in range(8)]
        kids = [(ch, c) for ch, c in kids if (d + 1, c) in index]
        first = index[(d + 1, kids[0][1])]
        # children in child-index order occupy consecutive slots (the popcount addressing of raycast.cl:151)
        assert [index[(d + 1, c)] for _, c in kids] == list(range(first, first + len(kids)))
        links[x] = (((1 << 20) + first - x) << 11) | sum(1 << ch for ch, _ in kids)
    return links, np.zeros(8, np.uint32), normals, colors


def _crosses_box(q, s, lo, hi):
    """Whether the ray q + u s (u > 0) passes through the box [lo, hi], in double precision."""
    with np.errstate(divide="ignore", invalid="ignore"):
        t0, t1 = (lo - q) / s, (hi - q) / s
    return np.nanmin(np.maximum(t0, t1)) > max(np.nanmax(np.minimum(t0, t1)), 0.0)


def test_shadow_ray_closed_form():
    """A floor lit by light 1 and a slab hovering above it, seen from straight above. Exactly the floor
    pixels whose ray towards light 1 (-normalize(2,-2,3), shading.cl:8) crosses the slab, found here in
    double precision, lose light 1's term; every other floor pixel keeps the unshadowed colour. The shadow
    ray starts a quarter of the hit voxel (2^-(depth+2)) back along the primary ray, as the extension
    defines it."""
    D, G = 4, 16
    floor_rgb, slab_rgb, up = (200, 150, 100), (90, 90, 90), (0, 127, 0)
    cells = {(i, 0, k): (up, floor_rgb) for i in range(G) for k in range(G)}
    slab = {(i, 3, k): (up, slab_rgb) for i in (4, 5, 6) for k in (3, 4)}
    cells.update(slab)
    links, far, normals, colors = _voxel_pool(D, cells)
    floor_ids = {x for x in range(len(links)) if links[x] == 0 and (colors[x] & 0xFF) == floor_rgb[0]}
    W, H = 128, 128
    vpi = camera_vpi(0.0, -90.0, (0.5, 2.0, 0.5), fov=28.0, aspect=1.0)
    lit = orc.render_plain(links, far, 0, normals, colors, 0, W, H, vpi, flags=0, want_hits=True)
    sh = orc.render_plain(links, far, 0, normals, colors, 0, W, H, vpi, flags=1, want_hits=True)
    np.testing.assert_array_equal(lit["codes"], sh["codes"])  # shadows never change what is hit
    # every hit faces up, so light 1 reaches it unless occluded: one shadow ray per hit
    assert sh["stats"][1] == int((lit["hit_node"] >= 0).sum()) and lit["stats"][1] == 0

    # closed-form ProcessHit colours of an up-facing normal: light 1 adds 1.2/sqrt(17), light 2 1.5/sqrt(22)
    l1, l2 = 0.6 * 2 / math.sqrt(17), 0.5 * 3 / math.sqrt(22)
    full = lambda rgb: np.array([min(255.0, ch * (0.1 + l1 + l2)) for ch in rgb])  # noqa: E731
    dark = lambda rgb: np.array([min(255.0, ch * (0.1 + l2)) for ch in rgb])  # noqa: E731
    s_dir = -np.array([2.0, -2.0, 3.0]) / math.sqrt(17)
    lo, hi = np.array([4, 3, 3]) / G, np.array([7, 4, 5]) / G
    px_lit = lit["image"].view(np.uint8).reshape(H, W, 4)[..., :3].astype(float)
    px_sh = sh["image"].view(np.uint8).reshape(H, W, 4)[..., :3].astype(float)
    shadowed = unshadowed = 0
    for y in range(H):
        for x in range(W):
            node = int(lit["hit_node"][y, x])
            if node < 0:  # background
                np.testing.assert_array_equal(px_sh[y, x], px_lit[y, x])
                continue
            if node not in floor_ids:  # the slab: lit, or shadowed by its own voxels on its far sides
                assert np.all(np.abs(px_lit[y, x] - full(slab_rgb)) <= 1.0)
                assert np.array_equal(px_sh[y, x], px_lit[y, x]) or np.all(np.abs(px_sh[y, x] - dark(slab_rgb)) <= 1)
                continue
            assert np.all(np.abs(px_lit[y, x] - full(floor_rgb)) <= 1.0), (x, y)
            o, d = _ray(vpi, x, y, W, H)
            t = (1.0 / G - o[1]) / d[1]  # the floor's top face
            q = o + d * (t - 2.0 ** -(D + 2))
            inside = [_crosses_box(q, s_dir, lo + e, hi - e) for e in (1e-4, -1e-4)]
            if inside[0] != inside[1]:
                continue  # grazes the slab's boundary: not decided at this precision
            if inside[0]:
                shadowed += 1
                assert np.all(np.abs(px_sh[y, x] - dark(floor_rgb)) <= 1.0), (x, y)
            else:
                unshadowed += 1
                np.testing.assert_array_equal(px_sh[y, x], px_lit[y, x], err_msg=str((x, y)))
    assert shadowed > 500 and unshadowed > 10000


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


ROOT_BLOCK_PARENT = (1 << 24) - 1


class _Blocks:
    """Headers and stored links of every block of a .tree file, read once."""

    def __init__(self, t):
        self.t, self.nib, self._d = t, t.nodes_in_block, {}

    def __getitem__(self, b):
        if b not in self._d:
            hdr, data = self.t.load_block(b)
            self._d[b] = (hdr, data[:4 * self.nib].view(np.uint32).copy())
        return self._d[b]


def _pack_cache(mask, fault, dup, far, ptr):
    """PackCacheNodeLink (octree-constants.h:41-44)."""
    return (int(ptr) << 11) | (int(far) << 10) | (int(dup) << 9) | (int(fault) << 8) | int(mask)


def _expected_pool(blocks, c):
    """The pool words and far pointers implied by the set of loaded blocks and their slots, derived from the
    .tree file alone: a block's own non-fault links become near pointers relative to the node holding them,
    its fault links stay as stored, and each root of a loaded child block turns its parent's pointer into a
    near or far pointer to the root's slot, flagged duplicate when the stored parent pointer was not a fault
    (gpu-octree-cache.cpp:119-141, 169-208). Unloading must restore exactly this state for the parent."""
    nib = blocks.nib
    slot_of = {b: s for b, s in ((b, c.block_slot(b)) for b in range(blocks.t.blocks)) if s >= 0}
    links = {}
    for b, s in slot_of.items():
        L = blocks[b][1].astype(np.int64)
        mask, fault, dup, ptr = L & 255, (L >> 8) & 1, (L >> 9) & 1, L >> 11
        n = np.arange(nib)
        near = (((1 << 20) + ptr - n) << 11) | (dup << 9) | mask
        links[s] = np.where((mask != 0) & (fault == 0), near, L).astype(np.uint32)
    far = {}
    for b, s in slot_of.items():
        hdr = blocks[b][0]
        if int(hdr[1]) == ROOT_BLOCK_PARENT:
            continue
        sp = slot_of[int(hdr[1])]  # a loaded block's parent is loaded
        for i in range(int(hdr[2])):
            pidx, cidx, pv = int(hdr[3 + 3 * i]), int(hdr[4 + 3 * i]), int(hdr[5 + 3 * i])
            target, at = s * nib + cidx, sp * nib + pidx
            dup = 1 - ((pv >> 8) & 1)
            if abs(at - target) >= 1 << 20:
                far[s * 8 + i] = target
                links[sp][pidx] = _pack_cache(pv & 255, 0, dup, 1, s * 8 + i)
            else:
                links[sp][pidx] = _pack_cache(pv & 255, 0, dup, 0, (1 << 20) + target - at)
    return slot_of, links, far


def _assert_pool_state(blocks, c):
    slot_of, exp_links, exp_far = _expected_pool(blocks, c)
    nib = blocks.nib
    links, far = c.links(), c.far()
    for s, words in exp_links.items():
        got = links[s * nib:(s + 1) * nib]
        bad = np.nonzero(got != words)[0]
        assert len(bad) == 0, f"slot {s}: {len(bad)} words differ, first at node {bad[0]}: " \
                              f"{int(got[bad[0]]):#x} != {int(words[bad[0]]):#x}"
    for i, v in exp_far.items():
        assert int(far[i]) == v, f"far pointer {i}"
    return slot_of, exp_far


def _requestable(blocks, loaded, parents=None):
    """Blocks not loaded that a fault link of a loaded block (of `parents`, when given) points to."""
    out = set()
    for b in (loaded if parents is None else parents):
        L = blocks[b][1]
        out.update(int(v >> 11) for v in L[((L & 255) != 0) & (((L >> 8) & 1) == 1)])
    return sorted(out - set(loaded))


def _transaction(c, blocks):
    c.StartRequestSession()
    c.StartRequestTransaction()
    for b in blocks:
        c.RequestBlock(b)
    c.EndRequestTransaction()
    c.EndRequestSession()


def test_load_unload_restores_parent_links():
    """Loading a block rewrites its parent's pointers to it; evicting it restores the parent's words to
    their values before the load, bit for bit (gpu-octree-cache.cpp:243-293), and the whole pool stays the
    state the loaded set implies."""
    _, path = terrain_tree(7, 128, 0, 4)
    t = orc.Tree(path)
    nib = t.nodes_in_block
    blocks = _Blocks(t)
    c = orc.Cache(t, 3)
    c.initial_fill()  # root + its first two BFS children
    loaded0, _ = _assert_pool_state(blocks, c)
    root_slot = loaded0[t.root_block]
    before = c.links().copy()
    target = _requestable(blocks, loaded0, [t.root_block])[0]
    hdr = blocks[target][0]
    assert int(hdr[1]) == t.root_block
    ptr_words = [root_slot * nib + int(hdr[3 + 3 * i]) for i in range(int(hdr[2]))]
    _transaction(c, [target])
    loaded, _ = _assert_pool_state(blocks, c)
    assert target in loaded and len(set(loaded0) - set(loaded)) == 1  # one block was evicted for it
    mid = c.links()
    assert all(mid[w] != before[w] for w in ptr_words)  # the parent now points at the loaded block
    # load the root's other children one at a time until the target is evicted in turn
    for b in _requestable(blocks, loaded0, [t.root_block]):
        if target not in loaded:
            break
        if b not in loaded:
            _transaction(c, [b])
            loaded, _ = _assert_pool_state(blocks, c)
    assert target not in loaded
    after = c.links()
    for w in ptr_words:  # the target's parent pointers are back to their words before the load
        assert after[w] == before[w], f"word {w}: {int(after[w]):#x} != {int(before[w]):#x}"
    assert c.counters()[1] >= 2  # evictions


def test_pool_state_follows_loads_and_evictions_with_far_pointers():
    """Random request transactions on a pool whose slots lie more than 2^20 nodes apart (far pointers):
    after every transaction each word of every loaded slot and every live far pointer is the one the set of
    loaded blocks implies."""
    _, path = terrain_tree(10, 8192, 0, 4)
    t = orc.Tree(path)
    blocks = _Blocks(t)
    c = orc.Cache(t, 300)
    c.initial_fill()
    loaded, far = _assert_pool_state(blocks, c)
    rng = np.random.default_rng(3)
    saw_far = len(far)
    for _ in range(12):
        cand = _requestable(blocks, loaded)
        pick = [int(b) for b in rng.choice(cand, size=min(len(cand), int(rng.integers(1, 40))), replace=False)]
        _transaction(c, pick)
        loaded, far = _assert_pool_state(blocks, c)
        saw_far = max(saw_far, len(far))
    uploaded, evicted = c.counters()[:2]
    assert evicted > 50 and saw_far > 0


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

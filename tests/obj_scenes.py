"""Seeded Wavefront OBJ scenes (with MTL libraries and PNM textures) for the
importer tests, and a walker that reads a built tree back as BFS node lists."""
from __future__ import annotations

import os

import numpy as np

import oracle as orc
from rayt_amd import api


def write_ppm(path, rgb, maxval=255, ascii=False):
    """rgb: uint8/uint16 array (h, w, 3)."""
    h, w, _ = rgb.shape
    if ascii:
        body = "\n".join(" ".join(str(int(v)) for v in row.reshape(-1)) for row in rgb)
        open(path, "w").write(f"P3\n# made by the tests\n{w} {h}\n{maxval}\n{body}\n")
        return
    data = rgb.astype(">u2").tobytes() if maxval > 255 else rgb.astype(np.uint8).tobytes()
    open(path, "wb").write(f"P6\n{w} {h}\n{maxval}\n".encode() + data)


def write_pgm(path, grey, maxval=255):
    h, w = grey.shape
    data = grey.astype(">u2").tobytes() if maxval > 255 else grey.astype(np.uint8).tobytes()
    open(path, "wb").write(f"P5\n# bump\n{w} {h}\n{maxval}\n".encode() + data)


def write_text(path, lines):
    open(path, "w").write("\n".join(lines) + "\n")


def random_textured_scene(d, seed=0, n_tris=40, quads=4):
    """A triangle soup plus a few quads over three materials: a flat colour, a
    colour texture (map_Kd, with Ka), and a bump-mapped one; vt on most faces,
    vn on some, fan-split polygons, unsupported lines mixed in. Returns the
    OBJ path."""
    rng = np.random.default_rng(seed)
    os.makedirs(d, exist_ok=True)
    tex = rng.integers(0, 256, size=(7, 9, 3))
    write_ppm(os.path.join(d, "albedo.ppm"), tex)
    bump = rng.integers(0, 256, size=(6, 5))
    write_pgm(os.path.join(d, "bump.pgm"), bump)
    write_text(os.path.join(d, "scene.mtl"), [
        "# materials", "newmtl flat", "Kd 0.8 0.3 0.2", "Ka 0.02 0.02 0.05", "Ks 1 1 1", "Ns 20", "illum 2",
        "newmtl textured", "Kd 1 0.9 0.8", "map_Kd -s 1 1 1 albedo.ppm", "d 1",
        "newmtl bumpy", "Kd 0.3 0.6 0.9", "bump bump.pgm -bm 0.7", "Ni 1.5",
    ])
    lines = ["# random scene", "mtllib scene.mtl", "o soup", "g group1"]
    # vertices: triangles of random size scattered in a unit-ish cube
    verts = []
    for _ in range(n_tris):
        c = rng.uniform(0.0, 1.0, 3)
        for _ in range(3):
            verts.append(c + rng.normal(0, 0.15, 3))
    for _ in range(quads):  # planar quads: corner + two edges
        o, a, b = rng.uniform(0, 1, 3), rng.normal(0, 0.3, 3), rng.normal(0, 0.3, 3)
        verts += [o, o + a, o + a + b, o + b]
    for v in verts:
        lines.append("v %.9g %.9g %.9g" % tuple(v))
    uv = rng.uniform(-1.5, 2.5, size=(len(verts), 2))
    for t in uv:
        lines.append("vt %.7g %.7g" % tuple(t))
    nrm = rng.normal(size=(len(verts), 3))
    for n in nrm:
        lines.append("vn %.7g %.7g %.7g" % tuple(n))
    mats = ["flat", "textured", "bumpy"]
    cur = None
    for t in range(n_tris):
        m = mats[t % 3]
        if m != cur:
            lines.append(f"usemtl {m}")
            cur = m
        i = 3 * t + 1
        kind = t % 4
        if kind == 0:
            lines.append(f"f {i} {i + 1} {i + 2}")
        elif kind == 1:
            lines.append(f"f {i}/{i} {i + 1}/{i + 1} {i + 2}/{i + 2}")
        elif kind == 2:
            lines.append(f"f {i}/{i}/{i} {i + 1}/{i + 1}/{i + 1} {i + 2}/{i + 2}/{i + 2}")
        else:
            lines.append(f"f {i}//{i}  {i + 1}//{i + 1} {i + 2}//{i + 2}")
        if t % 7 == 3:
            lines.append("s 1")
    base = 3 * n_tris + 1
    for q in range(quads):
        lines.append(f"usemtl {mats[q % 3]}")
        i = base + 4 * q
        lines.append(f"f {i}/{i} {i + 1}/{i + 1} {i + 2}/{i + 2} {i + 3}/{i + 3}")
    path = os.path.join(d, "scene.obj")
    write_text(path, lines)
    return path


def product_nodes(obj_path, level, nodes_in_block=64, threads=0, tree_path=None):
    """The product's tree of an OBJ, read back from its .tree file through the
    oracle's loader: (masks, data[n, 8]) in BFS order, children in bit order."""
    t = api.StoredOctree.obj(obj_path, level, nodes_in_block, threads)
    tree_path = tree_path or obj_path + f".{level}.{nodes_in_block}.tree"
    t.write(tree_path)
    t.close()
    return walk_tree(tree_path)


def walk_tree(tree_path):
    t = orc.Tree(tree_path)
    c = orc.Cache(t, t.blocks)
    c.initial_fill()
    links, far = c.links(), c.far()
    nrm, col = c.channel_bytes(1, 4).reshape(-1, 4), c.channel_bytes(2, 4).reshape(-1, 4)
    queue, masks, data = [c.root()], [], []
    i = 0
    while i < len(queue):
        n = queue[i]
        i += 1
        link = int(links[n])
        mask = link & 255
        masks.append(mask)
        data.append(np.concatenate([nrm[n], col[n]]))
        if mask:
            assert not (link >> 8) & 1, "fault link in a resident pool"
            ptr = link >> 11
            first = int(far[ptr]) if (link >> 10) & 1 else n + ptr - (1 << 20)
            queue.extend(first + k for k in range(bin(mask).count("1")))
    return np.array(masks, np.uint8), np.array(data, np.uint8).reshape(-1, 8)


def bfs_cells(masks):
    """(depth, integer cell coordinates ix, iy, iz) of every node of a BFS node list."""
    out = [(0, 0, 0, 0)]
    cells = []
    for i, m in enumerate(masks):
        d, x, y, z = out[i]
        cells.append(out[i])
        for ch in range(8):
            if m & (1 << ch):
                out.append((d + 1, 2 * x + (ch & 1), 2 * y + ((ch >> 1) & 1), 2 * z + ((ch >> 2) & 1)))
    assert len(out) == len(masks)
    return cells

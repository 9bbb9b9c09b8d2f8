"""Writes the golden vectors of tests/golden/: sphere5_n16.npz and terrain5_n16.npz.

The reference ships no golden vectors and cannot run here, so these vectors
come from the oracle (itself pinned by brute force and closed forms in
tests/test_cpu_oracle*.py). Committed, they freeze the builder's block
layout, the oracle and the GPU path against each other: a change to any of
them shows up as a fixture mismatch.

Scenes (16 nodes per block, so the trees span many blocks):
  * sphere5_n16: the sphere tree of scene sources/stored-octree-test.cpp:317-323
    (centre (0.3, 0.4, 0.65), radius 0.23, level 5; 60 blocks). Its channels
    are node indices, not normals / colours, so it is rendered without them
    (N.L shading of a zero normal, white, no shadow rays).
  * terrain5_n16: a depth-5 seeded terrain with 2 spheres and 2 boxes, with
    normals and colours, rendered with shadow rays.
Per scene: the SHA-256 of the .tree bytes; three cameras (view_proj_inv as
uint32 bits); per camera at 64x48: hit node, hit t (uint32 bits), RGBA8 pixel
and block feedback code of the fused render; and, for a pool of half the
blocks, the per-pass fault lists and the image of a full frame.

Run from the repository root: python tests/golden/make_golden.py
"""
import hashlib
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle as orc  # noqa: E402
from rayt_amd import api  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
NIB = 16
W, H = 64, 48
CAMERAS = [(30.0, -20.0, (0.1, 0.9, 1.3)), (-135.0, -10.0, (0.8, 0.7, 0.8)), (0.0, -89.0, (0.3, 1.2, 0.65))]
SCENES = {
    # name: (builder, uses normals/colours and shadow rays)
    "sphere5_n16": (lambda: api.StoredOctree.sphere((0.3, 0.4, 0.65), 0.23, 5, NIB), False),
    "terrain5_n16": (lambda: api.StoredOctree.terrain(api.TerrainSpec(depth=5, seed=3, spheres=2, boxes=2), NIB),
                     True),
}


def camera(yaw, pitch, pos):
    return api.Camera(yaw=yaw, pitch=pitch, position=pos, fov=60.0, aspect=W / H).view_proj_inv()


def tree_file(name):
    tree = SCENES[name][0]()
    path = os.path.join(tempfile.gettempdir(), f"rayt_golden_{name}.tree")
    tree.write(path)
    return tree, path


def channels(c, shaded):
    if not shaded:
        return None, None
    return c.channel_bytes(1, 4).view(np.uint32), c.channel_bytes(2, 4).view(np.uint32)


def make(name):
    tree, path = tree_file(name)
    shaded = SCENES[name][1]
    flags = 1 if shaded else 0
    digest = hashlib.sha256(open(path, "rb").read()).hexdigest()
    t = orc.Tree(path)
    c = orc.Cache(t, t.blocks)
    c.initial_fill()
    normals, colors = channels(c, shaded)
    vpis, nodes, ts, imgs, codes = [], [], [], [], []
    for yaw, pitch, pos in CAMERAS:
        vpi = np.asarray(camera(yaw, pitch, pos), np.float32).reshape(16)
        r = orc.render_plain(c.links(), c.far(), c.root(), normals, colors, NIB, W, H, vpi, flags=flags,
                             want_hits=True)
        vpis.append(vpi.view(np.uint32))
        nodes.append(r["hit_node"].astype(np.int32))
        ts.append(r["hit_t"].view(np.uint32))
        imgs.append(r["image"].astype(np.uint32))
        codes.append(r["codes"].astype(np.int32))
    small = orc.Cache(t, t.blocks // 2)
    small.initial_fill()
    fr = small.render_frame(W, H, vpis[0].view(np.float32), flags=flags, seed=7)
    counts = np.array([len(p) for p in fr["pass_faults"]], np.int32)
    faults = np.array([x for p in fr["pass_faults"] for x in p], np.int32)
    out = os.path.join(HERE, f"{name}.npz")
    np.savez_compressed(out, tree_sha256=np.frombuffer(bytes.fromhex(digest), np.uint8), root=np.int32(c.root()),
                        blocks=np.int32(t.blocks), flags=np.int32(flags), view_proj_inv_bits=np.stack(vpis),
                        hit_node=np.stack(nodes), hit_t_bits=np.stack(ts), image=np.stack(imgs),
                        codes=np.stack(codes), frame_pool=np.int32(t.blocks // 2), frame_pass_counts=counts,
                        frame_faults=faults, frame_image=fr["image"].astype(np.uint32),
                        frame_passes=np.int32(fr["passes"]))
    os.remove(path)
    print(f"wrote {out}: {t.blocks} blocks, tree {digest[:16]}..., frame faults per pass {counts.tolist()}")


if __name__ == "__main__":
    for n in SCENES:
        make(n)

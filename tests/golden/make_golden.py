"""Writes tests/golden/sphere5_n16.npz: golden vectors on a tiny tree.

The reference ships no golden vectors and cannot run here, so these vectors
come from the oracle (itself pinned by brute force and closed forms in
tests/test_cpu_oracle*.py). Committed, they freeze the builder's block
layout, the oracle and the GPU path against each other: a change to any of
them shows up as a fixture mismatch.

Contents: the sphere tree of scene sources/stored-octree-test.cpp:317-323
(centre (0.3, 0.4, 0.65), radius 0.23, level 5, 16 nodes per block: 60
blocks) as the SHA-256 of its .tree bytes; three cameras (view_proj_inv as
uint32 bits); per camera at 64x48: hit node, hit t (uint32 bits), RGBA8 pixel
and block feedback code of the fused render with shadow rays; and, for a pool
of 30 blocks, the per-pass fault lists and the image of a full frame (three
passes).

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

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sphere5_n16.npz")
CENTER, RADIUS, LEVEL, NIB = (0.3, 0.4, 0.65), 0.23, 5, 16
W, H = 64, 48
CAMERAS = [(30.0, -20.0, (0.1, 0.9, 1.3)), (-135.0, -10.0, (0.8, 0.7, 0.8)), (0.0, -89.0, (0.3, 1.2, 0.65))]
SMALL_POOL = 30


def camera(yaw, pitch, pos):
    return api.Camera(yaw=yaw, pitch=pitch, position=pos, fov=60.0, aspect=W / H).view_proj_inv()


def tree_file():
    tree = api.StoredOctree.sphere(CENTER, RADIUS, LEVEL, NIB)
    path = os.path.join(tempfile.gettempdir(), "rayt_golden_sphere5_n16.tree")
    tree.write(path)
    return tree, path


def main():
    tree, path = tree_file()
    digest = hashlib.sha256(open(path, "rb").read()).hexdigest()
    t = orc.Tree(path)
    c = orc.Cache(t, t.blocks)
    c.initial_fill()
    links, far = c.links(), c.far()
    normals = c.channel_bytes(1, 4).view(np.uint32)
    colors = c.channel_bytes(2, 4).view(np.uint32)
    vpis, nodes, ts, imgs, codes = [], [], [], [], []
    for yaw, pitch, pos in CAMERAS:
        vpi = np.asarray(camera(yaw, pitch, pos), np.float32).reshape(16)
        r = orc.render_plain(links, far, c.root(), normals, colors, NIB, W, H, vpi, flags=1, want_hits=True)
        vpis.append(vpi.view(np.uint32))
        nodes.append(r["hit_node"].astype(np.int32))
        ts.append(r["hit_t"].view(np.uint32))
        imgs.append(r["image"].astype(np.uint32))
        codes.append(r["codes"].astype(np.int32))
    small = orc.Cache(t, SMALL_POOL)
    small.initial_fill()
    fr = small.render_frame(W, H, vpis[0].view(np.float32), flags=1, seed=7)
    counts = np.array([len(p) for p in fr["pass_faults"]], np.int32)
    faults = np.array([x for p in fr["pass_faults"] for x in p], np.int32)
    np.savez_compressed(OUT, tree_sha256=np.frombuffer(bytes.fromhex(digest), np.uint8), root=np.int32(c.root()),
                        view_proj_inv_bits=np.stack(vpis), hit_node=np.stack(nodes), hit_t_bits=np.stack(ts),
                        image=np.stack(imgs), codes=np.stack(codes), frame_pass_counts=counts,
                        frame_faults=faults, frame_image=fr["image"].astype(np.uint32),
                        frame_passes=np.int32(fr["passes"]))
    os.remove(path)
    print(f"wrote {OUT}: tree {digest[:16]}..., {len(counts)} passes, faults per pass {counts.tolist()}")


if __name__ == "__main__":
    main()

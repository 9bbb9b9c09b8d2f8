"""Seeded synthetic workloads shared by the benchmark and the tests.

- the depth-11 scene of the single-view measurement: fBm terrain (4 octaves,
  heights in [0, 0.6]) plus 64 boxes with edges in [0.01, 0.1], seed 1, and
  its fixed 1080p camera;
- the multi-view mix: resolutions drawn from {256x256, 640x480, 1280x720,
  1920x1080, 3840x2160} with weights {0.3, 0.3, 0.2, 0.15, 0.05}, positions
  uniform in [0,1]^3, directions uniform on the sphere;
- the fly-through: a closed Catmull-Rom spline through 8 seeded control points
  above the terrain.

Cameras follow scene/camera.cpp (forward = (sin yaw cos pitch, sin pitch,
-cos yaw cos pitch)); all views use a 60 degree horizontal field of view.
"""
from __future__ import annotations

import math

import numpy as np

from rayt_amd import api

MIX_SIZES = [(256, 256), (640, 480), (1280, 720), (1920, 1080), (3840, 2160)]
MIX_WEIGHTS = [0.3, 0.3, 0.2, 0.15, 0.05]

# Pose (yaw, pitch, position) of the single-view 1080p measurement: from above one corner, looking across the
# terrain.
PLAIN_POSE = (35.0, -25.0, (0.05, 0.75, 1.25))


def plain_scene_spec(depth: int = 11) -> api.TerrainSpec:
    """fBm terrain + 64 boxes (edges 0.01-0.1), seed 1: the single-view scene at `depth`."""
    return api.TerrainSpec(depth=depth, seed=1, octaves=4, boxes=64, box_emin=0.01, box_emax=0.1)


def camera(yaw: float, pitch: float, pos, width: int, height: int) -> np.ndarray:
    """view_proj_inv of a 60-degree camera with the view's aspect ratio."""
    return api.Camera(yaw=yaw, pitch=pitch, position=pos, fov=60.0, aspect=width / height).view_proj_inv()


def view_mix(n: int, seed: int):
    """n views (vpi, width, height): seeded resolutions from the mix, positions uniform in [0,1]^3, directions
    uniform on the sphere."""
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(n):
        w, h = MIX_SIZES[rng.choice(len(MIX_SIZES), p=MIX_WEIGHTS)]
        pos = tuple(float(x) for x in rng.uniform(0.0, 1.0, 3))
        yaw = float(rng.uniform(-180.0, 180.0))
        pitch = math.degrees(math.asin(float(rng.uniform(-1.0, 1.0))))
        specs.append((camera(yaw, pitch, pos, w, h), w, h))
    return specs


def spline_poses(n: int, seed: int):
    """n camera poses (yaw, pitch, position) along a closed Catmull-Rom spline through 8 seeded control points
    above the terrain, each looking at a point of the terrain near the centre that moves with the path."""
    rng = np.random.default_rng(seed)
    ctrl = np.column_stack([rng.uniform(0.1, 0.9, 8), rng.uniform(0.55, 0.85, 8), rng.uniform(0.1, 0.9, 8)])

    def point(s):
        k = int(math.floor(s)) % 8
        u = s - math.floor(s)
        p0, p1, p2, p3 = (ctrl[(k + j - 1) % 8] for j in range(4))
        return 0.5 * ((2 * p1) + (-p0 + p2) * u + (2 * p0 - 5 * p1 + 4 * p2 - p3) * u * u
                      + (-p0 + 3 * p1 - 3 * p2 + p3) * u ** 3)

    poses = []
    for i in range(n):
        s = 8.0 * i / n
        p, q = point(s), point(s + 0.02)
        target = np.array([0.5, 0.25, 0.5]) + 2.0 * (q - p)
        v = target - p
        yaw = math.degrees(math.atan2(v[0], -v[2]))
        pitch = math.degrees(math.atan2(v[1], math.hypot(v[0], v[2])))
        poses.append((yaw, pitch, tuple(float(x) for x in p)))
    return poses

"""Benchmark: Mrays/s of the plain single-view render on the depth-11 SVO at 1080p.

Workload (one step = one GPURayTracer::RenderFrame): seeded depth-11 (2048^3)
fBm terrain (4 octaves, heights in [0, 0.6]) plus 64 random boxes with edges
in [0.01, 0.1], blocked into 8192-node blocks and fully resident in HBM; one
1920x1080 view, 60 degree horizontal FOV, no LOD, primary rays plus one shadow
ray per lit hit. Each frame runs the tracing pass, the feedback extraction and
the cache transaction that marks the hit blocks, as the reference does.

Prints one JSON line (see the field list at the bottom). The CPU baseline is
the oracle (a strict-fp32 C++ restatement of the same traversal) on the same
frame, on up to 16 host threads.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg: str):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--depth", type=int, default=11)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from rayt_amd import api

    W, H = args.width, args.height
    t = time.time()
    spec = api.TerrainSpec(depth=args.depth, seed=1, octaves=4, boxes=64, box_emin=0.01, box_emax=0.1)
    tree = api.StoredOctree.terrain(spec, 8192)
    info = tree.info()
    log(f"scene depth {args.depth}: {info.source_nodes} nodes, {info.blocks_count} blocks, built in {time.time() - t:.1f}s")

    ctx = api.Context(0)
    t = time.time()
    cache = api.GPUOctreeCacheManager(info.blocks_count, tree, ctx)
    cache.InitialFillCache()
    log(f"pool filled ({info.blocks_count} blocks) in {time.time() - t:.2f}s")

    cam = api.Camera(yaw=35.0, pitch=-25.0, position=(0.05, 0.75, 1.25), fov=60.0, aspect=W / H)
    vpi = cam.view_proj_inv()
    flags = api.L.RT_FLAG_SHADOWS
    lod = 0.0

    # Ray and fetch counts of this frame (counting kernel, outside the timed region).
    _, _, st = api.render_plain(ctx, cache, W, H, vpi, lod=lod, flags=flags, count=True)
    primary, shadow, fetches, hits = st
    rays = primary + shadow
    bytes_per_launch = 4 * fetches + 8 * hits + 8 * primary  # links+far pointers, normal+colour, pixel+code
    log(f"frame: {primary} primary + {shadow} shadow rays, {fetches} node fetches ({fetches / rays:.1f}/ray), {hits} hits")

    tr = api.GPURayTracer(cache, W, H)
    tr.set_frame_time_limit(1e9)
    tr.set_lod_voxel_size(lod)
    tr.set_flags(flags)
    tr.set_seed(1)
    for _ in range(args.warmup):
        tr.RenderFrame(vpi)
    ctx.WaitForAll()
    kernel_ms = 0.0
    t0 = time.perf_counter()
    for i in range(args.steps):
        tr.RenderFrame(vpi)
        kernel_ms += tr.timings()[0]
    ctx.WaitForAll()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    ms_per_step = elapsed / args.steps * 1e3
    value = rays * args.steps / elapsed / 1e6
    avg_kernel_ms = kernel_ms / args.steps
    achieved_gbs = bytes_per_launch / (avg_kernel_ms * 1e-3) / 1e9
    log(f"{args.steps} frames in {elapsed * 1e3:.1f} ms: {value:.0f} Mrays/s; trace kernel {avg_kernel_ms:.3f} ms/frame "
        f"({rays / avg_kernel_ms / 1e3:.0f} Mrays/s), {achieved_gbs:.0f} GB/s algorithmic")

    img = tr.out_image()
    import oracle as orc  # CPU baseline + one-frame check of the measured output

    cpu = None
    if not args.no_cpu_baseline:
        links = cache.channel(0).read_all(np.uint32)
        far = cache.far_pointers().read_all(np.uint32)
        normals = cache.channel_by_name("normals").read_all(np.uint32)
        colors = cache.channel_by_name("colors").read_all(np.uint32)
        threads = max(1, min(16, os.cpu_count() or 1))
        runs, tc = 0, 0.0
        r = None
        while tc < args.cpu_seconds or runs == 0:
            a = time.perf_counter()
            r = orc.render_plain(links, far, cache.root_node_index(), normals, colors, cache.nodes_in_block(), W, H,
                                 vpi, lod=lod, flags=flags, threads=threads)
            tc += time.perf_counter() - a
            runs += 1
        cpu_rays = r["stats"][0] + r["stats"][1]
        cpu = {"value": round(cpu_rays * runs / tc / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
               "sample": f"{runs} full {W}x{H} frame(s) of the same workload ({tc:.1f}s)"}
        match = bool(np.array_equal(img, r["image"]))
        log(f"CPU oracle: {cpu['value']} Mrays/s on {threads} threads; GPU frame identical to oracle: {match}")
        if not match:
            raise SystemExit("GPU image differs from the oracle")

    out = {
        "metric": "Mrays/s (1080p, depth-11 SVO) plain vs batched; p50/p99 request latency",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural octree)",
        "config": {"workload": "plain: depth-11 fBm terrain + 64 boxes, resident, 1920x1080, primary + shadow rays",
                   "width": W, "height": H, "depth": args.depth, "nodes_in_block": 8192, "lod_voxel_size": lod,
                   "blocks": info.blocks_count, "nodes": info.source_nodes, "rays_per_frame": rays,
                   "primary_rays": primary, "shadow_rays": shadow},
        "roofline": {"bound": "hbm", "achieved": round(achieved_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved_gbs / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "RenderPlainKernel", "kernel_ms": round(avg_kernel_ms, 4),
                     "algorithmic_bytes_per_launch": bytes_per_launch},
        "cpu_baseline": cpu,
        "kernel_mrays_s": round(rays / avg_kernel_ms / 1e3, 1),
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

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


MIX_SIZES = [(256, 256), (640, 480), (1280, 720), (1920, 1080), (3840, 2160)]
MIX_WEIGHTS = [0.3, 0.3, 0.2, 0.15, 0.05]


def view_mix(n: int, seed: int):
    """n views: seeded resolutions from the mix, positions uniform in [0,1]^3, directions uniform on the sphere."""
    import math
    import numpy as np
    from rayt_amd import api
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(n):
        w, h = MIX_SIZES[rng.choice(len(MIX_SIZES), p=MIX_WEIGHTS)]
        pos = tuple(float(x) for x in rng.uniform(0.0, 1.0, 3))
        yaw = float(rng.uniform(-180.0, 180.0))
        pitch = math.degrees(math.asin(float(rng.uniform(-1.0, 1.0))))
        vpi = api.Camera(yaw=yaw, pitch=pitch, position=pos, fov=60.0, aspect=w / h).view_proj_inv()
        specs.append((vpi, w, h))
    return specs


def measure_batched(ctx, cache, depth: int, steps: int, nviews: int = 64, seed: int = 7):
    """64-view mix in one persistent launch vs the plain kernel launched once per view."""
    import numpy as np
    from rayt_amd import api
    flags = api.L.RT_FLAG_SHADOWS
    specs = view_mix(nviews, seed)
    b = api.BatchRenderer(cache, nviews)
    b.config(tile_shape=0, lds_levels=12, wgs_per_cu=16, tree_depth=depth)
    views, total, offsets = b.make_views(specs, flags=flags)
    out = api.Buffer(ctx, 4 * total)
    codes = api.Buffer(ctx, 4 * total)
    st = b.render(views, nviews, out, codes, count=True)
    rays = st[0] + st[1]
    t = time.time()
    best, best_ms = b.autotune(views, nviews, out, codes, repeats=2)
    log(f"batched: {nviews} views, {total} pixels, {rays} rays; autotune {time.time() - t:.1f}s -> "
        f"tile {api.BatchRenderer.TILE_SHAPES[best[0]]}, {best[1]} LDS levels, {best[2]} waves/CU ({best_ms:.2f} ms)")
    ms = []
    for _ in range(steps):
        b.render(views, nviews, out, codes)
        ms.append(b.last_ms())
    batched_ms = float(np.median(ms))
    img = out.read_all(np.uint32)
    # the plain kernel, one launch per view, same views
    pout = api.Buffer(ctx, 4 * 3840 * 2160)
    pcodes = api.Buffer(ctx, 4 * 3840 * 2160)
    ctx.WaitForAll()
    reps = max(1, steps // 3)
    t0 = time.perf_counter()
    for _ in range(reps):
        for vpi, w, h in specs:
            api.render_plain(ctx, cache, w, h, vpi, lod=0.0, flags=flags, out=pout, codes=pcodes)
    ctx.WaitForAll()
    plain_ms = (time.perf_counter() - t0) / reps * 1e3
    mismatches = 0
    checked = list(range(0, nviews, max(1, nviews // 8)))
    for k in checked:
        vpi, w, h = specs[k]
        api.render_plain(ctx, cache, w, h, vpi, lod=0.0, flags=flags, out=pout, codes=pcodes)
        if not np.array_equal(pout.Read(0, 4 * w * h, np.uint32), img[offsets[k]:offsets[k] + w * h]):
            mismatches += 1
    res = {"views": nviews, "pixels": total, "rays": rays, "ms": round(batched_ms, 3),
           "mrays_s": round(rays / batched_ms / 1e3, 1), "plain_per_view_ms": round(plain_ms, 3),
           "plain_per_view_mrays_s": round(rays / plain_ms / 1e3, 1),
           "speedup_vs_plain": round(plain_ms / batched_ms, 3),
           "config": {"tile": "x".join(map(str, api.BatchRenderer.TILE_SHAPES[best[0]])), "lds_levels": best[1],
                      "wgs_per_cu": best[2]},
           "views_checked_bit_exact": len(checked) - mismatches}
    log(f"batched {res['mrays_s']} Mrays/s ({batched_ms:.2f} ms) vs plain per-view launches {res['plain_per_view_mrays_s']} "
        f"Mrays/s ({plain_ms:.2f} ms): {res['speedup_vs_plain']}x; view mismatches {mismatches}")
    if mismatches:
        raise SystemExit("batched output differs from the plain kernel")
    b.close()
    return res


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
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MB",
                    help="also run a device-to-device copy of MB MiB (3 times) as the byte-counter reference")
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
    if args.calibrate_copy:
        n = args.calibrate_copy << 20
        src, dst = api.Buffer(ctx, n), api.Buffer(ctx, n)
        for _ in range(3):
            dst.CopyFrom(src)
        ctx.WaitForAll()
        log(f"calibration copy: 3 x {n} bytes")
        del src, dst
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

    # Launch shape of the plain kernel: the fastest of a small sweep (outside the timed region).
    pout, pcodes = api.Buffer(ctx, 4 * W * H), api.Buffer(ctx, 4 * W * H)
    shapes = []
    for waves in (1, 4):
        for levels in (10, 12, 16):
            api.plain_config(ctx, levels, waves, args.depth)
            api.render_plain(ctx, cache, W, H, vpi, lod=lod, flags=flags, out=pout, codes=pcodes)
            tm = api.EventTimer(ctx)
            tm.start()
            for _ in range(5):
                api.render_plain(ctx, cache, W, H, vpi, lod=lod, flags=flags, out=pout, codes=pcodes)
            shapes.append((tm.stop() / 5, waves, levels))
    shapes.sort()
    plain_ms, plain_waves, plain_levels = shapes[0]
    api.plain_config(ctx, plain_levels, plain_waves, args.depth)
    log("plain launch shapes (ms, waves/WG, LDS levels): " + ", ".join(f"({m:.3f}, {w}, {l})" for m, w, l in shapes))

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
    batched = None if args.no_batched else measure_batched(ctx, cache, args.depth, max(3, args.steps // 3))
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
                   "primary_rays": primary, "shadow_rays": shadow,
                   "plain_launch": {"waves_per_workgroup": plain_waves, "lds_stack_levels": plain_levels}},
        "roofline": {"bound": "hbm", "achieved": round(achieved_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved_gbs / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "RenderPlainKernel", "kernel_ms": round(avg_kernel_ms, 4),
                     "algorithmic_bytes_per_launch": bytes_per_launch},
        "cpu_baseline": cpu,
        "kernel_mrays_s": round(rays / avg_kernel_ms / 1e3, 1),
        "batched": batched,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

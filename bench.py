"""Benchmark: Mrays/s of the plain single-view render on the depth-11 SVO at 1080p.

Workload (one step = one GPURayTracer::RenderFrame): seeded depth-11 (2048^3)
fBm terrain (4 octaves, heights in [0, 0.6]) plus 64 random boxes with edges
in [0.01, 0.1], blocked into 8192-node blocks and fully resident in HBM; one
1920x1080 view, 60 degree horizontal FOV, no LOD, primary rays plus one shadow
ray per lit hit. Each frame runs the tracing pass, the feedback extraction and
the cache transaction that marks the hit blocks, as the reference does.

Prints one JSON line (see the field list at the bottom). The CPU baseline is
the oracle (a strict-fp32 C++ restatement of the same traversal) on the same
frame, on every host core. Beside the headline come the same frame under a
moving camera, the 64-view batched mix, the served request stream (depths
12-15 sharing a 16 GiB block pool, depth 15 served from its .tree file), the
depth-14 out-of-core fly-through in an 8 GiB pool, an imported OBJ scene and
the reference's CPU-sized case.
"""
from __future__ import annotations

import argparse
import json
import os
import shutil
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg: str):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "render_plain_depth11_1080p_hbm_traffic.md")


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def measured_traffic():
    """HBM bytes per launch of the plain kernel on this workload, from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE runs (scaled by the copy-kernel calibration recorded in the same file), or None."""
    try:
        for line in open(TRAFFIC_PROFILE):
            line = line.strip().lstrip("-* ")
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1].strip())
    except OSError:
        pass
    return None


def measure_batched(ctx, cache, depth: int, steps: int, nviews: int = 64, seed: int = 7, shape=None):
    """64-view mix in one batched launch vs the plain kernel launched once per view.
    shape: (tile, LDS levels, workgroups per CU) instead of the autotuner's choice."""
    import numpy as np
    from rayt_amd import api
    flags = api.L.RT_FLAG_SHADOWS
    from rayt_amd.workloads import view_mix
    specs = view_mix(nviews, seed)
    b = api.BatchRenderer(cache, nviews)
    b.config(tile_shape=0, lds_levels=12, wgs_per_cu=16, tree_depth=depth)
    views, total, offsets = b.make_views(specs, flags=flags)
    out = api.Buffer(ctx, 4 * total)
    codes = api.Buffer(ctx, 4 * total)
    st = b.render(views, nviews, out, codes, count=True)
    rays = st[0] + st[1]
    t = time.time()
    if shape:
        best = shape
        b.config(tile_shape=best[0], lds_levels=best[1], wgs_per_cu=best[2], tree_depth=depth)
        b.render(views, nviews, out, codes)
        best_ms = b.last_ms()
    else:
        best, best_ms = b.autotune(views, nviews, out, codes, repeats=2)
    log(f"batched: {nviews} views, {total} pixels, {rays} rays; launch shape chosen in {time.time() - t:.1f}s -> "
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


def pool_shares(trees, budget: int):
    """Blocks of pool per scene under a shared byte budget, by water-filling: scenes smaller than an equal
    share of what is left are fully resident, the others split the rest equally (served out of core)."""
    sizes = [(t.info().blocks_count, t.info().bytes_in_block_content + 36) for t in trees]
    order = sorted(range(len(trees)), key=lambda i: sizes[i][0] * sizes[i][1])
    left, out = budget, [0] * len(trees)
    for k, i in enumerate(order):
        blocks, bb = sizes[i]
        share = left // (len(trees) - k)
        out[i] = min(blocks, share // bb)
        left -= out[i] * bb
    return out


def measure_served(ctx, scenes, rate: float, seconds: float, burst: int, seed: int = 9, check: int = 16,
                   pool_gb: float = 16.0, max_block_loads: int = 128, shape=(0, 12, 0), dense_rate: float = 0.0,
                   dense_seconds: float = 10.0, rounds_in_flight: int = 4):
    """Poisson request stream over several scenes through the native continuous-batching server.

    scenes: list of (tree, depth), sharing a block-pool budget of `pool_gb` GiB (pool_shares: scenes that
    do not fit are served out of core, loading blocks from the feedback of their requests). Each request
    is a seeded pose (position uniform in [0,1]^3, direction uniform on the sphere) and a resolution from
    the batched mix, on a uniformly drawn scene. Returns the open-loop run (p50/p99 latency at the offered
    rate), a second open loop at `dense_rate` (0: none) and a closed burst (all requests queued at once: the
    server's throughput at saturation).
    max_block_loads bounds the blocks loaded per out-of-core scene between rounds."""
    import numpy as np
    from rayt_amd import api
    from rayt_amd.workloads import view_mix
    flags = api.L.RT_FLAG_SHADOWS
    rng = np.random.default_rng(seed)
    budget = int(pool_gb * (1 << 30))
    srv = api.RenderServer(ctx, max_views_per_round=64, rounds_in_flight=rounds_in_flight, framebuffer_bytes=2 << 30,
                           pool_bytes=budget, tile_shape=shape[0], lds_levels=shape[1], wgs_per_cu=shape[2],
                           max_block_loads=max_block_loads)
    shares = pool_shares([t for t, _ in scenes], budget)
    partial = []
    for (tree, depth), share in zip(scenes, shares):
        srv.add_scene(tree, depth, max_blocks=share)
        partial.append(share < tree.info().blocks_count)
    pools = [{"depth": d, "blocks": t.info().blocks_count, "pool_blocks": sh} for (t, d), sh in zip(scenes, shares)]
    log(f"served scene pools: {pools}")

    def make(n):
        specs = view_mix(n, int(rng.integers(1 << 30)))
        return [(vpi, w, h, int(rng.integers(len(scenes)))) for vpi, w, h in specs]

    # warm-up: load every scene, first launches
    srv.submit_many([(s[0], s[1], s[2], k) for k, s in enumerate(make(len(scenes)))], flags=flags)
    srv.drain(600000)
    srv.poll(4096)

    # per-request check: served results of the first requests equal the plain single-view render
    chk = make(check)
    ids = srv.submit_many(chk, flags=flags, want_codes=True)
    srv.drain(600000)
    got = {c.id: (img, cod, c.fault_blocks) for c, img, cod in srv.poll(4096)}
    caches = {}
    bad = checked = 0
    # The served scenes go deeper than the frame's: the plain reference renders with a stack of unknown depth.
    api.plain_config(ctx, 16, 4, 0)
    for rid, (vpi, w, h, s) in zip(ids, chk):
        img, cod, nf = got[rid]
        if partial[s]:  # out of core: the result is consistent with its fault count
            bad += int((nf > 0) != bool(np.any(cod < 0)))
            continue
        checked += 1
        if s not in caches:
            tree, _ = scenes[s]
            caches[s] = api.GPUOctreeCacheManager(tree.info().blocks_count, tree, ctx)
            caches[s].InitialFillCache()
        out, codes, _ = api.render_plain(ctx, caches[s], w, h, vpi, flags=flags)
        if not (np.array_equal(img, out.read_all(np.uint32)) and np.array_equal(cod, codes.read_all(np.int32))):
            bad += 1
    del caches
    if bad:
        raise SystemExit(f"{bad} served requests differ from the plain render")

    # Measured requests are delivered in the server's pinned memory and released as soon as their
    # completion is seen (the client only reads them), so the host copy to a caller buffer is not timed.
    missing = [0]  # requests whose result lacked blocks of an out-of-core scene

    def collect(timeout_ms=0):
        got = srv.poll(1024, timeout_ms=timeout_ms, keep_pinned=True)
        for c, _, _ in got:
            if c.status != 0:
                raise SystemExit(f"served request {c.id} failed ({c.status})")
            missing[0] += int(c.fault_blocks > 0)
            srv.release(c.id)
        return len(got)

    def open_loop(rate, seconds):
        """Poisson arrivals at `rate` for `seconds`; returns the server's statistics of the run."""
        n = max(1, int(rate * seconds))
        reqs = make(n)
        gaps = rng.exponential(1.0 / rate, n)
        srv.reset_stats()
        missing[0] = 0
        t0 = time.perf_counter()
        due = t0
        done = 0
        next_log = t0 + 10.0
        for i, (vpi, w, h, s) in enumerate(reqs):
            due += gaps[i]
            while True:
                now = time.perf_counter()
                if now >= due:
                    break
                done += collect()
                time.sleep(min(0.0005, max(0.0, due - now)))
            srv.submit(vpi, w, h, scene=s, flags=flags, pinned=True)
            if time.perf_counter() > next_log:
                log(f"served at {rate:g} req/s: {i + 1}/{n} submitted")
                next_log += 10.0
        while done < n:
            done += collect(1000)
        st = srv.stats()
        wall = time.perf_counter() - t0
        res = {"rate_req_s": rate, "seconds": round(wall, 2), "requests": int(st.completed),
               "failed": int(st.failed), "p50_ms": round(st.p50_ms, 3), "p99_ms": round(st.p99_ms, 3),
               "mean_ms": round(st.mean_ms, 3), "max_ms": round(st.max_ms, 3),
               "mrays_s": round((st.primary_rays + st.shadow_rays) / wall / 1e6, 1),
               "rounds": int(st.rounds), "mean_views_per_round": round(st.mean_views_per_round, 2),
               "max_rounds_in_flight": int(st.max_rounds_in_flight), "scene_loads": int(st.scene_loads),
               "block_uploads": int(st.block_uploads), "load_transactions": int(st.load_transactions),
               "requests_missing_blocks": missing[0]}
        log(f"served open loop at {rate:g} req/s: {res}")
        return res

    open_stated = open_loop(rate, seconds)
    # the same stream ten times denser: requests arrive while rounds run, so rounds batch several of them
    open_dense = open_loop(dense_rate, dense_seconds) if dense_rate > 0 else None

    # closed burst: everything queued at once
    reqs = make(burst)
    srv.reset_stats()
    missing[0] = 0
    srv.submit_many(reqs, flags=flags, pinned=True)
    done = 0
    while done < burst:
        done += collect(1000)
    st = srv.stats()
    closed = {"requests": int(st.completed), "p50_ms": round(st.p50_ms, 3), "p99_ms": round(st.p99_ms, 3),
              "mrays_s": round(st.mrays_per_s, 1), "span_ms": round(st.span_ms, 2), "rounds": int(st.rounds),
              "mean_views_per_round": round(st.mean_views_per_round, 2),
              "max_rounds_in_flight": int(st.max_rounds_in_flight), "block_uploads": int(st.block_uploads),
              "requests_missing_blocks": missing[0]}
    log(f"served burst: {closed}")
    srv.close()
    return {"scene_depths": [d for _, d in scenes], "scene_pools": pools, "pool_budget_bytes": budget,
            "framebuffer_bytes": 2 << 30, "requests_checked_bit_exact": checked, "open_loop": open_stated,
            "open_loop_dense": open_dense, "burst": closed}


def measure_plain_moving(ctx, cache, tr, W: int, H: int, flags: int, lod: float, frames: int = 64):
    """The plain frame over a moving camera: the seeded 64-pose spline fly-through over the same resident scene
    (a new view every frame, so the tile order learned from earlier frames is always one frame old). Rays per
    frame come from the counting kernel, outside the timed region."""
    from rayt_amd import api
    from rayt_amd.workloads import camera, spline_poses
    vpis = [camera(y, p, pos, W, H) for y, p, pos in spline_poses(frames, seed=5)]
    rays = 0
    for v in vpis:
        _, _, st = api.render_plain(ctx, cache, W, H, v, lod=lod, flags=flags, count=True)
        rays += st[0] + st[1]
    for v in vpis[:4]:
        tr.RenderFrame(v)
    ms0, n0 = tr.trace_totals()
    t0 = time.perf_counter()
    for v in vpis:
        tr.RenderFrame(v)
    ms1, n1 = tr.trace_totals()
    wall = time.perf_counter() - t0
    res = {"workload": f"{frames}-pose seeded spline fly-through, same scene and resolution", "frames": frames,
           "rays": rays, "ms_per_frame": round(wall / frames * 1e3, 4), "mrays_s": round(rays / wall / 1e6, 1),
           "kernel_ms_per_frame": round((ms1 - ms0) / (n1 - n0), 4)}
    log(f"plain frames, moving camera: {res}")
    return res


def measure_out_of_core(ctx, tree, depth: int, frac: float, frames: int, W: int, H: int, lod: float = 1.0,
                        pool_gb: float = 0.0):
    """Fly-through over a block pool that holds `frac` of the tree's blocks (or `pool_gb` GiB of blocks when
    given): every frame runs passes, feedback
    and cache transactions (faulted blocks uploaded from host memory under LRU eviction) until no ray faults
    or the session fills the pool. LOD voxel size 1 (the application's default). The frames' parity with the
    oracle is the tests' (tests/test_gpu_frame.py: fault lists, images, pool words and LRU order per frame)."""
    from rayt_amd import api
    info = tree.info()
    # The launch shape of the tracing passes and the fused single-pass frames: one wave per workgroup (436-438
    # frames/s at depth 14 against 420-424 with four), and the smallest stack window that holds the whole
    # stack (depth - 1 levels) up to 10 levels, the most that leaves 8 waves per SIMD; a deeper tree keeps
    # its deepest 10 levels in LDS and the rarely used top ones in a private array (with four-wave
    # workgroups: 421-423 frames/s against 400-401 with the whole 13-level stack in a 14-level window).
    levels = min(l for l in (8, 10, 12, 14, 16, 22) if l >= depth - 1) if depth - 1 <= 10 else 10
    api.plain_config(ctx, levels, 1, depth)
    block_bytes = info.bytes_in_block_content + 36  # channels + far-pointer slots + page-table entry
    if pool_gb > 0:
        max_blocks = max(8, min(info.blocks_count, int(pool_gb * (1 << 30) / block_bytes)))
    else:
        max_blocks = max(8, int(info.blocks_count * frac))
    cache = api.GPUOctreeCacheManager(max_blocks, tree, ctx)
    cache.InitialFillCache()
    tr = api.GPURayTracer(cache, W, H)
    tr.set_frame_time_limit(1e9)
    tr.set_lod_voxel_size(lod)
    tr.set_flags(api.L.RT_FLAG_SHADOWS)
    tr.set_seed(3)
    from rayt_amd.workloads import spline_poses
    poses = spline_poses(frames, seed=11)
    passes = uploads = rays = 0
    ctx.WaitForAll()
    t0 = time.perf_counter()
    up0 = cache.counters()[0]
    for k, (yaw, pitch, pos) in enumerate(poses):
        vpi = api.Camera(yaw=yaw, pitch=pitch, position=pos, fov=60.0, aspect=W / H).view_proj_inv()
        tr.RenderFrame(vpi)
        stt = tr.frame_stats()
        passes += stt[0]
        rays += stt[4]
        if (k + 1) % 32 == 0:
            log(f"out-of-core: {k + 1} frames")
    ctx.WaitForAll()
    wall = time.perf_counter() - t0
    uploads = cache.counters()[0] - up0
    res = {"depth": depth, "blocks": info.blocks_count, "pool_blocks": max_blocks, "lod_voxel_size": lod,
           "pool_bytes": max_blocks * block_bytes, "frames": frames,
           "width": W, "height": H, "frames_per_s": round(frames / wall, 2), "mrays_s": round(rays / wall / 1e6, 1),
           "passes_per_frame": round(passes / frames, 2), "uploads_per_frame": round(uploads / frames, 1)}
    log(f"out-of-core fly-through: {res}")
    return res


def write_grid_obj(d: str, n: int) -> str:
    """A seeded heightfield mesh: (n+1)^2 vertices, n^2 quads (2 n^2 triangles after fan splitting), vt
    tiled 4x, one material with a colour texture and a bump map (both PNM)."""
    import numpy as np
    os.makedirs(d, exist_ok=True)
    rng = np.random.default_rng(0)
    tex = rng.integers(0, 256, size=(64, 64, 3)).astype(np.uint8)
    open(os.path.join(d, "albedo.ppm"), "wb").write(b"P6\n64 64\n255\n" + tex.tobytes())
    bump = rng.integers(0, 256, size=(32, 32)).astype(np.uint8)
    open(os.path.join(d, "bump.pgm"), "wb").write(b"P5\n32 32\n255\n" + bump.tobytes())
    open(os.path.join(d, "grid.mtl"), "w").write("newmtl g\nKa 0.05 0.05 0.05\nKd 0.9 0.9 0.9\nmap_Kd albedo.ppm\n"
                                                  "bump bump.pgm -bm 0.3\n")
    x = np.linspace(0.0, 1.0, n + 1)
    X, Z = np.meshgrid(x, x, indexing="ij")
    Y = 0.3 + 0.1 * np.sin(6 * X) * np.cos(5 * Z) + 0.05 * np.sin(17 * X + 3 * Z)
    lines = ["mtllib grid.mtl", "usemtl g"]
    lines += [f"v {a:.7f} {b:.7f} {c:.7f}" for a, b, c in zip(X.ravel(), Y.ravel(), Z.ravel())]
    lines += [f"vt {4 * a:.6f} {4 * c:.6f}" for a, c in zip(X.ravel(), Z.ravel())]
    i, j = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    a = (i * (n + 1) + j + 1).ravel()
    b, c, e = a + n + 1, a + n + 2, a + 1
    lines += [f"f {p}/{p} {q}/{q} {r}/{r} {s}/{s}" for p, q, r, s in zip(a, b, c, e)]
    path = os.path.join(d, "grid.obj")
    open(path, "w").write("\n".join(lines) + "\n")
    return path


def measure_obj_scene(ctx, depth: int, grid: int, steps: int, W: int, H: int):
    """An imported OBJ scene (ImportObjScene): the heightfield mesh voxelised to `depth` on the host (timed),
    then fully resident 1080p frames (primary + shadow rays) through the same frame path as the headline."""
    import tempfile
    from rayt_amd import api
    d = os.path.join(tempfile.gettempdir(), f"rayt_bench_obj_{grid}")
    path = write_grid_obj(d, grid)
    t = time.perf_counter()
    tree = api.StoredOctree.obj(path, depth, 8192)
    import_s = time.perf_counter() - t
    info = tree.info()
    log(f"OBJ scene: {2 * grid * grid} triangles -> depth {depth}, {info.source_nodes} nodes, {info.blocks_count} "
        f"blocks, imported in {import_s:.2f}s")
    api.plain_config(ctx, min(l for l in (8, 10, 12, 14, 16, 22) if l >= depth - 1), 4, depth)
    cache = api.GPUOctreeCacheManager(info.blocks_count, tree, ctx)
    cache.InitialFillCache()
    vpi = api.Camera(yaw=35.0, pitch=-25.0, position=(0.05, 0.75, 1.25), fov=60.0, aspect=W / H).view_proj_inv()
    tr = api.GPURayTracer(cache, W, H)
    tr.set_frame_time_limit(1e9)
    tr.set_lod_voxel_size(0.0)
    tr.set_flags(api.L.RT_FLAG_SHADOWS)
    tr.set_seed(1)
    for _ in range(3):
        tr.RenderFrame(vpi)
    ctx.WaitForAll()
    rays = 0
    t0 = time.perf_counter()
    for _ in range(steps):
        tr.RenderFrame(vpi)
        rays += tr.frame_stats()[4]
    ctx.WaitForAll()
    el = time.perf_counter() - t0
    res = {"triangles": 2 * grid * grid, "depth": depth, "nodes": info.source_nodes, "blocks": info.blocks_count,
           "import_s": round(import_s, 2), "import_mnodes_s": round(info.source_nodes / import_s / 1e6, 2),
           "host_threads": os.cpu_count(), "width": W, "height": H, "frame_ms": round(el / steps * 1e3, 3),
           "mrays_s": round(rays / el / 1e6, 1)}
    log(f"OBJ scene frames: {res}")
    del tr, cache, tree
    return res


def measure_cpu_plain_config(ctx, threads: int):
    """The reference's CPU-runnable case: depth-8 terrain + 8 spheres, 640x480, 60 degree FOV, fixed pose,
    primary rays only, fully resident, no LOD. The oracle on all threads and on one, and the GPU kernel."""
    import numpy as np
    from rayt_amd import api
    import oracle as orc
    spec = api.TerrainSpec(depth=8, seed=1, octaves=4, spheres=8, sphere_rmin=0.05, sphere_rmax=0.2)
    tree = api.StoredOctree.terrain(spec, 8192)
    cache = api.GPUOctreeCacheManager(tree.info().blocks_count, tree, ctx)
    cache.InitialFillCache()
    W, H = 640, 480
    vpi = api.Camera(yaw=35.0, pitch=-25.0, position=(0.05, 0.75, 1.25), fov=60.0, aspect=W / H).view_proj_inv()
    out, codes = api.Buffer(ctx, 4 * W * H), api.Buffer(ctx, 4 * W * H)
    for _ in range(3):
        api.render_plain(ctx, cache, W, H, vpi, out=out, codes=codes)
    tm = api.EventTimer(ctx)
    tm.start()
    for _ in range(20):
        api.render_plain(ctx, cache, W, H, vpi, out=out, codes=codes)
    gpu_ms = tm.stop() / 20
    links = cache.channel(0).read_all(np.uint32)
    far = cache.far_pointers().read_all(np.uint32)
    normals = cache.channel_by_name("normals").read_all(np.uint32)
    colors = cache.channel_by_name("colors").read_all(np.uint32)

    def cpu(th, secs):
        runs, tc, r = 0, 0.0, None
        while tc < secs or runs == 0:
            a = time.perf_counter()
            r = orc.render_plain(links, far, cache.root_node_index(), normals, colors, cache.nodes_in_block(), W, H,
                                 vpi, threads=th)
            tc += time.perf_counter() - a
            runs += 1
        return W * H * runs / tc / 1e6, r

    cpu_all, r = cpu(threads, 3.0)
    cpu_one, _ = cpu(1, 2.0)
    same = bool(np.array_equal(out.read_all(np.uint32).reshape(H, W), r["image"]))
    res = {"workload": "depth-8 terrain + 8 spheres, 640x480, primary rays, fixed pose", "gpu_mrays_s":
           round(W * H / gpu_ms / 1e3, 1), "gpu_ms": round(gpu_ms, 4), "cpu_mrays_s": round(cpu_all, 2),
           "cpu_threads": threads, "cpu_1thread_mrays_s": round(cpu_one, 2), "gpu_equals_oracle": same}
    log(f"CPU-plain config: {res}")
    if not same:
        raise SystemExit("GPU image differs from the oracle on the CPU-plain config")
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
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--serve-rate", type=float, default=200.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--serve-dense-rate", type=float, default=2000.0,
                    help="a second open loop at this rate (0: none), where rounds batch the requests that arrive")
    ap.add_argument("--serve-dense-seconds", type=float, default=10.0)
    ap.add_argument("--serve-burst", type=int, default=256, help="requests queued at once for the saturation run")
    ap.add_argument("--serve-rounds", type=int, default=4, help="rounds the server keeps in flight (1-4)")
    ap.add_argument("--serve-depths", default="12,13,14,15",
                    help="depths of the served scenes (fBm terrain, 6 octaves, seed = depth)")
    ap.add_argument("--serve-file-depth", type=int, default=15,
                    help="served scenes at least this deep are built straight into a .tree file under --scene-dir "
                         "and served from it (0: every scene in host RAM)")
    ap.add_argument("--scene-dir", default="/tmp", help="where file-backed scenes are written (deleted afterwards)")
    ap.add_argument("--serve-pool-gb", type=float, default=16.0,
                    help="block-pool budget shared by the served scenes; scenes beyond it are served out of core")
    ap.add_argument("--serve-shape", default="0,12,0", metavar="TILE,LEVELS,WGS",
                    help="launch shape of the server's rounds (WGS 0: one wave per tile)")
    ap.add_argument("--serve-max-loads", type=int, default=128,
                    help="blocks loaded per out-of-core scene between rounds (0: until its pool is full)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MB",
                    help="also run a device-to-device copy of MB MiB (3 times) as the byte-counter reference")
    ap.add_argument("--no-outofcore", action="store_true")
    ap.add_argument("--ooc-depth", type=int, default=14)
    ap.add_argument("--ooc-frac", type=float, default=0.25, help="pool size as a fraction of the tree's blocks")
    ap.add_argument("--ooc-pool-gb", type=float, default=8.0, help="pool size in GiB (0: --ooc-frac of the tree)")
    ap.add_argument("--ooc-frames", type=int, default=128)
    ap.add_argument("--no-cpu-plain", action="store_true")
    ap.add_argument("--no-plain-moving", action="store_true", help="skip the moving-camera plain frames")
    ap.add_argument("--no-obj", action="store_true")
    ap.add_argument("--obj-depth", type=int, default=11)
    ap.add_argument("--obj-grid", type=int, default=256, help="heightfield mesh of 2 * grid^2 triangles")
    ap.add_argument("--dump", default="", metavar="DIR", help="write the measured frame and its feedback codes as PNG")
    ap.add_argument("--plain-shape", default="1,10", metavar="WAVES,LEVELS|sweep",
                    help="plain kernel launch shape (the one the committed HBM-traffic profile measured), or "
                         "'sweep' to time 1/4 waves x 10/12/16 LDS levels and keep the fastest")
    ap.add_argument("--batch-shape", default="", metavar="TILE,LEVELS,WGS",
                    help="batched launch shape instead of the autotuner (tile 0=8x8, 1=16x4, 2=32x2)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from rayt_amd import api

    W, H = args.width, args.height
    t = time.time()
    from rayt_amd.workloads import PLAIN_POSE, camera, plain_scene_spec
    tree = api.StoredOctree.terrain(plain_scene_spec(args.depth), 8192)
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

    vpi = camera(*PLAIN_POSE, W, H)
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
    if args.plain_shape != "sweep":
        plain_waves, plain_levels = (int(x) for x in args.plain_shape.split(","))
    else:
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
        _, plain_waves, plain_levels = shapes[0]
        log("plain launch shapes (ms, waves/WG, LDS levels): " + ", ".join(f"({m:.3f}, {w}, {l})" for m, w, l in shapes))
    api.plain_config(ctx, plain_levels, plain_waves, args.depth)

    tr = api.GPURayTracer(cache, W, H)
    tr.set_frame_time_limit(1e9)
    tr.set_lod_voxel_size(lod)
    tr.set_flags(flags)
    tr.set_seed(1)
    for _ in range(args.warmup):
        tr.RenderFrame(vpi)
    ctx.WaitForAll()
    ms0, n0 = tr.trace_totals()  # finishes the last warm-up frame
    t0 = time.perf_counter()
    for i in range(args.steps):
        tr.RenderFrame(vpi)  # pipelined: the host half of frame i runs while frame i + 1 traces
    ms1, n1 = tr.trace_totals()  # finishes the last frame: every timed frame is complete here
    ctx.WaitForAll()
    t1 = time.perf_counter()
    assert n1 - n0 == args.steps
    elapsed = t1 - t0
    ms_per_step = elapsed / args.steps * 1e3
    value = rays * args.steps / elapsed / 1e6
    avg_kernel_ms = (ms1 - ms0) / args.steps
    achieved_gbs = bytes_per_launch / (avg_kernel_ms * 1e-3) / 1e9
    log(f"{args.steps} frames in {elapsed * 1e3:.1f} ms: {value:.0f} Mrays/s; trace kernel {avg_kernel_ms:.3f} ms/frame "
        f"({rays / avg_kernel_ms / 1e3:.0f} Mrays/s), {achieved_gbs:.0f} GB/s algorithmic")

    img = tr.out_image()
    moving = None if args.no_plain_moving else measure_plain_moving(ctx, cache, tr, W, H, flags, lod)
    if args.dump:
        from rayt_amd import imageio
        os.makedirs(args.dump, exist_ok=True)
        imageio.write_png(os.path.join(args.dump, "plain_frame.png"), img)
        imageio.write_png(os.path.join(args.dump, "plain_codes.png"), imageio.code_colors(tr.codes()))
    batch_shape = tuple(int(x) for x in args.batch_shape.split(",")) if args.batch_shape else None
    batched = None if args.no_batched else measure_batched(ctx, cache, args.depth, max(3, args.steps // 3),
                                                           shape=batch_shape)
    # Terrain scenes of the served and out-of-core runs: fBm terrain, 6 octaves, seed = depth (the out-of-core
    # depth-14 scene is the served one). Scenes of depth >= --serve-file-depth are built into a .tree file and
    # served from it (block reads through the file loader), the others in host RAM.
    built = {}
    files = []

    def terrain_scene(depth: int, to_file: bool):
        if depth in built:
            return built[depth]
        spec = api.TerrainSpec(depth=depth, seed=depth, octaves=6)
        if depth >= 13:
            os.environ["RAYT_BUILD_LOG"] = "1"  # the builder reports its progress (tens of seconds of work)
        t = time.time()
        need = int(1.2 * 16 * (1 << 30) * 4 ** (depth - 14))  # about 4x per level; depth 14 is ~16 GiB of blocks
        if to_file and shutil.disk_usage(args.scene_dir).free > need:
            path = os.path.join(args.scene_dir, f"rayt_terrain_d{depth}_s{depth}_o6.tree")
            files.append(path)
            tree = api.StoredOctree.terrain_file(spec, path, 8192)
            where = f"file {path} ({os.path.getsize(path) / 2**30:.1f} GiB)"
        else:
            tree = api.StoredOctree.terrain(spec, 8192)
            where = "host RAM"
        os.environ.pop("RAYT_BUILD_LOG", None)
        log(f"scene depth {depth} (6 octaves, seed {depth}): {tree.info().blocks_count} blocks in {where}, "
            f"built in {time.time() - t:.1f}s")
        built[depth] = (tree, where, round(time.time() - t, 1))
        return built[depth]

    served = None
    outofcore = None
    try:
        if not args.no_served:
            depths = [int(x) for x in args.serve_depths.split(",")]
            scenes = [(terrain_scene(d, args.serve_file_depth > 0 and d >= args.serve_file_depth)[0], d)
                      for d in depths]
            served = measure_served(ctx, scenes, args.serve_rate, args.serve_seconds, args.serve_burst,
                                    pool_gb=args.serve_pool_gb, max_block_loads=args.serve_max_loads,
                                    shape=tuple(int(x) for x in args.serve_shape.split(",")),
                                    dense_rate=args.serve_dense_rate, dense_seconds=args.serve_dense_seconds,
                                    rounds_in_flight=args.serve_rounds)
            served["scenes"] = [{"depth": d, "store": built[d][1], "build_s": built[d][2]} for d in depths]
            del scenes
        if not args.no_outofcore:
            otree, where, build_s = terrain_scene(args.ooc_depth, False)
            outofcore = measure_out_of_core(ctx, otree, args.ooc_depth, args.ooc_frac, args.ooc_frames, W, H,
                                            pool_gb=args.ooc_pool_gb)
            outofcore["build_s"] = build_s
            del otree
    finally:
        built.clear()
        for path in files:
            try:
                os.remove(path)
            except OSError:
                pass
    obj_scene = None if args.no_obj else measure_obj_scene(ctx, args.obj_depth, args.obj_grid, args.steps, W, H)
    host_cores = os.cpu_count() or 1
    cpu_plain = None if args.no_cpu_plain or args.no_cpu_baseline else measure_cpu_plain_config(ctx, host_cores)

    cpu = None
    if not args.no_cpu_baseline:
        import oracle as orc  # the CPU baseline, and the check of the measured frame against it
        links = cache.channel(0).read_all(np.uint32)
        far = cache.far_pointers().read_all(np.uint32)
        normals = cache.channel_by_name("normals").read_all(np.uint32)
        colors = cache.channel_by_name("colors").read_all(np.uint32)
        threads = host_cores
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
               "sample": f"{runs} full {W}x{H} frame(s) of the same workload ({tc:.1f}s)",
               "nproc": host_cores, "cpu_model": cpu_model()}
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
                     "frac": round(achieved_gbs / HBM_PEAK_GBS, 4), "traffic": measured_traffic(),
                     "kernel": "RenderPlainKernel", "kernel_ms": round(avg_kernel_ms, 4),
                     "algorithmic_bytes_per_launch": bytes_per_launch},
        "cpu_baseline": cpu,
        "kernel_mrays_s": round(rays / avg_kernel_ms / 1e3, 1),
        "plain_moving": moving,
        "batched": batched,
        "served": served,
        "out_of_core": outofcore,
        "obj_scene": obj_scene,
        "cpu_plain_config": cpu_plain,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

"""Host-side mirror of the reference renderer API over the C ABI.

Class and method names follow the reference so that code written against it
reads the same:
  Context               CL util/cl-context.h:7-23        (CLContext)
  Buffer                CL util/cl-buffer.h:9-35         (CLBuffer)
  StoredOctree          CPU data/stored-octree-loader.h  (StoredOctreeLoader)
  GPUOctreeCacheManager GPU data/gpu-octree-cache-manager.h:14-68
  GPUSort / GPUScan     GPU algorithms/gpu-sort.h, gpu-scan.h
  RayTracingFeedbackExtractor GPU_renderer/ray-tracing-feedback-extractor.h:14-66
  GPURayTracer          GPU_renderer/gpu-ray-tracer.h:11-55
  Camera                scene/camera.h:9-46
Errors raise RaytError (the reference crash()es).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from ._lib import RaytError, call, lib  # noqa: F401

BACKGROUND = np.array([0.0, np.float32(216 / 255.0), 1.0, 1.0], dtype=np.float32)


def _fp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class Context:
    """One device and one in-order HIP stream."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        call("rt_ctx_create", device, C.byref(h))
        self.h = h

    def WaitForAll(self):
        call("rt_ctx_finish", self.h)

    def compute_units(self) -> int:
        n = C.c_int()
        call("rt_ctx_compute_units", self.h, C.byref(n))
        return n.value

    def close(self):
        if self.h:
            lib().rt_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Buffer:
    """Device buffer; borrowed buffers (owned by a cache or tracer) are not freed."""

    def __init__(self, ctx: Context, size: int = 0, handle=None, owner=None):
        self.ctx = ctx
        self._owner = owner
        if handle is not None:
            self.h = C.c_void_p(handle) if isinstance(handle, int) else handle
            self.owned = False
        else:
            h = C.c_void_p()
            call("rt_buf_create", ctx.h, size, C.byref(h))
            self.h = h
            self.owned = True

    def size(self) -> int:
        return int(lib().rt_buf_size(self.h))

    def Write(self, offset: int, data: np.ndarray, blocking: bool = True):
        data = np.ascontiguousarray(data)
        call("rt_buf_write", self.h, offset, data.nbytes, data.ctypes.data, int(blocking), None, 0, None)

    def Read(self, offset: int, nbytes: int, dtype=np.uint8) -> np.ndarray:
        out = np.empty(nbytes // np.dtype(dtype).itemsize, dtype=dtype)
        call("rt_buf_read", self.h, offset, out.nbytes, out.ctypes.data, 1, None, 0, None)
        return out

    def read_all(self, dtype=np.uint8) -> np.ndarray:
        return self.Read(0, self.size(), dtype)

    def CopyFrom(self, src: "Buffer", src_start: int = 0, dest_start: int = 0, length: int | None = None):
        n = src.size() if length is None else length
        call("rt_buf_copy", self.h, dest_start, src.h, src_start, n, None, 0, None)

    def close(self):
        if self.owned and self.h:
            lib().rt_buf_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def buffer_from(ctx: Context, arr: np.ndarray) -> Buffer:
    arr = np.ascontiguousarray(arr)
    b = Buffer(ctx, max(arr.nbytes, 4))
    if arr.nbytes:
        b.Write(0, arr)
    return b


@dataclass
class TerrainSpec:
    depth: int = 8
    seed: int = 1
    octaves: int = 4
    height_lo: float = 0.0
    height_hi: float = 0.6
    spheres: int = 0
    sphere_rmin: float = 0.05
    sphere_rmax: float = 0.2
    boxes: int = 0
    box_emin: float = 0.01
    box_emax: float = 0.1
    threads: int = 0


class StoredOctree:
    """A block source: a .tree file or a tree built in host RAM."""

    def __init__(self, handle):
        self.h = handle

    @staticmethod
    def open(path: str) -> "StoredOctree":
        h = C.c_void_p()
        call("rt_tree_open", path.encode(), C.byref(h))
        return StoredOctree(h)

    @staticmethod
    def sphere(center, radius: float, level: int, nodes_in_block: int) -> "StoredOctree":
        h = C.c_void_p()
        call("rt_tree_build_sphere", float(center[0]), float(center[1]), float(center[2]), float(radius), level,
             nodes_in_block, C.byref(h))
        return StoredOctree(h)

    @staticmethod
    def _params(spec: TerrainSpec):
        return L.TerrainParams(spec.depth, spec.seed, spec.octaves, spec.height_lo, spec.height_hi, spec.spheres,
                               spec.sphere_rmin, spec.sphere_rmax, spec.boxes, spec.box_emin, spec.box_emax,
                               spec.threads)

    @staticmethod
    def terrain(spec: TerrainSpec, nodes_in_block: int = 8192) -> "StoredOctree":
        """The terrain scene blocked by the parallel builder into host RAM."""
        p = StoredOctree._params(spec)
        h = C.c_void_p()
        call("rt_tree_build_terrain", C.byref(p), nodes_in_block, C.byref(h))
        return StoredOctree(h)

    @staticmethod
    def terrain_file(spec: TerrainSpec, path: str, nodes_in_block: int = 8192) -> "StoredOctree":
        """The terrain scene blocked straight into the .tree file `path`, opened as a file-backed source."""
        p = StoredOctree._params(spec)
        h = C.c_void_p()
        call("rt_tree_build_terrain_file", C.byref(p), nodes_in_block, path.encode(), C.byref(h))
        return StoredOctree(h)

    @staticmethod
    def terrain_levelwise(spec: TerrainSpec, nodes_in_block: int = 8192) -> "StoredOctree":
        """The same blocks from the whole node tree generated in RAM first (depth <= 14)."""
        p = StoredOctree._params(spec)
        h = C.c_void_p()
        call("rt_tree_build_terrain_levelwise", C.byref(p), nodes_in_block, C.byref(h))
        return StoredOctree(h)

    @staticmethod
    def obj(obj_file_name: str, max_level: int, nodes_in_block: int, threads: int = 0) -> "StoredOctree":
        """A voxelised Wavefront OBJ scene (BuilderDataSource, scene sources/import-obj-scene.cpp:529-638)."""
        h = C.c_void_p()
        call("rt_tree_import_obj", obj_file_name.encode(), max_level, nodes_in_block, threads, C.byref(h))
        return StoredOctree(h)

    def write(self, path: str):
        call("rt_tree_write", self.h, path.encode())

    def info(self) -> L.TreeInfo:
        i = L.TreeInfo()
        call("rt_tree_info_get", self.h, C.byref(i))
        return i

    def channels(self) -> list[tuple[int, str]]:
        out = []
        for c in range(self.info().channels_count):
            b = C.c_int()
            name = C.create_string_buffer(256)
            call("rt_tree_channel", self.h, c, C.byref(b), name, 256)
            out.append((b.value, name.value.decode()))
        return out

    def load_block(self, index: int):
        info = self.info()
        hdr = np.zeros(76, np.uint8)
        payload = np.zeros(info.bytes_in_block_content, np.uint8)
        call("rt_tree_load_block", self.h, index, hdr.ctypes.data_as(L.U8P), payload.ctypes.data_as(L.U8P),
             payload.nbytes)
        return hdr, payload

    def close(self):
        if self.h:
            lib().rt_tree_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ImportObjScene(max_level: int, nodes_in_block: int, obj_file_name: str, out_tree_file_name: str) -> None:
    """ImportObjScene (scene sources/import-obj-scene.cpp:640-652): OBJ file -> .tree file."""
    t = StoredOctree.obj(obj_file_name, max_level, nodes_in_block)
    try:
        t.write(out_tree_file_name)
    finally:
        t.close()


class GPUOctreeCacheManager:
    def __init__(self, max_blocks_count: int, tree: StoredOctree, ctx: Context):
        h = C.c_void_p()
        call("rt_cache_create", ctx.h, tree.h, max_blocks_count, C.byref(h))
        self.h, self.ctx, self.tree = h, ctx, tree

    def InitialFillCache(self):
        call("rt_cache_initial_fill", self.h)

    def root_node_index(self) -> int:
        return call("rt_cache_root_node_index", self.h)

    def nodes_in_block(self) -> int:
        return call("rt_cache_nodes_in_block", self.h)

    def max_blocks(self) -> int:
        return call("rt_cache_max_blocks", self.h)

    def loaded_blocks(self) -> int:
        return call("rt_cache_loaded_blocks", self.h)

    def channel(self, index: int) -> Buffer | None:
        p = lib().rt_cache_channel(self.h, index)
        return Buffer(self.ctx, handle=p, owner=self) if p else None

    def channel_by_name(self, name: str) -> Buffer | None:
        p = lib().rt_cache_channel_by_name(self.h, name.encode())
        return Buffer(self.ctx, handle=p, owner=self) if p else None

    def far_pointers(self) -> Buffer:
        return Buffer(self.ctx, handle=lib().rt_cache_far_pointers(self.h), owner=self)

    def block_slot(self, block: int) -> int:
        return lib().rt_cache_block_slot(self.h, block)

    def slot_block(self, slot: int) -> int:
        return lib().rt_cache_slot_block(self.h, slot)

    def StartRequestSession(self):
        call("rt_cache_start_session", self.h)

    def StartRequestTransaction(self):
        call("rt_cache_start_transaction", self.h)

    def MarkBlockAsUsed(self, slot: int):
        call("rt_cache_mark_block_used", self.h, slot)

    def MarkParentAsUsed(self, slot: int):
        call("rt_cache_mark_parent_used", self.h, slot)

    def RequestBlock(self, block: int):
        call("rt_cache_request_block", self.h, block)

    def TransactionFilledCache(self) -> bool:
        return bool(call("rt_cache_transaction_filled", self.h))

    def SessionFilledCache(self) -> bool:
        return bool(call("rt_cache_session_filled", self.h))

    def EndRequestTransaction(self):
        call("rt_cache_end_transaction", self.h)

    def EndRequestSession(self):
        call("rt_cache_end_session", self.h)

    def lru_order(self) -> list[int]:
        n = call("rt_cache_lru_order", self.h, None, 0)
        out = (C.c_int * max(n, 1))()
        call("rt_cache_lru_order", self.h, out, n)
        return list(out[:n])

    def counters(self) -> list[int]:
        out = (C.c_uint64 * 4)()
        call("rt_cache_counters", self.h, out)
        return list(out)

    def pool_arrays(self) -> dict:
        """Host copies of the pool: links, far pointers and every other channel (bytes)."""
        d = {"links": self.channel(0).read_all(np.uint32), "far": self.far_pointers().read_all(np.uint32)}
        for i, (b, name) in enumerate(self.tree.channels()):
            if i == 0:
                continue
            d[name] = self.channel(i).read_all(np.uint8)
        return d

    def close(self):
        if self.h:
            lib().rt_cache_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GPUSort:
    """Ascending sort of int32 in place (GPU algorithms/gpu-sort.cpp:16-49)."""

    def __init__(self, length: int, ctx: Context):
        self.length, self.ctx = length, ctx
        self.temp = Buffer(ctx, max(4, 4 * length))

    def Sort(self, data: Buffer):
        call("rt_sort_i32", self.ctx.h, data.h, self.length, self.temp.h)


class GPUScan:
    """Inclusive prefix sum of int32 in place (GPU algorithms/gpu-scan.cpp:15-43)."""

    def __init__(self, length: int, ctx: Context):
        self.length, self.ctx = length, ctx
        self.temp = Buffer(ctx, max(4, 4 * length))

    def Scan(self, data: Buffer):
        call("rt_scan_i32", self.ctx.h, data.h, self.length, self.temp.h)


class RayTracingFeedbackExtractor:
    def __init__(self, frame_width: int, frame_height: int, ctx: Context, handle=None):
        self.ctx = ctx
        if handle is not None:
            self.h, self.owned = handle, False
        else:
            h = C.c_void_p()
            call("rt_feedback_create", ctx.h, frame_width * frame_height, C.byref(h))
            self.h, self.owned = h, True

    def set_hits_to_extract(self, v: int):
        call("rt_feedback_set_hits_to_extract", self.h, v)

    def set_seed(self, seed: int):
        call("rt_feedback_set_seed", self.h, seed)

    def Process(self, faults_and_hits: Buffer, nodes_in_block: int, block_level: bool = False):
        call("rt_feedback_process", self.h, faults_and_hits.h, nodes_in_block, int(block_level))

    def _list(self, which: int) -> list[int]:
        n = call("rt_feedback_list", self.h, which, None, 0)
        out = (C.c_int * max(n, 1))()
        call("rt_feedback_list", self.h, which, out, n)
        return list(out[:n])

    def fault_blocks(self):
        return self._list(0)

    def hit_blocks(self):
        return self._list(1)

    def duplicate_hit_blocks(self):
        return self._list(2)

    def stats(self):
        u, e = C.c_int(), C.c_int()
        call("rt_feedback_stats", self.h, C.byref(u), C.byref(e))
        return u.value, e.value

    def close(self):
        if self.owned and self.h:
            lib().rt_feedback_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GPURayTracer:
    def __init__(self, cache_manager: GPUOctreeCacheManager, frame_width: int, frame_height: int):
        h = C.c_void_p()
        call("rt_tracer_create", cache_manager.h, frame_width, frame_height, C.byref(h))
        self.h, self.cache, self.w, self.hgt = h, cache_manager, frame_width, frame_height

    def set_frame_time_limit(self, s: float):
        call("rt_tracer_set_frame_time_limit", self.h, s)

    def set_lod_voxel_size(self, v: float):
        call("rt_tracer_set_lod_voxel_size", self.h, v)

    def set_max_passes(self, n: int):
        call("rt_tracer_set_max_passes", self.h, n)

    def set_flags(self, flags: int):
        call("rt_tracer_set_flags", self.h, flags)

    def set_seed(self, seed: int):
        call("rt_tracer_set_seed", self.h, seed)

    def RenderFrame(self, view_proj_inv: np.ndarray):
        m = np.ascontiguousarray(view_proj_inv, dtype=np.float32).reshape(16)
        call("rt_tracer_render_frame", self.h, _fp(m))

    def out_image(self) -> np.ndarray:
        b = Buffer(self.cache.ctx, handle=lib().rt_tracer_out_image(self.h), owner=self)
        return b.read_all(np.uint32).reshape(self.hgt, self.w)

    def codes(self) -> np.ndarray:
        b = Buffer(self.cache.ctx, handle=lib().rt_tracer_codes(self.h), owner=self)
        return b.read_all(np.int32).reshape(self.hgt, self.w)

    def feedback(self) -> RayTracingFeedbackExtractor:
        return RayTracingFeedbackExtractor(0, 0, self.cache.ctx, handle=C.c_void_p(lib().rt_tracer_feedback(self.h)))

    def frame_stats(self) -> list[int]:
        out = (C.c_uint64 * 5)()
        call("rt_tracer_frame_stats", self.h, out)
        return list(out)

    def set_pipelining(self, enabled: bool):
        call("rt_tracer_set_pipelining", self.h, int(enabled))

    def trace_totals(self) -> tuple[float, int]:
        """(device ms in tracing passes summed over every finished frame, frames)."""
        ms, n = C.c_double(), C.c_uint64()
        call("rt_tracer_trace_totals", self.h, C.byref(ms), C.byref(n))
        return ms.value, n.value

    def timings(self) -> tuple[float, float]:
        """(device ms in tracing passes, host wall ms) of the last RenderFrame."""
        out = (C.c_double * 2)()
        call("rt_tracer_timings", self.h, out)
        return out[0], out[1]

    def pass_faults(self) -> list[list[int]]:
        n = call("rt_tracer_pass_faults", self.h, None, 0, None, 0)
        counts = (C.c_int * max(n, 1))()
        call("rt_tracer_pass_faults", self.h, counts, n, None, 0)
        total = sum(counts[:n])
        faults = (C.c_int * max(total, 1))()
        call("rt_tracer_pass_faults", self.h, counts, n, faults, total)
        out, k = [], 0
        for c in counts[:n]:
            out.append(list(faults[k:k + c]))
            k += c
        return out

    def close(self):
        if self.h:
            lib().rt_tracer_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Camera:
    """scene/camera.h:9-46: yaw/pitch/roll in degrees, horizontal FOV, near plane, aspect."""

    def __init__(self, yaw=0.0, pitch=0.0, roll=0.0, position=(0.0, 0.0, 0.0), fov=90.0, near=0.01,
                 aspect=4.0 / 3.0):
        self.yaw, self.pitch, self.roll = yaw, pitch, roll
        self.position = tuple(position)
        self.fov, self.near, self.aspect = fov, near, aspect

    def view_proj_inv(self) -> np.ndarray:
        pos = np.array(self.position, np.float32)
        out = np.zeros(16, np.float32)
        call("rt_camera_view_proj_inv", np.float32(self.yaw), np.float32(self.pitch), np.float32(self.roll),
             _fp(pos), np.float32(self.fov), np.float32(self.near), np.float32(self.aspect), _fp(out))
        return out


def plain_config(ctx: Context, lds_levels: int = 16, wg_waves: int = 4, tree_depth: int = 0):
    """Launch shape of the plain kernel on this context (rt_plain_config)."""
    call("rt_plain_config", ctx.h, lds_levels, wg_waves, tree_depth)


class EventTimer:
    """Device time between two events recorded on the context stream."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        self.a = self.b = None

    def start(self):
        self.a = C.c_void_p()
        call("rt_evt_record", self.ctx.h, C.byref(self.a))

    def stop(self) -> float:
        self.b = C.c_void_p()
        call("rt_evt_record", self.ctx.h, C.byref(self.b))
        arr = (C.c_void_p * 1)(self.b)
        call("rt_evt_wait", 1, arr)
        ns = C.c_double()
        call("rt_evt_elapsed_ns", self.a, self.b, C.byref(ns))
        lib().rt_evt_release(self.a)
        lib().rt_evt_release(self.b)
        return ns.value * 1e-6


def render_plain(ctx: Context, cache: GPUOctreeCacheManager, width: int, height: int, vpi: np.ndarray,
                 lod: float = 0.0, flags: int = 0, out: Buffer | None = None, codes: Buffer | None = None,
                 count: bool = False):
    """One fused single-pass render; returns (out, codes, stats or None)."""
    out = out or Buffer(ctx, 4 * width * height)
    codes = codes or Buffer(ctx, 4 * width * height)
    m = np.ascontiguousarray(vpi, np.float32).reshape(16)
    normals = cache.channel_by_name("normals")
    colors = cache.channel_by_name("colors")
    stats = (C.c_uint64 * 4)()
    call("rt_render_plain", ctx.h, out.h, codes.h, width, height, cache.channel(0).h, cache.far_pointers().h,
         cache.root_node_index(), normals.h if normals else None, colors.h if colors else None,
         cache.nodes_in_block(), _fp(m), np.float32(lod), _fp(BACKGROUND), flags | (L.RT_FLAG_COUNT if count else 0),
         stats if count else None)
    return out, codes, (list(stats) if count else None)


class BatchRenderer:
    """Many views in one launch (rt_batch_*): one wave per tile, or persistent waves."""

    TILE_SHAPES = {0: (8, 8), 1: (16, 4), 2: (32, 2)}

    def __init__(self, cache: GPUOctreeCacheManager, max_views: int = 256):
        h = C.c_void_p()
        call("rt_batch_create", cache.h, max_views, C.byref(h))
        self.h, self.cache, self.ctx = h, cache, cache.ctx

    def config(self, tile_shape: int = 0, lds_levels: int = 12, wgs_per_cu: int = 0, tree_depth: int = 11):
        call("rt_batch_config", self.h, tile_shape, lds_levels, wgs_per_cu, tree_depth)

    def add_scene(self, cache: GPUOctreeCacheManager) -> int:
        """Registers another resident scene; returns its index for make_views."""
        idx = call("rt_batch_add_scene", self.h, cache.h)
        self._scenes = getattr(self, "_scenes", []) + [cache]  # keep the caches alive
        return idx

    @staticmethod
    def make_views(specs, flags: int = 0, lod: float = 0.0):
        """specs: list of (view_proj_inv, width, height[, scene]). Returns (ctypes array, total pixels, offsets)."""
        arr = (L.View * len(specs))()
        off, offsets = 0, []
        for i, spec in enumerate(specs):
            vpi, w, h = spec[:3]
            arr[i].view_proj_inv[:] = [float(x) for x in np.asarray(vpi, np.float32).reshape(16)]
            arr[i].width, arr[i].height = w, h
            arr[i].lod_voxel_size = lod
            arr[i].flags = flags
            arr[i].scene = spec[3] if len(spec) > 3 else 0
            arr[i].pixel_offset = off
            offsets.append(off)
            off += w * h
        return arr, off, offsets

    def render(self, views, nviews: int, out: Buffer, codes: Buffer | None = None, count: bool = False):
        stats = (C.c_uint64 * 4)()
        call("rt_batch_render", self.h, C.cast(views, C.c_void_p), nviews, out.h, codes.h if codes else None,
             stats if count else None)
        return list(stats) if count else None

    def last_ms(self) -> float:
        ms = C.c_double()
        call("rt_batch_last_ms", self.h, C.byref(ms))
        return ms.value

    def autotune(self, views, nviews: int, out: Buffer, codes: Buffer | None = None, repeats: int = 3):
        """Picks the fastest (tile shape, LDS levels, workgroups/CU); returns (config tuple, ms)."""
        best = (C.c_int * 3)()
        ms = C.c_double()
        call("rt_batch_autotune", self.h, C.cast(views, C.c_void_p), nviews, out.h, codes.h if codes else None,
             repeats, best, C.byref(ms))
        return (best[0], best[1], best[2]), ms.value

    def close(self):
        if self.h:
            lib().rt_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RenderServer:
    """Continuous-batching request loop (rt_server_*): FIFO admission by framebuffer-arena and block-pool
    memory, rounds of batched launches overlapping on several streams, p50/p99 latency."""

    def __init__(self, ctx: Context, max_views_per_round: int = 64, rounds_in_flight: int = 2,
                 framebuffer_bytes: int = 1 << 30, pool_bytes: int = 16 << 30, tile_shape: int = 0,
                 lds_levels: int = 12, wgs_per_cu: int = 0, max_block_loads: int = 0):
        cfg = L.ServerConfig(max_views_per_round, rounds_in_flight, framebuffer_bytes, pool_bytes, tile_shape,
                             lds_levels, wgs_per_cu, max_block_loads)
        h = C.c_void_p()
        call("rt_server_create", ctx.h, C.byref(cfg), C.byref(h))
        self.h, self.ctx = h, ctx
        self._trees = []
        self._live = {}  # request id -> (request struct, image array, codes array)

    def add_scene(self, tree: StoredOctree, depth: int = 0, max_blocks: int = 0) -> int:
        """max_blocks > 0 and below the tree's block count: the scene is served out of core from a pool
        of that many blocks (rt_server_add_scene_pool)."""
        idx = call("rt_server_add_scene_pool", self.h, tree.h, depth, max_blocks)
        self._trees.append(tree)
        return idx

    def submit(self, vpi, width: int, height: int, scene: int = 0, flags: int = 0, lod: float = 0.0,
               want_image: bool = True, want_codes: bool = False, pinned: bool = False) -> int:
        return self.submit_many([(vpi, width, height, scene)], flags, lod, want_image, want_codes, pinned)[0]

    def submit_many(self, specs, flags: int = 0, lod: float = 0.0, want_image: bool = True,
                    want_codes: bool = False, pinned: bool = False) -> list[int]:
        """specs: list of (view_proj_inv, width, height, scene), queued together. Returns request ids.
        pinned: results are delivered in the server's pinned memory (see poll) instead of being copied."""
        n = len(specs)
        arr = (L.Request * n)()
        bufs = []
        for i, (vpi, width, height, scene) in enumerate(specs):
            img = np.empty(width * height, np.uint32) if want_image and not pinned else None
            cod = np.empty(width * height, np.int32) if want_codes and not pinned else None
            r = arr[i]
            r.view_proj_inv[:] = [float(x) for x in np.asarray(vpi, np.float32).reshape(16)]
            r.width, r.height, r.lod_voxel_size, r.flags, r.scene = width, height, lod, flags, scene
            r.outputs = (L.RT_OUT_IMAGE if want_image else 0) | (L.RT_OUT_CODES if want_codes else 0)
            r.delivery = L.RT_DELIVER_PINNED if pinned else L.RT_DELIVER_COPY
            r.out_image = img.ctypes.data if img is not None else None
            r.out_codes = cod.ctypes.data if cod is not None else None
            bufs.append((img, cod, width * height))
        ids = (C.c_uint64 * n)()
        call("rt_server_submit_many", self.h, C.cast(arr, C.c_void_p), n, ids)
        for i in range(n):
            self._live[ids[i]] = bufs[i]
        return list(ids)

    def poll(self, cap: int = 256, timeout_ms: int = 0, keep_pinned: bool = False):
        """Returns a list of (completion struct, image or None, codes or None). Pinned results are copied
        out and released, unless keep_pinned: then image/codes are views of the server's pinned memory,
        valid until release(completion.id)."""
        arr = (L.Completion * cap)()
        n = call("rt_server_poll", self.h, C.cast(arr, C.c_void_p), cap, timeout_ms)
        out = []
        for i in range(n):
            c = arr[i]
            img, cod, npx = self._live.pop(c.id, (None, None, 0))
            if c.image or c.codes:
                img = np.ctypeslib.as_array(C.cast(c.image, C.POINTER(C.c_uint32)), (npx,)) if c.image else None
                cod = np.ctypeslib.as_array(C.cast(c.codes, C.POINTER(C.c_int32)), (npx,)) if c.codes else None
                if not keep_pinned:
                    img = img.copy() if img is not None else None
                    cod = cod.copy() if cod is not None else None
                    self.release(c.id)
            out.append((c, img, cod))
        return out

    def release(self, rid: int):
        call("rt_server_release", self.h, rid)

    def drain(self, timeout_ms: int = 60000):
        call("rt_server_drain", self.h, timeout_ms)

    def stats(self) -> L.ServerStats:
        s = L.ServerStats()
        call("rt_server_stats_get", self.h, C.byref(s))
        return s

    def reset_stats(self):
        call("rt_server_reset_stats", self.h)

    def close(self):
        if self.h:
            lib().rt_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

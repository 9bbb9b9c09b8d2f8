"""ORACLE - test infrastructure only.

ctypes binding of oracle/_build/liboracle.so, the strict-fp32 CPU restatement
of the reference renderer's hot path (see oracle.cpp for the reference files
each function follows). Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it, always as the checker or the timed CPU baseline,
never as part of the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(ORACLE_DIR, "_build", "liboracle.so")

P = C.c_void_p
I = C.c_int
F = C.c_float
FP = C.POINTER(C.c_float)
U32P = C.POINTER(C.c_uint32)
I32P = C.POINTER(C.c_int32)
U64P = C.POINTER(C.c_uint64)
I64P = C.POINTER(C.c_int64)
U8P = C.POINTER(C.c_uint8)

_SIG = {
    "orc_tree_open": (P, [C.c_char_p]),
    "orc_tree_close": (None, [P]),
    "orc_tree_info": (I, [P, U32P]),
    "orc_tree_load_block": (I, [P, C.c_uint32, U32P, U8P, C.c_uint64]),
    "orc_cache_create": (P, [P, I]),
    "orc_cache_destroy": (None, [P]),
    "orc_cache_initial_fill": (I, [P]),
    "orc_cache_root": (I, [P]),
    "orc_cache_loaded": (I, [P]),
    "orc_cache_copy": (C.c_int64, [P, I, P, C.c_uint64]),
    "orc_cache_block_slot": (I, [P, I]),
    "orc_cache_lru_order": (I, [P, I32P, I]),
    "orc_cache_counters": (None, [P, U64P]),
    "orc_cache_op": (I, [P, I, I]),
    "orc_render_plain": (I, [U32P, U32P, I, U32P, U32P, I, I, I, FP, F, FP, I, U32P, I32P, FP, I32P, U32P, U64P, I,
                             C.c_uint64, C.c_uint64]),
    "orc_init_tracing_frame": (I, [FP, I, I, FP, F]),
    "orc_cache_render_plain": (I, [P, I, I, FP, F, FP, I, U32P, I32P, U64P, I]),
    "orc_raytracing_pass": (I, [I32P, FP, I, I, U32P, U32P, I, U32P, U32P, I, I, I, C.c_uint64, C.c_uint64]),
    "orc_finish_tracing_frame": (I, [U32P, FP, I, I, FP, U32P, U32P]),
    "orc_ln_fixed": (None, [FP, FP, C.c_int64]),
    "orc_raytrace_depth": (I, [U32P, U32P, I, I, I, FP, F, U32P, I32P, I, C.c_uint64, C.c_uint64]),
    "orc_raytrace_restart": (I, [U32P, I, I, I, FP, F, U32P, I32P, I]),
    "orc_legacy_links": (I, [U32P, U32P, C.c_uint64, U32P]),
    "orc_node_to_block": (None, [I, I, I32P]),
    "orc_feedback_extract": (I, [I32P, I, I64P, I32P, I32P, I32P, I, I32P]),
    "orc_render_frame": (I, [P, I, I, FP, F, FP, I, I, I, U64P, C.POINTER(C.c_int), U32P, I32P, I32P, I, I32P, I,
                             U64P, I]),
    "orc_camera_vpi": (None, [F, F, F, FP, F, F, F, FP]),
    "orc_check_sphere_pool": (I, [U32P, U32P, U8P, U8P, I, F, F, F, F, I, I]),
    "orc_traverse_check": (I, [P, U64P]),
    "orc_check_constant_depth": (C.c_int64, [P, I, U64P]),
    "orc_check_sphere_tree": (I, [P, F, F, F, F, I]),
    "orc_tri_box": (I, [FP, FP, FP, FP, FP]),
    "orc_linear_at": (F, [FP, I, I, F, F]),
    "orc_import_obj": (C.c_int64, [C.c_char_p, I]),
    "orc_import_obj_fetch": (None, [U8P, U8P]),
    "orc_import_obj_error": (C.c_char_p, []),
}

BACKGROUND = np.array([0.0, np.float32(216 / 255.0), 1.0, 1.0], dtype=np.float32)
NO_HIT = 0x7FFFFFFF

_lib = None


def build() -> str:
    """Compiles the oracle with its Makefile (strict IEEE fp32)."""
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        l = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIG.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def _count(a) -> int:
    """Element count of an array argument (the traversal's bounds on the pool arrays)."""
    return 0 if a is None else int(np.asarray(a).size)


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


class Tree:
    def __init__(self, path: str):
        self.h = lib().orc_tree_open(path.encode())
        if not self.h:
            raise RuntimeError(f"oracle cannot open {path}")
        info = np.zeros(4, np.uint32)
        lib().orc_tree_info(self.h, _p(info, U32P))
        self.nodes_in_block, self.blocks, self.root_block, self.channels = (int(x) for x in info)

    def load_block(self, index: int):
        hdr = np.zeros(27, np.uint32)
        data = np.zeros(16 << 20, np.uint8)
        rc = lib().orc_tree_load_block(self.h, index, _p(hdr, U32P), _p(data, U8P), data.nbytes)
        if rc:
            raise RuntimeError(f"oracle load_block {index} failed: {rc}")
        return hdr, data

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_tree_close(self.h)
            self.h = None


class Cache:
    """GPUOctreeCacheManager restated over host arrays."""

    def __init__(self, tree: Tree, max_blocks: int):
        self.tree = tree
        self.max_blocks = max_blocks
        self.h = lib().orc_cache_create(tree.h, max_blocks)

    def initial_fill(self):
        if lib().orc_cache_initial_fill(self.h):
            raise RuntimeError("oracle initial fill failed")

    def root(self) -> int:
        return lib().orc_cache_root(self.h)

    def loaded(self) -> int:
        return lib().orc_cache_loaded(self.h)

    def _copy(self, which: int, dtype, count: int):
        out = np.zeros(count, dtype)
        n = lib().orc_cache_copy(self.h, which, out.ctypes.data, out.nbytes)
        if n < 0:
            raise RuntimeError("oracle cache copy failed")
        return out

    def links(self):
        return self._copy(0, np.uint32, self.max_blocks * self.tree.nodes_in_block)

    def far(self):
        return self._copy(1, np.uint32, self.max_blocks * 8)

    def channel_bytes(self, k: int, bytes_in_node: int):
        """Channel k (1-based as in the tree header) as raw bytes."""
        return self._copy(1 + k, np.uint8, self.max_blocks * self.tree.nodes_in_block * bytes_in_node)

    def block_slot(self, block: int) -> int:
        return lib().orc_cache_block_slot(self.h, block)

    def lru_order(self):
        n = lib().orc_cache_lru_order(self.h, None, 0)
        out = np.zeros(max(n, 1), np.int32)
        lib().orc_cache_lru_order(self.h, _p(out, I32P), n)
        return [int(x) for x in out[:n]]

    def counters(self):
        out = np.zeros(4, np.uint64)
        lib().orc_cache_counters(self.h, _p(out, U64P))
        return [int(x) for x in out]

    def op(self, op: int, arg: int = 0) -> int:
        rc = lib().orc_cache_op(self.h, op, arg)
        if rc < 0:
            raise RuntimeError(f"oracle cache op {op} failed")
        return rc

    StartRequestSession = lambda self: self.op(0)  # noqa: E731
    StartRequestTransaction = lambda self: self.op(1)  # noqa: E731
    MarkBlockAsUsed = lambda self, s: self.op(2, s)  # noqa: E731
    MarkParentAsUsed = lambda self, s: self.op(3, s)  # noqa: E731
    RequestBlock = lambda self, b: self.op(4, b)  # noqa: E731
    TransactionFilledCache = lambda self: bool(self.op(5))  # noqa: E731
    SessionFilledCache = lambda self: bool(self.op(6))  # noqa: E731
    EndRequestTransaction = lambda self: self.op(7)  # noqa: E731
    EndRequestSession = lambda self: self.op(8)  # noqa: E731

    def render_plain(self, width, height, vpi, lod=0.0, flags=0, threads=0):
        """render_plain over this cache's pool as it stands (block-level codes), without copying the pool."""
        out = np.zeros(width * height, np.uint32)
        codes = np.zeros(width * height, np.int32)
        stats = np.zeros(4, np.uint64)
        m = np.ascontiguousarray(vpi, np.float32).reshape(16)
        if lib().orc_cache_render_plain(self.h, width, height, _p(m, FP), lod, _p(BACKGROUND, FP), flags,
                                        _p(out, U32P), _p(codes, I32P), _p(stats, U64P), threads) < 0:
            raise RuntimeError("oracle render failed")
        return {"image": out.reshape(height, width), "codes": codes.reshape(height, width),
                "stats": [int(x) for x in stats]}

    def render_frame(self, width, height, vpi, lod=0.0, flags=0, max_passes=0, hits_to_extract=4000, seed=0,
                     expected_faults=1 << 16, threads=0, max_pass_faults=1 << 22):
        P_ = width * height
        out = np.zeros(P_, np.uint32)
        codes = np.zeros(P_, np.int32)
        counts = np.full(1024, -1, np.int32)
        faults = np.zeros(max_pass_faults, np.int32)
        stats = np.zeros(4, np.uint64)
        rng = C.c_uint64(seed)
        ef = C.c_int(expected_faults)
        m = np.ascontiguousarray(vpi, np.float32).reshape(16)
        n = lib().orc_render_frame(self.h, width, height, _p(m, FP), lod, _p(BACKGROUND, FP), flags, max_passes,
                                   hits_to_extract, C.byref(rng), C.byref(ef), _p(out, U32P), _p(codes, I32P),
                                   _p(counts, I32P), len(counts), _p(faults, I32P), len(faults), _p(stats, U64P),
                                   threads)
        if n < 0:
            raise RuntimeError("oracle render_frame failed")
        nfb = int(np.count_nonzero(counts >= 0))
        per_pass, k = [], 0
        for c in counts[:nfb]:
            per_pass.append([int(x) for x in faults[k:k + c]])
            k += int(c)
        return {"image": out.reshape(height, width), "codes": codes.reshape(height, width), "passes": n,
                "pass_faults": per_pass, "stats": stats, "rng": rng.value, "expected_faults": ef.value}

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_cache_destroy(self.h)
            self.h = None


def render_plain(links, far, root, normals, colors, nodes_in_block, width, height, vpi, lod=0.0, flags=0,
                 threads=0, want_hits=False):
    """Fused single-pass render. Returns dict(image, codes, stats[, hit_t, hit_node, fetches])."""
    P_ = width * height
    out = np.zeros(P_, np.uint32)
    codes = np.zeros(P_, np.int32)
    stats = np.zeros(4, np.uint64)
    hit_t = np.zeros(P_, np.float32) if want_hits else None
    hit_node = np.zeros(P_, np.int32) if want_hits else None
    fetches = np.zeros(P_, np.uint32) if want_hits else None
    m = np.ascontiguousarray(vpi, np.float32).reshape(16)
    lib().orc_render_plain(_p(links, U32P), _p(far, U32P), root, _p(normals, U32P), _p(colors, U32P),
                           nodes_in_block, width, height, _p(m, FP), lod, _p(BACKGROUND, FP), flags, _p(out, U32P),
                           _p(codes, I32P), _p(hit_t, FP), _p(hit_node, I32P), _p(fetches, U32P), _p(stats, U64P),
                           threads, _count(links), _count(far))
    d = {"image": out.reshape(height, width), "codes": codes.reshape(height, width), "stats": [int(x) for x in stats]}
    if want_hits:
        d.update(hit_t=hit_t.reshape(height, width), hit_node=hit_node.reshape(height, width),
                 fetches=fetches.reshape(height, width))
    return d


def init_tracing_frame(width, height, vpi, lod=0.0):
    st = np.zeros(width * height * 12, np.float32)
    m = np.ascontiguousarray(vpi, np.float32).reshape(16)
    lib().orc_init_tracing_frame(_p(st, FP), width, height, _p(m, FP), lod)
    return st


def raytracing_pass(states, width, height, links, far, root, normals, colors, nodes_in_block=0, flags=0, threads=0):
    codes = np.zeros(width * height, np.int32)
    lib().orc_raytracing_pass(_p(codes, I32P), _p(states, FP), width, height, _p(links, U32P), _p(far, U32P), root,
                              _p(normals, U32P), _p(colors, U32P), nodes_in_block, flags, threads, _count(links),
                              _count(far))
    return codes


def finish_tracing_frame(states, width, height, normals, colors):
    out = np.zeros(width * height, np.uint32)
    lib().orc_finish_tracing_frame(_p(out, U32P), _p(states, FP), width, height, _p(BACKGROUND, FP),
                                   _p(normals, U32P), _p(colors, U32P))
    return out


def raytrace_depth(links, far, root, width, height, vpi, lod=0.0, threads=0):
    out = np.zeros(width * height, np.uint32)
    hn = np.zeros(width * height, np.int32)
    m = np.ascontiguousarray(vpi, np.float32).reshape(16)
    lib().orc_raytrace_depth(_p(links, U32P), _p(far, U32P), root, width, height, _p(m, FP), lod, _p(out, U32P),
                             _p(hn, I32P), threads, _count(links), _count(far))
    return out, hn


def ln_fixed(x):
    """The fixed-order natural logarithm of the log-depth grey (float32 in, float32 out)."""
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros_like(x)
    lib().orc_ln_fixed(_p(x, FP), _p(out, FP), x.size)
    return out


def raytrace_restart(legacy_links, root, width, height, vpi, lod=0.0, threads=0):
    out = np.zeros(width * height, np.uint32)
    hn = np.zeros(width * height, np.int32)
    m = np.ascontiguousarray(vpi, np.float32).reshape(16)
    lib().orc_raytrace_restart(_p(legacy_links, U32P), root, width, height, _p(m, FP), lod, _p(out, U32P),
                               _p(hn, I32P), threads)
    return out, hn


def legacy_links(links, far):
    out = np.zeros(len(links), np.uint32)
    rc = lib().orc_legacy_links(_p(links, U32P), _p(far, U32P), len(links), _p(out, U32P))
    if rc:
        raise ValueError("pool needs more than 23-bit node indices")
    return out


def node_to_block(data: np.ndarray, block_size: int) -> np.ndarray:
    d = np.ascontiguousarray(data, np.int32).copy()
    lib().orc_node_to_block(len(d), block_size, _p(d, I32P))
    return d


def feedback_extract(codes, expected_faults=1 << 16, hits_to_extract=4000, seed=0):
    c = np.ascontiguousarray(codes, np.int32).reshape(-1)
    cap = len(c) + 1
    f, h, dp = np.zeros(cap, np.int32), np.zeros(cap, np.int32), np.zeros(cap, np.int32)
    counts = np.zeros(3, np.int32)
    state = np.array([expected_faults, hits_to_extract, np.int64(np.uint64(seed).view(np.int64)), 0, 0], np.int64)
    lib().orc_feedback_extract(_p(c, I32P), len(c), _p(state, I64P), _p(f, I32P), _p(h, I32P), _p(dp, I32P), cap,
                               _p(counts, I32P))
    return {"faults": list(f[:counts[0]]), "hits": list(h[:counts[1]]), "dups": list(dp[:counts[2]]),
            "expected_faults": int(state[0]), "rng": int(np.int64(state[2]).view(np.uint64)),
            "unique_count": int(state[4])}


def camera_vpi(yaw, pitch, roll, pos, fov, near, aspect):
    out = np.zeros(16, np.float32)
    p = np.array(pos, np.float32)
    lib().orc_camera_vpi(yaw, pitch, roll, _p(p, FP), fov, near, aspect, _p(out, FP))
    return out


def check_sphere_pool(links, far, ch1, ch2, root, center, radius, level, faultless=True) -> int:
    return lib().orc_check_sphere_pool(_p(links, U32P), _p(far, U32P), _p(ch1, U8P), _p(ch2, U8P), root,
                                       float(center[0]), float(center[1]), float(center[2]), float(radius), level,
                                       int(faultless))


def traverse_check(tree: Tree):
    """StoredOctreeTraverser + CheckTree over a .tree file: (error code, nodes reachable from the root)."""
    n = C.c_uint64()
    rc = lib().orc_traverse_check(tree.h, C.byref(n))
    return rc, n.value


def check_constant_depth(tree: Tree, expected_depth: int):
    """DebugCheckConstantDepthTree: (depth errors, or minus the traverser's error; nodes checked)."""
    n = C.c_uint64()
    rc = lib().orc_check_constant_depth(tree.h, expected_depth, C.byref(n))
    return rc, n.value


def check_sphere_tree(tree: Tree, center, radius, level) -> int:
    """CheckTestOctreeSphereWithLoader over a .tree file; 0 = pass."""
    return lib().orc_check_sphere_tree(tree.h, float(center[0]), float(center[1]), float(center[2]), float(radius),
                                       level)


def tri_box(box_min, box_max, a, b, c) -> bool:
    """TriangleBoxIntersection (util/intersections.cpp:29-59) on float32 inputs."""
    arrs = [np.ascontiguousarray(x, np.float32) for x in (box_min, box_max, a, b, c)]
    return bool(lib().orc_tri_box(*[_p(x, FP) for x in arrs]))


def linear_at(img, fx, fy) -> float:
    """CImg 1.4.9 _linear_atXY on a one-channel float image of shape (h, w)."""
    im = np.ascontiguousarray(img, np.float32)
    return float(lib().orc_linear_at(_p(im, FP), im.shape[1], im.shape[0], float(fx), float(fy)))


class ObjImportError(RuntimeError):
    pass


def import_obj(path: str, max_level: int):
    """ImportObjScene's node tree: (masks uint8[n], data uint8[n, 8]) in BFS order, children in bit order."""
    n = lib().orc_import_obj(path.encode(), max_level)
    if n < 0:
        raise ObjImportError(lib().orc_import_obj_error().decode())
    masks = np.zeros(n, np.uint8)
    data = np.zeros((n, 8), np.uint8)
    lib().orc_import_obj_fetch(_p(masks, U8P), _p(data, U8P))
    return masks, data

"""Image dumps in place of the reference's on-screen display path
(CL util/rendered-image-drawer.cpp:23-28 and cl-texture-2d-reference.cpp:62-74
blit `out_image` into a GL texture). Frames, hit-id maps and feedback-code
maps are written as PNG or PPM files instead, for visual checks.

`out_image` words are RGBA8 with R in the low byte (uchar4 in memory).
"""
from __future__ import annotations

import struct
import zlib

import numpy as np

NO_HIT = 0x7FFFFFFF


def rgba_bytes(image: np.ndarray) -> np.ndarray:
    """(H, W) uint32 RGBA8 words -> (H, W, 4) uint8."""
    img = np.ascontiguousarray(image, np.uint32)
    return img.view(np.uint8).reshape(img.shape[0], img.shape[1], 4)


def write_png(path: str, image: np.ndarray) -> None:
    """Writes an (H, W) RGBA8 word image as an 8-bit RGBA PNG (top row first)."""
    px = rgba_bytes(image)
    h, w = px.shape[:2]
    raw = b"".join(b"\x00" + px[y].tobytes() for y in range(h))

    def chunk(tag: bytes, data: bytes) -> bytes:
        return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)

    with open(path, "wb") as f:
        f.write(b"\x89PNG\r\n\x1a\n")
        f.write(chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0)))
        f.write(chunk(b"IDAT", zlib.compress(raw, 6)))
        f.write(chunk(b"IEND", b""))


def write_ppm(path: str, image: np.ndarray) -> None:
    """Writes an (H, W) RGBA8 word image as a binary PPM (alpha dropped)."""
    px = rgba_bytes(image)
    h, w = px.shape[:2]
    with open(path, "wb") as f:
        f.write(f"P6\n{w} {h}\n255\n".encode())
        f.write(np.ascontiguousarray(px[:, :, :3]).tobytes())


def id_colors(ids: np.ndarray, none_value: int = -1) -> np.ndarray:
    """False-colour map of integer ids (hit nodes, blocks): a hashed colour per id, black where
    `ids == none_value`. Returns (H, W) RGBA8 words."""
    v = np.asarray(ids).astype(np.uint64)
    h = (v * np.uint64(0x9E3779B97F4A7C15)) >> np.uint64(40)
    rgb = (h & np.uint64(0xFFFFFF)).astype(np.uint32) | np.uint32(0xFF000000)
    return np.where(np.asarray(ids) == none_value, np.uint32(0xFF000000), rgb).astype(np.uint32)


def code_colors(codes: np.ndarray) -> np.ndarray:
    """Feedback codes (rendering data format: -block-1 for a fault, (unit+1)<<1|dup for a hit, NO_HIT):
    faults red, duplicate hits yellow-tinted, hits by unit colour, NO_HIT black. Returns RGBA8 words."""
    c = np.asarray(codes, np.int64)
    out = id_colors(np.where(c > 0, c >> 1, -1))
    out = np.where(c < 0, np.uint32(0xFF0000FF), out)
    out = np.where((c > 0) & (c != NO_HIT) & ((c & 1) == 1), out | np.uint32(0x0000C0C0), out)
    return np.where(c == NO_HIT, np.uint32(0xFF000000), out).astype(np.uint32)

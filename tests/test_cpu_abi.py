"""The ctypes mirror of every struct in include/rayt_hip.h has the C compiler's
size and field offsets (a mismatch would silently scramble arguments)."""
import ctypes as C
import os
import subprocess

import pytest

from rayt_amd import _lib as L

STRUCTS = {
    "rt_tree_info": L.TreeInfo,
    "rt_terrain_params": L.TerrainParams,
    "rt_view": L.View,
    "rt_server_config": L.ServerConfig,
    "rt_request": L.Request,
    "rt_completion": L.Completion,
    "rt_server_stats": L.ServerStats,
}


def test_struct_layouts_match_c(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{L.HEADER_PATH}"', "int main(void) {"]
    for cname, py in STRUCTS.items():
        lines.append(f'  printf("{cname} sizeof %zu\\n", sizeof({cname}));')
        for field, _ in py._fields_:
            lines.append(f'  printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    r = subprocess.run(["gcc", "-std=c11", "-o", str(exe), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = {}
    for ln in out:
        if ln.strip():
            cname, field, val = ln.split()
            got[(cname, field)] = int(val)
    for cname, py in STRUCTS.items():
        assert got[(cname, "sizeof")] == C.sizeof(py), cname
        for field, _ in py._fields_:
            assert got[(cname, field)] == getattr(py, field).offset, (cname, field)


def test_image_dumps_round_trip(tmp_path):
    """PNG / PPM dumps of out_image (the replacement of the GL display path) decode to the same pixels."""
    import zlib
    import numpy as np
    from rayt_amd import imageio
    rng = np.random.default_rng(0)
    img = rng.integers(0, 2**32, (5, 7), dtype=np.uint64).astype(np.uint32)
    p = tmp_path / "a.png"
    imageio.write_png(str(p), img)
    data = p.read_bytes()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    ln = int.from_bytes(data[33:37], "big")
    assert data[37:41] == b"IDAT"
    raw = zlib.decompress(data[41:41 + ln])
    rows = np.frombuffer(raw, np.uint8).reshape(5, 1 + 7 * 4)
    assert (rows[:, 0] == 0).all()
    np.testing.assert_array_equal(rows[:, 1:].reshape(5, 7, 4), imageio.rgba_bytes(img))
    q = tmp_path / "a.ppm"
    imageio.write_ppm(str(q), img)
    body = q.read_bytes().split(b"\n", 3)[3]
    np.testing.assert_array_equal(np.frombuffer(body, np.uint8).reshape(5, 7, 3), imageio.rgba_bytes(img)[:, :, :3])
    codes = np.array([[-1, 0x7FFFFFFF, (3 << 1) | 1, 3 << 1]], np.int32)
    cc = imageio.code_colors(codes)
    assert cc[0, 0] == 0xFF0000FF and cc[0, 1] == 0xFF000000 and cc[0, 2] != cc[0, 3]


def test_missing_library_fails_loudly():
    """Without the HIP library every entry point raises; nothing falls back to another implementation."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = ("from rayt_amd import api\n"
            "try:\n"
            "    api.Context()\n"
            "except RuntimeError as e:\n"
            "    print('raised:', e)\n")
    env = dict(os.environ, RAYT_HIP_LIB=os.path.join(root, "_no_such_dir", "librayt_hip.so"))
    r = subprocess.run([sys.executable, "-c", code], cwd=root, env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "raised:" in r.stdout and "is not built" in r.stdout

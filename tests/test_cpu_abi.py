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

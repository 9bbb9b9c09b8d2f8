"""Builds tests/cl_util_shim_main.cpp (the CL util shims driven as the
reference's callers drive CL util) against include/rayt_cl_util.h and
librayt_hip.so, into tests/_bin/. Run by __graft_entry__.build(); the GPU test
runs the binary."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cl_util_shim_main.cpp")
BIN = os.path.join(ROOT, "tests", "_bin", "cl_util_shim_test")


def compile_command(out: str, syntax_only: bool = False):
    cmd = ["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I" + os.path.join(ROOT, "include"), SRC]
    if syntax_only:
        return cmd + ["-fsyntax-only"]
    return cmd + ["-L" + os.path.join(ROOT, "rayt_amd"), "-lrayt_hip", "-Wl,-rpath,$ORIGIN/../../rayt_amd", "-o", out]


def build() -> str:
    os.makedirs(os.path.dirname(BIN), exist_ok=True)
    r = subprocess.run(compile_command(BIN), capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"building {SRC} failed\n{r.stdout}\n{r.stderr}")
    return BIN


if __name__ == "__main__":
    print(build())

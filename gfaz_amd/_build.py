"""Build the in-tree native libraries (``python -m gfaz_amd._build``).

* ``_lib/libgfaz_amd.so``  - HIP kernels for gfx950 plus the host C++ runtime and the
  C ABI of ``include/gfaz_amd.h`` (hipcc, -O3, packed-FP32 ops disabled).
* ``_lib/libgfaz_synth.so`` - the host-only synthetic pangenome generator (g++/OpenMP).

Each call compiles from scratch into a temporary name and renames on success, so a
failed build never leaves a half-written library behind.
"""
from __future__ import annotations

import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "_lib")
OBJ_DIR = os.path.join(LIB_DIR, "obj")
INCLUDE = os.path.join(ROOT, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX_HOST", "g++")
ZSTD_INC = "/opt/conda/include"
ZSTD_LIB = "/usr/lib/x86_64-linux-gnu/libzstd.so.1"

HIP_FLAGS = [
    "-std=c++17", "-O3", "-fPIC", "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]
HOST_FLAGS = ["-std=c++17", "-O3", "-fPIC", "-fopenmp", "-Wall", "-Wextra",
              "-Wno-unused-parameter"]


def _run(cmd: list[str]) -> None:
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def _jobs(cmds: list[list[str]]) -> None:
    procs = []
    for cmd in cmds:
        print(" ".join(cmd), flush=True)
        procs.append((cmd, subprocess.Popen(cmd)))
    failed = [cmd for cmd, p in procs if p.wait() != 0]
    if failed:
        raise subprocess.CalledProcessError(1, failed[0])


def build_synth() -> str:
    os.makedirs(LIB_DIR, exist_ok=True)
    out = os.path.join(LIB_DIR, "libgfaz_synth.so")
    tmp = out + ".tmp"
    _run([CXX, *HOST_FLAGS, "-shared", "-o", tmp, os.path.join(CSRC, "synth.cpp")])
    os.replace(tmp, out)
    return out


def build_codec() -> str:
    os.makedirs(OBJ_DIR, exist_ok=True)
    hip_srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    host_srcs = sorted(glob.glob(os.path.join(CSRC, "host", "*.cpp")))
    objs, cmds = [], []
    inc = [f"-I{INCLUDE}", f"-I{CSRC}", f"-I{ZSTD_INC}"]
    for src in hip_srcs:
        obj = os.path.join(OBJ_DIR, os.path.basename(src) + ".o")
        cmds.append([HIPCC, *HIP_FLAGS, *inc, "-c", src, "-o", obj])
        objs.append(obj)
    for src in host_srcs:
        obj = os.path.join(OBJ_DIR, "host_" + os.path.basename(src) + ".o")
        cmds.append([CXX, *HOST_FLAGS, *inc, "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__",
                     "-c", src, "-o", obj])
        objs.append(obj)
    _jobs(cmds)
    out = os.path.join(LIB_DIR, "libgfaz_amd.so")
    tmp = out + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-fopenmp", "-o", tmp, *objs,
          ZSTD_LIB, "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, out)
    shutil.rmtree(OBJ_DIR, ignore_errors=True)
    return out


def build_all() -> None:
    build_synth()
    build_codec()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    if "all" in what:
        build_all()
    else:
        if "synth" in what:
            build_synth()
        if "codec" in what:
            build_codec()

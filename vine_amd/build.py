"""Build recipe for libvine.so (HIP kernels for gfx950 + host C++ + C ABI).

Compiles every source under vine_amd/csrc in-tree with hipcc (device code for
gfx950 only) and g++ (host-only files), then links vine_amd/libvine.so.
Usage: python -m vine_amd.build   (or vine_amd.build.build()).
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(ROOT, "_build", "obj")
OUT = os.path.join(HERE, "libvine.so")

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

COMMON = ["-O3", "-fPIC", "-std=c++17", "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
HIP_FLAGS = COMMON + [
    "--offload-arch=" + ARCH,
    # Packed fp32 instructions are kept out of the device code.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
HOST_FLAGS = COMMON + ["-Wall", "-Wno-unused-function", "-pthread"]


def sources():
    out = []
    for name in sorted(os.listdir(CSRC)):
        path = os.path.join(CSRC, name)
        if name.endswith(".hip"):
            out.append((path, "hip"))
        elif name.endswith(".cpp"):
            out.append((path, "host"))
    return out


def _compile(item):
    path, kind = item
    obj = os.path.join(OBJ, os.path.basename(path) + ".o")
    if kind == "hip":
        cmd = [HIPCC] + HIP_FLAGS + ["-c", path, "-o", obj]
    else:
        cmd = [CXX] + HOST_FLAGS + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs + ["-pthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    shutil.move(tmp, OUT)
    # the C++ facade (include/vine/cspp.hpp) compiled by a plain host compiler
    # against the library, as a user program would be
    facade = os.path.join(HERE, "vine_facade_check")
    cmd = [CXX, "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"), "-o", facade,
           os.path.join(ROOT, "tests", "cpp", "facade_main.cpp"), "-L" + HERE, "-lvine", "-Wl,-rpath,$ORIGIN"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("facade build failed: %s\n%s" % (" ".join(cmd), r.stderr))
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)

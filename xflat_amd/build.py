"""Builds the in-tree HIP library xflat_amd/libxflat.so for gfx950 and the
host programs over its C-ABI (xflat_amd/bin/).

Every translation unit under xflat_amd/csrc is compiled by hipcc in
parallel (object files under xflat_amd/_build/), then linked into one shared
library.  Packed FP32 instructions are disabled for all device code.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libxflat.so")
HOST = os.path.join(HERE, "host")
BIN = os.path.join(HERE, "bin")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
ROCM_LIB = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib")
ARCH = os.environ.get("XFLAT_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
DEV_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wno-unused-command-line-argument"]


def sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith((".hip", ".cpp")))


def _compile(src: str) -> tuple[str, str]:
    obj = os.path.join(OBJ, src + ".o")
    cmd = [HIPCC] + COMMON + DEV_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stderr[-4000:]}")
    return obj, r.stderr


def build(jobs: int | None = None, verbose: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = []
        for obj, err in ex.map(_compile, srcs):
            objs.append(obj)
            if verbose and err.strip():
                print(err, file=sys.stderr)
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    build_host()
    return LIB


def build_host() -> list[str]:
    """Host programs over the C-ABI (xflat_amd/host/*.cpp -> xflat_amd/bin/)."""
    os.makedirs(BIN, exist_ok=True)
    out = []
    cxx = shutil.which("g++") or "c++"
    for src in sorted(f for f in os.listdir(HOST) if f.endswith(".cpp")):
        exe = os.path.join(BIN, src[:-4])
        cmd = [cxx, "-O2", "-std=c++17", "-Wall", "-I", INCLUDE, os.path.join(HOST, src), "-o", exe + ".tmp",
               "-L", HERE, "-lxflat", "-Wl,-rpath,$ORIGIN/..", f"-Wl,-rpath-link,{ROCM_LIB}"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"host build failed on {src}:\n{r.stderr[-4000:]}")
        os.replace(exe + ".tmp", exe)
        out.append(exe)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))

"""Build libttc.so (gfx950) in-tree with hipcc.

    python -m thecodecombo_7_amd.build        # or build() from Python

The library is plain HIP C++ with a C ABI (include/ttc.h); nothing here
depends on torch.  Objects are compiled in parallel, then linked.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
SOURCES = ["ttc_api.hip", "ttc_nblist.hip", "ttc_pair.hip", "ttc_pme.hip", "ttc_md.hip", "ttc_mc.hip",
           "ttc_batch.hip", "ttc_angles.hip"]
# host C inside the same library (no GPU code): the reference-format readers
# and the .trr writer
HOST_SOURCES = ["host/ttc_io.c"]
LIB = os.path.join(HERE, "libttc.so")
CC = os.environ.get("CC", "gcc")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    # keep hipcc from forming packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unknown-pragmas", "-Wno-unused-value",
    "-I" + os.path.join(ROOT, "include"),
]


def _obj(src: str) -> str:
    return os.path.join(CSRC, "_build", src.replace(".hip", ".o"))


def _compile(src: str) -> str:
    out = _obj(src)
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return out


def _compile_host(src: str) -> str:
    out = os.path.join(CSRC, "_build", os.path.basename(src).replace(".c", ".o"))
    cmd = [CC, "-O2", "-fPIC", "-std=gnu99", "-Wall", "-Wextra", "-I" + os.path.join(ROOT, "include"),
           "-c", os.path.join(HERE, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{CC} failed for {src}:\n{r.stderr[-6000:]}")
    return out


def build(verbose: bool = False) -> str:
    os.makedirs(os.path.join(CSRC, "_build"), exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]
    with ThreadPoolExecutor(max_workers=min(6, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    objs += [_compile_host(s) for s in HOST_SOURCES]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)

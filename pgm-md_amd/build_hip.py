"""Builds the HIP engine for gfx950 into pgm-md_amd/lib/libpgm_hip.so (C ABI: include/pgm_hip.h).

    python pgm-md_amd/build_hip.py [--jobs N]

hipcc cross-compiles without a GPU.  Packed FP32 instructions are disabled for all device code
(-Xclang -target-feature -Xclang -packed-fp32-ops).
"""

from __future__ import annotations

import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "lib")
LIB = os.path.join(LIB_DIR, "libpgm_hip.so")
EXPORTS_MAP = os.path.join(CSRC, "exports.map")
SOURCES = ["nblist.hip", "pairs.hip", "pme.hip", "solver.hip", "adjoint.hip", "devmath.hip", "engine.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Werror=shadow", "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]  # fmt: skip


def _compile(src: str, obj_dir: str) -> str:
    obj = os.path.join(obj_dir, src.replace(".hip", ".o"))
    cmd = [HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    noise = "is not a recognized feature for this target"
    err = "\n".join(ln for ln in (r.stdout + r.stderr).splitlines() if noise not in ln)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{err}")
    if err.strip():
        print(err, file=sys.stderr)
    return obj


def build(jobs: int = 5, verbose: bool = True) -> str:
    obj_dir = os.path.join(HERE, "_build")
    os.makedirs(obj_dir, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, obj_dir), SOURCES))
    tmp = LIB + ".tmp"
    # export the C ABI only: no C++ symbol of the library can interpose on (or be interposed by) a caller's
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, f"-Wl,--version-script={EXPORTS_MAP}", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=5)
    a = ap.parse_args()
    build(a.jobs)

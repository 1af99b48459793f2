"""Build libccgpu.so (all HIP sources of dalton_amd/csrc) for gfx950, in-tree.

    python -m dalton_amd.build

Every source is compiled to its own object in parallel (each translation
unit carries its own device code), then the objects are linked into the
shared library.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libccgpu.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed FP32 VALU ops are kept out of every code object
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def sources() -> list[str]:
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")))


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libccgpu.so: " + " ".join(cmd[-3:]))


def build(verbose: bool = True) -> str:
    inc = ["-I", CSRC, "-I", os.path.join(HERE, "..", "include")]
    jobs = max(1, min(len(sources()), len(os.sched_getaffinity(0))))
    with tempfile.TemporaryDirectory(prefix="ccgpu_build_") as tmpdir:
        objs = []
        cmds = []
        for src in sources():
            obj = os.path.join(tmpdir, os.path.basename(src) + ".o")
            objs.append(obj)
            cmds.append([HIPCC, *FLAGS, *inc, "-c", src, "-o", obj])
        with ThreadPoolExecutor(jobs) as ex:
            list(ex.map(lambda c: _run(c, verbose), cmds))
        tmp = LIB + ".tmp"
        _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp], verbose)
        os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()

"""Build ``nefi_amd/_lib/libnefi_wave.so`` from ``nefi_amd/csrc`` with hipcc for gfx950.

``python -m nefi_amd.build`` compiles in-tree; the library travels with the source tree, so a
GPU machine runs what was built here without a compiler step. Each source compiles to its own
object in parallel (every kernel and its launcher live in one translation unit, so no
relocatable device code is needed), then the objects are linked into the shared library.
"""

from __future__ import annotations

import os
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
SRC = PKG / "csrc"
OUT = PKG / "_lib" / "libnefi_wave.so"
SOURCES = ["nefi_wave.hip", "nefi_wave_2d.hip", "nefi_wave_2d64.hip", "nefi_filter.hip",
           "nefi_stepper.hip"]

COMPILE_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed-FP32 instructions are kept out of the device code (some of their forms have given wrong
    # results on gfx950 without an error)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
]
LINK_FLAGS = ["--offload-arch=gfx950", "-shared", "-fPIC"]


def hipcc() -> str:
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    cand = Path(rocm) / "bin" / "hipcc"
    return str(cand) if cand.exists() else "hipcc"


def _run(cmd: list[str]) -> subprocess.CompletedProcess:
    return subprocess.run(cmd, capture_output=True, text=True)


def build_library(verbose: bool = False, jobs: int | None = None) -> Path:
    OUT.parent.mkdir(parents=True, exist_ok=True)
    jobs = jobs or max(1, min(len(SOURCES), os.cpu_count() or 1))
    with tempfile.TemporaryDirectory(prefix="nefi_amd_build_") as tmpdir:
        objs = [Path(tmpdir) / (Path(s).stem + ".o") for s in SOURCES]
        cmds = [[hipcc(), *COMPILE_FLAGS, f"-I{ROOT / 'include'}", f"-I{SRC}", "-c",
                 str(SRC / s), "-o", str(o)] for s, o in zip(SOURCES, objs)]
        with ThreadPoolExecutor(jobs) as pool:
            results = list(pool.map(_run, cmds))
        for cmd, res in zip(cmds, results):
            if res.returncode != 0:
                sys.stderr.write(res.stdout + res.stderr)
                raise RuntimeError(f"hipcc failed ({res.returncode}): {' '.join(cmd)}")
            if verbose:
                sys.stderr.write(res.stderr)
        tmp = OUT.with_suffix(".so.tmp")
        link = [hipcc(), *LINK_FLAGS, *map(str, objs), "-o", str(tmp)]
        res = _run(link)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError(f"hipcc link failed ({res.returncode}): {' '.join(link)}")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build_library(verbose=True))

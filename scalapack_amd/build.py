"""Build the native libraries in-tree (hipcc for gfx950, g++ for host-only code).

    python -m scalapack_amd.build        # or scalapack_amd.build.build()

Products (next to this file, git-ignored, shipped with the source tree):
    libsla_blacs1x1.so   single-process 1 x 1 BLACS grid + TOOLS index helpers
    libscalapack_amd.so  the HIP kernels and the C-ABI of include/scalapack_amd.h
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
BUILD = PKG / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SLA_OFFLOAD_ARCH", "gfx950")

# (source, object stem, extra flags): the panel leaf's two storage types are
# separate objects of one source so that they compile in parallel
HIP_SOURCES = [(f, f[:-4], []) for f in
               ["gemm_f64.hip", "gemm_lowp.hip", "gemm_small.hip", "panel_tslu.hip", "swap.hip",
                "trsm.hip", "lu.hip", "mixed.hip", "aux.hip", "batched.hip", "mlp.hip", "capi.hip"]]
HIP_SOURCES += [("panel.hip", "panel_f64", ["-DSLA_PANEL_TYPE=64"]),
                ("panel.hip", "panel_f32", ["-DSLA_PANEL_TYPE=32"])]

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    # Packed FP32 instructions are kept out of every kernel of this library.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # Fully unroll the register-resident panel loops (their arrays must not
    # fall back to scratch memory).
    "-mllvm", "-pragma-unroll-threshold=1000000",
    f"-I{INCLUDE}",
    f"-I{CSRC}",
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")


def _hip_object(entry, incremental: bool = False) -> Path:
    src, stem, extra = entry
    obj = BUILD / (stem + ".o")
    if incremental and obj.exists():
        deps = [CSRC / src, *CSRC.glob("*.h"), *INCLUDE.glob("*.h")]
        if all(d.stat().st_mtime <= obj.stat().st_mtime for d in deps):
            return obj
    _run([HIPCC, *HIP_FLAGS, *extra, "-c", str(CSRC / src), "-o", str(obj)])
    return obj


def build(jobs: int | None = None, incremental: bool = False) -> None:
    """Compile every native component; raises on any compiler error.
    incremental: keep object files newer than their source and every header."""
    BUILD.mkdir(exist_ok=True)
    shim = PKG / "libsla_blacs1x1.so"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", f"-I{INCLUDE}",
          str(CSRC / "blacs1x1.cpp"), "-o", str(shim)])
    jobs = jobs or min(len(HIP_SOURCES), os.cpu_count() or 1)
    for stale in BUILD.glob("*.o"):  # objects of sources that no longer exist
        if stale.stem not in {e[1] for e in HIP_SOURCES}:
            stale.unlink()
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda src: _hip_object(src, incremental), HIP_SOURCES))
    lib = PKG / "libscalapack_amd.so"
    tmp = lib.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp),
          f"-L{PKG}", "-lsla_blacs1x1", "-Wl,-rpath,$ORIGIN", "-ldl"])
    os.replace(tmp, lib)


def clean() -> None:
    shutil.rmtree(BUILD, ignore_errors=True)
    for name in ("libscalapack_amd.so", "libsla_blacs1x1.so"):
        (PKG / name).unlink(missing_ok=True)


if __name__ == "__main__":
    build()
    print("built", PKG / "libscalapack_amd.so")

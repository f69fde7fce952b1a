"""Builds librusterize_hip.so (gfx950) in-tree.

Run as ``python -m rusterize_amd.build`` or through ``__graft_entry__.build()``.
The host parts are compiled with g++ and the device runtime + kernels with
hipcc for gfx950 only; every f64 expression is compiled with
-ffp-contract=off so no a*b+c is fused into an FMA (the reference's Rust code
never contracts), which keeps crossings and pixel-centre rounding exact.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "librusterize_hip.so"
ROOT = PKG.parent
INCLUDE = ROOT / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")

HOST_FLAGS = ["-std=c++17", "-O2", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall",
              "-Wextra", f"-I{INCLUDE}"]
# -packed-fp32-ops: keep the compiler from emitting packed FP32 instructions.
HIP_FLAGS = ["--offload-arch=gfx950", "-std=c++17", "-O3", "-fPIC", "-ffp-contract=off",
             "-fno-fast-math", "-fno-gpu-rdc", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-Wall", f"-I{INCLUDE}"]


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(str(c) for c in cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (res.stdout + res.stderr).splitlines()
                    if "packed-fp32-ops" not in l and "is not a recognized feature" not in l)
    if res.returncode != 0:
        raise RuntimeError(f"command failed ({res.returncode}): {' '.join(map(str, cmd))}\n{out}")
    if verbose and out.strip():
        print(out, flush=True)


# The tile kernel templates are instantiated once per output dtype, each in a
# translation unit of its own (rz_tile_inst.hip), so they compile in parallel.
TILE_DTYPES = [("u8", "uint8_t"), ("u16", "uint16_t"), ("u32", "uint32_t"), ("u64", "uint64_t"),
               ("i8", "int8_t"), ("i16", "int16_t"), ("i32", "int32_t"), ("i64", "int64_t"),
               ("f32", "float"), ("f64", "double")]


def build_library(verbose: bool = False) -> Path:
    """Compiles and links librusterize_hip.so; returns its path."""
    BUILD.mkdir(exist_ok=True)
    host_obj = BUILD / "rz_host.o"
    dev_obj = BUILD / "rz_runtime.o"
    tile_objs = [BUILD / f"rz_tile_{name}.o" for name, _ in TILE_DTYPES]
    jobs = [
        [HIPCC, *HIP_FLAGS, "-c", str(CSRC / "rz_runtime.hip"), "-o", str(dev_obj)],
        [CXX, *HOST_FLAGS, "-c", str(CSRC / "rz_host.cpp"), "-o", str(host_obj)],
    ] + [
        [HIPCC, *HIP_FLAGS, f"-DRZ_TILE_T={ctype}", f"-DRZ_TILE_NAME={name}", "-c",
         str(CSRC / "rz_tile_inst.hip"), "-o", str(obj)]
        for (name, ctype), obj in zip(TILE_DTYPES, tile_objs)
    ]
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    with ThreadPoolExecutor(max_workers=workers) as ex:
        for f in [ex.submit(_run, j, verbose) for j in jobs]:
            f.result()
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), str(host_obj),
          str(dev_obj), *map(str, tile_objs)], verbose)
    os.replace(tmp, LIB)
    return LIB


def build_oracle(verbose: bool = False) -> Path:
    """Builds the CPU oracle (test infrastructure) with its own Makefile."""
    oracle = ROOT / "oracle"
    _run(["make", "-s", "-C", str(oracle), "-B", "build/liboracle.so"], verbose)
    return oracle / "build" / "liboracle.so"


def main() -> int:
    verbose = "-v" in sys.argv
    if shutil.which(HIPCC) is None and not Path(HIPCC).exists():
        print("hipcc not found", file=sys.stderr)
        return 1
    print(build_library(verbose))
    print(build_oracle(verbose))
    return 0


if __name__ == "__main__":
    sys.exit(main())

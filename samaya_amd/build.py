"""Builds the native library samaya_amd/_lib/libsamaya_amd.so with hipcc for gfx950.

Every source is compiled in-tree (objects under samaya_amd/_build/), in parallel, then linked
into one shared library that exports the C ABI of include/samaya_pdhg.h, include/samaya_pdlp.h,
include/samaya_io.h and include/samaya_mip.h. Flags:
  --offload-arch=gfx950              CDNA4 only
  -ffp-contract=off                  no FMA contraction: the double-double verifier needs a*b+c
                                     rounded twice, and the FP64 step arithmetic matches the
                                     reference's operation by operation
  -Xclang -target-feature -Xclang -packed-fp32-ops
                                     no packed FP32 instructions in device code
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
SRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB_DIR = PKG / "_lib"
LIB = LIB_DIR / "libsamaya_amd.so"

SOURCES = [
    SRC / "gpu" / "pdhg_device.hip",
    SRC / "gpu" / "pdhg_shared.hip",
    SRC / "gpu" / "spmv.hip",
    SRC / "gpu" / "spmv_noalloc.hip",
    SRC / "host" / "lp_data.cpp",
    SRC / "host" / "pdlp.cpp",
    SRC / "host" / "verify.cpp",
    SRC / "host" / "mip.cpp",
    SRC / "host" / "generators.cpp",
    SRC / "host" / "capi.cpp",
    SRC / "host" / "mps.cpp",
    SRC / "host" / "capi_io.cpp",
]

COMMON = [
    "-O3", "-std=c++20", "-fPIC", "-pthread", "-ffp-contract=off", "-Wall", "-Wno-unused-result",
    f"-I{ROOT / 'include'}",
]
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _compile(src: Path) -> Path:
    obj = BUILD / (src.parent.name + "_" + src.stem + ".o")
    cmd = [hipcc(), *COMMON]
    if src.suffix == ".hip":
        cmd += ["-x", "hip", *HIP_FLAGS]
    else:
        cmd += ["-x", "c++"]
    cmd += ["-c", str(src), "-o", str(obj)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{proc.stdout}\n{proc.stderr}")
    return obj


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    LIB_DIR.mkdir(exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as pool:
        objs = list(pool.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), "-shared", "--offload-arch=gfx950", "-fPIC", "-pthread", "-Wl,--no-undefined", *map(str, objs),
           "-o", str(tmp)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{proc.stdout}\n{proc.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)

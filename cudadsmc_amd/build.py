"""Build the engine's shared library (HIP kernels for gfx950 + C ABI).

Output: cudadsmc_amd/lib/libdsmc_hip.so, built in-tree so it travels with the
source tree. Host C++ goes through g++, HIP sources through hipcc; both with
-ffp-contract=off (the bit-exact FP policy shared with the CPU oracle).
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
BUILD = PKG / "_build"
LIB = PKG / "lib" / "libdsmc_hip.so"

HOST_SOURCES = ["host/settings.cpp", "host/geometry.cpp", "host/mt_table.cpp", "host/capi_host.cpp"]
HIP_SOURCES = ["kernels/step.hip", "kernels/no.hip", "solver.hip"]

ARCH = os.environ.get("DSMC_OFFLOAD_ARCH", "gfx950")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", f"-I{ROOT / 'include'}", f"-I{CSRC}",
          f"-I{CSRC / 'host'}"]
HIP_FLAGS = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-fno-gpu-rdc"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def _compile(src: str) -> Path:
    s = CSRC / src
    obj = BUILD / (src.replace("/", "_") + ".o")
    if s.suffix == ".hip":
        cmd = [HIPCC, *COMMON, *HIP_FLAGS, "-x", "hip", "-c", str(s), "-o", str(obj)]
    else:
        cmd = ["g++", *COMMON, "-Wall", "-c", str(s), "-o", str(obj)]
    _run(cmd)
    return obj


def build(verbose: bool = False) -> Path:
    """Compile every source and link libdsmc_hip.so (always rebuilds)."""
    BUILD.mkdir(parents=True, exist_ok=True)
    LIB.parent.mkdir(parents=True, exist_ok=True)
    srcs = HOST_SOURCES + HIP_SOURCES
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *map(str, objs), "-o", str(tmp), "-pthread"])
    os.replace(tmp, LIB)
    # the reference-style command-line driver, linked against the library
    app = PKG / "bin" / "dsmc_main"
    app.parent.mkdir(parents=True, exist_ok=True)
    _run(["g++", "-O2", "-std=c++17", f"-I{ROOT / 'include'}", str(PKG / "apps" / "dsmc_main.cpp"),
          f"-L{LIB.parent}", "-ldsmc_hip", f"-Wl,-rpath,{LIB.parent}", "-Wl,-rpath,$ORIGIN/../lib",
          "-o", str(app)])
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)

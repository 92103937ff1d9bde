"""Build recipe for the gfx950 engine library ``ladds_amd/libladds_amd.so``.

Every HIP/C++ source under ``ladds_amd/csrc`` is compiled with ``hipcc`` for
gfx950 only, then linked into one shared library whose exported symbols are
exactly the C ABI of ``include/ladds_amd.h``.  Packed FP32 instructions are
disabled (``-Xclang -target-feature -Xclang -packed-fp32-ops``).

Run ``python ladds_amd/build.py`` (or ``__graft_entry__.build()``); it does not import
the package, whose import loads the library this script builds.
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
LIB = PKG / "libladds_amd.so"
ADAPTER_CHECK = ROOT / "tests" / "cpp" / "adapter_check"
ARCH = "gfx950"

SOURCES = [
    "api.cpp",
    "ephemeris.cpp",
    "population.cpp",
    "propagate.hip",
    "export.hip",
    "conjunction.hip",
    "compact.hip",
    "tiles.hip",
    "adjoint.hip",
    "noalloc.cpp",
]

COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-fvisibility=hidden",
    "-Wall",
    "-Wno-unused-result",
    f"-I{ROOT / 'include'}",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(path).exists():
        raise RuntimeError("hipcc not found; the engine needs ROCm's hipcc to build")
    return path


def _compile(src: str) -> Path:
    obj = BUILD / (Path(src).stem + ".o")
    cmd = [hipcc(), *COMMON, "-x", "hip", "-c", str(CSRC / src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = False) -> Path:
    """Compile all sources and link the shared library.  Always rebuilds."""
    BUILD.mkdir(exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-o", str(tmp), *map(str, objs), "-lpthread"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    build_adapter_check(verbose)
    return LIB


def build_adapter_check(verbose: bool = False) -> Path:
    """Host-only C++ program that exercises the C++ adapter (include/ladds_amd.hpp)
    against the library; run by tests/test_cpp_adapter.py."""
    src = ADAPTER_CHECK.with_suffix(".cpp")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", f"-I{ROOT / 'include'}", str(src), "-o",
           str(ADAPTER_CHECK), f"-L{PKG}", "-lladds_amd", "-Wl,-rpath,$ORIGIN/../../ladds_amd"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"adapter check build failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    if verbose:
        print(f"built {ADAPTER_CHECK}")
    return ADAPTER_CHECK


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)

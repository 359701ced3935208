"""Build of the HIP library `noa_amd/_build/libnoa_hip.so` (gfx950) with hipcc, in-tree.

Usage: ``python -m noa_amd.build`` (or ``noa_amd.build.build()``). Every call recompiles.
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = BUILD / "libnoa_hip.so"
INCLUDE = PKG.parent / "include"

ARCH = "gfx950"

# -ffp-contract=off: never fuse a*b+c, so results follow the reference's expression rounding.
# -packed-fp32-ops disabled: no v_pk_*_f32 generation.
HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++20",
    "-fPIC",
    "-ffp-contract=off",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
    f"-I{INCLUDE}",
]

SOURCES = ["api.cpp", "extract.hip", "rasterize.hip", "insert.hip", "insert_extract.hip", "phase_shift.hip",
           "spectrum.hip", "ctf.hip", "bspline.hip", "fft.cpp", "fft_stage.hip", "mrc.cpp",
           "mrc_codec.hip"]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(path).exists():
        raise RuntimeError("hipcc not found: the HIP library cannot be built")
    return path


def _compile(src: Path, obj: Path, verbose: bool) -> None:
    cmd = [hipcc(), *HIP_FLAGS, "-c", str(src), "-o", str(obj)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src.name}:\n{res.stdout}\n{res.stderr}")
    if verbose and res.stderr.strip():
        print(res.stderr, file=sys.stderr)


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    BUILD.mkdir(parents=True, exist_ok=True)
    objs = []
    jobs = jobs or min(4, os.cpu_count() or 1)
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as pool:
        futs = []
        for name in SOURCES:
            src = CSRC / name
            obj = BUILD / (src.stem + ("_host" if src.suffix == ".cpp" else "") + ".o")
            objs.append(obj)
            futs.append(pool.submit(_compile, src, obj, verbose))
        for f in futs:
            f.result()
    tmp = LIB.with_suffix(".so.tmp")
    rocm = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), f"-L{rocm / 'lib'}", "-lrocfft",
           f"-Wl,-rpath,{rocm / 'lib'}", "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


TESTS = PKG.parent / "tests"
FACADE_TEST = TESTS / "_build" / "test_facade"


def build_tests(verbose: bool = False) -> Path:
    """The C++ facade test program (host code only: g++ against include/noa_hip.hpp and the built library)."""
    FACADE_TEST.parent.mkdir(parents=True, exist_ok=True)
    rocm = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
    cmd = ["g++", "-std=c++20", "-O1", "-Wall", "-Wextra", f"-I{INCLUDE}", f"-I{rocm / 'include'}",
           "-D__HIP_PLATFORM_AMD__", str(TESTS / "cpp" / "test_facade.cpp"), "-o", str(FACADE_TEST),
           f"-L{BUILD}", "-lnoa_hip", f"-L{rocm / 'lib'}", "-lamdhip64",
           f"-Wl,-rpath,$ORIGIN/../../noa_amd/_build:{rocm / 'lib'}"]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"g++ failed on the facade test:\n{res.stdout}\n{res.stderr}")
    return FACADE_TEST


if __name__ == "__main__":
    print(build(verbose=True))
    print(build_tests(verbose=True))

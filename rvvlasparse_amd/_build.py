"""Build librvvla_hip.so in-tree (gfx950 device code, OpenMP host code).

HIP sources (*.hip and capi.cpp) go through hipcc with --offload-arch=gfx950;
the host converters and generators go through g++ -fopenmp; g++ links the
shared library against libamdhip64.  Objects land in _build/ at the repo
root; the library lands next to this file.
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
BUILD = ROOT / "_build"
LIB = PKG / "librvvla_hip.so"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))

HIP_HOST = ("capi.cpp", "batch_host.cpp", "cxx_api.cpp", "autotune.cpp")  # host code that calls the HIP runtime
HIP_SOURCES = sorted(CSRC.glob("*.hip")) + [CSRC / f for f in HIP_HOST]
HOST_SOURCES = sorted(p for p in CSRC.glob("*.cpp") if p.name not in HIP_HOST)

HIPCC_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=gfx950",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
GXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-Wall", "-Wno-unused-function"]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in res.stdout.splitlines() if "packed-fp32-ops" not in l
                    and "is not a recognized feature" not in l)
    if res.returncode != 0:
        raise RuntimeError(f"build step failed ({res.returncode}): {' '.join(cmd)}\n{out}")


def _compile(src: Path) -> Path:
    obj = BUILD / (src.name + ".o")
    inc = [f"-I{INCLUDE}", f"-I{CSRC}"]
    if src in HIP_SOURCES:
        hipcc = str(ROCM / "bin" / "hipcc")
        lang = ["-x", "hip"] if src.suffix == ".cpp" else []
        _run([hipcc, *HIPCC_FLAGS, *inc, *lang, "-c", str(src), "-o", str(obj)])
    else:
        _run(["g++", *GXX_FLAGS, *inc, "-c", str(src), "-o", str(obj)])
    return obj


def build(verbose: bool = True, jobs: int | None = None) -> Path:
    """Compile every source and link librvvla_hip.so; returns its path."""
    BUILD.mkdir(exist_ok=True)
    sources = [s for s in HIP_SOURCES if s.exists()] + HOST_SOURCES
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, sources))
    tmp = LIB.with_suffix(".so.tmp")
    _run(["g++", "-shared", "-fopenmp", "-o", str(tmp), *map(str, objs),
          f"-L{ROCM / 'lib'}", "-lamdhip64", f"-Wl,-rpath,{ROCM / 'lib'}"])
    os.replace(tmp, LIB)
    # the C++ test programs (the drop-in example, the file-I/O check),
    # linked against the library
    for demo in sorted((ROOT / "tests" / "cxx").glob("*.cpp")):
        _run(["g++", "-O2", "-std=c++17", f"-I{INCLUDE}", str(demo), f"-L{PKG}", "-lrvvla_hip",
              "-Wl,-rpath,$ORIGIN/../rvvlasparse_amd", "-o", str(BUILD / demo.stem)])
    # the accuracy and performance programs (C-ABI users), into rvvlasparse_amd/bin
    bindir = PKG / "bin"
    bindir.mkdir(exist_ok=True)
    for app in sorted((CSRC.parent / "apps").glob("*.cpp")):
        _run(["g++", "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", f"-I{INCLUDE}",
              f"-I{ROCM / 'include'}", str(app), f"-L{PKG}", "-lrvvla_hip",
              f"-L{ROCM / 'lib'}", "-lamdhip64", "-Wl,-rpath,$ORIGIN/..",
              f"-Wl,-rpath,{ROCM / 'lib'}", "-o", str(bindir / app.stem)])
    if verbose:
        print(f"built {LIB} from {len(sources)} sources", flush=True)
    return LIB


def clean() -> None:
    shutil.rmtree(BUILD, ignore_errors=True)
    LIB.unlink(missing_ok=True)


if __name__ == "__main__":
    build()
    sys.exit(0)

"""Builds the C++ drop-in program (tests/cxx/api_drop_in.cpp) against
include/ and the in-tree library with g++, into a directory of the
caller's choosing."""
from __future__ import annotations

import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def build_drop_in(out_dir: Path) -> Path:
    out = Path(out_dir) / "api_drop_in"
    pkg = ROOT / "rvvlasparse_amd"
    cmd = ["g++", "-O2", "-std=c++17", f"-I{ROOT / 'include'}", str(ROOT / "tests" / "cxx" / "api_drop_in.cpp"),
           f"-L{pkg}", "-lrvvla_hip", f"-Wl,-rpath,{pkg}", "-o", str(out)]
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    assert res.returncode == 0, res.stdout[-4000:]
    return out

"""The C++ drop-in API (sparse_matrix.h: convert<T, Fmt, true>, sparse_mv<T,
Fmt, true, stage>) driven by a program written the way reference users
write theirs (tests/cxx/api_drop_in.cpp, built by the package build)."""
from __future__ import annotations

import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

BIN = Path(__file__).resolve().parents[1] / "_build" / "api_drop_in"


def test_cxx_drop_in_all_formats_and_stages():
    assert BIN.exists(), f"{BIN} not built (run the package build)"
    print("running", BIN, flush=True)
    res = subprocess.run([str(BIN), "120"], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                         text=True, timeout=100)
    print(res.stdout)
    assert res.returncode == 0, res.stdout[-3000:]
    assert res.stdout.strip().endswith("OK (0 failures)")

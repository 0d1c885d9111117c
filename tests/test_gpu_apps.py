"""The accuracy and performance programs (rvvlasparse_amd/bin, built from
rvvlasparse_amd/apps/*.cpp): every format and precision on a generated
matrix and on a Matrix Market file through the C-ABI."""
from __future__ import annotations

import subprocess
from pathlib import Path

import numpy as np
import pytest

import rvvlasparse_amd as R

pytestmark = pytest.mark.gpu

BIN = Path(__file__).resolve().parents[1] / "rvvlasparse_amd" / "bin"


def _run(args, timeout=240):
    print("run", args, flush=True)
    r = subprocess.run([str(BIN / args[0]), *args[1:]], capture_output=True, text=True,
                       timeout=timeout)
    return r.returncode, r.stdout, r.stderr


def test_accuracy_program_generated():
    rc, out, err = _run(["rvvla_accuracy", "random:20000:12"])
    assert rc == 0, out + err
    lines = [l for l in out.splitlines() if "err/bound" in l]
    assert len(lines) == 8 * 4, out  # 8 formats x 4 precisions
    assert out.strip().endswith("OK (0 failures)")


def test_accuracy_program_mtx(tmp_path):
    a = R.gen_powerlaw(6000, 1.8, 800, 60000, 64, seed=9)
    p = tmp_path / "pl.mtx"
    rows = np.repeat(np.arange(a.m), np.diff(a.rst))
    with open(p, "w") as f:
        f.write(f"%%MatrixMarket matrix coordinate real general\n{a.m} {a.n} {a.nnz}\n")
        for i, j, v in zip(rows, a.col, a.val):
            f.write(f"{i + 1} {j + 1} {float(v)!r}\n")
    rc, out, err = _run(["rvvla_accuracy", str(p), "--alpha", "2", "--beta", "0"])
    assert rc == 0, out + err
    assert "OK (0 failures)" in out


def test_perf_program():
    rc, out, err = _run(["rvvla_perf", "lap:256", "--iters", "20", "--dtypes", "f64,bf16"])
    assert rc == 0, out + err
    rows = [l.split() for l in out.splitlines()[2:] if l and not l.startswith("matrix")]
    assert len(rows) == 16, out
    for r in rows:
        assert float(r[2]) > 0 and float(r[3]) > 0

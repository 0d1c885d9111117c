"""The C++ drop-in API (sparse_matrix.h: convert<T, Fmt, true>, sparse_mv<T,
Fmt, true, stage>) driven by a program written the way reference users
write theirs (tests/cxx/api_drop_in.cpp, built here with g++ against
include/ and the library).  The program checks itself against a plain CSR
loop; every product it makes is also dumped and checked here against the
oracle (A x, and A^T x through the transposed CSR), row by row within
4 (len + 2) u (|alpha| sum|a||x| + |beta||y0|)."""
from __future__ import annotations

import subprocess

import numpy as np
import pytest

from cxx_build import build_drop_in
from helpers import assert_within_bound
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _read(path):
    raw = path.read_bytes()
    m, n, nnz, es, tr = np.frombuffer(raw, np.int64, 5)
    alpha, beta = np.frombuffer(raw, np.float64, 2, 40)
    o = 56
    rst = np.frombuffer(raw, np.int32, m + 1, o); o += 4 * (m + 1)
    col = np.frombuffer(raw, np.int32, nnz, o); o += 4 * nnz
    vt = np.float64 if es == 8 else np.float32
    val = np.frombuffer(raw, vt, nnz, o); o += es * nnz
    nb, ny = (m, n) if tr else (n, m)
    b = np.frombuffer(raw, vt, nb, o); o += es * nb
    y0 = np.frombuffer(raw, vt, ny, o); o += es * ny
    y = np.frombuffer(raw, vt, ny, o); o += es * ny
    assert o == len(raw)
    return int(m), int(n), rst, col, val, b, y0, y, float(alpha), float(beta), bool(tr), vt


def _transpose(m, n, rst, col, val):
    rows = np.repeat(np.arange(m, dtype=np.int32), np.diff(rst))
    order = np.lexsort((rows, col))
    trst = np.zeros(n + 1, np.int32)
    trst[1:] = np.cumsum(np.bincount(col, minlength=n))
    return trst, rows[order].astype(np.int32), np.ascontiguousarray(val[order])


def test_cxx_drop_in_all_formats_and_stages(tmp_path):
    exe = build_drop_in(tmp_path)
    dump = tmp_path / "dump"
    dump.mkdir()
    print("running", exe, flush=True)
    res = subprocess.run([str(exe), "120", str(dump)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                         text=True, timeout=100)
    print(res.stdout)
    assert res.returncode == 0, res.stdout[-3000:]
    assert res.stdout.strip().endswith("OK (0 failures)")
    files = sorted(dump.glob("*.bin"))
    fmts = {f.name.split("_")[1] for f in files}
    assert fmts == {"csr", "hcsr", "sell", "csr5", "lav", "vnec", "vhcc", "cvr"}, fmts
    for f in files:
        m, n, rst, col, val, b, y0, y, alpha, beta, tr, vt = _read(f)
        if tr:
            rst, col, val = _transpose(m, n, rst, col, val)
        y_ref = O.csr_spmv(rst, col, val, b, alpha, beta, y0)
        assert_within_bound(y, y_ref, rst, col, val, b, alpha, beta, y0, vt, f.name)

"""Matrix files: the Matrix Market reader against scipy.io.mmread on the
same files (real / integer / pattern, general / symmetric / skew-symmetric,
comments, blank lines, duplicates, empty rows and columns), the binary CSR
round trip, the C++ wrappers, the converter module, and refusals."""
from __future__ import annotations

import subprocess
import sys
import zlib
from pathlib import Path

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp

import rvvlasparse_amd as R

ROOT = Path(__file__).resolve().parents[1]


def _canonical(M) -> sp.csr_matrix:
    c = sp.csr_matrix(M, dtype=np.float64)
    c.sum_duplicates()
    c.sort_indices()
    return c


def _same(a: R.CSR, ref: sp.csr_matrix, exact=True):
    assert (a.m, a.n) == ref.shape
    assert np.array_equal(a.rst, ref.indptr)
    assert np.array_equal(a.col, ref.indices)
    if exact:  # same decimal strings, duplicates summed in file order
        assert np.array_equal(a.val, ref.data)
    else:  # mirrored duplicates may be summed in another order
        assert np.allclose(a.val, ref.data, rtol=4e-16 * 8, atol=1e-300)


def _write(tmp_path, name, text):
    p = tmp_path / name
    p.write_text(text)
    return p


def _rand_coo(m, n, k, seed, sym=None, integer=False):
    rng = np.random.default_rng(seed)
    i = rng.integers(0, m, k)
    j = rng.integers(0, n, k)
    if sym:
        i, j = np.maximum(i, j), np.minimum(i, j)
        if sym == "skew-symmetric":
            keep = i != j
            i, j = i[keep], j[keep]
    v = rng.integers(-9, 10, len(i)).astype(float) if integer else rng.normal(size=len(i))
    return i, j, v


def _mm_text(m, n, i, j, v, field="real", sym="general", comments=True):
    lines = [f"%%MatrixMarket matrix coordinate {field} {sym}"]
    if comments:
        lines += ["% generated for the reader test", "%"]
    lines.append(f"{m} {n} {len(i)}")
    for a, b, x in zip(i, j, v):
        if field == "pattern":
            lines.append(f"{a + 1} {b + 1}")
        elif field == "integer":
            lines.append(f"{a + 1} {b + 1} {int(x)}")
        else:
            lines.append(f"{a + 1} {b + 1} {float(x)!r}")
    return "\n".join(lines) + "\n"


@pytest.mark.parametrize("field", ["real", "integer", "pattern"])
@pytest.mark.parametrize("sym", ["general", "symmetric", "skew-symmetric"])
def test_mtx_matches_scipy(tmp_path, field, sym):
    if field == "pattern" and sym == "skew-symmetric":
        pytest.skip("pattern skew-symmetric is not a valid Matrix Market combination")
    m = 300
    n = 300 if sym != "general" else 250
    i, j, v = _rand_coo(m, n, 2000, seed=zlib.crc32(f"{field}/{sym}".encode()),
                        sym=None if sym == "general" else sym, integer=field == "integer")
    p = _write(tmp_path, "a.mtx", _mm_text(m, n, i, j, v, field, sym))
    a, code = R.read_mtx(p)
    _same(a, _canonical(scipy.io.mmread(str(p))), exact=sym == "general")
    assert code == "MC" + {"real": "R", "integer": "I", "pattern": "P"}[field] + \
        {"general": "G", "symmetric": "S", "skew-symmetric": "K"}[sym]


def test_duplicates_are_summed_and_rows_sorted(tmp_path):
    text = ("%%MatrixMarket matrix coordinate real general\n"
            "4 5 6\n"
            "3 5 1.5\n"
            "1 2 2.0\n\n"
            "3 1 -1.0\n"
            "3 5 0.25\n"
            "% a comment between entries\n"
            "1 2 0.5\n"
            "2 4 3.0\n")
    a, _ = R.read_mtx(_write(tmp_path, "d.mtx", text))
    assert np.array_equal(a.rst, [0, 1, 2, 4, 4])  # row 4 empty
    assert np.array_equal(a.col, [1, 3, 0, 4])
    assert np.array_equal(a.val, [2.5, 3.0, -1.0, 1.75])


def test_large_file_parallel_parse(tmp_path):
    """Large enough to be split among threads at line boundaries."""
    m = n = 20000
    i, j, v = _rand_coo(m, n, 200_000, seed=7)
    p = _write(tmp_path, "big.mtx", _mm_text(m, n, i, j, v, comments=False))
    a, _ = R.read_mtx(p)
    _same(a, _canonical(sp.coo_matrix((v, (i, j)), shape=(m, n))))


def test_binary_round_trip(tmp_path):
    a = R.gen_powerlaw(5000, 1.8, 500, 40000, 64, seed=3)
    p = tmp_path / "a.bin"
    R.write_bin(p, a, "MCRS")
    b, code = R.read_bin(p)
    assert code == "MCRS"
    assert (b.m, b.n) == (a.m, a.n)
    for f in ("rst", "col", "val"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f
    e = R.CSR(3, 4, np.zeros(4, np.int32), np.zeros(0, np.int32), np.zeros(0))
    R.write_bin(tmp_path / "e.bin", e)
    b, _ = R.read_bin(tmp_path / "e.bin")
    assert (b.m, b.n, b.nnz) == (3, 4, 0)


def test_converter_module(tmp_path):
    i, j, v = _rand_coo(50, 60, 300, seed=11)
    src = _write(tmp_path, "c.mtx", _mm_text(50, 60, i, j, v))
    dst = tmp_path / "c.bin"
    r = subprocess.run([sys.executable, "-m", "rvvlasparse_amd.mtx2bin", str(src), str(dst)],
                       cwd=ROOT, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    a, _ = R.read_mtx(src)
    b, _ = R.read_bin(dst)
    assert np.array_equal(a.rst, b.rst) and np.array_equal(a.col, b.col)
    assert np.array_equal(a.val, b.val)


@pytest.mark.parametrize("text,code", [
    ("%%MatrixMarket matrix array real general\n2 2\n1\n2\n3\n4\n", -4),
    ("%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 0\n", -4),
    ("%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1.0\n", -10),   # too few lines
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1.0\n", -10),   # row out of range
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1\n", -10),       # missing value
    ("not a matrix market file\n", -10),
    ("%%MatrixMarket matrix coordinate real symmetric\n2 3 1\n1 1 1.0\n", -10),  # not square
])
def test_refusals(tmp_path, text, code):
    with pytest.raises(R.RvvlaError) as e:
        R.read_mtx(_write(tmp_path, "bad.mtx", text))
    assert e.value.code == code


def test_missing_and_corrupt_files(tmp_path):
    with pytest.raises(R.RvvlaError) as e:
        R.read_mtx(tmp_path / "nope.mtx")
    assert e.value.code == -23
    a = R.gen_laplacian5(8)
    p = tmp_path / "t.bin"
    R.write_bin(p, a)
    raw = p.read_bytes()
    (tmp_path / "short.bin").write_bytes(raw[:-10])
    with pytest.raises(R.RvvlaError) as e:
        R.read_bin(tmp_path / "short.bin")
    assert e.value.code == -10
    bad = bytearray(raw)
    off = 40 + 4 * (a.m + 1)  # first column index
    bad[off:off + 4] = np.int32(a.n + 5).tobytes()
    (tmp_path / "badcol.bin").write_bytes(bytes(bad))
    with pytest.raises(R.RvvlaError) as e:
        R.read_bin(tmp_path / "badcol.bin")
    assert e.value.code == -10


def test_cxx_wrappers(tmp_path):
    """include/mtx_io.h: read_mtx / write_bin / read_bin on spMtxCRS<T>."""
    exe = ROOT / "_build" / "mtx_io_check"
    assert exe.exists(), "built by rvvlasparse_amd/_build.py"
    i, j, v = _rand_coo(40, 40, 200, seed=5, sym="symmetric")
    src = _write(tmp_path, "s.mtx", _mm_text(40, 40, i, j, v, "real", "symmetric"))
    r = subprocess.run([str(exe), str(src), str(tmp_path / "s.bin")], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    a, code = R.read_mtx(src)
    assert r.stdout.split() == [str(a.m), str(a.n), str(a.nnz), code]

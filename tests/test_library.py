"""Host-side tests of librvvla_hip.so: exported symbols, converters against
the oracle, parameter parsing, generators.  No GPU calls."""
from __future__ import annotations

import re
from pathlib import Path

import numpy as np
import pytest

import rvvlasparse_amd as R
from helpers import laplacian, rand_csr
from oracle import oracle as O

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        names |= set(re.findall(r"^\s*(?:[\w\s\*]+?)\b(rvvla_\w+)\s*\(", h.read_text(), re.M))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    syms = declared_symbols()
    assert len(syms) >= 15
    missing = [s for s in syms if not hasattr(R.lib, s)]
    assert not missing, missing


def test_strerror_and_names():
    assert R.lib.rvvla_strerror(0) == b"success"
    assert b"unknown format" in R.lib.rvvla_strerror(-1)
    assert [R.lib.rvvla_format_name(i).decode() for i in range(8)] == list(R.FORMATS)
    assert R.lib.rvvla_format_name(8) is None


def _csr(rst, col, val, n):
    return R.CSR(len(rst) - 1, n, rst, col, val)


@pytest.mark.parametrize("dt", [R.F64, R.F32])
@pytest.mark.parametrize("Rb,Cb,p", [(7, 5, 2.0), (13, 64, 0.5), (4, 4, 1.0), (100, 100, 3.0)])
def test_hcsr_converter_matches_oracle(dt, Rb, Cb, p):
    rst, col, val = rand_csr(90, 120, 0.08, Rb * 31 + Cb, empty_every=11, long_rows=(40,))
    a = _csr(rst, col, val, 120)
    L = R.convert("hcsr", a, dt, f"R={Rb},C={Cb},p={p}")
    o, err = O.hcsr_convert(90, 120, rst, col, a.values_for(dt), Rb, Cb, p)
    assert err == 0
    for f in ("block_ptr", "row_offset", "row_ptr", "col_idx", "values"):
        assert np.array_equal(L.array(f), o.array(f)), f
    for k in ("R", "C", "b_m", "b_n", "nz"):
        assert L.scalar(k) == o.scalars[k]


def test_hcsr_converter_golden_and_defaults():
    d = np.load(ROOT / "tests" / "golden" / "hcsr_ref.npz")
    for name in sorted({k.split("/")[0] for k in d.files}):
        m, n, Rb, Cb = d[f"{name}/shape"].tolist()
        v = d[f"{name}/val"]
        dt = R.F64 if v.dtype == np.float64 else R.F32
        a = R.CSR(m, n, d[f"{name}/rst"], d[f"{name}/col"], v.astype(np.float64))
        L = R.convert("hcsr", a, dt, f"R={Rb};C={Cb};p={float(d[f'{name}/p'][0])}")
        for f in ("block_ptr", "row_offset", "row_ptr", "col_idx"):
            assert np.array_equal(L.array(f), d[f"{name}/{f}"]), (name, f)
    # reference defaults: R = C = 2000 for FP64, 4000 for FP32, p = 2
    rst, col, val = laplacian(10)
    a = _csr(rst, col, val, 100)
    assert R.convert("hcsr", a, R.F64).scalar("R") == 2000
    assert R.convert("hcsr", a, R.F32).scalar("C") == 4000


def test_hcsr_converter_errors():
    rst, col, val = rand_csr(10, 10, 0.3, 0)
    a = _csr(rst, col, val, 10)
    with pytest.raises(R.RvvlaError) as e:
        R.convert("hcsr", a, R.F64, "R=0")
    assert e.value.code == -10
    with pytest.raises(R.RvvlaError) as e:
        R.convert("hcsr", a, R.F64, "C=70000")
    assert e.value.code == -10
    with pytest.raises(R.RvvlaError) as e:
        R.convert("hcsr", a, R.F64, "R=2,bogus")
    assert e.value.code == -10
    with pytest.raises(R.RvvlaError) as e:
        R.convert("csr", R.CSR(2, 2, np.array([0, 1, 2]), np.array([0, 5]), np.ones(2)))
    assert e.value.code == -10


@pytest.mark.parametrize("sigma", [1, 4, 64, 1000])
@pytest.mark.parametrize("dt", [R.F64, R.F32])
def test_sell_builder_matches_oracle(sigma, dt):
    m, n = 300, 260
    rst, col, val = rand_csr(m, n, 0.05, sigma, empty_every=9, long_rows=(17, 200))
    a = _csr(rst, col, val, n)
    L = R.convert("sell", a, dt, f"sigma={sigma}")
    o = O.sell_build(m, n, rst, col, a.values_for(dt), 64, sigma)
    for f in ("vCl", "vCs", "vBs", "vSBs", "vCol", "vPerm", "vVal"):
        assert np.array_equal(L.array(f), o.array(f)), f
    assert L.scalar("C") == 64 and L.scalar("Sigma") == sigma


def test_sell_default_sigma():
    """Default sigma: 1 (natural order, identity permutation) when 64-row
    slices in natural order pad by at most 1 % of the entries, else 16384."""
    lap = R.gen_laplacian5(64)  # boundary rows are shorter: tiny padding
    L = R.convert("sell", lap, R.F64)
    assert L.scalar("Sigma") == 1
    assert np.array_equal(L.array("vPerm"), np.arange(lap.m, dtype=np.uint32))
    # very uneven rows: a full sort by length (sigma = m) at least halves the
    # padded slots, so it is taken
    pl = R.gen_powerlaw(20000, 1.8, 2000, 200000, 256, seed=1)
    L = np.diff(pl.rst)
    natural = sum(L[r:r + 64].max() * len(L[r:r + 64]) for r in range(0, pl.m, 64))
    srt = np.sort(L)[::-1]
    full = sum(srt[r] * min(64, pl.m - r) for r in range(0, pl.m, 64))
    assert full <= 0.5 * natural
    assert R.convert("sell", pl, R.F64).scalar("Sigma") == pl.m
    # moderately uneven rows (lengths 10..13 at random): windows of 16384
    rng = np.random.default_rng(5)
    lens = rng.integers(10, 14, 5000)
    rst = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    col = np.concatenate([np.arange(l) for l in lens]).astype(np.int32)
    mod = R.CSR(5000, 100, rst, col, np.ones(rst[-1]))
    assert R.convert("sell", mod, R.F64).scalar("Sigma") == 16384
    # padding just over / under 1 %: rows of 100 entries, one row per slice longer
    m, n = 64 * 50, 4000
    for extra, want in ((1, 1), (2, 16384)):
        lens = np.full(m, 100)
        lens[::64] += extra  # pads 63 * extra slots per slice
        rst = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
        col = np.concatenate([np.arange(l) for l in lens]).astype(np.int32)
        a = R.CSR(m, n, rst, col, np.ones(rst[-1]))
        pad = 50 * 63 * extra
        assert (pad <= 0.01 * a.nnz) == (want == 1)
        assert R.convert("sell", a, R.F64).scalar("Sigma") == want


def test_generators():
    lap = R.gen_laplacian5(9)
    rst, col, val = laplacian(9)
    assert np.array_equal(lap.rst, rst) and np.array_equal(lap.col, col)
    assert np.array_equal(lap.val, val)
    rr = R.gen_random_rows(1000, 16, seed=0)
    ln = rr.row_lengths()
    assert np.all(ln == 16)
    for i in range(0, 1000, 97):
        c = rr.col[rr.rst[i]:rr.rst[i + 1]]
        assert np.all(np.diff(c) > 0) and i in c
    assert rr.val.min() >= -1 and rr.val.max() < 1
    rr2 = R.gen_random_rows(1000, 16, seed=0)
    assert np.array_equal(rr.col, rr2.col) and np.array_equal(rr.val, rr2.val)
    pl = R.gen_powerlaw(20000, 1.8, 2000, 300000, 64, seed=1)
    ln = pl.row_lengths()
    assert ln.min() >= 1 and ln.max() <= 2000
    assert abs(pl.nnz - 300000) < 0.02 * 300000
    rows = np.repeat(np.arange(pl.m), ln)
    assert np.all(np.diff(pl.col)[np.diff(rows) == 0] > 0)
    # half of each row (capped by the window) lies within +-band of the diagonal;
    # uniform draws may land there as well
    near = np.abs(pl.col.astype(np.int64) - rows) <= 64
    r = np.arange(pl.m)
    window = np.minimum(pl.m - 1, r + 64) - np.maximum(0, r - 64) + 1
    expect = np.minimum((ln + 1) // 2, window).sum()
    assert expect <= near.sum() <= expect + 0.01 * pl.nnz
    u = R.gen_uniform(10000, -1, 1, seed=4)
    assert -1 <= u.min() and u.max() < 1 and abs(u.mean()) < 0.05


def test_read_probe_refuses_bad_buffers():
    """The read probe checks its arguments before touching the device."""
    import ctypes
    bad = [(None, 1024), (ctypes.c_void_p(16), 8), (ctypes.c_void_p(16), 1000),
           (ctypes.c_void_p(8), 1024)]
    for buf, nbytes in bad:
        assert R.lib.rvvla_read_probe(buf, nbytes, ctypes.c_void_p(64), None) == -10  # RVVLA_ERR_INVALID_ARG

"""Pins the oracle: bit-equality with the reference's own scalar library
(compiled from its sources into oracle/_ref), the committed golden vectors,
the worked HCSR example, closed forms and dense brute force."""
from __future__ import annotations

import numpy as np
import pytest

from helpers import laplacian, rand_csr
from oracle import oracle as O

GOLD = __import__("pathlib").Path(__file__).resolve().parent / "golden"
need_ref = pytest.mark.skipif(O.ref is None, reason="oracle/_ref (compiled reference) not built")


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (0.75, -0.5), (-2.0, 1.0)])
@need_ref
def test_csr_oracle_equals_reference(dt, alpha, beta):
    rst, col, val = rand_csr(500, 400, 0.03, 11, empty_every=7, long_rows=(3, 250))
    rng = np.random.default_rng(5)
    x = rng.uniform(-1, 1, 400).astype(dt)
    y0 = rng.uniform(-1, 1, 500).astype(dt)
    v = val.astype(dt)
    a = O.csr_spmv(rst, col, v, x, alpha, beta, y0)
    b = O.csr_spmv(rst, col, v, x, alpha, beta, y0, use_ref=True)
    assert a.dtype == dt and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def test_csr_oracle_golden():
    d = np.load(GOLD / "csr_ref.npz")
    for name in ("rand_empty", "lap12"):
        for dt in ("float64", "float32"):
            y = O.csr_spmv(d[f"{name}/rst"], d[f"{name}/col"], d[f"{name}/{dt}/val"],
                           d[f"{name}/{dt}/x"], 0.75, -0.5, d[f"{name}/{dt}/y0"])
            assert np.array_equal(y.view(np.uint8), d[f"{name}/{dt}/y"].view(np.uint8)), (name, dt)


def test_csr_laplacian_closed_forms():
    g = 16
    rst, col, val = laplacian(g)
    y = O.csr_spmv(rst, col, val, np.ones(g * g))
    gy, gx = np.divmod(np.arange(g * g), g)
    expect = (gy == 0).astype(float) + (gy == g - 1) + (gx == 0) + (gx == g - 1)
    assert np.array_equal(y, expect)
    # a linear field is harmonic: A x vanishes on interior rows
    xl = 3.0 * gx + 5.0 * gy + 1.0
    y = O.csr_spmv(rst, col, val, xl)
    interior = (gy > 0) & (gy < g - 1) & (gx > 0) & (gx < g - 1)
    assert np.all(y[interior] == 0.0)


def test_csr_dense_bruteforce():
    rng = np.random.default_rng(3)
    dense = rng.uniform(-1, 1, (23, 17)) * (rng.uniform(size=(23, 17)) < 0.4)
    rst = np.zeros(24, np.int32)
    rst[1:] = np.cumsum((dense != 0).sum(1))
    col = np.nonzero(dense)[1].astype(np.int32)
    val = dense[dense != 0]
    x = rng.uniform(-1, 1, 17)
    y0 = rng.uniform(-1, 1, 23)
    y = O.csr_spmv(rst, col, val, x, 1.25, 0.5, y0)
    np.testing.assert_allclose(y, 1.25 * dense @ x + 0.5 * y0, rtol=0, atol=1e-14)


def test_csr_dd_and_absdot():
    # values whose naive sum cancels: dd must recover the exact result
    rst = np.array([0, 4], np.int32)
    col = np.array([0, 1, 2, 3], np.int32)
    val = np.array([1e16, 1.0, -1e16, 1.0])
    x = np.ones(4)
    assert O.csr_dot_dd(rst, col, val, x)[0] == 2.0
    assert O.csr_absdot(rst, col, val, -x)[0] == 2e16 + 2.0


def test_hcsr_worked_example():
    d = np.load(GOLD / "hcsr_ref.npz")
    rst, col, val = d["worked6_p1/rst"], d["worked6_p1/col"], d["worked6_p1/val"]
    o, err = O.hcsr_convert(6, 6, rst, col, val, 3, 3, 1.0)
    assert err == 0
    assert o.array("block_ptr").view(np.int32).tolist() == [0, 3, 5, 7, 10]
    ro = o.array("row_offset")
    assert [(int(v) >> 30, int(v) & 0x3FFFFFFF) for v in ro] == [(2, 0), (2, 3), (2, 5), (2, 7), (0, 10)]
    assert o.array("row_ptr").tolist() == [0, 0, 1, 0, 2, 0, 2, 1, 1, 2]
    assert o.array("col_idx").tolist() == [0, 1, 2, 1, 2, 0, 1, 0, 1, 2]
    assert o.array("values").view(np.float64).tolist() == [1, 2, 4, 3, 5, 6, 9, 7, 8, 10]
    o, _ = O.hcsr_convert(6, 6, rst, col, val, 3, 3, 0.5)
    assert [int(v) >> 30 for v in o.array("row_offset")[:4]] == [1, 1, 1, 1]
    assert o.array("row_ptr").tolist() == [0, 2, 3, 3, 0, 1, 1, 2, 0, 1, 1, 2, 0, 0, 2, 3]


def test_hcsr_oracle_golden():
    d = np.load(GOLD / "hcsr_ref.npz")
    names = sorted({k.split("/")[0] for k in d.files})
    for name in names:
        m, n, R, C = d[f"{name}/shape"].tolist()
        p = float(d[f"{name}/p"][0])
        v = d[f"{name}/val"]
        o, err = O.hcsr_convert(m, n, d[f"{name}/rst"], d[f"{name}/col"], v, R, C, p)
        assert err == 0
        for f in ("block_ptr", "row_offset", "row_ptr", "col_idx"):
            assert np.array_equal(o.array(f), d[f"{name}/{f}"]), (name, f)
        assert np.array_equal(o.array("values").view(v.dtype), d[f"{name}/values"])
        y = O.hcsr_spmv(o, d[f"{name}/x"], 1.5, 0.0)
        assert np.array_equal(y.view(np.uint8), d[f"{name}/y_hcsr"].view(np.uint8)), name


@need_ref
@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("R,C,p", [(7, 5, 2.0), (13, 64, 0.5), (4, 4, 1.0), (100, 100, 3.0), (1, 1, 2.0)])
def test_hcsr_oracle_equals_reference(dt, R, C, p):
    rst, col, val = rand_csr(90, 120, 0.08, R * 31 + C, empty_every=11, long_rows=(40,))
    v = val.astype(dt)
    o, e1 = O.hcsr_convert(90, 120, rst, col, v, R, C, p)
    r, e2 = O.hcsr_convert(90, 120, rst, col, v, R, C, p, use_ref=True)
    assert e1 == e2 == 0
    for f in ("block_ptr", "row_offset", "row_ptr", "col_idx", "values"):
        assert np.array_equal(o.array(f), r.array(f)), f
    x = np.random.default_rng(1).uniform(-1, 1, 120).astype(dt)
    y0 = np.random.default_rng(2).uniform(-1, 1, 90).astype(dt)
    a = O.hcsr_spmv(o, x, 0.5, 2.0, y0)
    b = O.hcsr_spmv(r, x, 0.5, 2.0, y0)
    assert np.array_equal(a.view(np.uint8), b.view(np.uint8))


@need_ref
def test_hcsr_error_codes_match_reference():
    rst, col, val = rand_csr(10, 10, 0.3, 0)
    for R, C in ((0, 3), (3, 0)):
        assert O.hcsr_convert(10, 10, rst, col, val, R, C, 2.0)[1] == -10
        assert O.hcsr_convert(10, 10, rst, col, val, R, C, 2.0, use_ref=True)[1] == -10


@need_ref
@pytest.mark.parametrize("sigma", [1, 4, 64, 1000])
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_sell_layout_drives_reference_kernel(sigma, dt):
    """The oracle SELL layout fed to the reference scalar SELL kernel gives
    exactly the reference scalar CSR result (same per-row summation order;
    padding adds zeros)."""
    m, n = 300, 260
    rst, col, val = rand_csr(m, n, 0.05, sigma, empty_every=9, long_rows=(17, 200))
    v = val.astype(dt)
    o = O.sell_build(m, n, rst, col, v, 64, sigma)
    arrays = {k: o.array(k) for k in ("vCl", "vCs", "vBs", "vSBs", "vCol", "vPerm")}
    arrays = {k: (a.view(np.int32) if k in ("vCl", "vCs", "vBs", "vSBs") else a) for k, a in arrays.items()}
    arrays["vVal"] = o.array("vVal").view(dt)
    rng = np.random.default_rng(9)
    x = rng.uniform(-1, 1, n).astype(dt)
    y0 = rng.uniform(-1, 1, m).astype(dt)
    y_ref = O.csr_spmv(rst, col, v, x, 1.5, -0.25, y0, use_ref=True)
    y_sell_ref = O.sell_spmv(arrays, m, n, x, 1.5, -0.25, y0, use_ref=True, sigma=sigma)
    y_sell_orc = O.sell_spmv(arrays, m, n, x, 1.5, -0.25, y0)
    assert np.array_equal(y_sell_ref.view(np.uint8), y_ref.view(np.uint8))
    assert np.array_equal(y_sell_orc.view(np.uint8), y_ref.view(np.uint8))
    # invariants: perm is a bijection; lengths descend inside each sigma window
    perm = arrays["vPerm"].astype(np.int64)
    assert np.array_equal(np.sort(perm), np.arange(m))
    ln = np.diff(rst)[perm]
    for w in range(0, m, sigma):
        seg = ln[w:w + sigma]
        assert np.all(seg[:-1] >= seg[1:])

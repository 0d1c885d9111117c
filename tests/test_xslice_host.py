"""Column-sliced CSR image, checked on the host (host-validation uploads):
the slice masks, bases and 64-row prefixes, the slice-major re-lay of the
entries (each (row, slice) pair holds exactly that row's entries in that
slice, in CSR order, bit for bit), the workgroup blocks (block b in slice
b mod P, covering the slice's pairs in order under the adaptive-block
limits), and the sliced product emulated from those arrays against the
oracle."""
from __future__ import annotations

import numpy as np
import pytest

import rvvlasparse_amd as R
from helpers import assert_within_bound, rand_csr
from oracle import oracle as O

LONG, CAP, ROWS = 1024, 2048, 256


@pytest.fixture(autouse=True)
def host_validation():
    R.check(R.lib.rvvla_set_host_validation(1))
    yield
    R.check(R.lib.rvvla_set_host_validation(0))


def _matrix():
    rst, col, val = rand_csr(4000, 30000, 0.0015, 7, empty_every=11,
                             long_rows=(9, 3000, 3001))
    return R.CSR(4000, 30000, rst, col, val)


def _check_blocks(ne, plen, lo_row, hi_row):
    """Non-empty (first, end) blocks: long-row blocks first, longest first,
    then stream blocks in row order; together they tile [lo_row, hi_row)."""
    is_long = np.array([hi - lo == 1 and plen[lo] > LONG for lo, hi in ne], bool)
    nl = int(is_long.sum())
    assert np.all(is_long[:nl]) and not np.any(is_long[nl:])
    lens = plen[ne[:nl, 0]]
    assert np.all(lens[:-1] >= lens[1:])
    st = ne[nl:]
    assert np.all(st[1:, 0] > st[:-1, 0])
    order = ne[np.argsort(ne[:, 0], kind="stable")]
    if hi_row > lo_row:
        assert order[0, 0] == lo_row and order[-1, 1] == hi_row
        assert np.array_equal(order[1:, 0], order[:-1, 1])
    for lo, hi in st:
        assert hi - lo <= ROWS and plen[lo:hi].sum() <= CAP and plen[lo:hi].max() <= LONG


def test_plain_blocks():
    a = _matrix()
    M = R.convert("csr", a, R.F64, "xslices=0").upload()
    blk = M.view("blocks").astype(np.int64).reshape(-1, 2)
    _check_blocks(blk, np.diff(a.rst), 0, a.m)
    assert len(M.view("xs.base")) == 0


def _check_image(a: R.CSR, M: R.DeviceMatrix, P: int):
    w = -(-a.n // P)
    mask = M.view("xs.mask")
    base = M.view("xs.base").astype(np.int64)
    pref = M.view("xs.pref").astype(np.int64)
    xrst = M.view("xs.rst").astype(np.int64)
    xcol = M.view("xs.col").astype(np.int64)
    xval = M.view("xs.val").view(np.float64 if M.dtype == R.F64 else np.float32)
    blk = M.view("xs.blk").astype(np.int64).reshape(-1, 2)
    m, nq = a.m, -(-a.m // 64)
    assert mask.shape == (m,) and base.shape == (P + 1,) and pref.shape == (P * nq,)
    lens = np.diff(a.rst)
    rows = np.repeat(np.arange(m), lens)
    sl = a.col.astype(np.int64) // w
    want = np.zeros(m, np.int64)
    np.bitwise_or.at(want, rows, 1 << sl)
    assert np.array_equal(mask.astype(np.int64), want)
    for p in range(P):
        has = (want >> p) & 1
        assert base[p + 1] - base[p] == has.sum()
        cum = np.concatenate([[0], np.cumsum(has)])
        assert np.array_equal(pref[p * nq:(p + 1) * nq], cum[np.arange(nq) * 64])
    assert xrst[0] == 0 and xrst[-1] == a.nnz and len(xrst) == base[P] + 1
    vals = a.values_for(M.dtype)
    for r in range(m):
        e0, e1 = a.rst[r], a.rst[r + 1]
        for p in range(P):
            if not (want[r] >> p) & 1:
                continue
            k = base[p] + pref[p * nq + r // 64] + int(
                sum((want[r0] >> p) & 1 for r0 in range(r // 64 * 64, r)))
            sel = np.nonzero(sl[e0:e1] == p)[0] + e0
            assert np.array_equal(xcol[xrst[k]:xrst[k + 1]], a.col[sel]), (r, p)
            assert np.array_equal(xval[xrst[k]:xrst[k + 1]], vals[sel]), (r, p)
    # blocks: with X = min(P, 8) XCDs, entry b of the launch runs on XCD
    # b mod X and must hold a slice p with p mod X == b mod X; slices run in
    # phases of X (all of phase 0 first); each slice's non-empty blocks tile
    # its pairs and keep the adaptive limits
    plen = np.diff(xrst)
    X = min(P, 8)
    assert len(blk) % X == 0
    ne = np.nonzero(blk[:, 1] > blk[:, 0])[0]
    sl_of = np.searchsorted(base, blk[ne, 0], side="right") - 1
    assert np.all(blk[ne, 1] <= base[sl_of + 1])
    assert np.all(sl_of % X == ne % X)
    assert np.all(np.diff(sl_of // X) >= 0)
    for p in range(P):
        _check_blocks(blk[ne[sl_of == p]], plen, base[p], base[p + 1])
    return base, pref, xrst, xcol, xval


@pytest.mark.parametrize("P", [2, 3, 8, 12, 16])
@pytest.mark.parametrize("dtype", [R.F64, R.F32])
def test_sliced_image(P, dtype):
    a = _matrix()
    M = R.convert("csr", a, dtype, f"xslices={P}").upload()
    base, pref, xrst, xcol, xval = _check_image(a, M, P)
    # the sliced product: partial per pair, then slice-order sums per row
    dt = np.float64 if dtype == R.F64 else np.float32
    x = R.gen_uniform(a.n, -1, 1, seed=5).astype(dt)
    part = O.csr_spmv(xrst.astype(np.int32), xcol.astype(np.int32), xval, x, 1.0, 0.0,
                      np.zeros(len(xrst) - 1, dt))
    mask = M.view("xs.mask").astype(np.int64)
    nq = -(-a.m // 64)
    y = np.zeros(a.m, dt)
    for r in range(a.m):
        s = dt(0)
        for p in range(P):
            if (mask[r] >> p) & 1:
                q0 = r // 64 * 64
                k = base[p] + pref[p * nq + r // 64] + int(((mask[q0:r] >> p) & 1).sum())
                s = dt(s + part[k])
        y[r] = s
    y_ref = O.csr_spmv(a.rst, a.col, a.values_for(dtype), x, 1.0, 0.0, np.zeros(a.m, dt))
    assert_within_bound(y, y_ref, a.rst, a.col, a.values_for(dtype), x, 1.0, 0.0,
                        np.zeros(a.m, dt), dt, f"sliced P={P}")


def test_sliced_16bit_values():
    a = _matrix()
    M = R.convert("csr", a, R.BF16, "xslices=8").upload()
    xv = M.view("xs.val")
    assert xv.dtype == np.uint16 and xv.shape == (a.nnz,)


def test_auto_choice():
    """Auto: sliced when x exceeds one L2 and columns are far from the
    diagonal; plain for a banded matrix or a small x; xslices=1 disables;
    LDS-sliced with enough entries per pair and per workgroup."""
    n = 600_000
    g = R.gen_random_rows(n, 2, seed=1)  # diagonal + one uniform column
    assert len(R.convert("csr", g, R.F64).upload().view("xs.base")) == 9
    big = R.gen_random_rows(4_400_000, 2, seed=1)  # FP64 x over 8 x 4 MiB: two per XCD
    assert len(R.convert("csr", big, R.F64).upload().view("xs.base")) == 17
    assert len(R.convert("csr", big, R.F32).upload().view("xs.base")) == 9
    assert len(R.convert("csr", g, R.F64, "xslices=1").upload().view("xs.base")) == 0
    lap = R.gen_laplacian5(800)  # n = 640000, banded
    assert len(R.convert("csr", lap, R.F64).upload().view("xs.base")) == 0
    small = _matrix()
    assert len(R.convert("csr", small, R.F64).upload().view("xs.base")) == 0
    # several entries per (row, 16384-column slice) pair and at least 16384
    # entries per workgroup: the LDS-sliced image instead
    pl = R.gen_powerlaw(1_000_000, 1.8, 20000, 16_000_000, 1024, seed=3)
    M = R.convert("csr", pl, R.F64).upload()
    assert len(M.view("xs.base")) == 0 and len(M.view("xl.rowptr")) == pl.m + 1
    # 32-bit x of 1 M columns (4 MiB) fits one L2: neither sliced image
    M = R.convert("csr", pl, R.F32).upload()
    assert len(M.view("xs.base")) == 0 and len(M.view("xl.rowptr")) == 0
    assert len(R.convert("csr", pl, R.F64, "xlds=0").upload().view("xl.rowptr")) == 0
    assert len(R.convert("csr", pl, R.F64, "xslices=8").upload().view("xl.rowptr")) == 0


def test_empty_and_bad_request():
    e = R.CSR(5, 7, np.zeros(6, np.int32), np.zeros(0, np.int32), np.zeros(0))
    M = R.convert("csr", e, R.F64, "xslices=4").upload()
    assert np.array_equal(M.view("xs.base"), np.zeros(5, np.uint32))
    assert M.view("xs.blk").size % 8 == 0
    with pytest.raises(R.RvvlaError) as ex:
        R.convert("csr", _matrix(), R.F64, "xslices=17").upload()
    assert ex.value.code == -10


def test_csr_sell_image_choice():
    """CSR keeps a natural-order SELL-64 image (its products run the SELL
    kernel) for matrices of at least 65536 rows of at most 64 entries whose
    64-row slices pad by at most 1 %; not for uneven or longer rows, small
    matrices or sell=0."""
    lap = R.gen_laplacian5(300)  # 90000 rows of 3..5 entries
    assert len(R.convert("csr", lap, R.F64).upload().view("sell.cl")) == -(-lap.m // 64)
    assert len(R.convert("csr", lap, R.F64, "sell=0").upload().view("sell.cl")) == 0
    uneven = R.gen_random_rows_var(70000, 1, 64, seed=4)
    assert len(R.convert("csr", uneven, R.F64).upload().view("sell.cl")) == 0
    small = R.gen_laplacian5(100)  # 10000 rows
    assert len(R.convert("csr", small, R.F32).upload().view("sell.cl")) == 0
    assert len(R.convert("csr", small, R.F32, "sell=1").upload().view("sell.cl")) == -(-small.m // 64)
    # even rows longer than one SELL piece (65 entries, more than 64 steps a
    # slice): not chosen by default, so long rows keep the CSR kernels
    even = R.gen_random_rows(66000, 65, seed=5)
    for dt in (R.F64, R.BF16):
        assert len(R.convert("csr", even, dt).upload().view("sell.cl")) == 0
    assert len(R.convert("csr", even, R.F64, "sell=1").upload().view("sell.cl")) == -(-even.m // 64)
    # the identity permutation of the CSR image is not kept on the device
    assert len(R.convert("csr", lap, R.F64).upload().view("sell.vPerm")) == 0

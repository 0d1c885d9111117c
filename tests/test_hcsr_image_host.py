"""HCSR device image, built in host memory (no GPU): one unit of work per
64-row slice of a block-row (or several, for slices with more than 64 rows
of work), each with dense runs of 16-bit in-block columns laid 64 lanes
wide and a tail of (column, lane, value) entries ordered by lane.  The
product is emulated from the uploaded arrays as k_hcsr_slices and
k_hcsr_combine read them, and compared with the oracle; every row must be
written by exactly one unit or split-slice sum, every partial slot written
once, and no unit may hold more than 64 rows of work."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import rvvlasparse_amd as R
from helpers import assert_within_bound, laplacian, rand_csr
from oracle import oracle as O

UNIT_ROWS = 64


@pytest.fixture(autouse=True)
def host_validation():
    R.check(R.lib.rvvla_set_host_validation(1))
    yield
    R.check(R.lib.rvvla_set_host_validation(0))


def _emulate(M, m, x):
    unit = M.view("hc.unit").view(np.int32).reshape(-1, 8)
    dxoff = M.view("hc.dxoff").view(np.int32).astype(np.int64)
    dcol = M.view("hc.dcol").view(np.uint16)
    dval = M.view("hc.dval").view(np.float64)  # FP64 uploads: raw 8-byte words
    tcol = M.view("hc.tcol").view(np.int32)
    trow = M.view("hc.trow").view(np.uint8)
    tval = M.view("hc.tval").view(np.float64)
    split = M.view("hc.split").view(np.int32).reshape(-1, 4)
    assert len(tcol) % 64 == 0 and len(dcol) == 64 * len(dxoff)
    assert np.all(dcol[-64:] == 0xFFFF), "the image ends with an empty slot row"
    y = np.zeros(m)
    covered = np.zeros(m, np.int64)
    part = {}
    for row0, nrows, s0, s1, te0, te1, out, _ in unit.astype(np.int64):
        assert 0 < nrows <= 64 and s0 <= s1
        assert te0 % 64 == 0 and te0 <= te1
        work = (s1 - s0) + (te1 - te0 + 63) // 64
        assert work <= UNIT_ROWS, "unit holds more than 64 rows of work"
        s = np.zeros(64)
        c = dcol[s0 * 64:s1 * 64].reshape(-1, 64).astype(np.int64)
        v = dval[s0 * 64:s1 * 64].reshape(-1, 64)
        full = c != 0xFFFF
        assert not full[:, nrows:].any(), "entry in a lane past the slice"
        s += np.where(full, v * x[dxoff[s0:s1, None] + np.where(full, c, 0)], 0.0).sum(0)
        lanes = trow[te0:te1].astype(np.int64)
        assert np.all(np.diff(lanes) >= 0), "tail not ordered by lane"
        assert np.all(lanes < max(nrows, 1))
        np.add.at(s, lanes, tval[te0:te1] * x[tcol[te0:te1]])
        if out < 0:
            y[row0:row0 + nrows] = s[:nrows]
            covered[row0:row0 + nrows] += 1
        else:
            assert out not in part, "partial slot written twice"
            part[int(out)] = s
    for row0, nrows, p0, p1 in split.astype(np.int64):
        assert p1 - p0 >= 2
        y[row0:row0 + nrows] = sum(part.pop(p) for p in range(p0, p1))[:nrows]
        covered[row0:row0 + nrows] += 1
    assert not part, "partial slots nobody sums"
    assert np.all(covered == 1), "rows covered other than once"
    return y


def _check(a, M, what):
    x = np.random.default_rng(5).uniform(-1, 1, a.n)
    y = _emulate(M, a.m, x)
    y_ref = O.csr_spmv(a.rst, a.col, a.val, x, 1.0, 0.0, np.zeros(a.m))
    assert_within_bound(y, y_ref, a.rst, a.col, a.val, x, 1.0, 0.0, np.zeros(a.m), np.float64,
                        what)


def _one_long_row():
    rng = np.random.default_rng(4)
    m, n = 130, 30000
    lens = rng.integers(0, 4, m)
    lens[70] = 20000
    rst = np.zeros(m + 1, np.int32)
    rst[1:] = np.cumsum(lens)
    col = np.concatenate([np.sort(rng.choice(n, k, replace=False)) for k in lens]).astype(np.int32)
    return R.CSR(m, n, rst, col, rng.normal(size=len(col)))


def _cases():
    rst, c, v = rand_csr(2600, 2300, 0.003, 2, empty_every=11, long_rows=(7, 1500, 1501))
    return {
        "lap": R.CSR(*(lambda t: (len(t[0]) - 1, 80 * 80, *t))(laplacian(80))),
        "rand_long": R.CSR(2600, 2300, rst, c, v),
        "powerlaw": R.gen_powerlaw(6000, 1.8, 2000, 60000, 128, seed=4),
        "one_long_row": _one_long_row(),
        "empty": R.CSR(5, 7, np.zeros(6, np.int32), np.zeros(0, np.int32), np.zeros(0)),
    }


CASES = _cases()


@pytest.mark.parametrize("params", ["", "R=300,C=400,p=2", "R=16,C=16,p=2", "R=600,C=64,p=1",
                                    "R=100,C=7,p=0.5"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_hcsr_image_emulated(name, params):
    a = CASES[name]
    L = R.convert("hcsr", a, R.F64, params)
    M = L.upload()
    _check(a, M, f"hcsr image {name} {params}")


def test_hcsr_image_dense_and_split():
    """The 80x80 Laplacian with R = C = 80 lays every row's entries in dense
    runs (no tail), and the 20,000-entry row is cut into several units."""
    M = R.convert("hcsr", CASES["lap"], R.F64, "R=80,C=80,p=2").upload()
    assert M.view("hc.tcol").size == 0 and M.view("hc.split").size == 0
    M = R.convert("hcsr", CASES["one_long_row"], R.F64, "R=128,C=1000,p=2").upload()
    split = M.view("hc.split").view(np.int32).reshape(-1, 4)
    assert len(split) == 1 and split[0, 3] - split[0, 2] >= 20000 // (64 * UNIT_ROWS)


def test_hcsr_image_column_ffff():
    """With C = 65536 the in-block column 0xFFFF is an entry, not an empty
    slot: runs holding it go to the tail."""
    m, n = 64, 70000
    rst = np.arange(0, 2 * m + 1, 2, dtype=np.int32)
    col = np.tile(np.array([100, 65535], np.int32), m)
    a = R.CSR(m, n, rst, col, np.linspace(1, 2, 2 * m))
    M = R.convert("hcsr", a, R.F64, "R=64,C=65536,p=2").upload()
    assert M.view("hc.tcol").size >= 2 * m
    _check(a, M, "hcsr image, column 0xFFFF")


def test_hcsr_image_unsorted_coo():
    """COO-kind blocks uploaded with their entries out of row order: the
    image orders each slice's tail by lane whatever the block order."""
    a = CASES["rand_long"]
    L = R.convert("hcsr", a, R.F64, "R=300,C=400,p=2")
    bp, ro = L.array("block_ptr").astype(np.int64), L.array("row_offset")
    rp, ci = L.array("row_ptr").copy(), L.array("col_idx").copy()
    va = L.array("values", np.float64).copy()
    rng = np.random.default_rng(9)
    shuffled = 0
    for b in range(len(bp) - 1):
        if (int(ro[b]) >> 30) != 2 or bp[b + 1] - bp[b] < 2:
            continue
        e0, e1, off = bp[b], bp[b + 1], int(ro[b]) & 0x3FFFFFFF
        p = rng.permutation(e1 - e0)
        rp[off:off + e1 - e0] = rp[off:off + e1 - e0][p]
        ci[e0:e1], va[e0:e1] = ci[e0:e1][p], va[e0:e1][p]
        shuffled += 1
    assert shuffled > 0
    h = C.c_void_p()
    bp32 = bp.astype(np.int32)
    R.check(R.lib.rvvla_hcsr_upload(R.F64, a.m, a.n, a.nnz, 300, 400, int(L.scalar("b_m")),
                                    int(L.scalar("b_n")), bp32.ctypes.data, ro.ctypes.data,
                                    rp.shape[0], rp.ctypes.data, ci.ctypes.data, va.ctypes.data,
                                    C.byref(h)))
    M = R.DeviceMatrix(h)
    _check(a, M, "hcsr image, unsorted COO blocks")

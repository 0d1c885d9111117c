"""HCSR device image, built in host memory (no GPU): the block-rows cut into
units of 256 local rows (their COO-kind blocks ordered by local row, with
the entry offset of every unit boundary) and the block-rows walked entry by
entry.  The product is emulated from the uploaded arrays exactly as the two
kernels read them, and compared with the oracle; every row must be covered
by exactly one unit or entry-by-entry block-row."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import rvvlasparse_amd as R
from helpers import assert_within_bound, laplacian, rand_csr
from oracle import oracle as O

UNIT = 256


@pytest.fixture(autouse=True)
def host_validation():
    R.check(R.lib.rvvla_set_host_validation(1))
    yield
    R.check(R.lib.rvvla_set_host_validation(0))


def _emulate(M, m, R_, C_, x):
    """y = A x from the uploaded HCSR arrays, unit by unit as k_hcsr_units
    reads them, then the entry-by-entry block-rows as k_hcsr_flat does."""
    v = {f: M.view(f) for f in ("brow_ptr", "nb_col", "nb_ptr", "nb_roff", "rp32", "coo16",
                                "col16", "val", "ebc16", "flat_rows", "u_br", "u_lr0", "csplit")}
    val = v["val"].view(np.float64)  # FP64 uploads: raw 8-byte words
    v["brow_ptr"], v["nb_ptr"] = v["brow_ptr"].view(np.int32), v["nb_ptr"].view(np.int32)
    y = np.zeros(m)
    covered = np.zeros(m, np.int64)
    for br, lr0 in zip(v["u_br"], v["u_lr0"]):
        r0 = int(br) * R_
        rows = min(R_, m - r0)
        nrows = min(UNIT, rows - int(lr0))
        unit = int(lr0) // UNIT
        s = np.zeros(nrows)
        acc = np.zeros(nrows)
        for k in range(v["brow_ptr"][br], v["brow_ptr"][br + 1]):
            xb = int(v["nb_col"][k]) * C_
            e0 = int(v["nb_ptr"][k])
            ro = int(v["nb_roff"][k])
            off = ro & 0x3FFFFFFF
            if ro >> 30 == 1:
                rp = v["rp32"][off:off + rows + 1].astype(np.int64)
                for t in range(nrows):
                    lo, hi = e0 + rp[lr0 + t], e0 + rp[lr0 + t + 1]
                    s[t] += np.dot(val[lo:hi], x[xb + v["col16"][lo:hi].astype(np.int64)])
            else:
                assert ro >> 30 == 2
                a0 = e0 + int(v["csplit"][off + unit])
                a1 = e0 + int(v["csplit"][off + unit + 1])
                lr = v["coo16"][a0:a1].astype(np.int64) - int(lr0)
                assert np.all((lr >= 0) & (lr < nrows)), "COO entry outside its unit"
                np.add.at(acc, lr, val[a0:a1] * x[xb + v["col16"][a0:a1].astype(np.int64)])
        y[r0 + lr0:r0 + lr0 + nrows] = s + acc
        covered[r0 + lr0:r0 + lr0 + nrows] += 1
    for br in v["flat_rows"]:
        r0 = int(br) * R_
        rows = min(R_, m - r0)
        E0 = int(v["nb_ptr"][v["brow_ptr"][br]])
        E1 = int(v["nb_ptr"][v["brow_ptr"][br + 1]])
        cols = v["ebc16"][E0:E1].astype(np.int64) * C_ + v["col16"][E0:E1]
        np.add.at(y, r0 + v["coo16"][E0:E1].astype(np.int64), val[E0:E1] * x[cols])
        covered[r0:r0 + rows] += 1
    assert np.all(covered == 1), "rows covered other than once"
    return y


def _check(a, M, R_, C_, what):
    x = np.random.default_rng(5).uniform(-1, 1, a.n)
    y = _emulate(M, a.m, R_, C_, x)
    y_ref = O.csr_spmv(a.rst, a.col, a.val, x, 1.0, 0.0, np.zeros(a.m))
    assert_within_bound(y, y_ref, a.rst, a.col, a.val, x, 1.0, 0.0, np.zeros(a.m), np.float64,
                        what)


def _cases():
    rst, c, v = rand_csr(2600, 2300, 0.003, 2, empty_every=11, long_rows=(7, 1500, 1501))
    return {
        "lap": R.CSR(*(lambda t: (len(t[0]) - 1, 80 * 80, *t))(laplacian(80))),
        "rand_long": R.CSR(2600, 2300, rst, c, v),
        "powerlaw": R.gen_powerlaw(6000, 1.8, 2000, 60000, 128, seed=4),
        "empty": R.CSR(5, 7, np.zeros(6, np.int32), np.zeros(0, np.int32), np.zeros(0)),
    }


CASES = _cases()


@pytest.mark.parametrize("params", ["", "R=300,C=400,p=2", "R=16,C=16,p=2", "R=600,C=64,p=1"])
@pytest.mark.parametrize("name", sorted(CASES))
def test_hcsr_units_emulated(name, params):
    a = CASES[name]
    L = R.convert("hcsr", a, R.F64, params)
    R_, C_ = int(L.scalar("R")), int(L.scalar("C"))
    M = L.upload()
    _check(a, M, R_, C_, f"hcsr units {name} {params}")


def test_hcsr_units_unsorted_coo():
    """COO-kind blocks uploaded with their entries out of row order: the
    upload orders them by local row (values and columns moved along) so
    that each unit's entries are one range."""
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
    _check(a, M, 300, 400, "hcsr units, unsorted COO blocks")

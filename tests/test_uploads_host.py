"""Uploads without a GPU: rvvla_set_host_validation(1) makes every upload
build its device image in host memory, so the upload paths (index
validation, image construction for all eight formats, edge cases such as an
all-empty matrix or a single row longer than the column count of a lane
group) run here.  Products on such a matrix are refused with
RVVLA_ERR_NO_DEVICE; corrupted direct uploads are refused with
RVVLA_ERR_INVALID_ARG."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import rvvlasparse_amd as R
from helpers import rand_csr

FORMATS = ["csr", "hcsr", "sell", "csr5", "lav", "vnec", "vhcc", "cvr"]
NO_DEVICE, INVALID_ARG = -21, -10


@pytest.fixture(autouse=True)
def host_validation():
    R.check(R.lib.rvvla_set_host_validation(1))
    yield
    R.check(R.lib.rvvla_set_host_validation(0))


def _cases():
    rng = np.random.default_rng(4)
    col = np.sort(rng.choice(20000, 9000, replace=False)).astype(np.int32)
    rst, c, v = rand_csr(900, 700, 0.01, 1, empty_every=7, long_rows=(3, 650, 651))
    return {
        "empty": R.CSR(5, 7, np.zeros(6, np.int32), np.zeros(0, np.int32), np.zeros(0)),
        "zero_rows": R.CSR(0, 3, np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0)),
        "one_long_row": R.CSR(1, 20000, np.array([0, 9000], np.int32), col,
                              rng.normal(size=9000)),
        "rand_long": R.CSR(900, 700, rst, c, v),
        "lap": R.gen_laplacian5(24),
        "trailing_empty": R.CSR(4, 4, np.array([0, 2, 2, 2, 2], np.int32),
                                np.array([0, 3], np.int32), np.array([1.0, -2.0])),
    }


CASES = _cases()


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("fmt", FORMATS)
@pytest.mark.parametrize("dtype", [R.F64, R.F32, R.BF16])
def test_upload_every_format(fmt, name, dtype):
    a = CASES[name]
    M = R.convert(fmt, a, dtype).upload()
    assert (M.m, M.n, M.nnz) == (a.m, a.n, a.nnz)
    assert M.format == R.FORMATS[fmt] and M.dtype == dtype
    assert M.device_bytes >= 0
    assert R.lib.rvvla_spmv(M.handle, 1.0, None, 0.0, None, None) in (NO_DEVICE, INVALID_ARG)
    if a.m and a.n:
        buf = np.zeros(max(a.m, a.n), np.float64)
        assert R.lib.rvvla_spmv(M.handle, 1.0, buf.ctypes.data, 0.0, buf.ctypes.data,
                                None) == NO_DEVICE


def test_csr_upload_validates():
    a = CASES["rand_long"]
    R.csr_upload(a, R.F64)  # accepted
    bad_col = a.col.copy()
    bad_col[17] = a.n
    bad_rst = a.rst.copy()
    bad_rst[5], bad_rst[6] = bad_rst[6], bad_rst[5] - 1
    for rst, col in ((a.rst, bad_col), (bad_rst, a.col)):
        h = C.c_void_p()
        vals = a.values_for(R.F64)
        rc = R.lib.rvvla_csr_upload(R.F64, a.m, a.n, a.nnz, rst.ctypes.data, col.ctypes.data,
                                    vals.ctypes.data, C.byref(h))
        assert rc == INVALID_ARG
        assert not h.value


def _hcsr_fields(a, dtype):
    L = R.convert("hcsr", a, dtype, "R=64,C=128,p=2")
    arr = {f: L.array(f) for f in ("block_ptr", "row_offset", "row_ptr", "col_idx")}
    vals = L.array("values", np.float64 if dtype == R.F64 else np.float32)
    return L, arr, vals


def _hcsr_upload(a, dtype, L, arr, vals):
    h = C.c_void_p()
    rc = R.lib.rvvla_hcsr_upload(dtype, a.m, a.n, a.nnz, 64, 128, int(L.scalar("b_m")),
                                 int(L.scalar("b_n")), arr["block_ptr"].ctypes.data,
                                 arr["row_offset"].ctypes.data, arr["row_ptr"].shape[0],
                                 arr["row_ptr"].ctypes.data, arr["col_idx"].ctypes.data,
                                 vals.ctypes.data, C.byref(h))
    return rc, h


def test_hcsr_upload_validates():
    a = CASES["rand_long"]
    L, arr, vals = _hcsr_fields(a, R.F64)
    rc, h = _hcsr_upload(a, R.F64, L, arr, vals)
    assert rc == 0
    R.DeviceMatrix(h)
    bad = dict(arr)
    bad["col_idx"] = arr["col_idx"].copy()
    bad["col_idx"][3] = 128  # local column outside the 128-wide block
    rc, h = _hcsr_upload(a, R.F64, L, bad, vals)
    assert rc == INVALID_ARG and not h.value
    bad = dict(arr)
    bad["block_ptr"] = arr["block_ptr"].copy()
    bad["block_ptr"][-1] += 1  # entry count disagrees with nnz
    rc, h = _hcsr_upload(a, R.F64, L, bad, vals)
    assert rc == INVALID_ARG and not h.value


def test_products_refused_in_host_mode():
    a = CASES["lap"]
    M = R.csr_upload(a, R.F32)
    X = np.zeros((a.n, 4), np.float32)
    assert R.lib.rvvla_spmm(M.handle, 4, 1.0, X.ctypes.data, 4, 0.0, X.ctypes.data, 4,
                            None) == NO_DEVICE
    x = np.zeros(a.n, np.float32)
    assert R.lib.rvvla_spmv_host(M.handle, 1.0, x.ctypes.data, 0.0, x.ctypes.data) == NO_DEVICE


@pytest.mark.parametrize("name", ["rand_long", "lap", "empty"])
@pytest.mark.parametrize("dtype", [R.F64, R.F32, R.BF16])
def test_mat_read_matches_layout(name, dtype):
    """rvvla_mat_read returns the uploaded SELL / CSR arrays unchanged."""
    a = CASES[name]
    L = R.convert("sell", a, dtype)
    M = L.upload()
    for f in ("vCl", "vCs", "vBs", "vSBs", "vCol", "vPerm"):
        assert np.array_equal(M.read(f), L.array(f)), f
    v = M.read("vVal")
    if dtype == R.F64:
        assert np.array_equal(v.view(np.float64), L.array("vVal", np.float64))
    elif dtype == R.F32:
        assert np.array_equal(v.view(np.float32), L.array("vVal", np.float32))
    else:
        assert v.dtype == np.uint16 and v.size == L.array("vCol").size
    C_ = R.convert("csr", a, dtype).upload()
    assert np.array_equal(C_.read("rst"), a.rst.astype(np.uint32))
    assert np.array_equal(C_.read("col"), a.col.astype(np.uint32))
    with pytest.raises(R.RvvlaError) as e:
        C_.read("vCl")
    assert e.value.code == -23


def test_device_conversion_needs_a_device():
    a = CASES["lap"]
    h = C.c_void_p()
    rc = R.lib.rvvla_sell_from_device(R.F64, a.m, a.n, a.nnz, a.rst.ctypes.data,
                                      a.col.ctypes.data, a.val.ctypes.data, 0, None, C.byref(h))
    assert rc == NO_DEVICE and not h.value


@pytest.mark.parametrize("params", ["R=16,C=16,p=2", "R=8,C=24,p=1"])
def test_hcsr_small_blocks_upload(params):
    """Block-rows with many blocks get per-entry rows and block columns."""
    for name in ("rand_long", "one_long_row", "lap", "empty"):
        a = CASES[name]
        M = R.convert("hcsr", a, R.F64, params).upload()
        assert (M.m, M.n, M.nnz) == (a.m, a.n, a.nnz)

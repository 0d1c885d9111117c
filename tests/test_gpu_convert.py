"""Device-side conversion: CSR -> SELL-C-sigma (rvvla_sell_from_device),
CSR -> HCSR (rvvla_hcsr_from_device), CSR -> CSR5 (rvvla_csr5_from_device),
CSR -> VNEC (rvvla_vnec_from_device), CSR -> LAV (rvvla_lav_from_device),
CSR -> VHCC (rvvla_vhcc_from_device) and CSR -> CVR (rvvla_cvr_from_device).  The arrays read back from the
device equal the host builders' bit for bit (SELL: for the default sigma
rule and explicit sigmas; HCSR: the reference layout for several block
sizes and p; CSR5: tile pointers, descriptors, empty offsets and transposed
entries for several sigmas; VNEC: merge-path coordinates, touched-column
lists and ranks; LAV: chunk tables, row order, masks, lanes and
sparse_part; VHCC: partitions, row-end masks and targets, hand-over flags,
padded entries and empty rows; CVR: streams, packs, flush masks, records,
final-flush rows and empty rows, refusals included), in every precision; the products match the
oracle; a bad device CSR is refused."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import rvvlasparse_amd as R
from helpers import assert_within_bound
from oracle import oracle as O
from test_gpu_parity import _device, case, inputs

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

FIELDS = ("vCl", "vCs", "vBs", "vSBs", "vCol", "vPerm", "vVal")


def _device_csr(a, dtype):
    dev = _device()
    rst = torch.from_numpy(a.rst.astype(np.int32)).to(dev)
    col = torch.from_numpy(a.col.astype(np.int32)).to(dev)
    vt = torch.float64 if dtype == R.F64 else torch.float32
    val = torch.from_numpy(np.asarray(a.val, np.float64)).to(vt).to(dev)
    return rst, col, val


@pytest.mark.parametrize("name", ["lap64", "powerlaw", "rand_long", "one_long_row", "empty",
                                  "random16_c1"])
@pytest.mark.parametrize("sigma", [0, 1, 64, 16384])
@pytest.mark.parametrize("dtype", [R.F64, R.F32, R.BF16, R.F16])
def test_device_sell_equals_host(name, sigma, dtype):
    a = case(name)
    print("device sell", name, sigma, dtype, flush=True)
    rst, col, val = _device_csr(a, dtype)
    D = R.sell_from_device(dtype, a.m, a.n, rst, col, val, sigma)
    torch.cuda.synchronize()
    H = R.convert("sell", a, dtype, f"sigma={sigma}" if sigma > 0 else "").upload()
    for f in FIELDS:
        assert np.array_equal(D.read(f), H.read(f)), f


@pytest.mark.parametrize("dtype", [R.F64, R.F32])
def test_device_sell_product(dtype):
    a = case("powerlaw")
    rst, col, val = _device_csr(a, dtype)
    D = R.sell_from_device(dtype, a.m, a.n, rst, col, val, 0)
    dev = _device()
    x64, y064 = inputs(a, 21)
    dt = np.float64 if dtype == R.F64 else np.float32
    io = torch.float64 if dtype == R.F64 else torch.float32
    x = torch.from_numpy(x64).to(io).to(dev)
    y = torch.from_numpy(y064).to(io).to(dev)
    D.spmv(x, y, 1.25, -0.5)
    torch.cuda.synchronize()
    v = a.values_for(dtype)
    y_ref = O.csr_spmv(a.rst, a.col, v, x64.astype(dt), 1.25, -0.5, y064.astype(dt))
    assert_within_bound(y.cpu().numpy(), y_ref, a.rst, a.col, v, x64.astype(dt), 1.25, -0.5,
                        y064.astype(dt), dt, "device-built SELL")


def test_device_sell_refuses_bad_csr():
    a = case("rand_long")
    rst, col, val = _device_csr(a, R.F64)
    bad = col.clone()
    bad[10] = a.n
    h = C.c_void_p()
    rc = R.lib.rvvla_sell_from_device(R.F64, a.m, a.n, a.nnz, rst.data_ptr(), bad.data_ptr(),
                                      val.data_ptr(), 0, None, C.byref(h))
    assert rc == -10 and not h.value
    badr = rst.clone()
    badr[5] = badr[6] + 1
    rc = R.lib.rvvla_sell_from_device(R.F64, a.m, a.n, a.nnz, badr.data_ptr(), col.data_ptr(),
                                      val.data_ptr(), 0, None, C.byref(h))
    assert rc == -10 and not h.value


# ---------------------------------------------------------------- HCSR
HCSR_FIELDS = ("block_ptr", "row_offset", "row_ptr", "col_idx")


def _round16(v32, dtype):
    """Host rounding of FP32 values to the 16-bit storage (nearest even)."""
    if dtype == R.F16:
        return v32.astype(np.float16).view(np.uint16)
    u = v32.view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def _hcsr_params(params):
    kv = dict(p.split("=") for p in params.split(",")) if params else {}
    return int(kv.get("R", 0)), int(kv.get("C", 0)), float(kv.get("p", 0.0))


@pytest.mark.parametrize("name", ["lap64", "powerlaw", "rand_long", "one_long_row", "empty",
                                  "random16_c1"])
@pytest.mark.parametrize("params", ["", "R=300,C=400,p=2", "R=16,C=16,p=2", "R=64,C=32,p=1",
                                    "R=7,C=5,p=0.5"])
@pytest.mark.parametrize("dtype", [R.F64, R.F32, R.BF16])
def test_device_hcsr_equals_host(name, params, dtype):
    """The reference layout built on the device equals the host converter's
    (which is pinned against the reference's crs_to_HCSR) bit for bit."""
    a = case(name)
    print("device hcsr", name, params, dtype, flush=True)
    rst, col, val = _device_csr(a, dtype)
    Rb, Cb, p = _hcsr_params(params)
    D = R.hcsr_from_device(dtype, a.m, a.n, rst, col, val, Rb, Cb, p)
    torch.cuda.synchronize()
    L = R.convert("hcsr", a, dtype, params)
    for f in HCSR_FIELDS:
        h = L.array(f)
        d = D.read(f).view(h.dtype) if h.size else D.read(f)
        assert d.shape == h.shape and np.array_equal(d, h), f
    hv = L.array("values", np.float64 if dtype == R.F64 else np.float32)
    dv = D.read("values")
    if dtype in (R.F64, R.F32):
        assert np.array_equal(dv.view(hv.dtype), hv)
    else:
        assert np.array_equal(dv, _round16(hv, dtype))


@pytest.mark.parametrize("name", ["lap64", "powerlaw", "rand_long", "one_long_row", "empty"])
@pytest.mark.parametrize("params", ["", "R=300,C=400,p=2", "R=16,C=16,p=2"])
@pytest.mark.parametrize("dtype", [R.F64, R.F32])
def test_device_hcsr_product(name, params, dtype):
    a = case(name)
    rst, col, val = _device_csr(a, dtype)
    D = R.hcsr_from_device(dtype, a.m, a.n, rst, col, val, *_hcsr_params(params))
    dev = _device()
    x64, y064 = inputs(a, 23)
    dt = np.float64 if dtype == R.F64 else np.float32
    io = torch.float64 if dtype == R.F64 else torch.float32
    x = torch.from_numpy(x64).to(io).to(dev)
    y = torch.from_numpy(y064).to(io).to(dev)
    D.spmv(x, y, 0.75, -1.5)
    torch.cuda.synchronize()
    v = a.values_for(dtype)
    y_ref = O.csr_spmv(a.rst, a.col, v, x64.astype(dt), 0.75, -1.5, y064.astype(dt))
    assert_within_bound(y.cpu().numpy(), y_ref, a.rst, a.col, v, x64.astype(dt), 0.75, -1.5,
                        y064.astype(dt), dt, f"device-built HCSR {name} {params}")


def test_device_hcsr_refuses_bad_input():
    a = case("rand_long")
    rst, col, val = _device_csr(a, R.F64)
    bad = col.clone()
    bad[10] = a.n
    h = C.c_void_p()
    rc = R.lib.rvvla_hcsr_from_device(R.F64, a.m, a.n, a.nnz, rst.data_ptr(), bad.data_ptr(),
                                      val.data_ptr(), 300, 400, 2.0, None, C.byref(h))
    assert rc == -10 and not h.value
    rc = R.lib.rvvla_hcsr_from_device(R.F64, a.m, a.n, a.nnz, rst.data_ptr(), col.data_ptr(),
                                      val.data_ptr(), 300, 70000, 2.0, None, C.byref(h))
    assert rc == -10 and not h.value  # C above 65536 (16-bit col_idx)


# ---------------------------------------------------------------- CSR5
CSR5_FIELDS = ("_csr5_tile_pointer", "_csr5_tile_descriptor", "_csr5_tile_empty_offset_pointer",
               "_csr5_tile_empty_offset", "_csr_column_index")


def _csr5_long_case():
    """Rows spanning many tiles, a row from the full tiles into the tail and
    long and short tail rows, with empty rows between (sigma 4: 256-entry
    tiles)."""
    T = 256
    lens = [3, 10 * T - 3, 5, 9 * T + 7, 0, 0, 12, 1, 9, 0, 3 * T + 1, 17 * T + 11, 0]
    lens.append(T - sum(lens) % T + T // 2)
    lens += [1, 0, 8, 9, 40]
    rng = np.random.default_rng(3)
    rst = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    col = np.concatenate([np.sort(rng.choice(40000, l, replace=False)) for l in lens]).astype(np.int32)
    return R.CSR(len(lens), 40000, rst, col, rng.normal(size=int(rst[-1])))


def _case5(name):
    return _csr5_long_case() if name == "long_segments" else case(name)


@pytest.mark.parametrize("name", ["lap64", "powerlaw", "rand_long", "one_long_row", "empty",
                                  "random16_c1", "long_segments"])
@pytest.mark.parametrize("sigma", [0, 4, 12, 16, 64])
@pytest.mark.parametrize("dtype", [R.F64, R.F32, R.BF16])
def test_device_csr5_equals_host(name, sigma, dtype):
    """The CSR5 layout built on the device equals the host builder's bit for
    bit: tile pointers (with the empty-row flag), descriptors, empty
    offsets, transposed columns and values."""
    a = _case5(name)
    print("device csr5", name, sigma, dtype, flush=True)
    rst, col, val = _device_csr(a, dtype)
    D = R.csr5_from_device(dtype, a.m, a.n, rst, col, val, sigma)
    torch.cuda.synchronize()
    L = R.convert("csr5", a, dtype, f"sigma={sigma}" if sigma > 0 else "")
    for f in CSR5_FIELDS:
        h = L.array(f)
        d = D.read(f)
        d = d.view(h.dtype) if d.size else d
        assert d.shape == h.shape and np.array_equal(d, h), f
    hv = L.array("_csr_value", np.float64 if dtype == R.F64 else np.float32)
    dv = D.read("_csr_value")
    if dtype in (R.F64, R.F32):
        assert np.array_equal(dv.view(hv.dtype), hv)
    else:
        assert np.array_equal(dv, _round16(hv, dtype))


@pytest.mark.parametrize("name", ["lap64", "powerlaw", "rand_long", "one_long_row", "empty",
                                  "long_segments"])
@pytest.mark.parametrize("sigma", [0, 4, 12])
@pytest.mark.parametrize("dtype", [R.F64, R.F32])
def test_device_csr5_product(name, sigma, dtype):
    a = _case5(name)
    rst, col, val = _device_csr(a, dtype)
    D = R.csr5_from_device(dtype, a.m, a.n, rst, col, val, sigma)
    dev = _device()
    x64, y064 = inputs(a, 29)
    dt = np.float64 if dtype == R.F64 else np.float32
    io = torch.float64 if dtype == R.F64 else torch.float32
    x = torch.from_numpy(x64).to(io).to(dev)
    y = torch.from_numpy(y064).to(io).to(dev)
    D.spmv(x, y, -0.5, 1.25)
    torch.cuda.synchronize()
    v = a.values_for(dtype)
    y_ref = O.csr_spmv(a.rst, a.col, v, x64.astype(dt), -0.5, 1.25, y064.astype(dt))
    assert_within_bound(y.cpu().numpy(), y_ref, a.rst, a.col, v, x64.astype(dt), -0.5, 1.25,
                        y064.astype(dt), dt, f"device-built CSR5 {name} sigma={sigma}")


def test_device_csr5_refuses_bad_input():
    a = case("rand_long")
    rst, col, val = _device_csr(a, R.F64)
    bad = col.clone()
    bad[10] = -1
    h = C.c_void_p()
    rc = R.lib.rvvla_csr5_from_device(R.F64, a.m, a.n, a.nnz, rst.data_ptr(), bad.data_ptr(),
                                      val.data_ptr(), 16, None, C.byref(h))
    assert rc == -10 and not h.value
    rc = R.lib.rvvla_csr5_from_device(R.F64, a.m, a.n, a.nnz, rst.data_ptr(), col.data_ptr(),
                                      val.data_ptr(), 65, None, C.byref(h))
    assert rc == -3 and not h.value  # sigma above 64


# ---------------------------------------------------------------- VNEC
@pytest.mark.parametrize("name", ["lap64", "powerlaw", "rand_long", "one_long_row", "empty",
                                  "random16_c1"])
@pytest.mark.parametrize("threads", [0, 97])
@pytest.mark.parametrize("dtype", [R.F64, R.F32])
def test_device_vnec_equals_host(name, threads, dtype):
    """Merge-path coordinates, touched-column lists and ranks built on the
    device equal the host builder's."""
    a = case(name)
    cap = 1024 if dtype == R.F64 else 2048
    if threads and -(-(a.m + a.nnz) // threads) > cap:
        threads = 0  # partitions larger than one workgroup stages: the default count
    print("device vnec", name, threads, dtype, flush=True)
    rst, col, val = _device_csr(a, dtype)
    D = R.vnec_from_device(dtype, a.m, a.n, rst, col, val, threads)
    torch.cuda.synchronize()
    L = R.convert("vnec", a, dtype, f"threads={threads}" if threads else "")
    st = L.array("thread_coord_start").view(np.int32).reshape(-1, 2)
    en = L.array("thread_coord_end").view(np.int32).reshape(-1, 2)
    cx = np.concatenate([st[:, 0], en[-1:, 0]])
    cy = np.concatenate([st[:, 1], en[-1:, 1]])
    assert np.array_equal(D.read("cx").view(np.int32), cx)
    assert np.array_equal(D.read("cy").view(np.int32), cy)
    assert np.array_equal(D.read("nec_ptr").view(np.int32), L.array("nec_ptr").view(np.int32))
    ux = D.read("use_x_indices")
    assert np.array_equal(ux.view(np.int32) if ux.size else ux,
                          L.array("use_x_indices").view(np.int32))
    assert np.array_equal(D.read("ecr_indices").astype(np.int64),
                          L.array("ecr_indices").view(np.int32).astype(np.int64))


@pytest.mark.parametrize("name", ["lap64", "powerlaw", "rand_long", "one_long_row", "empty"])
@pytest.mark.parametrize("dtype", [R.F64, R.F32, R.BF16])
def test_device_vnec_product(name, dtype):
    a = case(name)
    rst, col, val = _device_csr(a, dtype)
    D = R.vnec_from_device(dtype, a.m, a.n, rst, col, val)
    dev = _device()
    x64, y064 = inputs(a, 31)
    io = torch.float64 if dtype == R.F64 else torch.float32
    x = torch.from_numpy(x64).to(io).to(dev)
    y = torch.from_numpy(y064).to(io).to(dev)
    D.spmv(x, y, 1.0, 0.5)
    torch.cuda.synchronize()
    if dtype == R.BF16:
        x32, y032 = x64.astype(np.float32), y064.astype(np.float32)
        y64 = O.csr_spmv(a.rst, a.col, a.val, x32.astype(np.float64), 1.0, 0.5,
                         y032.astype(np.float64))
        if np.linalg.norm(y64) > 0:
            assert np.linalg.norm(y.cpu().numpy() - y64) / np.linalg.norm(y64) <= 5e-3
        return
    dt = np.float64 if dtype == R.F64 else np.float32
    v = a.values_for(dtype)
    y_ref = O.csr_spmv(a.rst, a.col, v, x64.astype(dt), 1.0, 0.5, y064.astype(dt))
    assert_within_bound(y.cpu().numpy(), y_ref, a.rst, a.col, v, x64.astype(dt), 1.0, 0.5,
                        y064.astype(dt), dt, f"device-built VNEC {name}")


# ---------------------------------------------------------------- LAV
LAV_FIELDS = ("chunk_offsets", "chunk_offsets_full_mask", "segment_ptr", "out_order", "mask",
              "col_id", "sparse_part.Rst", "sparse_part.Col")


def _lav_params(params):
    kv = dict(p.split("=") for p in params.split(",")) if params else {}
    return float(kv.get("T", -1.0)), int(kv.get("segment_len", 0))


@pytest.mark.parametrize("name", ["lap64", "powerlaw", "rand_long", "one_long_row", "empty",
                                  "random16_c1"])
@pytest.mark.parametrize("params", ["", "T=0", "T=0.5", "segment_len=1000", "T=0.1,segment_len=300"])
@pytest.mark.parametrize("dtype", [R.F64, R.F32, R.BF16])
def test_device_lav_equals_host(name, params, dtype):
    """Column ranking, sparse_part, row order, chunk tables, masks and lanes
    built on the device equal the host builder's bit for bit."""
    a = case(name)
    print("device lav", name, params, dtype, flush=True)
    rst, col, val = _device_csr(a, dtype)
    D = R.lav_from_device(dtype, a.m, a.n, rst, col, val, *_lav_params(params))
    torch.cuda.synchronize()
    L = R.convert("lav", a, dtype, params)
    for f in LAV_FIELDS:
        h = L.array(f)
        d = D.read(f)
        d = d.view(h.dtype) if d.size else d
        assert d.shape == h.shape and np.array_equal(d, h), f
    vt = np.float64 if dtype == R.F64 else np.float32
    for f in ("val", "sparse_part.Val"):
        hv = L.array(f, vt)
        dv = D.read(f)
        if dtype in (R.F64, R.F32):
            assert np.array_equal(dv.view(hv.dtype) if dv.size else dv, hv), f
        else:
            assert np.array_equal(dv, _round16(hv, dtype)), f


@pytest.mark.parametrize("name", ["lap64", "powerlaw", "rand_long", "one_long_row", "empty"])
@pytest.mark.parametrize("params", ["", "T=0", "segment_len=1000"])
@pytest.mark.parametrize("dtype", [R.F64, R.F32])
def test_device_lav_product(name, params, dtype):
    a = case(name)
    rst, col, val = _device_csr(a, dtype)
    D = R.lav_from_device(dtype, a.m, a.n, rst, col, val, *_lav_params(params))
    dev = _device()
    x64, y064 = inputs(a, 37)
    dt = np.float64 if dtype == R.F64 else np.float32
    io = torch.float64 if dtype == R.F64 else torch.float32
    x = torch.from_numpy(x64).to(io).to(dev)
    y = torch.from_numpy(y064).to(io).to(dev)
    D.spmv(x, y, 1.5, -0.25)
    torch.cuda.synchronize()
    v = a.values_for(dtype)
    y_ref = O.csr_spmv(a.rst, a.col, v, x64.astype(dt), 1.5, -0.25, y064.astype(dt))
    assert_within_bound(y.cpu().numpy(), y_ref, a.rst, a.col, v, x64.astype(dt), 1.5, -0.25,
                        y064.astype(dt), dt, f"device-built LAV {name} {params}")


# ---------------------------------------------------------------- VHCC
VHCC_FIELDS = ("start_vec", "rbase", "eor", "row_arr", "fwd_row", "pass", "col", "val",
               "empty_rows")


@pytest.mark.parametrize("name", ["lap64", "powerlaw", "rand_long", "one_long_row", "empty",
                                  "random16_c1"])
@pytest.mark.parametrize("threads", [0, 1, 7, 100000])
@pytest.mark.parametrize("dtype", [R.F64, R.F32, R.BF16, R.F16])
def test_device_vhcc_equals_host(name, threads, dtype):
    """Partitions, row-end masks, row targets (overflow redirects included),
    hand-over flags, padded entries and empty rows built on the device equal
    convert + upload bit for bit."""
    a = case(name)
    print("device vhcc", name, threads, dtype, flush=True)
    rst, col, val = _device_csr(a, dtype)
    D = R.vhcc_from_device(dtype, a.m, a.n, rst, col, val, threads)
    torch.cuda.synchronize()
    H = R.convert("vhcc", a, dtype, f"threads={threads}" if threads else "").upload()
    for f in VHCC_FIELDS:
        d, h = D.read(f), H.read(f)
        assert d.shape == h.shape and np.array_equal(d, h), f


def test_device_vhcc_continued_rows():
    """A row spanning several partitions: the middle partitions hand the
    partial on (pass), the last forwards it to the row; the product matches
    the oracle."""
    a = case("one_long_row")
    rst, col, val = _device_csr(a, R.F64)
    D = R.vhcc_from_device(R.F64, a.m, a.n, rst, col, val, 64)
    H = R.convert("vhcc", a, R.F64, "threads=64").upload()
    assert np.array_equal(D.read("pass"), H.read("pass"))
    assert D.read("pass").astype(np.int64).sum() > 0
    dev = _device()
    x64, y064 = inputs(a, 41)
    x = torch.from_numpy(x64).to(dev)
    y = torch.from_numpy(y064).to(dev)
    D.spmv(x, y, 1.0, 0.5)
    torch.cuda.synchronize()
    y_ref = O.csr_spmv(a.rst, a.col, a.val, x64, 1.0, 0.5, y064)
    assert_within_bound(y.cpu().numpy(), y_ref, a.rst, a.col, a.val, x64, 1.0, 0.5, y064,
                        np.float64, "device-built VHCC one_long_row")


@pytest.mark.parametrize("name", ["lap64", "powerlaw", "rand_long", "one_long_row", "empty"])
@pytest.mark.parametrize("dtype", [R.F64, R.F32])
def test_device_vhcc_product(name, dtype):
    a = case(name)
    rst, col, val = _device_csr(a, dtype)
    D = R.vhcc_from_device(dtype, a.m, a.n, rst, col, val)
    dev = _device()
    x64, y064 = inputs(a, 43)
    dt = np.float64 if dtype == R.F64 else np.float32
    io = torch.float64 if dtype == R.F64 else torch.float32
    x = torch.from_numpy(x64).to(io).to(dev)
    y = torch.from_numpy(y064).to(io).to(dev)
    D.spmv(x, y, -1.0, 2.0)
    torch.cuda.synchronize()
    v = a.values_for(dtype)
    y_ref = O.csr_spmv(a.rst, a.col, v, x64.astype(dt), -1.0, 2.0, y064.astype(dt))
    assert_within_bound(y.cpu().numpy(), y_ref, a.rst, a.col, v, x64.astype(dt), -1.0, 2.0,
                        y064.astype(dt), dt, f"device-built VHCC {name}")


def test_device_vhcc_refuses_bad_csr():
    a = case("lap64")
    rst, col, val = _device_csr(a, R.F64)
    bad = col.clone()
    bad[5] = a.n
    h = C.c_void_p()
    rc = R.lib.rvvla_vhcc_from_device(R.F64, a.m, a.n, a.nnz, rst.data_ptr(), bad.data_ptr(),
                                      val.data_ptr(), 0, None, C.byref(h))
    assert rc == -10 and not h.value


# ---------------------------------------------------------------- CVR
CVR_FIELDS = ("pack_start", "rec_start", "flush", "rec_rows", "tail_rows", "col", "val",
              "empty_rows", "split_rows", "split_ptr")


@pytest.mark.parametrize("name", ["lap64", "powerlaw", "rand_long", "one_long_row", "empty",
                                  "random16_c1"])
@pytest.mark.parametrize("streams", [0, 1, 3, 8, 50])
@pytest.mark.parametrize("dtype", [R.F64, R.F32, R.BF16, R.F16])
def test_device_cvr_equals_host(name, streams, dtype):
    """Stream boundaries, packs, flush masks, records, final-flush rows and
    empty rows built on the device equal convert + upload bit for bit; the
    layouts the host builder refuses are refused with the same code."""
    a = case(name)
    print("device cvr", name, streams, dtype, flush=True)
    rst, col, val = _device_csr(a, dtype)
    params = f"streams={streams}" if streams else ""
    try:
        H = R.convert("cvr", a, dtype, params).upload()
    except R.RvvlaError as e:
        assert e.code == -4
        with pytest.raises(R.RvvlaError) as d:
            R.cvr_from_device(dtype, a.m, a.n, rst, col, val, streams)
        assert d.value.code == -4
        return
    D = R.cvr_from_device(dtype, a.m, a.n, rst, col, val, streams)
    torch.cuda.synchronize()
    for f in CVR_FIELDS:
        d, h = D.read(f), H.read(f)
        assert d.shape == h.shape and np.array_equal(d, h), f


@pytest.mark.parametrize("name,streams", [("lap64", 0), ("powerlaw", 8), ("rand_long", 0),
                                          ("rand_long", 1), ("empty", 0), ("random16_c1", 0)])
@pytest.mark.parametrize("dtype", [R.F64, R.F32])
def test_device_cvr_product(name, streams, dtype):
    a = case(name)
    rst, col, val = _device_csr(a, dtype)
    D = R.cvr_from_device(dtype, a.m, a.n, rst, col, val, streams)
    dev = _device()
    x64, y064 = inputs(a, 47)
    dt = np.float64 if dtype == R.F64 else np.float32
    io = torch.float64 if dtype == R.F64 else torch.float32
    x = torch.from_numpy(x64).to(io).to(dev)
    y = torch.from_numpy(y064).to(io).to(dev)
    D.spmv(x, y, 0.75, -1.0)
    torch.cuda.synchronize()
    v = a.values_for(dtype)
    y_ref = O.csr_spmv(a.rst, a.col, v, x64.astype(dt), 0.75, -1.0, y064.astype(dt))
    assert_within_bound(y.cpu().numpy(), y_ref, a.rst, a.col, v, x64.astype(dt), 0.75, -1.0,
                        y064.astype(dt), dt, f"device-built CVR {name} streams={streams}")


def test_device_cvr_refuses_bad_csr():
    a = case("lap64")
    rst, col, val = _device_csr(a, R.F64)
    bad = rst.clone()
    bad[7] = bad[8] + 1  # row pointers out of order
    h = C.c_void_p()
    rc = R.lib.rvvla_cvr_from_device(R.F64, a.m, a.n, a.nnz, bad.data_ptr(), col.data_ptr(),
                                     val.data_ptr(), 0, None, C.byref(h))
    assert rc == -10 and not h.value

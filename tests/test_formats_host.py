"""LAV, CVR and VHCC layouts built by the library, run through the reference
kernels: the reference's own compiled scalar kernels for LAV and CVR
(oracle/_ref), and the reference's vector VHCC multiply restated lane by
lane (oracle) for VHCC, whose RVV source cannot be built here.  Plus the
layout invariants each format relies on."""
from __future__ import annotations

import numpy as np
import pytest

import rvvlasparse_amd as R
from helpers import assert_within_bound, laplacian, rand_csr
from oracle import oracle as O

need_ref = pytest.mark.skipif(O.ref is None, reason="oracle/_ref (compiled reference) not built")

CASES = {
    "rand_empty": (lambda: rand_csr(300, 260, 0.05, 1, empty_every=9, long_rows=(17, 200)), 260),
    "lap16": (lambda: laplacian(16), 256),
    "skewed": (lambda: rand_csr(700, 900, 0.01, 3, empty_every=4, long_rows=(0, 1, 650)), 900),
    "tiny": (lambda: rand_csr(3, 4, 0.6, 5), 4),
}


def layout_arrays(L, names, dt):
    out = {}
    for k in names:
        a = L.array(k)
        if k in ("vPack_vec_vals", "val", "_vals_arr", "sparse_part.Val"):
            a = a.view(dt)
        elif a.dtype == np.uint32 and k in ("vPack_vec_record", "writeback_pos_record",
                                            "chunk_offsets", "chunk_offsets_full_mask",
                                            "sparse_part.Rst", "sparse_part.Col", "_thr_info"):
            a = a.view(np.int32)
        out[k] = a
    return out


def _inputs(n, m, dt, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1, 1, n).astype(dt), rng.uniform(-1, 1, m).astype(dt)


def _cvr_segments(rst, streams):
    """Segments per row: rows longer than max(64, ceil(nnz / (64 streams)))
    entries are cut into pieces of that many."""
    nnz = int(rst[-1])
    cap = max(64, -(-nnz // (64 * streams)))
    return -(-np.diff(rst).astype(np.int64) // cap)


def test_cvr_segment_counts():
    """Records per row for rows of 3, 20,000 and 7 entries (hand-counted):
    one stream cuts at ceil(20,010 / 64) = 313 entries (64 segments), 40
    streams at 64 (313 segments)."""
    lens = np.array([3, 20000, 7])
    rst = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    col = np.concatenate([np.arange(k) for k in lens]).astype(np.int32)
    a = R.CSR(3, 20000, rst, col, np.ones(len(col)))
    for streams, want in ((1, [1, 64, 1]), (40, [1, 313, 1])):
        L = R.convert("cvr", a, R.F64, f"streams={streams}")
        rec = L.array("vPack_vec_record").view(np.int32).reshape(-1, 2)[:-1, 1]
        assert list(np.bincount(rec, minlength=3)) == want
        assert list(_cvr_segments(rst, streams)) == want


@need_ref
@pytest.mark.parametrize("lanes", [8, 64])
@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_cvr_single_stream_on_reference_kernel(name, lanes, dt):
    """Single stream: the reference scalar CVR kernel reproduces the scalar
    CSR result bit for bit on every row taken whole (one lane, entry order);
    a row cut into segments is one sum per segment, each added into y by
    its own record, so it meets the summation bound with two more
    roundings per segment."""
    (rst, col, val), n = CASES[name][0](), CASES[name][1]
    m = len(rst) - 1
    L = R.convert("cvr", R.CSR(m, n, rst, col, val), R.F64 if dt == np.float64 else R.F32,
                  f"SIMD_LEN={lanes},streams=1")
    arr = layout_arrays(L, ("vPack_vec_vals", "vPack_vec_cols", "vPack_vec_record",
                            "writeback_pos_record"), dt)
    sc = {k: L.scalar(k) for k in ("nItems", "numRows", "numCols", "SIMD_LEN", "Nthreads")}
    x, y0 = _inputs(n, m, dt, 1)
    y = O.cvr_spmv_ref(arr, sc, x, 1.5, -0.5, y0)
    y_ref = O.csr_spmv(rst, col, val.astype(dt), x, 1.5, -0.5, y0, use_ref=True)
    seg = _cvr_segments(rst, 1)
    whole = seg <= 1
    assert np.array_equal(y[whole], y_ref[whole])
    u = 2.0 ** -53 if dt == np.float64 else 2.0 ** -24
    v = val.astype(dt)
    scale = 1.5 * O.csr_absdot(rst, col, v, x) + 0.5 * np.abs(y0.astype(np.float64))
    ln = np.diff(rst).astype(np.float64)
    err = np.abs(y.astype(np.float64) - y_ref)
    assert np.all(err[~whole] <= (ln + 2 * seg + 2)[~whole] * u * scale[~whole])


@pytest.mark.parametrize("streams", [1, 3, 17])
@pytest.mark.parametrize("name", sorted(CASES))
def test_cvr_invariants(name, streams):
    (rst, col, val), n = CASES[name][0](), CASES[name][1]
    m = len(rst) - 1
    L = R.convert("cvr", R.CSR(m, n, rst, col, val), R.F64, f"streams={streams}")
    rec = L.array("vPack_vec_record").view(np.int32).reshape(-1, 2)
    rows = rec[:-1, 1]
    # every non-empty row recorded once per segment
    assert np.array_equal(np.sort(rows), np.repeat(np.arange(m), _cvr_segments(rst, streams)))
    vals = L.array("vPack_vec_vals").view(np.float64)
    assert np.isclose(np.sort(np.abs(vals[vals != 0])), np.sort(np.abs(val[val != 0]))).all()
    nstreams = int(L.scalar("Nthreads"))
    lanes = int(L.scalar("SIMD_LEN"))
    wb = L.array("writeback_pos_record").view(np.int32).reshape(nstreams, lanes + 2)
    assert np.all(np.diff(wb[:, 0]) >= 0) and np.all(np.diff(wb[:, 1]) >= 0)


@need_ref
@pytest.mark.parametrize("lanes,T,seg", [(8, 0.2, 40), (64, 0.2, 1 << 20), (64, 0.0, 32),
                                         (64, 1.0, 64), (8, 0.5, 7)])
@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_lav_on_reference_kernel(name, lanes, T, seg, dt):
    (rst, col, val), n = CASES[name][0](), CASES[name][1]
    m = len(rst) - 1
    L = R.convert("lav", R.CSR(m, n, rst, col, val), R.F64 if dt == np.float64 else R.F32,
                  f"SIMD_Lanes={lanes},T={T},segment_len={seg}")
    names = ("chunk_offsets", "chunk_offsets_full_mask", "segment_ptr", "out_order", "mask", "val",
             "col_id", "sparse_part.Rst", "sparse_part.Col", "sparse_part.Val")
    arr = layout_arrays(L, names, dt)
    sc = {k: L.scalar(k) for k in ("m", "n", "nz", "SIMD_Lanes", "segment_count", "chunk_count")}
    x, y0 = _inputs(n, m, dt, 2)
    v = val.astype(dt)
    y = O.lav_spmv_ref(arr, sc, x, 1.5, -0.5, y0)
    y_ref = O.csr_spmv(rst, col, v, x, 1.5, -0.5, y0)
    # the reference LAV adds alpha*val*x entry by entry into y: two more
    # roundings per entry than the CSR sum
    u = 2.0 ** -53 if dt == np.float64 else 2.0 ** -24
    ln = np.diff(rst).astype(np.float64)
    scale = 1.5 * O.csr_absdot(rst, col, v, x) + 0.5 * np.abs(y0.astype(np.float64))
    assert np.all(np.abs(y - y_ref) <= 8 * (ln + 2) * u * scale)
    # the dense part holds exactly the entries of the (1-T) most populated columns
    sp_nz = int(sc["nz"]) - int(L.scalar("sparse_nz"))
    assert np.count_nonzero(arr["val"]) <= sp_nz


@pytest.mark.parametrize("vec,parts", [(8, 1), (8, 5), (16, 3), (64, 1), (64, 4), (64, 64)])
@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_vhcc_on_restated_reference_kernel(name, vec, parts, dt):
    (rst, col, val), n = CASES[name][0](), CASES[name][1]
    m = len(rst) - 1
    L = R.convert("vhcc", R.CSR(m, n, rst, col, val), R.F64 if dt == np.float64 else R.F32,
                  f"vec_size={vec},threads={parts}")
    arr = layout_arrays(L, ("_veceor_ptr", "_scan_mask", "_row_arr", "_col_arr", "_vals_arr",
                            "_thr_info"), dt)
    arr["_scan_mask"] = arr["_scan_mask"].astype(np.uint8)
    sc = {k: L.scalar(k) for k in ("num_rows", "vec_size", "num_threads", "pad_rows")}
    x, y0 = _inputs(n, m, dt, 3)
    v = val.astype(dt)
    y = O.vhcc_spmv(arr, sc, x, 1.5, -0.5, y0)
    y_ref = O.csr_spmv(rst, col, v, x, 1.5, -0.5, y0)
    assert_within_bound(y, y_ref, rst, col, v, x, 1.5, -0.5, y0, dt, f"vhcc {name}")
    # every entry sits in vector order; every non-empty row ends exactly once
    ra = L.array("_row_arr")
    real = ra[ra < m]
    nonempty = np.nonzero(np.diff(rst))[0]
    assert set(real.tolist()) == set(nonempty.tolist())

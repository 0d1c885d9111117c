"""LDS-sliced CSR image, built in host memory (no GPU): x cut into slices of
W = 16384 columns, the entries slice-major with 16-bit columns inside the
slice (bit 15 on the first entry of each pair), one pair per (row, slice)
with entries, the workgroup units (an equal share of the entries per
workgroup), and the combine blocks whose partials are ordered by
(slice, row).  The product is emulated from the uploaded arrays exactly
as the two kernels read them and compared with the oracle; the structure
is checked against its definition."""
from __future__ import annotations

import numpy as np
import pytest

import rvvlasparse_amd as R
from helpers import assert_within_bound, laplacian, rand_csr
from oracle import oracle as O

CUS, W = 256, 16384


@pytest.fixture(autouse=True)
def host_validation():
    R.check(R.lib.rvvla_set_host_validation(1))
    yield
    R.check(R.lib.rvvla_set_host_validation(0))


def _view(M, dtype):
    vt = np.float64 if dtype == R.F64 else np.float32
    v = {}
    for f in ("pst", "wpos", "rowptr", "cblk", "units", "wg"):
        a = M.view("xl." + f)
        v[f] = a.view(np.int32).astype(np.int64) if a.size else np.zeros(0, np.int64)
    raw = M.view("xl.col").astype(np.int64)
    v["col"], v["head"] = raw & 0x7FFF, (raw >> 15) & 1
    v["lpos"] = M.view("xl.lpos").astype(np.int64)
    val = M.view("xl.val")
    v["val"] = val.view(vt) if val.size else np.zeros(0, vt)
    return v


def _comb(dtype):
    """Combine block limits: 256 rows / 2048 partials (FP64), 512 / 4096."""
    return (256, 2048) if dtype == R.F64 else (512, 4096)


def _emulate(v, m, n, W, x, alpha, beta, y0, COMB_PAIRS):
    """The slice pass (a partial per pair into its slot) and the combine
    (partials to their row-major places, each row's added in order)."""
    npairs = len(v["wpos"])
    part = np.full(max(npairs, 1), np.nan)
    units = v["units"].reshape(-1, 3)
    for p, g0, g1 in units:
        xs = x[p * W:min(n, p * W + W)]
        for g in range(g0, g1):
            lo, hi = v["pst"][g], v["pst"][g + 1]
            assert np.all(v["col"][lo:hi] < len(xs)), "column outside its slice"
            part[v["wpos"][g]] = np.dot(v["val"][lo:hi].astype(np.float64), xs[v["col"][lo:hi]])
    assert not np.isnan(part[:npairs]).any(), "every slot written"
    rp = v["rowptr"]
    y = np.zeros(m)
    for r0, r1, base, cnt in v["cblk"].reshape(-1, 4):
        assert base == rp[r0] and cnt == rp[r1] - rp[r0]
        if cnt > COMB_PAIRS:
            assert r1 == r0 + 1
            y[r0] = part[base:base + cnt].sum()
            continue
        prod = np.full(cnt, np.nan)
        prod[v["lpos"][base:base + cnt]] = part[base:base + cnt]
        for r in range(r0, r1):
            y[r] = prod[rp[r] - base:rp[r + 1] - base].sum()
    return alpha * y + beta * y0


CASES = {
    # uniform columns over a wide x: many slices, short pairs
    "wide_random": lambda: (*rand_csr(2000, 200000, 0.0004, 3, empty_every=11), 200000),
    # long rows (35000 entries) crossing every slice
    "long_rows": lambda: (*rand_csr(300, 70000, 0.01, 5, long_rows=(7, 120, 250)), 70000),
    # a single slice
    "lap32": lambda: (*laplacian(32), 1024),
}


def _many_pairs():
    """One row with 9000 pairs (one entry in each of 9000 slices) among
    short rows: a combine block of its own (more than 4096 partials),
    summed in place."""
    n = 9000 * W
    cols = [np.array([5]), np.arange(9000, dtype=np.int64) * W + 7, np.array([n - 1])]
    rst = np.array([0, 1, 9001, 9002], np.int32)
    return rst, np.concatenate(cols).astype(np.int32), np.linspace(-1, 1, 9002), n


@pytest.mark.parametrize("name", list(CASES) + ["many_pairs"])
@pytest.mark.parametrize("dtype", [R.F64, R.F32])
def test_xlds_layout_and_product(name, dtype):
    rst, col, val, n = CASES[name]() if name != "many_pairs" else _many_pairs()
    a = R.CSR(len(rst) - 1, n, rst, col, val)
    M = R.convert("csr", a, dtype, "xlds=1").upload()
    v = _view(M, dtype)
    m, npairs = a.m, len(v["wpos"])
    COMB_ROWS, COMB_PAIRS = _comb(dtype)
    nwg = CUS  # one workgroup per CU (its LDS holds one x slice and the accumulators)
    # entries and pairs
    pst = v["pst"]
    assert pst[0] == 0 and pst[-1] == a.nnz and np.all(np.diff(pst) > 0)
    heads = np.zeros(a.nnz, np.int64)
    heads[pst[:-1]] = 1
    assert np.array_equal(v["head"], heads), "bit 15 marks exactly the first entry of each pair"
    assert sorted(v["wpos"].tolist()) == list(range(npairs))
    # units: contiguous per workgroup, tiling every pair, slices ascending
    units = v["units"].reshape(-1, 3)
    wg = v["wg"]
    assert len(wg) == nwg + 1 and wg[0] == 0 and wg[-1] == len(units) and np.all(np.diff(wg) >= 0)
    assert units[0, 1] == 0 and units[-1, 2] == npairs and np.all(units[1:, 1] == units[:-1, 2])
    assert np.all(np.diff(units[:, 0]) >= 0) and np.all(units[:, 2] > units[:, 1])
    share = [pst[units[wg[w + 1] - 1, 2]] - pst[units[wg[w], 1]] for w in range(nwg) if wg[w + 1] > wg[w]]
    longest = np.diff(pst).max()
    assert max(share) <= -(-a.nnz // nwg) + longest, "an equal share of the entries per workgroup"
    # combine blocks: cover the rows in order, within the limits
    cb = v["cblk"].reshape(-1, 4)[:, :2]
    rp = v["rowptr"]
    assert cb[0, 0] == 0 and cb[-1, 1] == m and np.all(cb[1:, 0] == cb[:-1, 1])
    for r0, r1 in cb:
        cnt = rp[r1] - rp[r0]
        assert (r1 - r0 <= COMB_ROWS and cnt <= COMB_PAIRS) or (r1 == r0 + 1)
        if cnt <= COMB_PAIRS:
            assert sorted(v["lpos"][rp[r0]:rp[r1]].tolist()) == list(range(cnt))
    # each pair's entries are its row's entries in its slice; the slots of a
    # block are ordered by (slice, row)
    slice_of_pair = np.repeat(units[:, 0], units[:, 2] - units[:, 1])
    slot_pair = np.empty(npairs, np.int64)
    slot_pair[v["wpos"]] = np.arange(npairs)
    for r0, r1 in cb:
        base, cnt = rp[r0], rp[r1] - rp[r0]
        if cnt > COMB_PAIRS or cnt == 0:
            continue
        place = v["lpos"][base:base + cnt]  # row-major place of each slot
        row_of_place = np.repeat(np.arange(r0, r1), np.diff(rp[r0:r1 + 1]))
        rows = row_of_place[place]
        sl = slice_of_pair[slot_pair[base:base + cnt]]
        key = sl * (m + 1) + rows
        assert np.all(np.diff(key) > 0), "slots ordered by (slice, row)"
        for k in range(0, cnt, max(1, cnt // 20)):
            r, p = rows[k], sl[k]
            g = slot_pair[base + k]
            c = col[rst[r]:rst[r + 1]]
            assert np.array_equal(v["col"][pst[g]:pst[g + 1]], c[(c >= p * W) & (c < (p + 1) * W)] - p * W)
    # product
    rng = np.random.default_rng(5)
    dt = np.float64 if dtype == R.F64 else np.float32
    x = rng.uniform(-1, 1, n).astype(dt)
    y0 = rng.uniform(-1, 1, m).astype(dt)
    vals = a.values_for(dtype)
    y = _emulate(v, m, n, W, x.astype(np.float64), 1.25, -0.5, y0.astype(np.float64), COMB_PAIRS)
    y_ref = O.csr_spmv(a.rst, a.col, vals, x, 1.25, -0.5, y0)
    assert_within_bound(y.astype(dt), y_ref, a.rst, a.col, vals, x, 1.25, -0.5, y0, dt,
                        f"xlds {name}")


def test_xlds_empty_matrix():
    a = R.CSR(5, 7, np.zeros(6, np.int32), np.zeros(0, np.int32), np.zeros(0))
    M = R.convert("csr", a, R.F64, "xlds=1").upload()
    assert M.view("xl.wpos").size == 0
    assert np.array_equal(M.view("xl.rowptr").view(np.int32), np.zeros(6, np.int32))
    assert np.array_equal(M.view("xl.cblk").view(np.int32), np.array([0, 5, 0, 0], np.int32))

"""LDS-sliced CSR image, built in host memory (no GPU): x cut into slices of
W columns (16384 doubles or 32768 floats: 128 KiB of LDS), the entries
slice-major with 16-bit columns inside the slice, each slice's entries
cut into tasks of 512 entries (512 slots each, values word-interleaved), bit 15 on the
first entry of every piece (a (row, slice) pair's entries inside one task),
one partial slot per piece in stream order, the workgroup units (an equal
share of the entries per workgroup), the combine blocks with their runs of
slots per slice and the row-major place of each slot.  The product is
emulated from the uploaded arrays as the two kernels read them and compared
with the oracle; the structure is checked against its definition."""
from __future__ import annotations

import numpy as np
import pytest

import rvvlasparse_amd as R
from helpers import assert_within_bound, laplacian, rand_csr
from oracle import oracle as O

CUS, TASK = 256, 512


@pytest.fixture(autouse=True)
def host_validation():
    R.check(R.lib.rvvla_set_host_validation(1))
    yield
    R.check(R.lib.rvvla_set_host_validation(0))


def _w(dtype):
    return 16384 if dtype == R.F64 else 32768


def _cap(dtype, cwg=1024):
    """Combine blocks for cwg-thread workgroups: at most 2 cwg rows and
    64 cwg bytes of partials (FP64 partials for FP64, FP32 otherwise)."""
    return 2 * cwg, 64 * cwg // (8 if dtype == R.F64 else 4)


def _view(M, dtype):
    vt = np.float64 if dtype == R.F64 else np.float32
    v = {}
    for f in ("task", "units", "wg", "rowptr", "cblk", "off"):
        a = M.view("xl." + f)
        v[f] = a.view(np.int32).astype(np.int64) if a.size else np.zeros(0, np.int64)
    raw = M.view("xl.col").astype(np.int64)
    v["col"], v["mark"] = raw & 0x7FFF, (raw >> 15) & 1
    v["lpos"] = M.view("xl.lpos").astype(np.int64)
    val = M.view("xl.val")
    val = val.view(vt) if val.size else np.zeros(0, vt)
    # inside each task the values are word-interleaved: entry 8 l + k of a
    # task at slot ((k // vpw) * 64 + l) * vpw + k % vpw
    vpw = 2 if dtype == R.F64 else 4
    at = np.arange(len(val))
    r = at % TASK
    lane, k = r // 8, r % 8
    v["val"] = val[at - r + ((k // vpw) * 64 + lane) * vpw + k % vpw] if len(val) else val
    return v


def _slice_pass(v, n, W, x):
    """Each task's pieces summed into their slots (slot = marks before)."""
    tasks = v["task"].reshape(-1, 4)
    npieces = len(v["lpos"])
    part = np.full(max(npieces, 1), np.nan)
    for eb, ec, pc0, p in tasks:
        assert eb % TASK == 0 and 0 < ec <= TASK
        xs = x[p * W:min(n, p * W + W)]
        cols, vals, marks = v["col"][eb:eb + ec], v["val"][eb:eb + ec], v["mark"][eb:eb + ec]
        assert marks[0] == 1, "every task starts a piece"
        assert np.all(cols < len(xs)), "column outside its slice"
        prod = vals.astype(np.float64) * xs[cols]
        starts = np.flatnonzero(marks)
        sums = np.add.reduceat(prod, starts)
        assert np.all(np.isnan(part[pc0:pc0 + len(starts)])), "each slot written once"
        part[pc0:pc0 + len(starts)] = sums
    assert not np.isnan(part[:npieces]).any(), "every slot written"
    return part


def _combine(v, m, S, part, alpha, beta, y0, cap):
    rp, off = v["rowptr"], v["off"].reshape(-1, S) if S else None
    y = np.zeros(m)
    for b, (r0, r1) in enumerate(v["cblk"].reshape(-1, 2)):
        base, cnt = rp[r0], rp[r1] - rp[r0]
        runs = [np.arange(off[b, p], off[b + 1, p]) for p in range(S)]
        slots = np.concatenate(runs) if runs else np.zeros(0, np.int64)
        assert len(slots) == cnt, "the runs hold the block's pieces"
        if cnt > cap:
            assert r1 == r0 + 1
            y[r0] = part[slots].sum()
            continue
        prod = np.full(cnt, np.nan)
        prod[v["lpos"][slots]] = part[slots]
        for r in range(r0, r1):
            y[r] = prod[rp[r] - base:rp[r + 1] - base].sum()
    return alpha * y + beta * y0


CASES = {
    # uniform columns over a wide x: many slices, short pairs
    "wide_random": lambda: (*rand_csr(2000, 200000, 0.0004, 3, empty_every=11), 200000),
    # long rows (35000 entries) crossing every slice, pieces cut by tasks
    "long_rows": lambda: (*rand_csr(300, 70000, 0.01, 5, long_rows=(7, 120, 250)), 70000),
    # a single slice
    "lap32": lambda: (*laplacian(32), 1024),
}


def _many_pieces(W):
    """One row with 9000 pairs (one entry in each of 9000 slices) among
    short rows: more pieces than a combine block holds, summed in place."""
    n = 9000 * W
    cols = [np.array([5]), np.arange(9000, dtype=np.int64) * W + 7, np.array([n - 1])]
    rst = np.array([0, 1, 9001, 9002], np.int32)
    return rst, np.concatenate(cols).astype(np.int32), np.linspace(-1, 1, 9002), n


def _check_structure(a, v, dtype, rst, col, cwg=1024):
    W = _w(dtype)
    S = (a.n + W - 1) // W
    m, nnz = a.m, a.nnz
    rows_cap, cap = _cap(dtype, cwg)
    tasks = v["task"].reshape(-1, 4)
    npieces = len(v["lpos"])
    marks = v["mark"]
    # tasks: per slice, consecutive 512-entry cuts of the slice's segment
    for p in range(S):
        tp = tasks[tasks[:, 3] == p]
        if len(tp) == 0:
            continue
        assert tp[0, 0] % TASK == 0 and np.all(np.diff(tp[:, 0]) == TASK) and np.all(tp[:-1, 1] == TASK)
    assert tasks[:, 1].sum() == nnz, "every entry in one task"
    assert np.array_equal(tasks[:, 2], np.concatenate([[0], np.cumsum(
        [marks[eb:eb + ec].sum() for eb, ec, _, _ in tasks])[:-1]])), "slots numbered by marks"
    assert npieces == sum(marks[eb:eb + ec].sum() for eb, ec, _, _ in tasks)
    in_task = np.zeros(len(marks), bool)
    for eb, ec, _, _ in tasks:
        in_task[eb:eb + ec] = True
    assert not marks[~in_task].any() and np.all(v["col"][~in_task] == 0), "padding is zero"
    assert np.all(v["val"][~in_task] == 0)
    # each slice's entries are its rows' entries in CSR order
    for p in range(min(S, 4)):
        tp = tasks[tasks[:, 3] == p]
        got = np.concatenate([v["col"][eb:eb + ec] for eb, ec, _, _ in tp]) if len(tp) else np.zeros(0)
        exp = [c[(c >= p * W) & (c < (p + 1) * W)] - p * W for c in
               (col[rst[r]:rst[r + 1]] for r in range(m))]
        assert np.array_equal(got, np.concatenate(exp) if exp else np.zeros(0))
    # units: contiguous per workgroup, tiling every task, slices ascending
    units = v["units"].reshape(-1, 3)
    wg = v["wg"]
    # FP64 splits its tasks over four workgroups a CU, the 32-bit-x forms one
    nwg = CUS * (4 if dtype == R.F64 else 1)
    assert len(wg) == nwg + 1 and wg[0] == 0 and wg[-1] == len(units) and np.all(np.diff(wg) >= 0)
    assert units[0, 1] == 0 and units[-1, 2] == len(tasks) and np.all(units[1:, 1] == units[:-1, 2])
    assert np.all(units[:, 2] > units[:, 1])
    for u0, u1 in zip(wg[:-1], wg[1:]):
        assert np.all(np.diff(units[u0:u1, 0]) > 0), "one unit per slice in a workgroup"
    for p, t0, t1 in units:
        assert np.all(tasks[t0:t1, 3] == p)
    share = [tasks[units[wg[w], 1]:units[wg[w + 1] - 1, 2], 1].sum() for w in range(nwg) if wg[w + 1] > wg[w]]
    assert max(share) <= -(-nnz // nwg) + TASK, "an equal share of the entries per workgroup"
    # combine blocks: cover the rows in order, within the limits; the places
    cb = v["cblk"].reshape(-1, 2)
    rp = v["rowptr"]
    assert rp[0] == 0 and rp[-1] == npieces and np.all(np.diff(rp) >= 0)
    assert cb[0, 0] == 0 and cb[-1, 1] == m and np.all(cb[1:, 0] == cb[:-1, 1])
    off = v["off"].reshape(-1, S)
    assert off.shape[0] == len(cb) + 1 and np.all(np.diff(off, axis=0) >= 0)
    for b, (r0, r1) in enumerate(cb):
        cnt = rp[r1] - rp[r0]
        assert (r1 - r0 <= rows_cap and cnt <= cap) or (r1 == r0 + 1)
        if cnt <= cap:
            slots = np.concatenate([np.arange(off[b, p], off[b + 1, p]) for p in range(S)])
            assert sorted(v["lpos"][slots].tolist()) == list(range(cnt))


@pytest.mark.parametrize("cwg", [256, 512, 1024])
@pytest.mark.parametrize("name", list(CASES) + ["many_pieces"])
@pytest.mark.parametrize("dtype", [R.F64, R.F32])
def test_xlds_layout_and_product(name, dtype, cwg):
    W = _w(dtype)
    rst, col, val, n = CASES[name]() if name != "many_pieces" else _many_pieces(W)
    a = R.CSR(len(rst) - 1, n, rst, col, val)
    params = "xlds=1" if cwg == 1024 else f"xlds=1,xlcomb={cwg}"
    M = R.convert("csr", a, dtype, params).upload()
    v = _view(M, dtype)
    _check_structure(a, v, dtype, rst, col, cwg)
    rng = np.random.default_rng(5)
    dt = np.float64 if dtype == R.F64 else np.float32
    x = rng.uniform(-1, 1, n).astype(dt)
    y0 = rng.uniform(-1, 1, a.m).astype(dt)
    vals = a.values_for(dtype)
    S = (n + W - 1) // W
    part = _slice_pass(v, n, W, x.astype(np.float64))
    y = _combine(v, a.m, S, part, 1.25, -0.5, y0.astype(np.float64), _cap(dtype, cwg)[1])
    y_ref = O.csr_spmv(a.rst, a.col, vals, x, 1.25, -0.5, y0)
    assert_within_bound(y.astype(dt), y_ref, a.rst, a.col, vals, x, 1.25, -0.5, y0, dt,
                        f"xlds {name}")


def test_xlds_pieces_cut_by_tasks():
    """A 2000-entry row inside one slice is one pair cut into pieces at the
    512-entry task boundaries; each piece's first entry is marked."""
    n = 16384
    rst = np.array([0, 3, 2003, 2004], np.int32)
    col = np.concatenate([[0, 1, 2], np.arange(2000) * 8, [n - 1]]).astype(np.int32)
    a = R.CSR(3, n, rst, col, np.linspace(-1, 1, len(col)))
    v = _view(R.convert("csr", a, R.F64, "xlds=1").upload(), R.F64)
    tasks = v["task"].reshape(-1, 4)
    assert np.array_equal(tasks[:, :2], [[0, 512], [512, 512], [1024, 512], [1536, 468]])
    # pieces: row 0 (1), row 1 cut at 512, 1024, 1536 (4), row 2 (1)
    assert np.array_equal(v["rowptr"], [0, 1, 5, 6])
    assert np.flatnonzero(v["mark"][:2004]).tolist() == [0, 3, 512, 1024, 1536, 2003]


def test_xlds_combine_width_refused():
    rst, col, val = laplacian(32)
    a = R.CSR(len(rst) - 1, 1024, rst, col, val)
    with pytest.raises(R.RvvlaError):
        R.convert("csr", a, R.F64, "xlds=1,xlcomb=300").upload()


def test_xlds_empty_matrix():
    a = R.CSR(5, 7, np.zeros(6, np.int32), np.zeros(0, np.int32), np.zeros(0))
    M = R.convert("csr", a, R.F64, "xlds=1").upload()
    assert M.view("xl.lpos").size == 0 and M.view("xl.task").size == 0
    assert np.array_equal(M.view("xl.rowptr").view(np.int32), np.zeros(6, np.int32))
    assert np.array_equal(M.view("xl.cblk").view(np.int32), np.array([0, 5], np.int32))

"""GPU parity: every format and precision against the CPU oracle on the same
seeded inputs, through the C-ABI (librvvla_hip.so).

FP64 / FP32: |y - y_oracle| <= 4*(len+2)*u*(|alpha|*sum|a||x| + |beta|*|y0|)
per row (u = 2^-53 / 2^-24).  16-bit values (FP32 x, y and accumulation):
per row |y - y_fp64| <= (u_v + (4*len+10)*2^-24) * (|alpha|*sum|a||x| +
|beta|*|y0|) against the FP64 oracle on the unrounded matrix, u_v = 2^-8
(bf16, 8-bit significand) or 2^-11 (fp16, 11-bit significand), the unit
roundoff of round-to-nearest value storage, plus FP32 arithmetic.
"""
from __future__ import annotations

import numpy as np
import pytest

import rvvlasparse_amd as R
from helpers import assert_within_bound, laplacian, rand_csr
from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

FORMATS_READY = ["csr", "hcsr", "sell", "csr5", "lav", "vnec", "vhcc", "cvr"]


def _device():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch.device("cuda:0")


def _csr(t, n=None):
    rst, col, val = t
    return R.CSR(len(rst) - 1, n if n is not None else int(col.max()) + 1 if len(col) else 1,
                 rst, col, val)


_CACHE: dict = {}


def case(name) -> R.CSR:
    if name not in _CACHE:
        if name == "rand_long":
            _CACHE[name] = _csr(rand_csr(3000, 2600, 0.004, 1, empty_every=13,
                                         long_rows=(5, 1700, 1701)), 2600)
        elif name == "lap64":
            _CACHE[name] = _csr(laplacian(64), 4096)
        elif name == "powerlaw":
            _CACHE[name] = R.gen_powerlaw(30000, 1.8, 5000, 450000, 256, seed=2)
        elif name == "random16_c1":
            _CACHE[name] = R.gen_random_rows(65536, 16, seed=0)
        elif name == "empty":
            _CACHE[name] = R.CSR(5, 7, np.zeros(6, np.int32), np.zeros(0, np.int32), np.zeros(0))
        elif name == "one_long_row":
            rng = np.random.default_rng(4)
            col = np.sort(rng.choice(100000, 60000, replace=False)).astype(np.int32)
            _CACHE[name] = R.CSR(1, 100000, np.array([0, 60000], np.int32), col,
                                 rng.normal(size=60000))
        else:
            raise KeyError(name)
    return _CACHE[name]


CASES = ["rand_long", "lap64", "powerlaw", "random16_c1", "empty", "one_long_row"]


def run_gpu(fmt, dtype, a, x64, y064, alpha, beta, params=""):
    dev = _device()
    L = R.convert(fmt, a, dtype, params)
    M = L.upload()
    io = torch.float64 if dtype == R.F64 else torch.float32
    x = torch.from_numpy(x64).to(io).to(dev)
    y = torch.from_numpy(y064).to(io).to(dev)
    M.spmv(x, y, alpha, beta)
    torch.cuda.synchronize()
    return y.cpu().numpy()


def inputs(a, seed=0):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1, 1, a.n), rng.uniform(-1, 1, a.m)


@pytest.mark.parametrize("fmt", FORMATS_READY)
@pytest.mark.parametrize("dtype", [R.F64, R.F32])
@pytest.mark.parametrize("name", CASES)
def test_spmv_drop_in_precisions(fmt, dtype, name):
    a = case(name)
    x64, y064 = inputs(a, 1)
    dt = np.float64 if dtype == R.F64 else np.float32
    for alpha, beta in ((1.0, 0.0), (0.5, -1.5)):
        y = run_gpu(fmt, dtype, a, x64, y064, alpha, beta)
        v, x, y0 = a.values_for(dtype), x64.astype(dt), y064.astype(dt)
        y_ref = O.csr_spmv(a.rst, a.col, v, x, alpha, beta, y0)
        assert y.dtype == dt
        assert_within_bound(y, y_ref, a.rst, a.col, v, x, alpha, beta, y0, dt,
                            f"{fmt}/{name}/{dt.__name__}/a={alpha},b={beta}")


@pytest.mark.parametrize("fmt", FORMATS_READY)
@pytest.mark.parametrize("dtype", [R.BF16, R.F16])
@pytest.mark.parametrize("name", ["rand_long", "powerlaw", "lap64", "one_long_row"])
def test_spmv_16bit_values(fmt, dtype, name):
    a = case(name)
    x64, y064 = inputs(a, 2)
    alpha, beta = 1.0, 0.5
    y = run_gpu(fmt, dtype, a, x64, y064, alpha, beta)
    # the FP64 oracle on the unrounded FP64 matrix and the FP32-rounded x, y0
    x32, y032 = x64.astype(np.float32), y064.astype(np.float32)
    y64 = O.csr_spmv(a.rst, a.col, a.val, x32.astype(np.float64), alpha, beta,
                     y032.astype(np.float64))
    uv = 2.0 ** -8 if dtype == R.BF16 else 2.0 ** -11
    ln = np.diff(a.rst).astype(np.float64)
    scale = abs(alpha) * O.csr_absdot(a.rst, a.col, a.val, x32.astype(np.float64)) \
        + abs(beta) * np.abs(y032.astype(np.float64))
    bound = (uv + (4 * ln + 10) * 2.0 ** -24) * scale
    err = np.abs(y.astype(np.float64) - y64)
    assert np.all(err <= bound), f"{fmt}/{name}: max err/bound {np.max(err / np.maximum(bound, 1e-300)):.3f}"
    if np.linalg.norm(y64) > 0:
        rel = np.linalg.norm(y - y64) / np.linalg.norm(y64)
        assert rel <= (5e-3 if dtype == R.BF16 else 1e-3), rel


@pytest.mark.parametrize("P", [3, 8, 16])
@pytest.mark.parametrize("dtype", [R.F64, R.F32])
@pytest.mark.parametrize("name", ["rand_long", "powerlaw", "one_long_row", "empty", "lap64"])
def test_csr_column_slices(P, dtype, name):
    """The column-sliced CSR image (one slice per XCD, partials combined in
    slice order) against the oracle, with and without beta."""
    a = case(name)
    x64, y064 = inputs(a, 6)
    dt = np.float64 if dtype == R.F64 else np.float32
    for alpha, beta in ((1.0, 0.0), (-0.5, 2.0)):
        y = run_gpu("csr", dtype, a, x64, y064, alpha, beta, f"xslices={P}")
        v, x, y0 = a.values_for(dtype), x64.astype(dt), y064.astype(dt)
        y_ref = O.csr_spmv(a.rst, a.col, v, x, alpha, beta, y0)
        assert_within_bound(y, y_ref, a.rst, a.col, v, x, alpha, beta, y0, dt,
                            f"csr xslices={P}/{name}/a={alpha},b={beta}")


@pytest.mark.parametrize("dtype", [R.BF16, R.F16])
def test_csr_column_slices_16bit(dtype):
    a = case("powerlaw")
    x64, y064 = inputs(a, 7)
    y = run_gpu("csr", dtype, a, x64, y064, 1.0, 0.5, "xslices=8")
    y_plain = run_gpu("csr", dtype, a, x64, y064, 1.0, 0.5, "xslices=0")
    x32, y032 = x64.astype(np.float32), y064.astype(np.float32)
    y64 = O.csr_spmv(a.rst, a.col, a.val, x32.astype(np.float64), 1.0, 0.5,
                     y032.astype(np.float64))
    for out in (y, y_plain):
        rel = np.linalg.norm(out - y64) / np.linalg.norm(y64)
        assert rel <= (5e-3 if dtype == R.BF16 else 1e-3), rel


def _compensation_case(dtype):
    """A 50,000-entry row (row 217) among rows of 1..3 entries: its first
    half of products positive, its second half negative, so a left-to-right
    FP32 sum climbs past 2^15 and rounds on every addition there.  Values
    have 7 fraction bits and x 2, on [1, 2): every product is a multiple of
    2^-9 below 4, exact in FP32 and in the 16-bit storage, so the FP64 sum
    is exact (the oracle's row is the exact sum), and so is the FP32 sum of
    any 8191 or fewer of them: a compensated sum whose partials each cover
    at most that many entries is exact, a plain one is not."""
    del dtype  # the same values are exact in BF16 and FP16
    rng = np.random.default_rng(31)
    m, n, big = 600, 120000, 50000
    lens = rng.integers(1, 4, m)
    lens[217] = big
    cols = [np.sort(rng.choice(n, k, replace=False)) for k in lens]
    rst = np.zeros(m + 1, np.int32)
    rst[1:] = np.cumsum(lens)
    col = np.concatenate(cols).astype(np.int32)
    v = np.round(rng.uniform(1.0, 2.0, len(col)) * 128) / 128
    sign = rng.choice([-1.0, 1.0], len(col))
    e0 = rst[217]
    sign[e0:e0 + big // 2] = 1.0
    sign[e0 + big // 2:e0 + big] = -1.0
    a = R.CSR(m, n, rst, col, v * sign)
    x = np.round(rng.uniform(1.0, 2.0, n) * 4) / 4
    return a, x


def _compensated_bound(y_exact, absdot):
    """One rounding of the result, plus a margin far below one rounding of
    a plain sum's running value: 4 u |y| + 2^-32 sum|a||x| (u = 2^-24)."""
    return 4 * 2.0 ** -24 * np.abs(y_exact) + 2.0 ** -32 * absdot


@pytest.mark.parametrize("dtype", [R.BF16, R.F16])
@pytest.mark.parametrize("fmt,params", [("csr", ""), ("csr", "xlds=1"), ("csr", "xslices=8"),
                                        ("csr", "sell=1"),
                                        ("hcsr", ""),
                                        ("sell", ""), ("csr5", ""), ("lav", ""), ("vnec", ""),
                                        ("vhcc", ""), ("cvr", "")])
def test_16bit_long_row_compensated(fmt, params, dtype):
    """Rows longer than 1024 entries are summed with TwoSum compensation in
    the 16-bit value forms: the long row's error against its exact sum stays
    within _compensated_bound, which a plain FP32 left-to-right sum of the
    row, or of its 512-entry partials, misses by far (next test); every row
    also meets 4 * 2^-24 * sum|a||x|."""
    a, x = _compensation_case(dtype)
    print("16-bit compensation", fmt, params, dtype, flush=True)
    y0 = np.zeros(a.m)
    y = run_gpu(fmt, dtype, a, x, y0, 1.0, 0.0, params)
    y_ref = O.csr_spmv(a.rst, a.col, a.val, x, 1.0, 0.0)
    scale = O.csr_absdot(a.rst, a.col, a.val, x)
    err = np.abs(y.astype(np.float64) - y_ref)
    lim = _compensated_bound(y_ref[217], scale[217])
    assert err[217] <= lim, f"{fmt} {params}: long-row error {err[217]:.3g} > {lim:.3g}"
    assert np.all(err <= 4 * 2.0 ** -24 * scale)


def test_16bit_compensation_bound_power():
    """The bound of the compensation test against FP32 summations of its
    long row: a plain left-to-right sum (a per-row loop without TwoSum)
    misses it by more than 1000x, and a plain left-to-right sum of exact
    512-entry partials by more than 10x; TwoSum-compensated sums of exact
    512-entry or 8-entry partials meet it."""
    a, x = _compensation_case(R.BF16)
    e0, e1 = a.rst[217], a.rst[218]
    t = a.val[e0:e1] * x[a.col[e0:e1]]
    t32 = t.astype(np.float32)
    assert np.array_equal(t32.astype(np.float64), t), "products exact in FP32"
    v32 = a.val.astype(np.float32)
    assert not np.any(v32.view(np.uint32) & 0xFFFF), "values exact in BF16"
    assert np.array_equal(v32.astype(np.float16).astype(np.float64), a.val), "and in FP16"
    exact = t.sum()
    lim = _compensated_bound(exact, np.abs(t).sum())
    seq = np.float32(0)
    for q in t32:
        seq = np.float32(seq + q)
    assert abs(float(seq) - exact) > 1000 * lim
    for chunk in (512, 8):
        parts = np.array([np.float32(t[i:i + chunk].sum()) for i in range(0, len(t), chunk)])
        plain = np.float32(0)
        for q in parts:
            plain = np.float32(plain + q)
        if chunk == 512:
            assert abs(float(plain) - exact) > 10 * lim
        comp = np.float32(parts.astype(np.float64).sum())
        assert abs(float(comp) - exact) <= lim


@pytest.mark.parametrize("name", CASES)
@pytest.mark.parametrize("params", ["sell=1", "sell=0"])
@pytest.mark.parametrize("dtype", [R.F64, R.F32, R.BF16, R.F16])
def test_csr_sell_image(name, params, dtype):
    """CSR products with and without the matrix's SELL-64 image, against
    the oracle."""
    a = case(name)
    x64, y064 = inputs(a, 15)
    dt = np.float64 if dtype == R.F64 else np.float32
    y = run_gpu("csr", dtype, a, x64, y064, 1.0, 0.5, params)
    if dtype in (R.BF16, R.F16):
        x32, y032 = x64.astype(np.float32), y064.astype(np.float32)
        y64 = O.csr_spmv(a.rst, a.col, a.val, x32.astype(np.float64), 1.0, 0.5, y032.astype(np.float64))
        if np.linalg.norm(y64) > 0:
            rel = np.linalg.norm(y - y64) / np.linalg.norm(y64)
            assert rel <= (5e-3 if dtype == R.BF16 else 1e-3), rel
        return
    v, x, y0 = a.values_for(dtype), x64.astype(dt), y064.astype(dt)
    y_ref = O.csr_spmv(a.rst, a.col, v, x, 1.0, 0.5, y0)
    assert_within_bound(y, y_ref, a.rst, a.col, v, x, 1.0, 0.5, y0, dt, f"csr {params} {name}")


@pytest.mark.parametrize("dtype", [R.F64, R.F32])
def test_csr_sell_image_nonfinite_x(dtype):
    """The SELL image of CSR reads only each row's own entries: with x[0] =
    Inf, an empty row (whose padding points at column 0) gives beta*y and a
    short row of a wider slice whose last column holds Inf gives that Inf,
    as the reference CSR product does (no 0*Inf NaN from padding)."""
    n = 64
    rows = [[], [0], [3, 5], [1, 2, 7, 9, 11]] + [[(r * 7) % n, (r * 7 + 1) % n] for r in range(60)]
    rst = np.zeros(len(rows) + 1, np.int32)
    rst[1:] = np.cumsum([len(r) for r in rows])
    col = np.concatenate([np.sort(np.array(r, np.int64)) for r in rows if r]).astype(np.int32)
    a = R.CSR(len(rows), n, rst, col, np.linspace(0.5, 1.5, len(col)))
    x64 = np.linspace(-1, 1, n)
    x64[0] = np.inf
    y064 = np.full(a.m, 2.0)
    dt = np.float64 if dtype == R.F64 else np.float32
    y = run_gpu("csr", dtype, a, x64, y064, 1.0, 0.5, "sell=1")
    y_ref = O.csr_spmv(a.rst, a.col, a.values_for(dtype), x64.astype(dt), 1.0, 0.5, y064.astype(dt))
    assert np.array_equal(np.isnan(y), np.isnan(y_ref)) and not np.isnan(y).any()
    assert y[0] == 1.0 and np.isinf(y[1])
    fin = np.isfinite(y_ref)
    assert np.array_equal(np.isinf(y), np.isinf(y_ref))
    assert np.allclose(y[fin], y_ref[fin])


def _nonfinite_case():
    """3000 x 3000, no empty rows: 1..40 entries a row, columns drawn from
    1..n-1 except that every 97th row also holds column 0; two rows of 1700
    entries (long-row paths) that avoid column 0; positive values."""
    rng = np.random.default_rng(21)
    m = n = 3000
    rows = []
    for r in range(m):
        k = 1700 if r in (500, 2222) else int(rng.integers(1, 41))
        c = rng.choice(np.arange(1, n), k, replace=False)
        if r % 97 == 0:
            c[0] = 0
        rows.append(np.sort(c))
    rst = np.zeros(m + 1, np.int32)
    rst[1:] = np.cumsum([len(c) for c in rows])
    col = np.concatenate(rows).astype(np.int32)
    return R.CSR(m, n, rst, col, rng.uniform(0.5, 1.5, len(col)))


@pytest.mark.parametrize("fmt", FORMATS_READY)
@pytest.mark.parametrize("dtype", [R.F64, R.F32])
def test_nonfinite_x_all_formats(fmt, dtype):
    """x[0] = +Inf: the rows that hold column 0 (positive values) give +Inf
    and every other row is finite and within the bound of the oracle, as in
    the reference CSR product: no format lets a padding slot, a lane past its
    stream or a batch past its partition multiply the Inf (0 * Inf = NaN)."""
    if "nonfinite" not in _CACHE:
        _CACHE["nonfinite"] = _nonfinite_case()
    a = _CACHE["nonfinite"]
    x64, _ = inputs(a, 5)
    x64[0] = np.inf
    dt = np.float64 if dtype == R.F64 else np.float32
    touches = np.zeros(a.m, bool)
    rows = np.repeat(np.arange(a.m), np.diff(a.rst))
    touches[rows[a.col == 0]] = True
    for alpha, beta in ((1.0, 0.0), (0.5, -1.5)):
        y064 = np.random.default_rng(6).uniform(-1, 1, a.m)
        y = run_gpu(fmt, dtype, a, x64, y064, alpha, beta)
        assert not np.isnan(y).any(), f"{fmt}: NaN in rows {np.flatnonzero(np.isnan(y))[:8]}"
        assert np.array_equal(np.isposinf(y), touches), f"{fmt}: Inf rows differ"
        v, x, y0 = a.values_for(dtype), x64.astype(dt), y064.astype(dt)
        keep = ~touches
        xf = x.copy()
        xf[0] = 0  # the rows kept never read it
        y_ref = O.csr_spmv(a.rst, a.col, v, xf, alpha, beta, y0)
        sub = np.flatnonzero(keep)
        rs = np.concatenate([[0], np.cumsum(np.diff(a.rst)[sub])]).astype(np.int32)
        idx = np.concatenate([np.arange(a.rst[r], a.rst[r + 1]) for r in sub])
        assert_within_bound(y[sub], y_ref[sub], rs, a.col[idx], v[idx], xf, alpha, beta, y0[sub],
                            dt, f"{fmt}/nonfinite/{dt.__name__}/a={alpha},b={beta}")


@pytest.mark.parametrize("cwg", [256, 512, 1024])
@pytest.mark.parametrize("name", CASES + ["wide"])
@pytest.mark.parametrize("dtype", [R.F64, R.F32])
def test_csr_lds_slices(name, dtype, cwg):
    """The LDS-sliced CSR image (x slices held in LDS, a partial per
    (row, slice) pair, partials added in slice order) against the oracle,
    including long pairs summed by a whole wavefront and empty rows, with
    each combine workgroup width."""
    a = case(name) if name != "wide" else _wide()
    print("csr xlds", name, dtype, cwg, flush=True)
    x64, y064 = inputs(a, 8)
    dt = np.float64 if dtype == R.F64 else np.float32
    params = f"xlds=1,xlcomb={cwg}"
    for alpha, beta in ((1.0, 0.0), (-0.5, 2.0)):
        y = run_gpu("csr", dtype, a, x64, y064, alpha, beta, params)
        v, x, y0 = a.values_for(dtype), x64.astype(dt), y064.astype(dt)
        y_ref = O.csr_spmv(a.rst, a.col, v, x, alpha, beta, y0)
        assert_within_bound(y, y_ref, a.rst, a.col, v, x, alpha, beta, y0, dt,
                            f"csr xlds/{name}/a={alpha},b={beta}")
        # the order of the additions is fixed by the data: the same bits again
        assert np.array_equal(y, run_gpu("csr", dtype, a, x64, y064, alpha, beta, params))


def _wide():
    """Uniform columns over n = 1.5 M (92 FP64 slices), rows of 0..40 entries
    and a few of 20000."""
    if "wide" not in _CACHE:
        rng = np.random.default_rng(12)
        m, n = 20000, 1_500_000
        lens = rng.integers(0, 41, m)
        lens[[3, 9000, 19999]] = 20000
        rst = np.zeros(m + 1, np.int32)
        rst[1:] = np.cumsum(lens)
        col = np.concatenate([np.sort(rng.choice(n, k, replace=False)) for k in lens]).astype(np.int32)
        _CACHE["wide"] = R.CSR(m, n, rst, col, rng.uniform(-1, 1, len(col)))
    return _CACHE["wide"]


@pytest.mark.parametrize("cwg", [256, 1024])
@pytest.mark.parametrize("dtype", [R.F64, R.F32])
def test_csr_lds_slices_many_pairs(dtype, cwg):
    """A row with an entry in each of 5000 slices of 16384 columns (more
    partials than a combine block holds: summed by the whole workgroup)
    between short rows, against the oracle."""
    W, S = 16384, 5000
    n = W * S
    rng = np.random.default_rng(13)
    cols = [np.array([3, 70000]), np.arange(S, dtype=np.int64) * W + rng.integers(0, W, S),
            np.array([n - 2])]
    rst = np.array([0, 2, 2 + S, 3 + S], np.int32)
    a = R.CSR(3, n, rst, np.concatenate(cols).astype(np.int32), rng.uniform(-1, 1, S + 3))
    print("csr xlds many pairs", dtype, flush=True)
    x64, y064 = inputs(a, 14)
    dt = np.float64 if dtype == R.F64 else np.float32
    y = run_gpu("csr", dtype, a, x64, y064, 0.5, -1.0, f"xlds=1,xlcomb={cwg}")
    v, x, y0 = a.values_for(dtype), x64.astype(dt), y064.astype(dt)
    y_ref = O.csr_spmv(a.rst, a.col, v, x, 0.5, -1.0, y0)
    assert_within_bound(y, y_ref, a.rst, a.col, v, x, 0.5, -1.0, y0, dt, "csr xlds many pairs")


def test_csr_lds_slices_full_size():
    """The power-law matrix of the benchmark (4.2 M rows, 64 M entries): the
    image chosen by default (LDS-sliced, FP64) against the oracle."""
    a = R.gen_powerlaw(4194304, 1.8, 50000, 64_000_000, 1024, seed=0)
    print("csr xlds full size", a.m, a.nnz, flush=True)
    x64, y064 = inputs(a, 10)
    y = run_gpu("csr", R.F64, a, x64, y064, 1.0, 0.5, "")
    y_ref = O.csr_spmv(a.rst, a.col, a.val, x64, 1.0, 0.5, y064)
    assert_within_bound(y, y_ref, a.rst, a.col, a.val, x64, 1.0, 0.5, y064, np.float64,
                        "csr default image, power-law 4.2 M rows")


@pytest.mark.parametrize("dtype", [R.BF16, R.F16])
def test_csr_lds_slices_16bit(dtype):
    a = case("powerlaw")
    x64, y064 = inputs(a, 9)
    y = run_gpu("csr", dtype, a, x64, y064, 1.0, 0.5, "xlds=1")
    x32, y032 = x64.astype(np.float32), y064.astype(np.float32)
    y64 = O.csr_spmv(a.rst, a.col, a.val, x32.astype(np.float64), 1.0, 0.5,
                     y032.astype(np.float64))
    rel = np.linalg.norm(y - y64) / np.linalg.norm(y64)
    assert rel <= (5e-3 if dtype == R.BF16 else 1e-3), rel


@pytest.mark.parametrize("params", ["R=16,C=16,p=2", "R=64,C=32,p=1", "R=300,C=400,p=2"])
@pytest.mark.parametrize("dtype", [R.F64, R.F32])
@pytest.mark.parametrize("name", ["rand_long", "powerlaw", "lap64"])
def test_hcsr_block_sizes(params, dtype, name):
    """HCSR with small blocks (block-rows of more than 32 non-empty blocks
    take the entry-by-entry path) and larger ones, against the oracle."""
    a = case(name)
    x64, y064 = inputs(a, 41)
    y = run_gpu("hcsr", dtype, a, x64, y064, 1.0, 0.5, params)
    dt = np.float64 if dtype == R.F64 else np.float32
    v, x, y0 = a.values_for(dtype), x64.astype(dt), y064.astype(dt)
    y_ref = O.csr_spmv(a.rst, a.col, v, x, 1.0, 0.5, y0)
    assert_within_bound(y, y_ref, a.rst, a.col, v, x, 1.0, 0.5, y0, dt, f"hcsr {params} {name}")


@pytest.mark.parametrize("sigma", [4, 8, 12, 16, 32])
@pytest.mark.parametrize("dtype", [R.F64, R.F32, R.BF16])
@pytest.mark.parametrize("name", ["rand_long", "powerlaw", "lap64", "random16_c1"])
def test_csr5_sigmas(sigma, dtype, name):
    """CSR5 tile kernels: the LDS-staged ones (sigma 8, 16) and the general
    one (other sigmas), against the oracle."""
    a = case(name)
    x64, y064 = inputs(a, 31)
    y = run_gpu("csr5", dtype, a, x64, y064, 0.75, -1.0, f"sigma={sigma}")
    if dtype == R.BF16:
        x32, y032 = x64.astype(np.float32), y064.astype(np.float32)
        y64 = O.csr_spmv(a.rst, a.col, a.val, x32.astype(np.float64), 0.75, -1.0,
                         y032.astype(np.float64))
        assert np.linalg.norm(y - y64) / np.linalg.norm(y64) <= 5e-3
        return
    dt = np.float64 if dtype == R.F64 else np.float32
    v, x, y0 = a.values_for(dtype), x64.astype(dt), y064.astype(dt)
    y_ref = O.csr_spmv(a.rst, a.col, v, x, 0.75, -1.0, y0)
    assert_within_bound(y, y_ref, a.rst, a.col, v, x, 0.75, -1.0, y0, dt,
                        f"csr5 sigma={sigma} {name}")


def _band(m, half):
    """m x m band matrix, entries |i - j| <= half, values from a fixed seed."""
    rows, cols = [], []
    for i in range(m):
        j = np.arange(max(0, i - half), min(m, i + half + 1))
        rows.append(np.full(len(j), i))
        cols.append(j)
    col = np.concatenate(cols).astype(np.int32)
    rst = np.concatenate([[0], np.cumsum([len(c) for c in cols])]).astype(np.int32)
    return R.CSR(m, m, rst, col, np.random.default_rng(m).normal(size=len(col)))


@pytest.mark.parametrize("params", ["T=0", "", "T=0.5"])
@pytest.mark.parametrize("name", ["lap64", "band", "random16_c1"])
@pytest.mark.parametrize("dtype", [R.F64, R.F32])
def test_lav_sparse_part_forms(params, name, dtype):
    """LAV with the sparse_part rows added by the dense kernels (no
    sparse_part row longer than 32 entries; with T=0 the sparse_part is
    empty and the band's long chunks run as pieces) and with a separate
    sparse_part multiply (long sparse_part rows), against the oracle."""
    a = _band(3000, 70) if name == "band" else case(name)
    x64, y064 = inputs(a, 17)
    dt = np.float64 if dtype == R.F64 else np.float32
    for alpha, beta in ((1.0, 0.0), (0.75, -0.5)):
        y = run_gpu("lav", dtype, a, x64, y064, alpha, beta, params)
        v, x, y0 = a.values_for(dtype), x64.astype(dt), y064.astype(dt)
        y_ref = O.csr_spmv(a.rst, a.col, v, x, alpha, beta, y0)
        assert_within_bound(y, y_ref, a.rst, a.col, v, x, alpha, beta, y0, dt,
                            f"lav {params} {name}/a={alpha},b={beta}")


@pytest.mark.parametrize("sigma", [4, 8, 16])
@pytest.mark.parametrize("dtype", [R.F64, R.F32])
def test_csr5_long_segments(sigma, dtype):
    """Rows spanning many CSR5 tiles: one ending exactly at a tile boundary,
    ones ending inside a later tile, one running from the full tiles into
    the tail, and tail rows on either side of the per-thread limit, with
    empty rows between them (the segment ranges are found at upload and the
    long ones are summed by a wavefront each)."""
    T = 64 * sigma
    lens = [3, 10 * T - 3, 5, 9 * T + 7, 0, 0, 12, 1, 9, 0, 3 * T + 1, 17 * T + 11, 0]
    # the last row starts in the full tiles and ends in the tail, then tail
    # rows of 1, 8, 9 and 40 entries (the tail holds fewer than T entries)
    head = sum(lens)
    lens.append(T - head % T + T // 2)
    lens += [1, 0, 8, 9, 40]
    assert sum(lens) % T != 0
    rng = np.random.default_rng(sigma)
    n = 40000
    rst = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    col = np.concatenate([np.sort(rng.choice(n, l, replace=False)) for l in lens]).astype(np.int32)
    a = R.CSR(len(lens), n, rst, col, rng.normal(size=int(rst[-1])))
    x64, y064 = inputs(a, 7)
    y = run_gpu("csr5", dtype, a, x64, y064, 1.25, 0.5, f"sigma={sigma}")
    dt = np.float64 if dtype == R.F64 else np.float32
    v, x, y0 = a.values_for(dtype), x64.astype(dt), y064.astype(dt)
    y_ref = O.csr_spmv(a.rst, a.col, v, x, 1.25, 0.5, y0)
    assert_within_bound(y, y_ref, a.rst, a.col, v, x, 1.25, 0.5, y0, dt,
                        f"csr5 long segments sigma={sigma}")


@pytest.mark.parametrize("dtype", [R.F64, R.F32])
def test_direct_uploads_match_converted(dtype):
    """rvvla_csr_upload / rvvla_hcsr_upload / rvvla_sell_upload (reference
    container fields) give the same result as convert + upload."""
    import ctypes as C
    a = case("rand_long")
    x64, y064 = inputs(a, 3)
    dev = _device()
    io = torch.float64 if dtype == R.F64 else torch.float32
    x = torch.from_numpy(x64).to(io).to(dev)
    outs = []
    M = R.csr_upload(a, dtype)
    y = torch.from_numpy(y064).to(io).to(dev)
    M.spmv(x, y, 1.0, 0.25)
    outs.append(y.cpu().numpy())
    L = R.convert("hcsr", a, dtype, "R=256,C=512,p=2")
    vals = L.array("values", np.float64 if dtype == R.F64 else np.float32)
    arr = {f: L.array(f) for f in ("block_ptr", "row_offset", "row_ptr", "col_idx")}
    h = C.c_void_p()
    R.check(R.lib.rvvla_hcsr_upload(dtype, a.m, a.n, a.nnz, 256, 512, int(L.scalar("b_m")),
                                    int(L.scalar("b_n")), arr["block_ptr"].ctypes.data,
                                    arr["row_offset"].ctypes.data, arr["row_ptr"].shape[0],
                                    arr["row_ptr"].ctypes.data, arr["col_idx"].ctypes.data,
                                    vals.ctypes.data, C.byref(h)))
    Mh = R.DeviceMatrix(h)
    y = torch.from_numpy(y064).to(io).to(dev)
    Mh.spmv(x, y, 1.0, 0.25)
    outs.append(y.cpu().numpy())
    dt = np.float64 if dtype == R.F64 else np.float32
    y_ref = O.csr_spmv(a.rst, a.col, a.values_for(dtype), x64.astype(dt), 1.0, 0.25,
                       y064.astype(dt))
    for y in outs:
        assert_within_bound(y, y_ref, a.rst, a.col, a.values_for(dtype), x64.astype(dt), 1.0,
                            0.25, y064.astype(dt), dt, "direct upload")


@pytest.mark.parametrize("dtype", [R.F64, R.F32])
def test_hcsr_upload_unsorted_coo_blocks(dtype):
    """A directly uploaded HCSR whose COO-kind blocks list their entries out
    of row order (allowed by the format) gives the same product: the upload
    orders those blocks by local row for the row-unit kernel."""
    import ctypes as C
    a = case("rand_long")
    x64, y064 = inputs(a, 8)
    L = R.convert("hcsr", a, dtype, "R=300,C=400,p=2")
    vt = np.float64 if dtype == R.F64 else np.float32
    bp, ro = L.array("block_ptr").astype(np.int64), L.array("row_offset")
    rp, ci, va = L.array("row_ptr").copy(), L.array("col_idx").copy(), L.array("values", vt).copy()
    rng = np.random.default_rng(3)
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
    bp32 = bp.astype(np.int32)  # kept alive across the call
    R.check(R.lib.rvvla_hcsr_upload(dtype, a.m, a.n, a.nnz, 300, 400, int(L.scalar("b_m")),
                                    int(L.scalar("b_n")), bp32.ctypes.data,
                                    ro.ctypes.data, rp.shape[0], rp.ctypes.data, ci.ctypes.data,
                                    va.ctypes.data, C.byref(h)))
    M = R.DeviceMatrix(h)
    dev = _device()
    io = torch.float64 if dtype == R.F64 else torch.float32
    x = torch.from_numpy(x64).to(io).to(dev)
    y = torch.from_numpy(y064).to(io).to(dev)
    M.spmv(x, y, 1.0, -0.5)
    dt = vt
    v = a.values_for(dtype)
    y_ref = O.csr_spmv(a.rst, a.col, v, x64.astype(dt), 1.0, -0.5, y064.astype(dt))
    assert_within_bound(y.cpu().numpy(), y_ref, a.rst, a.col, v, x64.astype(dt), 1.0, -0.5,
                        y064.astype(dt), dt, "hcsr unsorted COO")


def _transpose_csr(a: R.CSR) -> R.CSR:
    """CSR of A^T: entries of each column of A in row order."""
    order = np.argsort(a.col, kind="stable")
    rows = np.repeat(np.arange(a.m, dtype=np.int32), np.diff(a.rst))
    cnt = np.bincount(a.col, minlength=a.n)
    rst = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int32)
    return R.CSR(a.n, a.m, rst, rows[order], a.val[order])


@pytest.mark.parametrize("dtype", [R.F64, R.F32])
@pytest.mark.parametrize("name", ["rand_long", "powerlaw", "one_long_row", "empty", "lap64"])
def test_spmv_transpose_csr(dtype, name):
    """SPARSE_OPERATION_TRANSPOSE: y = alpha*A^T*x + beta*y on CSR, against the
    oracle's CSR product with the transpose (also the staged host path)."""
    a = case(name)
    t = _transpose_csr(a)
    dev = _device()
    rng = np.random.default_rng(12)
    x64, y064 = rng.uniform(-1, 1, a.m), rng.uniform(-1, 1, a.n)
    dt = np.float64 if dtype == R.F64 else np.float32
    io = torch.float64 if dtype == R.F64 else torch.float32
    M = R.convert("csr", a, dtype).upload()
    for alpha, beta in ((1.0, 0.0), (0.5, -2.0)):
        x = torch.from_numpy(x64).to(io).to(dev)
        y = torch.from_numpy(y064).to(io).to(dev)
        M.spmv_t(x, y, alpha, beta)
        torch.cuda.synchronize()
        v, xx, y0 = t.values_for(dtype), x64.astype(dt), y064.astype(dt)
        y_ref = O.csr_spmv(t.rst, t.col, v, xx, alpha, beta, y0)
        assert_within_bound(y.cpu().numpy(), y_ref, t.rst, t.col, v, xx, alpha, beta, y0, dt,
                            f"A^T {name} a={alpha} b={beta}")
    yh = y064.astype(dt).copy()
    xh = x64.astype(dt)  # kept alive across the call
    R.check(R.lib.rvvla_spmv_host_op(M.handle, 1, 1.0, xh.ctypes.data, 0.0, yh.ctypes.data))
    y_ref = O.csr_spmv(t.rst, t.col, t.values_for(dtype), x64.astype(dt), 1.0, 0.0,
                       np.zeros(a.n, dt))
    assert_within_bound(yh, y_ref, t.rst, t.col, t.values_for(dtype), x64.astype(dt), 1.0, 0.0,
                        np.zeros(a.n, dt), dt, f"A^T host {name}")


@pytest.mark.parametrize("dtype", [R.BF16, R.F16])
def test_spmv_transpose_16bit(dtype):
    a = case("rand_long")
    t = _transpose_csr(a)
    dev = _device()
    rng = np.random.default_rng(13)
    x64 = rng.uniform(-1, 1, a.m).astype(np.float32).astype(np.float64)
    M = R.convert("csr", a, dtype).upload()
    x = torch.from_numpy(x64).float().to(dev)
    y = torch.zeros(a.n, dtype=torch.float32, device=dev)
    M.spmv_t(x, y)
    y64 = O.csr_spmv(t.rst, t.col, t.val, x64, 1.0, 0.0, np.zeros(a.n))
    rel = np.linalg.norm(y.cpu().numpy() - y64) / np.linalg.norm(y64)
    assert rel <= (5e-3 if dtype == R.BF16 else 1e-3), rel


@pytest.mark.parametrize("fmt", [f for f in FORMATS_READY if f != "csr"])
@pytest.mark.parametrize("dtype", [R.F64, R.F32])
@pytest.mark.parametrize("name", ["rand_long", "lap64", "powerlaw", "empty"])
def test_spmv_transpose_formats(fmt, dtype, name):
    """SPARSE_OPERATION_TRANSPOSE on every format converted from CSR: the
    CSR image of A^T built from the kept source, against the oracle's CSR
    product with the transpose (device and staged host paths)."""
    a = case(name)
    t = _transpose_csr(a)
    dev = _device()
    rng = np.random.default_rng(14)
    x64, y064 = rng.uniform(-1, 1, a.m), rng.uniform(-1, 1, a.n)
    dt = np.float64 if dtype == R.F64 else np.float32
    io = torch.float64 if dtype == R.F64 else torch.float32
    print("transpose", fmt, name, dtype, flush=True)
    M = R.convert(fmt, a, dtype).upload()
    x = torch.from_numpy(x64).to(io).to(dev)
    y = torch.from_numpy(y064).to(io).to(dev)
    M.spmv_t(x, y, 0.5, -2.0)
    torch.cuda.synchronize()
    v, xx, y0 = t.values_for(dtype), x64.astype(dt), y064.astype(dt)
    y_ref = O.csr_spmv(t.rst, t.col, v, xx, 0.5, -2.0, y0)
    assert_within_bound(y.cpu().numpy(), y_ref, t.rst, t.col, v, xx, 0.5, -2.0, y0, dt,
                        f"{fmt} A^T {name}")
    yh = y0.copy()
    R.check(R.lib.rvvla_spmv_host_op(M.handle, 1, 0.5, xx.ctypes.data, -2.0, yh.ctypes.data))
    assert_within_bound(yh, y_ref, t.rst, t.col, v, xx, 0.5, -2.0, y0, dt, f"{fmt} A^T host {name}")


@pytest.mark.parametrize("fmt", ["hcsr", "sell", "cvr"])
def test_spmv_transpose_formats_16bit(fmt):
    a = case("rand_long")
    t = _transpose_csr(a)
    dev = _device()
    x64 = np.random.default_rng(15).uniform(-1, 1, a.m).astype(np.float32).astype(np.float64)
    M = R.convert(fmt, a, R.BF16).upload()
    x = torch.from_numpy(x64).float().to(dev)
    y = torch.zeros(a.n, dtype=torch.float32, device=dev)
    M.spmv_t(x, y)
    y64 = O.csr_spmv(t.rst, t.col, t.val, x64, 1.0, 0.0, np.zeros(a.n))
    rel = np.linalg.norm(y.cpu().numpy() - y64) / np.linalg.norm(y64)
    assert rel <= 5e-3, rel


def test_spmv_transpose_needs_source():
    """A matrix uploaded from format arrays (no CSR kept) answers the
    transposed product with RVVLA_ERR_UNSUPPORTED_SPMV."""
    a = case("lap64")
    dev = _device()
    rst, col, val = (torch.from_numpy(v).to(dev) for v in (a.rst, a.col, a.val))
    D = R.hcsr_from_device(R.F64, a.m, a.n, rst, col, val, 0, 0, 0.0)
    x = torch.zeros(a.m, dtype=torch.float64, device=dev)
    y = torch.zeros(a.n, dtype=torch.float64, device=dev)
    with pytest.raises(R.RvvlaError) as e:
        D.spmv_t(x, y)
    assert e.value.code == -4


def test_spmv_host_path():
    """rvvla_spmv_host: the std::vector ALL_STAGES form."""
    _device()
    a = case("lap64")
    x64, y064 = inputs(a, 4)
    M = R.convert("sell", a, R.F64).upload()
    y = y064.copy()
    M.spmv_host(x64, y, 2.0, -1.0)
    y_ref = O.csr_spmv(a.rst, a.col, a.val, x64, 2.0, -1.0, y064)
    assert_within_bound(y, y_ref, a.rst, a.col, a.val, x64, 2.0, -1.0, y064, np.float64, "host")


@pytest.mark.parametrize("dtype", [R.F64, R.F32])
@pytest.mark.parametrize("k", [1, 5, 16, 64, 70])
def test_spmm_csr(dtype, k):
    a = case("rand_long")
    dev = _device()
    rng = np.random.default_rng(k)
    X64 = rng.uniform(-1, 1, (a.n, k))
    Y064 = rng.uniform(-1, 1, (a.m, k))
    io = torch.float64 if dtype == R.F64 else torch.float32
    dt = np.float64 if dtype == R.F64 else np.float32
    M = R.csr_upload(a, dtype)
    X = torch.from_numpy(X64).to(io).to(dev)
    Y = torch.from_numpy(Y064).to(io).to(dev)
    M.spmm(X, Y, 0.5, 2.0)
    Y = Y.cpu().numpy()
    v = a.values_for(dtype)
    for j in range(k):
        y_ref = O.csr_spmv(a.rst, a.col, v, X64[:, j].astype(dt), 0.5, 2.0, Y064[:, j].astype(dt))
        assert_within_bound(Y[:, j], y_ref, a.rst, a.col, v, X64[:, j].astype(dt), 0.5, 2.0,
                            Y064[:, j].astype(dt), dt, f"spmm col {j}")

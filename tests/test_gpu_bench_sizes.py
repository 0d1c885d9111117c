"""Parity at the benchmark's own sizes: the 2048 x 2048 5-point Laplacian in
every format, FP64 and FP32; the 4.2M-row, 64M-entry power-law matrix in
every format in FP64 and in CSR in FP32, BF16 and FP16; and the 512-matrix
batch.  Every FP64/FP32 row against the oracle within
4 (len + 2) u (|alpha| sum|a||x| + |beta| |y0|); the 16-bit forms against
the FP64 oracle normwise (5e-3 BF16, 1e-3 FP16) and row by row within the
value rounding plus the FP32 summation."""
from __future__ import annotations

import numpy as np
import pytest

import rvvlasparse_amd as R
from helpers import assert_within_bound
from oracle import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

FORMATS = ["csr", "hcsr", "sell", "csr5", "lav", "vnec", "vhcc", "cvr"]
_CACHE = {}


def _mat(name):
    if name not in _CACHE:
        if name == "laplacian":
            a = R.gen_laplacian5(2048)
        else:
            a = R.gen_powerlaw(4194304, 1.8, 50000, 64_000_000, 1024, seed=0)
        x = R.gen_uniform(a.n, -1.0, 1.0, seed=0, purpose=3)
        y0 = R.gen_uniform(a.m, -1.0, 1.0, seed=1, purpose=3)
        _CACHE[name] = (a, x, y0)
    return _CACHE[name]


def _product(fmt, dtype, a, x64, y064, alpha, beta, params=""):
    dev = torch.device("cuda:0")
    io = torch.float64 if dtype == R.F64 else torch.float32
    M = R.convert(fmt, a, dtype, params).upload()
    x = torch.from_numpy(x64).to(io).to(dev)
    y = torch.from_numpy(y064).to(io).to(dev)
    M.spmv(x, y, alpha, beta)
    torch.cuda.synchronize()
    del M
    return y.cpu().numpy()


@pytest.mark.parametrize("dtype", [R.F64, R.F32])
@pytest.mark.parametrize("fmt", FORMATS + ["hcsr:R=2048,C=2048"])
def test_laplacian_2048(fmt, dtype):
    a, x64, y064 = _mat("laplacian")
    fmt, _, params = fmt.partition(":")
    print("laplacian 2048", fmt, params, dtype, flush=True)
    dt = np.float64 if dtype == R.F64 else np.float32
    v, x, y0 = a.values_for(dtype), x64.astype(dt), y064.astype(dt)
    for alpha, beta in ((1.0, 0.0), (0.5, -1.5)):
        y = _product(fmt, dtype, a, x64, y064, alpha, beta, params)
        y_ref = O.csr_spmv(a.rst, a.col, v, x, alpha, beta, y0)
        assert_within_bound(y, y_ref, a.rst, a.col, v, x, alpha, beta, y0, dt,
                            f"laplacian {fmt} {params} {dt.__name__} a={alpha} b={beta}")


@pytest.mark.parametrize("fmt", FORMATS)
def test_powerlaw_f64(fmt):
    a, x, y0 = _mat("powerlaw")
    print("powerlaw f64", fmt, flush=True)
    y = _product(fmt, R.F64, a, x, y0, 1.0, 0.0)
    y_ref = O.csr_spmv(a.rst, a.col, a.val, x, 1.0, 0.0)
    assert_within_bound(y, y_ref, a.rst, a.col, a.val, x, 1.0, 0.0, np.zeros(a.m), np.float64,
                        f"powerlaw {fmt} f64")


@pytest.mark.parametrize("dtype", [R.F32, R.BF16, R.F16])
def test_powerlaw_csr_reduced(dtype):
    a, x64, y064 = _mat("powerlaw")
    print("powerlaw csr", dtype, flush=True)
    x32, y032 = x64.astype(np.float32), y064.astype(np.float32)
    y = _product("csr", dtype, a, x64, y064, 1.0, 0.5)
    if dtype == R.F32:
        v = a.values_for(dtype)
        y_ref = O.csr_spmv(a.rst, a.col, v, x32, 1.0, 0.5, y032)
        assert_within_bound(y, y_ref, a.rst, a.col, v, x32, 1.0, 0.5, y032, np.float32,
                            "powerlaw csr f32")
        return
    # against the FP64 oracle on the unrounded matrix and the FP32 x, y0
    y64 = O.csr_spmv(a.rst, a.col, a.val, x32.astype(np.float64), 1.0, 0.5,
                     y032.astype(np.float64))
    rel = np.linalg.norm(y - y64) / np.linalg.norm(y64)
    assert rel <= (5e-3 if dtype == R.BF16 else 1e-3), rel
    # per row: the values' rounding (relative 2^-8 or 2^-11; FP16 values
    # below 2^-14 are subnormal, rounded to within 2^-25 absolute) plus the
    # FP32 summation and the result's rounding
    uv = 2.0 ** -8 if dtype == R.BF16 else 2.0 ** -11
    ln = np.diff(a.rst).astype(np.float64)
    x64r = x32.astype(np.float64)
    scale = O.csr_absdot(a.rst, a.col, a.val, x64r)
    bound = uv * scale + 4 * (ln + 2) * 2.0 ** -24 * (scale + 0.5 * np.abs(y032))
    if dtype == R.F16:
        bound += 2.0 ** -25 * O.csr_absdot(a.rst, a.col, np.ones(a.nnz), x64r)
    err = np.abs(y - y64)
    assert np.all(err <= bound), int(np.argmax(err - bound))


def test_batch_512():
    """The benchmark's batch (same seeds): 512 matrices, n log-uniform in
    [1e3, 2e5], 8..64 entries per row, k in 1..16, FP32, autotuned."""
    from test_gpu_batch import _copy_from, _copy_to
    rng = np.random.default_rng(11)
    ns = np.exp(rng.uniform(np.log(1e3), np.log(2e5), 512)).astype(int)
    ks = [int(k) for k in rng.integers(1, 17, 512)]
    mats = [R.gen_random_rows_var(int(n), 8, 64, seed=100 + i) for i, n in enumerate(ns)]
    print("batch 512", sum(a.nnz for a in mats), flush=True)
    B = R.Batch(mats, ks, R.F32)
    dev = torch.device("cuda:0")
    xr = np.random.default_rng(3)
    Xs = []
    for i, (a, k) in enumerate(zip(mats, ks)):
        xp, _, _, ld = B.block(i)
        X = xr.uniform(-1, 1, (a.n, k)).astype(np.float32)
        _copy_to(xp, X, ld, torch.float32, dev)
        Xs.append(X)
    torch.cuda.synchronize()
    B.run(1.0, 0.0, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    for i, (a, k) in enumerate(zip(mats, ks)):
        _, yp, _, ld = B.block(i)
        Y = _copy_from(yp, (a.m, k), ld, torch.float32)
        v = a.val.astype(np.float32)
        for j in range(k):
            x = np.ascontiguousarray(Xs[i][:, j])
            y_ref = O.csr_spmv(a.rst, a.col, v, x, 1.0, 0.0)
            assert_within_bound(Y[:, j], y_ref, a.rst, a.col, v, x, 1.0, 0.0, np.zeros(a.m),
                                np.float32, f"batch 512: matrix {i} rhs {j}")

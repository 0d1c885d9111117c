"""Batched uneven SpMM in one persistent launch: every matrix's Y_i against
the oracle column by column, for the autotuned and for both forced item
kernels, and graph replays (with eager work between them) against the
eager result."""
from __future__ import annotations

import numpy as np
import pytest

import rvvlasparse_amd as R
from helpers import assert_within_bound
from oracle import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _batch_inputs(count=24, seed=0):
    rng = np.random.default_rng(seed)
    ns = np.exp(rng.uniform(np.log(300), np.log(6000), count)).astype(int)
    mats = [R.gen_random_rows_var(int(n), 8, 64, seed=seed * 1000 + i) for i, n in enumerate(ns)]
    ks = [int(k) for k in rng.integers(1, 17, count)]
    return mats, ks


def _expected_ld(k, dtype):
    """Row stride of X_i / Y_i: k, rounded up to whole 16-byte loads when k > 2."""
    v = 2 if dtype == R.F64 else 4
    return k if k <= 2 else -(-k // v) * v


def _run_and_check(B, mats, ks, dtype, alpha, beta, seed=1):
    dev = torch.device("cuda:0")
    tdt = torch.float64 if dtype == R.F64 else torch.float32
    dt = np.float64 if dtype == R.F64 else np.float32
    rng = np.random.default_rng(seed)
    Xs, Y0s = [], []
    for i, (a, k) in enumerate(zip(mats, ks)):
        xp, yp, _, ld = B.block(i)
        assert ld == _expected_ld(k, dtype)
        X = rng.uniform(-1, 1, (a.n, k)).astype(dt)
        Y0 = rng.uniform(-1, 1, (a.m, k)).astype(dt)
        torch.cuda.synchronize()
        _copy_to(xp, X, ld, tdt, dev)
        _copy_to(yp, Y0, ld, tdt, dev)
        Xs.append(X)
        Y0s.append(Y0)
    torch.cuda.synchronize()
    B.run(alpha, beta, torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    for i, (a, k) in enumerate(zip(mats, ks)):
        _, yp, _, ld = B.block(i)
        Y = _copy_from(yp, (a.m, k), ld, tdt)
        v = a.val.astype(dt)
        for j in range(k):
            y_ref = O.csr_spmv(a.rst, a.col, v, Xs[i][:, j], alpha, beta, Y0s[i][:, j])
            assert_within_bound(Y[:, j], y_ref, a.rst, a.col, v, Xs[i][:, j], alpha, beta,
                                Y0s[i][:, j], dt, f"batch mat {i} rhs {j}")
    return Xs


def _copy_to(ptr, arr, ld, tdt, dev):
    t = torch.from_numpy(np.ascontiguousarray(arr)).to(dev)
    torch.cuda.synchronize()
    # device-to-device copy into the batch's buffer through a torch view of it
    _view(ptr, arr.shape, ld, tdt).copy_(t)


def _view(ptr, shape, ld, tdt):
    """A torch tensor aliasing the rows x cols block at ptr whose rows are ld
    elements apart (no ownership)."""
    rows, cols = shape

    class _Holder:
        pass

    h = _Holder()
    es = 8 if tdt == torch.float64 else 4
    h.__cuda_array_interface__ = {
        "shape": (rows, cols), "strides": (ld * es, es),
        "typestr": "<f8" if tdt == torch.float64 else "<f4",
        "data": (int(ptr), False), "version": 2}
    return torch.as_tensor(h, device="cuda")


def _copy_from(ptr, shape, ld, tdt):
    return _view(ptr, shape, ld, tdt).cpu().numpy()


@pytest.mark.parametrize("dtype", [R.F32, R.F64])
def test_batch_autotuned(dtype):
    mats, ks = _batch_inputs(20, seed=3)
    print("batch autotuned", len(mats), flush=True)
    B = R.Batch(mats, ks, dtype)
    _run_and_check(B, mats, ks, dtype, 1.0, 0.0)
    _run_and_check(B, mats, ks, dtype, 0.5, -1.25, seed=2)
    keys = {B.tune_key(i) for i in range(len(mats))}
    assert len(keys) >= 1


@pytest.mark.parametrize("params", ["format=0,param=1", "format=0,param=4", "format=0,param=64",
                                    "format=1,param=1", "format=1,param=4096"])
def test_batch_forced_item_kernels(params):
    mats, ks = _batch_inputs(12, seed=5)
    B = R.Batch(mats, ks, R.F32, params)
    _run_and_check(B, mats, ks, R.F32, 1.5, 0.5)


@pytest.mark.parametrize("params", ["format=0,param=2", "format=0,param=16"])
def test_batch_forced_csr_f64(params):
    """FP64: two right-hand sides per lane when k is even, one otherwise."""
    mats, ks = _batch_inputs(12, seed=6)
    B = R.Batch(mats, ks, R.F64, params)
    _run_and_check(B, mats, ks, R.F64, -1.0, 0.25)


def test_batch_graph_replays():
    mats, ks = _batch_inputs(10, seed=9)
    B = R.Batch(mats, ks, R.F32)
    Xs = _run_and_check(B, mats, ks, R.F32, 1.0, 0.0)
    eager = [_copy_from(B.block(i)[1], (a.m, k), B.block(i)[3], torch.float32)
             for i, (a, k) in enumerate(zip(mats, ks))]
    B.capture(1.0, 0.0)
    s = torch.cuda.current_stream().cuda_stream
    for rep in range(4):
        # eager work on the same stream between replays
        junk = torch.randn(1 << 20, device="cuda").sum()
        torch.cuda.synchronize()
        for i, (a, k) in enumerate(zip(mats, ks)):
            _view(B.block(i)[1], (a.m, k), B.block(i)[3], torch.float32).fill_(float(rep + 7))
        torch.cuda.synchronize()
        B.replay(s)
        torch.cuda.synchronize()
        for i, (a, k) in enumerate(zip(mats, ks)):
            got = _copy_from(B.block(i)[1], (a.m, k), B.block(i)[3], torch.float32)
            np.testing.assert_array_equal(got, eager[i])
        del junk


@pytest.mark.parametrize("dtype", [R.F32, R.F64])
def test_batch_padding_columns(dtype):
    """NaN in the padding columns of X_i and Y_i: Y_i's real columns match
    the oracle (the padding feeds only dropped lanes, and beta never reads
    Y's padding) and Y's padding is left as it was."""
    mats, ks = _batch_inputs(8, seed=12)
    ks = [3, 5, 7, 9, 11, 13, 15, 6]
    B = R.Batch(mats, ks, dtype, "format=0,param=0")
    tdt = torch.float64 if dtype == R.F64 else torch.float32
    for i, (a, k) in enumerate(zip(mats, ks)):
        xp, yp, _, ld = B.block(i)
        if ld > k:
            _view(xp, (a.n, ld), ld, tdt)[:, k:].fill_(float("nan"))
            _view(yp, (a.m, ld), ld, tdt)[:, k:].fill_(float("nan"))
    torch.cuda.synchronize()
    _run_and_check(B, mats, ks, dtype, 1.0, 0.5, seed=4)
    for i, (a, k) in enumerate(zip(mats, ks)):
        _, yp, _, ld = B.block(i)
        if ld > k:
            pad = _copy_from(yp, (a.m, ld), ld, tdt)[:, k:]
            assert np.isnan(pad).all(), f"batch mat {i}: Y padding written"

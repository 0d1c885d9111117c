"""Differentiable SpMM: the hand-written backward (dX = A^T dY over the CSC
permutation, dtheta = sampled dense product) against FP64 central finite
differences and against the gradient assembled from the CPU oracle's scalar
CSR path, on a 2048 x 2048 instance with 32 distinct uniform columns per row
and X of 64 columns (the construction of the differentiable-SpMM workload at
small size).  The loss is sum(W * Y), linear in theta and in X, so central
differences are exact up to rounding."""
from __future__ import annotations

import numpy as np
import pytest

import rvvlasparse_amd as R
from oracle import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _instance(n=2048, k=64, seed=7):
    a = R.gen_random_rows(n, 32, seed=seed, include_diag=False, normal=True, stdev=32 ** -0.5)
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, k))
    W = rng.normal(size=(n, k))
    return a, X, W


def _oracle_grads(a, X, W):
    """dtheta_e = <W[row_e], X[col_e]>, dX = A^T W, from the oracle's CSR
    multiply (A^T built on the host, one oracle SpMV per column)."""
    rows = np.repeat(np.arange(a.m), np.diff(a.rst))
    dtheta = np.einsum("ij,ij->i", W[rows], X[a.col])
    order = np.lexsort((rows, a.col))
    t_rst = np.zeros(a.n + 1, np.int32)
    t_rst[1:] = np.cumsum(np.bincount(a.col, minlength=a.n))
    t_col = rows[order].astype(np.int32)
    t_val = a.val[order]
    dX = np.stack([O.csr_spmv(t_rst, t_col, t_val, W[:, j]) for j in range(W.shape[1])], axis=1)
    return dtheta, dX


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_spmm_forward_backward_against_oracle(dtype):
    from rvvlasparse_amd.autograd import SparseStructure, spmm
    a, X, W = _instance()
    dev = torch.device("cuda:0")
    A = SparseStructure(a.m, a.n, a.rst, a.col, dtype)
    theta = torch.tensor(a.val, dtype=dtype, device=dev, requires_grad=True)
    Xt = torch.tensor(X, dtype=dtype, device=dev, requires_grad=True)
    Wt = torch.tensor(W, dtype=dtype, device=dev)
    print(f"spmm autograd {dtype} n={a.m} nnz={a.nnz} k={X.shape[1]}", flush=True)
    Y = spmm(A, theta, Xt)
    (Y * Wt).sum().backward()
    torch.cuda.synchronize()
    # forward against the oracle, column by column
    Y_ref = np.stack([O.csr_spmv(a.rst, a.col, a.val, X[:, j]) for j in range(X.shape[1])], axis=1)
    tol = 1e-12 if dtype == torch.float64 else 2e-5
    np.testing.assert_allclose(Y.detach().cpu().numpy(), Y_ref, rtol=0, atol=tol * np.abs(Y_ref).max())
    dtheta_ref, dX_ref = _oracle_grads(a, X, W)
    np.testing.assert_allclose(theta.grad.cpu().numpy(), dtheta_ref, rtol=0,
                               atol=tol * np.abs(dtheta_ref).max())
    np.testing.assert_allclose(Xt.grad.cpu().numpy(), dX_ref, rtol=0, atol=tol * np.abs(dX_ref).max())


def test_spmm_central_finite_differences_fp64():
    from rvvlasparse_amd.autograd import SparseStructure, spmm
    a, X, W = _instance()
    dev = torch.device("cuda:0")
    A = SparseStructure(a.m, a.n, a.rst, a.col, torch.float64)
    theta = torch.tensor(a.val, device=dev, requires_grad=True)
    Xt = torch.tensor(X, device=dev, requires_grad=True)
    Wt = torch.tensor(W, device=dev)

    def loss(th, x):
        return float((spmm(A, th, x) * Wt).sum())

    (spmm(A, theta, Xt) * Wt).sum().backward()
    g_theta, g_X = theta.grad.cpu().numpy(), Xt.grad.cpu().numpy()
    rng = np.random.default_rng(3)
    h = 1e-3
    with torch.no_grad():
        for e in rng.choice(a.nnz, 20, replace=False):
            tp, tm = theta.detach().clone(), theta.detach().clone()
            tp[e] += h
            tm[e] -= h
            fd = (loss(tp, Xt.detach()) - loss(tm, Xt.detach())) / (2 * h)
            assert abs(fd - g_theta[e]) <= 1e-7 * max(1.0, abs(g_theta[e])), (e, fd, g_theta[e])
        for i, j in zip(rng.integers(0, a.n, 20), rng.integers(0, X.shape[1], 20)):
            xp, xm = Xt.detach().clone(), Xt.detach().clone()
            xp[i, j] += h
            xm[i, j] -= h
            fd = (loss(theta.detach(), xp) - loss(theta.detach(), xm)) / (2 * h)
            assert abs(fd - g_X[i, j]) <= 1e-7 * max(1.0, abs(g_X[i, j])), (i, j, fd, g_X[i, j])


@pytest.mark.parametrize("k", [1, 3, 64, 100])
def test_spmm_transpose_and_sddmm_odd_widths(k):
    """Widths that are not a multiple of the vector width or exceed 64."""
    from rvvlasparse_amd.autograd import SparseStructure, spmm
    a = R.gen_random_rows(500, 12, seed=k, include_diag=True, normal=True)
    rng = np.random.default_rng(k)
    X = rng.normal(size=(a.n, k))
    W = rng.normal(size=(a.m, k))
    dev = torch.device("cuda:0")
    A = SparseStructure(a.m, a.n, a.rst, a.col, torch.float64)
    theta = torch.tensor(a.val, device=dev, requires_grad=True)
    Xt = torch.tensor(X, device=dev, requires_grad=True)
    (spmm(A, theta, Xt) * torch.tensor(W, device=dev)).sum().backward()
    dtheta_ref, dX_ref = _oracle_grads(a, X, W)
    np.testing.assert_allclose(theta.grad.cpu().numpy(), dtheta_ref, rtol=0, atol=1e-12)
    np.testing.assert_allclose(Xt.grad.cpu().numpy(), dX_ref, rtol=0, atol=1e-12)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("k", [64, 16, 8, 5, 1])
@pytest.mark.parametrize("which", ["both", "dX", "dtheta"])
def test_spmm_backward_fused_against_oracle(dtype, k, which):
    """rvvla_spmm_backward (one pass over the columns for dX and dtheta; k = 5
    takes the two separate passes) against the oracle's gradients, with
    either output alone too; rows of A with no entries in a column range and
    empty columns included (n = 1500, 3..12 entries per row)."""
    import ctypes as C
    from rvvlasparse_amd.autograd import SparseStructure, lib
    a = R.gen_random_rows_var(1500, 3, 12, seed=k)
    rng = np.random.default_rng(k)
    X = rng.normal(size=(a.n, k))
    W = rng.normal(size=(a.m, k))
    dev = torch.device("cuda:0")
    A = SparseStructure(a.m, a.n, a.rst, a.col, dtype)
    theta = torch.tensor(a.val, dtype=dtype, device=dev)
    Xt = torch.tensor(X, dtype=dtype, device=dev)
    Wt = torch.tensor(W, dtype=dtype, device=dev)
    dX = torch.full((a.n, k), float("nan"), dtype=dtype, device=dev) if which != "dtheta" else None
    dth = torch.full((a.nnz,), float("nan"), dtype=dtype, device=dev) if which != "dX" else None
    print(f"fused backward {dtype} k={k} {which}", flush=True)
    R.check(lib.rvvla_spmm_backward(A.handle, theta.data_ptr(), k, Wt.data_ptr(), k, Xt.data_ptr(), k,
                                    dX.data_ptr() if dX is not None else None, k,
                                    dth.data_ptr() if dth is not None else None, None))
    torch.cuda.synchronize()
    dtheta_ref, dX_ref = _oracle_grads(a, X, W)
    tol = 1e-12 if dtype == torch.float64 else 2e-5
    if dX is not None:
        np.testing.assert_allclose(dX.cpu().numpy(), dX_ref, rtol=0, atol=tol * np.abs(dX_ref).max())
    if dth is not None:
        np.testing.assert_allclose(dth.cpu().numpy(), dtheta_ref, rtol=0,
                                   atol=tol * np.abs(dtheta_ref).max())

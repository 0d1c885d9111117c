"""Differentiable SpMM on the gfx950 kernels: Y = A(theta) X.

``SparseStructure`` holds a CSR pattern resident on the device (plus the
CSC permutation used by the transposed product); ``spmm(A, theta, X)`` is a
``torch.autograd.Function`` whose backward computes
  dX     = A(theta)^T dY        (gather over the cached CSC permutation, no atomics)
  dtheta = <dY[row_e, :], X[col_e, :]> for every entry e   (sampled dense product)
with hand-written HIP kernels: rvvla_spmm_values forward, and one pass over
A's columns for both gradients (rvvla_spmm_backward: each dY row gathered
once for dX and dtheta).  FP32 and FP64.  No reference counterpart: the
reference multiplies one vector with fixed values.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import CSR, F32, F64, check, csr_upload, lib

_i64, _i32, _d, _vp = C.c_int64, C.c_int, C.c_double, C.c_void_p
for _name, _args in {
    "rvvla_prepare_transpose": [_vp],
    "rvvla_spmm_values": [_vp, _vp, _i32, _d, _vp, _i64, _d, _vp, _i64, _vp],
    "rvvla_spmm_transpose": [_vp, _vp, _i32, _d, _vp, _i64, _d, _vp, _i64, _vp],
    "rvvla_sddmm": [_vp, _i32, _d, _vp, _i64, _vp, _i64, _vp, _vp],
    "rvvla_spmm_backward": [_vp, _vp, _i32, _vp, _i64, _vp, _i64, _vp, _i64, _vp, _vp],
}.items():
    _f = getattr(lib, _name)
    _f.restype, _f.argtypes = _i32, _args


class SparseStructure:
    """A CSR pattern (int32 row pointers / columns) resident on the device."""

    def __init__(self, m: int, n: int, rst: np.ndarray, col: np.ndarray,
                 dtype: torch.dtype = torch.float32):
        if dtype not in (torch.float32, torch.float64):
            raise TypeError("SparseStructure: float32 or float64")
        self.m, self.n = int(m), int(n)
        self.nnz = int(col.shape[0])
        self.dtype = dtype
        self._M = csr_upload(CSR(m, n, rst, col, np.zeros(self.nnz)),
                             F64 if dtype == torch.float64 else F32)
        check(lib.rvvla_prepare_transpose(self._M.handle))
        self.rst = torch.from_numpy(np.ascontiguousarray(rst, np.int64))
        self.col = torch.from_numpy(np.ascontiguousarray(col, np.int64))

    @property
    def handle(self) -> int:
        return self._M.handle


def _stream(t: torch.Tensor) -> int:
    return torch.cuda.current_stream(t.device).cuda_stream


def _check_operands(A: SparseStructure, theta: torch.Tensor, X: torch.Tensor):
    if theta.dtype != A.dtype or X.dtype != A.dtype:
        raise TypeError("theta and X must have the structure's dtype")
    if theta.shape != (A.nnz,) or X.dim() != 2 or X.shape[0] != A.n:
        raise ValueError("theta must be (nnz,) and X (n, k)")
    if not (theta.is_cuda and X.is_cuda):
        raise ValueError("theta and X must be on the HIP device")


class _SpMM(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, X, A):
        _check_operands(A, theta, X)
        theta = theta.contiguous()
        X = X.contiguous()
        k = X.shape[1]
        Y = torch.empty((A.m, k), dtype=X.dtype, device=X.device)
        check(lib.rvvla_spmm_values(A.handle, theta.data_ptr(), k, 1.0, X.data_ptr(), X.stride(0),
                                    0.0, Y.data_ptr(), Y.stride(0), _stream(X)))
        ctx.save_for_backward(theta, X)
        ctx.A = A
        return Y

    @staticmethod
    def backward(ctx, dY):
        theta, X = ctx.saved_tensors
        A = ctx.A
        dY = dY.contiguous()
        k = X.shape[1]
        dtheta = torch.empty_like(theta) if ctx.needs_input_grad[0] else None
        dX = torch.empty_like(X) if ctx.needs_input_grad[1] else None
        if dtheta is not None or dX is not None:
            check(lib.rvvla_spmm_backward(A.handle, theta.data_ptr(), k, dY.data_ptr(),
                                          dY.stride(0), X.data_ptr(), X.stride(0),
                                          dX.data_ptr() if dX is not None else None,
                                          dX.stride(0) if dX is not None else k,
                                          dtheta.data_ptr() if dtheta is not None else None,
                                          _stream(dY)))
        return dtheta, dX, None


def spmm(A: SparseStructure, theta: torch.Tensor, X: torch.Tensor) -> torch.Tensor:
    """Y = A(theta) X, differentiable in theta (the values, CSR order) and X."""
    return _SpMM.apply(theta, X, A)


def spmv(A: SparseStructure, theta: torch.Tensor, x: torch.Tensor) -> torch.Tensor:
    """y = A(theta) x, differentiable (a one-column SpMM)."""
    return spmm(A, theta, x.reshape(-1, 1)).reshape(-1)

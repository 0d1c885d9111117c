"""rvvlasparse_amd -- MI355X (gfx950) SpMV in the eight RVVLASparse formats.

Python view of the C-ABI in include/rvvla_hip.h (librvvla_hip.so, built
in-tree by ``rvvlasparse_amd._build``).  Importing fails loudly when the
library is missing: there is no CPU fallback.

Device vectors are passed as torch tensors on a HIP device (or raw device
pointers); torch only provides memory and streams here.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "librvvla_hip.so"

F64, F32, BF16, F16 = 0, 1, 2, 3
DTYPES = {"f64": F64, "f32": F32, "bf16": BF16, "f16": F16}
FORMATS = {"csr": 0, "hcsr": 1, "sell": 2, "csr5": 3, "lav": 4, "vnec": 5, "vhcc": 6, "cvr": 7}
FORMAT_NAMES = {v: k for k, v in FORMATS.items()}


class RvvlaError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"rvvla error {code}: {msg}")
        self.code = code


if not LIB_PATH.exists():
    raise ImportError(
        f"{LIB_PATH} is not built; run `python -m rvvlasparse_amd._build` "
        "(the HIP library is required: there is no CPU fallback)")

# One HIP runtime per process: torch ships its own libamdhip64.so.7.  Loading
# torch first makes the library's libamdhip64.so.7 dependency resolve to that
# copy; loading the library first would bring in a second HIP/HSA runtime
# (/opt/rocm) that then finds no device.
try:
    import torch as _torch  # noqa: F401
except ImportError:
    _torch = None

lib = C.CDLL(str(LIB_PATH))


class GenCSR(C.Structure):
    _fields_ = [("m", C.c_int64), ("n", C.c_int64), ("nnz", C.c_int64),
                ("rst", C.POINTER(C.c_int32)), ("col", C.POINTER(C.c_int32)),
                ("val", C.POINTER(C.c_double))]


_i64, _i32, _d, _vp, _cp = C.c_int64, C.c_int, C.c_double, C.c_void_p, C.c_char_p
_sigs = {
    "rvvla_strerror": (_cp, [_i32]),
    "rvvla_version": (_i32, []),
    "rvvla_format_name": (_cp, [_i32]),
    "rvvla_convert": (_i32, [_i32, _i32, _i64, _i64, _i64, _vp, _vp, _vp, _cp, C.POINTER(_vp)]),
    "rvvla_layout_array": (_i32, [_vp, _cp, C.POINTER(_vp), C.POINTER(_i64), C.POINTER(_i32)]),
    "rvvla_layout_scalar": (_i32, [_vp, _cp, C.POINTER(_d)]),
    "rvvla_layout_free": (_i32, [_vp]),
    "rvvla_upload": (_i32, [_vp, C.POINTER(_vp)]),
    "rvvla_csr_upload": (_i32, [_i32, _i64, _i64, _i64, _vp, _vp, _vp, C.POINTER(_vp)]),
    "rvvla_hcsr_upload": (_i32, [_i32, _i64, _i64, _i64, _i64, _i64, _i64, _i64, _vp, _vp, _i64,
                                 _vp, _vp, _vp, C.POINTER(_vp)]),
    "rvvla_sell_upload": (_i32, [_i32, _i64, _i64, _i64, _i64, _i32, _i32, _vp, _vp, _vp, _vp,
                                 _i64, _vp, _vp, _vp, C.POINTER(_vp)]),
    "rvvla_set_host_validation": (_i32, [_i32]),
    "rvvla_mat_view": (_i32, [_vp, _cp, C.POINTER(_vp), C.POINTER(_i64), C.POINTER(_i32)]),
    "rvvla_mat_read": (_i32, [_vp, _cp, _vp, _i64, C.POINTER(_i64), C.POINTER(_i32)]),
    "rvvla_sell_from_device": (_i32, [_i32, _i64, _i64, _i64, _vp, _vp, _vp, _i32, _vp,
                                      C.POINTER(_vp)]),
    "rvvla_hcsr_from_device": (_i32, [_i32, _i64, _i64, _i64, _vp, _vp, _vp, _i64, _i64, _d,
                                      _vp, C.POINTER(_vp)]),
    "rvvla_csr5_from_device": (_i32, [_i32, _i64, _i64, _i64, _vp, _vp, _vp, _i32, _vp,
                                      C.POINTER(_vp)]),
    "rvvla_vnec_from_device": (_i32, [_i32, _i64, _i64, _i64, _vp, _vp, _vp, _i64, _vp,
                                      C.POINTER(_vp)]),
    "rvvla_lav_from_device": (_i32, [_i32, _i64, _i64, _i64, _vp, _vp, _vp, _d, _i64, _vp,
                                     C.POINTER(_vp)]),
    "rvvla_vhcc_from_device": (_i32, [_i32, _i64, _i64, _i64, _vp, _vp, _vp, _i64, _vp,
                                      C.POINTER(_vp)]),
    "rvvla_cvr_from_device": (_i32, [_i32, _i64, _i64, _i64, _vp, _vp, _vp, _i64, _vp,
                                     C.POINTER(_vp)]),
    "rvvla_mat_free": (_i32, [_vp]),
    "rvvla_mat_info": (_i32, [_vp, C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_i64),
                              C.POINTER(_i32), C.POINTER(_i32), C.POINTER(_i64)]),
    "rvvla_spmv": (_i32, [_vp, _d, _vp, _d, _vp, _vp]),
    "rvvla_spmv_noalloc": (_i32, [_vp, _d, _vp, _d, _vp, _vp]),
    "rvvla_spmv_op": (_i32, [_vp, _i32, _d, _vp, _d, _vp, _vp]),
    "rvvla_spmv_host_op": (_i32, [_vp, _i32, _d, _vp, _d, _vp]),
    "rvvla_spmm": (_i32, [_vp, _i32, _d, _vp, _i64, _d, _vp, _i64, _vp]),
    "rvvla_read_probe": (_i32, [_vp, _i64, _vp, _vp]),
    "rvvla_spmv_host": (_i32, [_vp, _d, _vp, _d, _vp]),
    "rvvla_gen_laplacian5": (_i32, [_i64, C.POINTER(GenCSR)]),
    "rvvla_gen_random_rows": (_i32, [_i64, _i32, _i32, _i32, _d, C.c_uint64, C.POINTER(GenCSR)]),
    "rvvla_gen_powerlaw": (_i32, [_i64, _d, _i64, _i64, _i64, C.c_uint64, C.POINTER(GenCSR)]),
    "rvvla_gen_random_rows_var": (_i32, [_i64, _i32, _i32, C.c_uint64, C.POINTER(GenCSR)]),
    "rvvla_gen_uniform": (_i32, [_i64, _d, _d, C.c_uint64, C.c_uint64, _vp]),
    "rvvla_gen_free": (_i32, [C.POINTER(GenCSR)]),
    "rvvla_mtx_read": (_i32, [_cp, C.POINTER(GenCSR), C.c_char_p]),
    "rvvla_bin_write": (_i32, [_cp, C.POINTER(GenCSR), C.c_char_p]),
    "rvvla_bin_read": (_i32, [_cp, C.POINTER(GenCSR), C.c_char_p]),
}
for _name, (_res, _args) in _sigs.items():
    _f = getattr(lib, _name)
    _f.restype, _f.argtypes = _res, _args


def check(code: int) -> None:
    if code != 0:
        raise RvvlaError(code, lib.rvvla_strerror(code).decode())


def _ptr(a) -> int | None:
    """Address of a numpy array, a torch tensor or an int."""
    if a is None:
        return None
    if isinstance(a, int):
        return a
    if isinstance(a, np.ndarray):
        return a.ctypes.data
    return a.data_ptr()


# --------------------------------------------------------------- matrices
class CSR:
    """Host CSR matrix (int32 Rst/Col, float64 values) -- spMtxCRS fields."""

    def __init__(self, m: int, n: int, rst: np.ndarray, col: np.ndarray, val: np.ndarray):
        self.m, self.n = int(m), int(n)
        self.rst = np.ascontiguousarray(rst, dtype=np.int32)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)

    @property
    def nnz(self) -> int:
        return int(self.col.shape[0])

    def row_lengths(self) -> np.ndarray:
        return np.diff(self.rst)

    def values_for(self, dtype: int) -> np.ndarray:
        return self.val if dtype == F64 else self.val.astype(np.float32)


def _take(g: GenCSR) -> CSR:
    try:
        m, n, nnz = g.m, g.n, g.nnz
        rst = np.ctypeslib.as_array(g.rst, shape=(m + 1,)).copy()
        col = np.ctypeslib.as_array(g.col, shape=(max(nnz, 1),))[:nnz].copy()
        val = np.ctypeslib.as_array(g.val, shape=(max(nnz, 1),))[:nnz].copy()
    finally:
        lib.rvvla_gen_free(C.byref(g))
    return CSR(m, n, rst, col, val)


def gen_laplacian5(grid: int) -> CSR:
    g = GenCSR()
    check(lib.rvvla_gen_laplacian5(grid, C.byref(g)))
    return _take(g)


def gen_random_rows(n: int, k: int, seed: int = 0, include_diag: bool = True,
                    normal: bool = False, stdev: float = 1.0) -> CSR:
    g = GenCSR()
    check(lib.rvvla_gen_random_rows(n, k, int(include_diag), int(normal), stdev, seed, C.byref(g)))
    return _take(g)


def gen_powerlaw(n: int, alpha: float = 1.8, maxlen: int = 50000, target_nnz: int = 64_000_000,
                 band: int = 1024, seed: int = 0) -> CSR:
    g = GenCSR()
    check(lib.rvvla_gen_powerlaw(n, alpha, maxlen, target_nnz, band, seed, C.byref(g)))
    return _take(g)


def gen_uniform(count: int, lo: float = -1.0, hi: float = 1.0, seed: int = 0,
                purpose: int = 3) -> np.ndarray:
    v = np.empty(count, dtype=np.float64)
    check(lib.rvvla_gen_uniform(count, lo, hi, seed, purpose, _ptr(v)))
    return v


_ELEM = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}


class Layout:
    """Host-side converted matrix: convert<T, Fmt, true>(dist, src, params)."""

    def __init__(self, fmt: str | int, a: CSR, dtype: int = F64, params: str = ""):
        self.format = FORMATS[fmt] if isinstance(fmt, str) else int(fmt)
        self.dtype = dtype
        self.csr = a
        self._vals = a.values_for(dtype)
        h = C.c_void_p()
        check(lib.rvvla_convert(self.format, dtype, a.m, a.n, a.nnz, _ptr(a.rst), _ptr(a.col),
                                _ptr(self._vals), params.encode(), C.byref(h)))
        self._h = h

    def array(self, name: str, dtype=None) -> np.ndarray:
        """Copy of one array of the converted structure (raw unsigned view
        unless `dtype` is given)."""
        d, cnt, eb = C.c_void_p(), C.c_int64(), C.c_int()
        check(lib.rvvla_layout_array(self._h, name.encode(), C.byref(d), C.byref(cnt), C.byref(eb)))
        raw = _ELEM[eb.value]
        if cnt.value == 0:
            out = np.empty(0, dtype=raw)
        else:
            buf = (C.c_char * (cnt.value * eb.value)).from_address(d.value)
            out = np.frombuffer(buf, dtype=raw).copy()
        return out.view(dtype) if dtype is not None else out

    def scalar(self, name: str) -> float:
        v = C.c_double()
        check(lib.rvvla_layout_scalar(self._h, name.encode(), C.byref(v)))
        return v.value

    def upload(self) -> "DeviceMatrix":
        h = C.c_void_p()
        check(lib.rvvla_upload(self._h, C.byref(h)))
        return DeviceMatrix(h)

    def __del__(self):
        if getattr(self, "_h", None) and lib is not None:  # lib is None at interpreter exit
            lib.rvvla_layout_free(self._h)
            self._h = None


class DeviceMatrix:
    """A device-resident matrix (rvvla_mat_t)."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle
        m, n, nnz, b = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        f, d = C.c_int(), C.c_int()
        check(lib.rvvla_mat_info(handle, C.byref(m), C.byref(n), C.byref(nnz), C.byref(f),
                                 C.byref(d), C.byref(b)))
        self.m, self.n, self.nnz = m.value, n.value, nnz.value
        self.format, self.dtype, self.device_bytes = f.value, d.value, b.value

    @property
    def handle(self) -> int:
        return self._h.value

    def spmv(self, x, y, alpha: float = 1.0, beta: float = 0.0, stream=None) -> None:
        """y = alpha*A*x + beta*y on device vectors, ordered on `stream`
        (a torch.cuda.Stream, a raw hipStream_t int, or None = torch's
        current stream when torch tensors are passed)."""
        check(lib.rvvla_spmv(self._h, alpha, _ptr(x), beta, _ptr(y), _stream(stream, y)))

    def spmv_noalloc(self, x, y, alpha: float = 1.0, beta: float = 0.0, stream=None) -> None:
        """The same product with no device allocation (rvvla_spmv_noalloc):
        the per-product buffers live in the library's 16 MiB device array;
        overlapping calls on one device race."""
        check(lib.rvvla_spmv_noalloc(self._h, alpha, _ptr(x), beta, _ptr(y), _stream(stream, y)))

    def spmv_t(self, x, y, alpha: float = 1.0, beta: float = 0.0, stream=None) -> None:
        """y = alpha*A^T*x + beta*y (SPARSE_OPERATION_TRANSPOSE; CSR matrices)."""
        check(lib.rvvla_spmv_op(self._h, 1, alpha, _ptr(x), beta, _ptr(y), _stream(stream, y)))

    def spmm(self, X, Y, alpha: float = 1.0, beta: float = 0.0, stream=None) -> None:
        k = int(X.shape[1])
        check(lib.rvvla_spmm(self._h, k, alpha, _ptr(X), int(X.stride(0)), beta, _ptr(Y),
                             int(Y.stride(0)), _stream(stream, Y)))

    def view(self, field: str) -> np.ndarray:
        """Copy of one array of a matrix uploaded in host-validation mode."""
        d, cnt, eb = C.c_void_p(), C.c_int64(), C.c_int()
        check(lib.rvvla_mat_view(self._h, field.encode(), C.byref(d), C.byref(cnt), C.byref(eb)))
        raw = _ELEM[eb.value]
        if cnt.value == 0:
            return np.empty(0, dtype=raw)
        buf = (C.c_char * (cnt.value * eb.value)).from_address(d.value)
        return np.frombuffer(buf, dtype=raw).copy()

    def read(self, field: str) -> np.ndarray:
        """Copy of one array of the uploaded image (device or host-validation)."""
        cnt, eb = C.c_int64(), C.c_int()
        check(lib.rvvla_mat_read(self._h, field.encode(), None, 0, C.byref(cnt), C.byref(eb)))
        out = np.empty(cnt.value, dtype=_ELEM[eb.value])
        check(lib.rvvla_mat_read(self._h, field.encode(), out.ctypes.data if out.size else None,
                                 out.nbytes, C.byref(cnt), C.byref(eb)))
        return out

    def spmv_host(self, b: np.ndarray, y: np.ndarray, alpha: float = 1.0, beta: float = 0.0):
        check(lib.rvvla_spmv_host(self._h, alpha, _ptr(b), beta, _ptr(y)))

    def __del__(self):
        if getattr(self, "_h", None) and lib is not None:  # lib is None at interpreter exit
            lib.rvvla_mat_free(self._h)
            self._h = None


def _stream(stream, like):
    if stream is None:
        if hasattr(like, "device") and getattr(like, "is_cuda", False):
            import torch
            return torch.cuda.current_stream(like.device).cuda_stream
        return None
    if isinstance(stream, int):
        return stream
    return stream.cuda_stream


def sell_from_device(dtype: int, m: int, n: int, rst, col, val, sigma: int = 0,
                     stream=None) -> DeviceMatrix:
    """SELL-C-sigma image built on the device from device CSR arrays (torch
    tensors or raw addresses): int32 rst/col, float64 (F64) or float32
    values.  sigma <= 0: the default rule."""
    h = C.c_void_p()
    nnz = int(col.numel()) if hasattr(col, "numel") else int(len(col))
    check(lib.rvvla_sell_from_device(dtype, m, n, nnz, _ptr(rst), _ptr(col), _ptr(val), sigma,
                                     _stream(stream, val), C.byref(h)))
    return DeviceMatrix(h)


def hcsr_from_device(dtype: int, m: int, n: int, rst, col, val, R: int = 0, C_: int = 0,
                     p: float = 0.0, stream=None) -> DeviceMatrix:
    """HCSR built on the device from device CSR arrays (torch tensors or raw
    addresses): int32 rst/col, float64 (F64) or float32 values.  R, C_, p <= 0:
    the reference defaults."""
    h = C.c_void_p()
    nnz = int(col.numel()) if hasattr(col, "numel") else int(len(col))
    check(lib.rvvla_hcsr_from_device(dtype, m, n, nnz, _ptr(rst), _ptr(col), _ptr(val), R, C_, p,
                                     _stream(stream, val), C.byref(h)))
    return DeviceMatrix(h)


def csr5_from_device(dtype: int, m: int, n: int, rst, col, val, sigma: int = 0,
                     stream=None) -> DeviceMatrix:
    """CSR5 (omega = 64) built on the device from device CSR arrays (torch
    tensors or raw addresses).  sigma <= 0: the default."""
    h = C.c_void_p()
    nnz = int(col.numel()) if hasattr(col, "numel") else int(len(col))
    check(lib.rvvla_csr5_from_device(dtype, m, n, nnz, _ptr(rst), _ptr(col), _ptr(val), sigma,
                                     _stream(stream, val), C.byref(h)))
    return DeviceMatrix(h)


def vnec_from_device(dtype: int, m: int, n: int, rst, col, val, nparts: int = 0,
                     stream=None) -> DeviceMatrix:
    """VNEC built on the device from device CSR arrays (torch tensors or raw
    addresses).  nparts <= 0: the default partition count."""
    h = C.c_void_p()
    nnz = int(col.numel()) if hasattr(col, "numel") else int(len(col))
    check(lib.rvvla_vnec_from_device(dtype, m, n, nnz, _ptr(rst), _ptr(col), _ptr(val), nparts,
                                     _stream(stream, val), C.byref(h)))
    return DeviceMatrix(h)


def cvr_from_device(dtype: int, m: int, n: int, rst, col, val, nstreams: int = 0,
                    stream=None) -> DeviceMatrix:
    """CVR (SIMD_LEN 64) built on the device from device CSR arrays (torch
    tensors or raw addresses).  nstreams <= 0: the default stream count."""
    h = C.c_void_p()
    nnz = int(col.numel()) if hasattr(col, "numel") else int(len(col))
    check(lib.rvvla_cvr_from_device(dtype, m, n, nnz, _ptr(rst), _ptr(col), _ptr(val), nstreams,
                                    _stream(stream, val), C.byref(h)))
    return DeviceMatrix(h)


def vhcc_from_device(dtype: int, m: int, n: int, rst, col, val, nparts: int = 0,
                     stream=None) -> DeviceMatrix:
    """VHCC (vec_size 64) built on the device from device CSR arrays (torch
    tensors or raw addresses).  nparts <= 0: the default partition count."""
    h = C.c_void_p()
    nnz = int(col.numel()) if hasattr(col, "numel") else int(len(col))
    check(lib.rvvla_vhcc_from_device(dtype, m, n, nnz, _ptr(rst), _ptr(col), _ptr(val), nparts,
                                     _stream(stream, val), C.byref(h)))
    return DeviceMatrix(h)


def lav_from_device(dtype: int, m: int, n: int, rst, col, val, T: float = -1.0,
                    segment_len: int = 0, stream=None) -> DeviceMatrix:
    """LAV built on the device from device CSR arrays (torch tensors or raw
    addresses).  T < 0, segment_len <= 0: the convert defaults."""
    h = C.c_void_p()
    nnz = int(col.numel()) if hasattr(col, "numel") else int(len(col))
    check(lib.rvvla_lav_from_device(dtype, m, n, nnz, _ptr(rst), _ptr(col), _ptr(val), T,
                                    segment_len, _stream(stream, val), C.byref(h)))
    return DeviceMatrix(h)


def csr_upload(a: CSR, dtype: int = F64) -> DeviceMatrix:
    h = C.c_void_p()
    vals = a.values_for(dtype)
    check(lib.rvvla_csr_upload(dtype, a.m, a.n, a.nnz, _ptr(a.rst), _ptr(a.col), _ptr(vals),
                               C.byref(h)))
    return DeviceMatrix(h)


def convert(fmt: str | int, a: CSR, dtype: int = F64, params: str = "") -> Layout:
    return Layout(fmt, a, dtype, params)


def autotune(a: CSR, dtype: int = F64, params: str = "") -> tuple[str, str]:
    """The fastest (format name, convertParams text) for `a` on the current
    device among the autotuner's candidates (rvvla_autotune); params
    "formats=a+b", "reps=N", "cache=<path>"."""
    vals = a.values_for(dtype)
    fmt, out = C.c_int(), C.create_string_buffer(256)
    check(lib.rvvla_autotune(dtype, a.m, a.n, a.nnz, _ptr(a.rst), _ptr(a.col), _ptr(vals),
                             params.encode(), C.byref(fmt), out, 256))
    name = [k for k, v in FORMATS.items() if v == fmt.value][0]
    return name, out.value.decode()


for _name, (_res, _args) in {
    "rvvla_batch_create": (_i32, [_i32, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _cp, C.POINTER(_vp)]),
    "rvvla_batch_block": (_i32, [_vp, _i32, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_i64),
                                 C.POINTER(_i32)]),
    "rvvla_batch_tune_key": (_i32, [_vp, _i32, _cp, _i32]),
    "rvvla_batch_run": (_i32, [_vp, _d, _d, _vp]),
    "rvvla_batch_capture": (_i32, [_vp, _d, _d]),
    "rvvla_batch_replay": (_i32, [_vp, _vp]),
    "rvvla_batch_free": (_i32, [_vp]),
    "rvvla_autotune": (_i32, [_i32, _i64, _i64, _i64, _vp, _vp, _vp, _cp, C.POINTER(_i32), _cp,
                              _i32]),
}.items():
    _f = getattr(lib, _name)
    _f.restype, _f.argtypes = _res, _args


class Batch:
    """Many CSR matrices (each with its own 1..16 right-hand sides) multiplied
    in one persistent launch (rvvla_batch_*).  X_i and Y_i live in the
    batch's device buffers, row-major with row stride ld_i >= k_i;
    ``block(i)`` returns their device addresses and ld_i."""

    def __init__(self, mats: list, ks: list, dtype: int = F32, params: str = ""):
        cnt = len(mats)
        self.mats, self.ks, self.dtype = mats, list(ks), dtype
        vt = np.float64 if dtype == F64 else np.float32
        self._vals = [np.ascontiguousarray(a.val, vt) for a in mats]
        ms = np.array([a.m for a in mats], np.int64)
        ns = np.array([a.n for a in mats], np.int64)
        rp = (C.c_void_p * cnt)(*[a.rst.ctypes.data for a in mats])
        cp = (C.c_void_p * cnt)(*[a.col.ctypes.data for a in mats])
        vp = (C.c_void_p * cnt)(*[v.ctypes.data for v in self._vals])
        kk = np.array(ks, np.int32)
        h = C.c_void_p()
        check(lib.rvvla_batch_create(dtype, cnt, ms.ctypes.data, ns.ctypes.data, rp, cp, vp,
                                     kk.ctypes.data, params.encode(), C.byref(h)))
        self._h = h

    def block(self, i: int):
        """(X_i address, Y_i address, item kernel format, row stride ld_i)."""
        x, y, f, ld = C.c_void_p(), C.c_void_p(), C.c_int(), C.c_int64()
        check(lib.rvvla_batch_block(self._h, i, C.byref(x), C.byref(y), C.byref(ld), C.byref(f)))
        return x.value, y.value, f.value, ld.value

    def tune_key(self, i: int) -> str:
        buf = C.create_string_buffer(512)
        check(lib.rvvla_batch_tune_key(self._h, i, buf, 512))
        return buf.value.decode()

    def run(self, alpha=1.0, beta=0.0, stream=None):
        check(lib.rvvla_batch_run(self._h, alpha, beta, stream))

    def capture(self, alpha=1.0, beta=0.0):
        check(lib.rvvla_batch_capture(self._h, alpha, beta))

    def replay(self, stream=None):
        check(lib.rvvla_batch_replay(self._h, stream))

    def __del__(self):
        if getattr(self, "_h", None) and lib is not None:  # lib is None at interpreter exit
            lib.rvvla_batch_free(self._h)
            self._h = None


def read_mtx(path: str):
    """Matrix Market coordinate file -> (CSR with float64 values, MM type
    code such as "MCRG")."""
    g, code = GenCSR(), C.create_string_buffer(4)
    check(lib.rvvla_mtx_read(str(path).encode(), C.byref(g), code))
    return _take(g), code.raw.decode()


def read_bin(path: str):
    """Binary CSR file -> (CSR, MM type code)."""
    g, code = GenCSR(), C.create_string_buffer(4)
    check(lib.rvvla_bin_read(str(path).encode(), C.byref(g), code))
    return _take(g), code.raw.decode()


def write_bin(path: str, a: "CSR", matcode: str = "MCRG") -> None:
    val = np.ascontiguousarray(a.val, np.float64)
    rst = np.ascontiguousarray(a.rst, np.int32)
    col = np.ascontiguousarray(a.col, np.int32)
    g = GenCSR(a.m, a.n, a.nnz, rst.ctypes.data_as(C.POINTER(C.c_int32)),
               col.ctypes.data_as(C.POINTER(C.c_int32)),
               val.ctypes.data_as(C.POINTER(C.c_double)))
    check(lib.rvvla_bin_write(str(path).encode(), C.byref(g), matcode.encode()[:4].ljust(4)))


def gen_random_rows_var(n: int, kmin: int, kmax: int, seed: int = 0) -> CSR:
    """n x n, row lengths uniform in [kmin, kmax], distinct uniform columns,
    values N(0,1) (rvvla_gen_random_rows with per-row lengths)."""
    g = GenCSR()
    check(lib.rvvla_gen_random_rows_var(n, kmin, kmax, seed, C.byref(g)))
    return _take(g)


def io_dtype(dtype: int):
    """numpy dtype of x and y for a value dtype (FP32 for the 16-bit forms)."""
    return np.float64 if dtype == F64 else np.float32

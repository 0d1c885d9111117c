"""ctypes view of the CPU checkers -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load
this module.  ``orc`` is the restatement in oracle.cpp (liboracle.so);
``ref`` is the reference's own scalar library compiled from its sources
(_ref/libref.so), present only where the reference sources were available
at build time.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_SO = HERE / "liboracle.so"
REF_SO = HERE / "_ref" / "libref.so"

_i64, _i32, _d, _f, _vp, _cp = C.c_int64, C.c_int, C.c_double, C.c_float, C.c_void_p, C.c_char_p


def build(ref: bool = True) -> None:
    """make -C oracle (liboracle.so, and _ref/libref.so when the reference is present)."""
    subprocess.run(["make", "-s", "-C", str(HERE)] + ([] if ref else ["liboracle.so"]),
                   check=True)


def _load(path: Path, sigs: dict):
    lib = C.CDLL(str(path))
    for name, (res, args) in sigs.items():
        f = getattr(lib, name)
        f.restype, f.argtypes = res, args
    return lib


def _p(a):
    return None if a is None else a.ctypes.data


_ORC_SIGS = {
    "orc_csr_spmv_f64": (_i32, [_i64, _vp, _vp, _vp, _vp, _d, _d, _vp]),
    "orc_csr_spmv_f32": (_i32, [_i64, _vp, _vp, _vp, _vp, _f, _f, _vp]),
    "orc_csr_absdot": (_i32, [_i64, _vp, _vp, _vp, _vp, _vp]),
    "orc_csr_dot_dd": (_i32, [_i64, _vp, _vp, _vp, _vp, _vp]),
    "orc_hcsr_convert_f64": (_vp, [_i64, _i64, _i64, _vp, _vp, _vp, _i64, _i64, _d, C.POINTER(_i32)]),
    "orc_hcsr_convert_f32": (_vp, [_i64, _i64, _i64, _vp, _vp, _vp, _i64, _i64, _d, C.POINTER(_i32)]),
    "orc_hcsr_spmv_f64": (_i32, [_vp, _vp, _d, _d, _vp]),
    "orc_hcsr_spmv_f32": (_i32, [_vp, _vp, _f, _f, _vp]),
    "orc_sell_build_f64": (_vp, [_i64, _i64, _i64, _vp, _vp, _vp, _i32, _i32]),
    "orc_sell_build_f32": (_vp, [_i64, _i64, _i64, _vp, _vp, _vp, _i32, _i32]),
    "orc_sell_spmv_f64": (_i32, [_i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _d, _d, _vp]),
    "orc_sell_spmv_f32": (_i32, [_i64, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _f, _f, _vp]),
    "orc_obj_array": (_i32, [_vp, _cp, C.POINTER(_vp), C.POINTER(_i64), C.POINTER(_i32)]),
    "orc_obj_scalar": (_i32, [_vp, _cp, C.POINTER(_d)]),
    "orc_obj_free": (None, [_vp]),
    "orc_csr5_build_f64": (_vp, [_i64, _i64, _i64, _vp, _vp, _vp, _i32, _i32]),
    "orc_csr5_build_f32": (_vp, [_i64, _i64, _i64, _vp, _vp, _vp, _i32, _i32]),
    "orc_csr5_spmv_f64": (_i32, [_vp, _vp, _d, _d, _vp]),
    "orc_csr5_spmv_f32": (_i32, [_vp, _vp, _f, _f, _vp]),
    "orc_obj_new": (_vp, []),
    "orc_obj_set_array": (_i32, [_vp, _cp, _vp, _i64, _i32]),
    "orc_obj_set_scalar": (_i32, [_vp, _cp, _d]),
    "orc_vhcc_spmv_f64": (_i32, [_i64, _i32, _i32, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _d, _d, _vp]),
    "orc_vhcc_spmv_f32": (_i32, [_i64, _i32, _i32, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _f, _f, _vp]),
    "orc_vnec_ird": (C.c_float, [_i32, _i32, _i32, _vp, _vp, _i32, _i32, _vp]),
    "orc_num_threads": (_i32, []),
    "orc_set_threads": (None, [_i32]),
}

_REF_SIGS = {
    "ref_csr_spmv_f64": (_i32, [_i64, _i64, _i64, _vp, _vp, _vp, _vp, _d, _d, _vp]),
    "ref_csr_spmv_f32": (_i32, [_i64, _i64, _i64, _vp, _vp, _vp, _vp, _f, _f, _vp]),
    "ref_hcsr_convert_f64": (_vp, [_i64, _i64, _i64, _vp, _vp, _vp, _i32, _i32, _d, C.POINTER(_i32)]),
    "ref_hcsr_convert_f32": (_vp, [_i64, _i64, _i64, _vp, _vp, _vp, _i32, _i32, _d, C.POINTER(_i32)]),
    "ref_hcsr_array_f64": (_i32, [_vp, _cp, C.POINTER(_vp), C.POINTER(_i64), C.POINTER(_i32)]),
    "ref_hcsr_array_f32": (_i32, [_vp, _cp, C.POINTER(_vp), C.POINTER(_i64), C.POINTER(_i32)]),
    "ref_hcsr_dims_f64": (_i32, [_vp, _vp]),
    "ref_hcsr_dims_f32": (_i32, [_vp, _vp]),
    "ref_hcsr_spmv_f64": (_i32, [_vp, _vp, _d, _d, _vp]),
    "ref_hcsr_spmv_f32": (_i32, [_vp, _vp, _f, _f, _vp]),
    "ref_hcsr_free_f64": (None, [_vp]),
    "ref_hcsr_free_f32": (None, [_vp]),
    "ref_cvr_spmv_f64": (_i32, [_i64, _i64, _i64, _i32, _i32, _vp, _vp, _i64, _vp, _vp, _vp, _d, _d, _vp]),
    "ref_cvr_spmv_f32": (_i32, [_i64, _i64, _i64, _i32, _i32, _vp, _vp, _i64, _vp, _vp, _vp, _f, _f, _vp]),
    "ref_lav_spmv_f64": (_i32, [_i64, _i64, _i64, _i32, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64,
                                _vp, _i64, _vp, _i64, _vp, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _d,
                                _d, _vp]),
    "ref_lav_spmv_f32": (_i32, [_i64, _i64, _i64, _i32, _i64, _i64, _vp, _i64, _vp, _i64, _vp, _i64,
                                _vp, _i64, _vp, _i64, _vp, _vp, _i64, _i64, _vp, _vp, _vp, _vp, _f,
                                _f, _vp]),
    "ref_sell_spmv_f64": (_i32, [_i64, _i64, _i64, _i32, _i32, _vp, _vp, _vp, _vp, _i64, _vp, _vp,
                                 _vp, _vp, _d, _d, _vp]),
    "ref_sell_spmv_f32": (_i32, [_i64, _i64, _i64, _i32, _i32, _vp, _vp, _vp, _vp, _i64, _vp, _vp,
                                 _vp, _vp, _f, _f, _vp]),
    "ref_csr_bench_new_f64": (_vp, [_i64, _i64, _i64, _vp, _vp, _vp, _vp]),
    "ref_csr_bench_new_f32": (_vp, [_i64, _i64, _i64, _vp, _vp, _vp, _vp]),
    "ref_hcsr_bench_new_f64": (_vp, [_i64, _i64, _i64, _vp, _vp, _vp, _vp, _i32, _i32, _d]),
    "ref_csr_bench_run_f64": (_i32, [_vp]),
    "ref_bench_run": (_i32, [_vp]),
    "ref_bench_y": (_vp, [_vp]),
    "ref_csr_bench_free": (None, [_vp]),
}

orc = _load(ORACLE_SO, _ORC_SIGS) if ORACLE_SO.exists() else None
ref = _load(REF_SO, _REF_SIGS) if REF_SO.exists() else None

_ELEM = {1: np.uint8, 2: np.uint16, 4: np.uint32, 8: np.uint64}


def _sfx(dt) -> str:
    return "f64" if np.dtype(dt) == np.float64 else "f32"


# --------------------------------------------------------------- CSR
def csr_spmv(rst, col, val, x, alpha=1.0, beta=0.0, y=None, use_ref=False):
    """y = alpha*A*x + beta*y with the scalar CSR restatement (or the real
    reference when use_ref).  dtype follows `val` (float64 or float32)."""
    dt = val.dtype
    m = rst.shape[0] - 1
    y = np.zeros(m, dtype=dt) if y is None else np.array(y, dtype=dt, copy=True)
    x = np.ascontiguousarray(x, dtype=dt)
    if use_ref:
        getattr(ref, "ref_csr_spmv_" + _sfx(dt))(m, x.shape[0], col.shape[0], _p(rst), _p(col),
                                                 _p(val), _p(x), alpha, beta, _p(y))
    else:
        getattr(orc, "orc_csr_spmv_" + _sfx(dt))(m, _p(rst), _p(col), _p(val), _p(x), alpha, beta,
                                                 _p(y))
    return y


def csr_absdot(rst, col, val, x):
    m = rst.shape[0] - 1
    out = np.empty(m, dtype=np.float64)
    # keep the converted copies alive for the duration of the call
    v64 = np.ascontiguousarray(val, np.float64)
    x64 = np.ascontiguousarray(x, np.float64)
    orc.orc_csr_absdot(m, _p(rst), _p(col), _p(v64), _p(x64), _p(out))
    return out


def csr_dot_dd(rst, col, val, x):
    m = rst.shape[0] - 1
    out = np.empty(m, dtype=np.float64)
    v64 = np.ascontiguousarray(val, np.float64)
    x64 = np.ascontiguousarray(x, np.float64)
    orc.orc_csr_dot_dd(m, _p(rst), _p(col), _p(v64), _p(x64), _p(out))
    return out


def spmv_bound(rst, col, val, x, alpha, beta, y0, unit_roundoff):
    """Per-row bound 4*(len+2)*u*(|alpha|*sum|a||x| + |beta|*|y0|): the
    4*len*eps*sum|a||x| row bound, widened by the two extra roundings of
    the alpha/beta update."""
    ln = np.diff(rst).astype(np.float64)
    scale = abs(alpha) * csr_absdot(rst, col, val, x)
    if y0 is not None and beta != 0.0:
        scale = scale + abs(beta) * np.abs(np.asarray(y0, np.float64))
    return 4.0 * (ln + 2.0) * unit_roundoff * scale


# --------------------------------------------------------------- converted objects
class Obj:
    """A structure built by the oracle or by the compiled reference."""

    def __init__(self, handle, getter, freer, scalars):
        self._h, self._get, self._free = handle, getter, freer
        self.scalars = scalars

    def array(self, name):
        d, cnt, eb = C.c_void_p(), C.c_int64(), C.c_int()
        rc = self._get(self._h, name.encode(), C.byref(d), C.byref(cnt), C.byref(eb))
        if rc:
            raise KeyError(name)
        if cnt.value == 0:
            return np.empty(0, dtype=_ELEM[eb.value])
        buf = (C.c_char * (cnt.value * eb.value)).from_address(d.value)
        return np.frombuffer(buf, dtype=_ELEM[eb.value]).copy()

    def __del__(self):
        if getattr(self, "_h", None):
            self._free(self._h)
            self._h = None


def _orc_scalars(h, names):
    out = {}
    for k in names:
        v = C.c_double()
        if orc.orc_obj_scalar(h, k.encode(), C.byref(v)) == 0:
            out[k] = v.value
    return out


HCSR_SCALARS = ("R", "C", "m", "n", "nz", "b_m", "b_n")


def hcsr_convert(m, n, rst, col, val, R, C_, p, use_ref=False):
    """(Obj, err) -- err is 0, -10 (invalid_argument), -11 (length_error)
    or -12 (logic_error) as the reference converter raises them."""
    err = C.c_int()
    sfx = _sfx(val.dtype)
    if use_ref:
        h = getattr(ref, "ref_hcsr_convert_" + sfx)(m, n, col.shape[0], _p(rst), _p(col), _p(val),
                                                   R, C_, p, C.byref(err))
        if not h:
            return None, err.value
        dims = np.zeros(7, dtype=np.int64)
        getattr(ref, "ref_hcsr_dims_" + sfx)(h, _p(dims))
        sc = dict(zip(HCSR_SCALARS, map(float, dims)))
        o = Obj(h, getattr(ref, "ref_hcsr_array_" + sfx), getattr(ref, "ref_hcsr_free_" + sfx), sc)
        o.kind = ("ref", sfx)
        return o, 0
    h = getattr(orc, "orc_hcsr_convert_" + sfx)(m, n, col.shape[0], _p(rst), _p(col), _p(val),
                                               R, C_, p, C.byref(err))
    if not h:
        return None, err.value
    o = Obj(h, orc.orc_obj_array, orc.orc_obj_free, _orc_scalars(h, HCSR_SCALARS))
    o.kind = ("orc", sfx)
    return o, 0


def hcsr_spmv(obj, x, alpha=1.0, beta=0.0, y=None):
    src, sfx = obj.kind
    dt = np.float64 if sfx == "f64" else np.float32
    m = int(obj.scalars["m"])
    y = np.zeros(m, dtype=dt) if y is None else np.array(y, dtype=dt, copy=True)
    x = np.ascontiguousarray(x, dtype=dt)
    fn = getattr(ref if src == "ref" else orc, f"{src}_hcsr_spmv_{sfx}")
    fn(obj._h, _p(x), alpha, beta, _p(y))
    return y


SELL_SCALARS = ("m", "n", "nz", "cnt_b", "C", "Sigma")


def sell_build(m, n, rst, col, val, C_=64, sigma=1):
    h = getattr(orc, "orc_sell_build_" + _sfx(val.dtype))(m, n, col.shape[0], _p(rst), _p(col),
                                                         _p(val), C_, sigma)
    return Obj(h, orc.orc_obj_array, orc.orc_obj_free, _orc_scalars(h, SELL_SCALARS))


def sell_spmv(arrays: dict, m, n, x, alpha=1.0, beta=0.0, y=None, use_ref=False, C_=64, sigma=1):
    """SELL multiply on a dict of the seven SELL arrays (any builder's)."""
    vv = arrays["vVal"]
    dt = vv.dtype
    y = np.zeros(m, dtype=dt) if y is None else np.array(y, dtype=dt, copy=True)
    x = np.ascontiguousarray(x, dtype=dt)
    a = {k: np.ascontiguousarray(v) for k, v in arrays.items()}
    nb = a["vCl"].shape[0]
    if use_ref:
        getattr(ref, "ref_sell_spmv_" + _sfx(dt))(
            m, n, nb, C_, sigma, _p(a["vCl"]), _p(a["vCs"]), _p(a["vBs"]), _p(a["vSBs"]),
            a["vCol"].shape[0], _p(a["vCol"]), _p(a["vPerm"]), _p(a["vVal"]), _p(x), alpha, beta,
            _p(y))
    else:
        getattr(orc, "orc_sell_spmv_" + _sfx(dt))(
            m, nb, _p(a["vCl"]), _p(a["vCs"]), _p(a["vBs"]), _p(a["vSBs"]), _p(a["vCol"]),
            _p(a["vPerm"]), _p(a["vVal"]), _p(x), alpha, beta, _p(y))
    return y


# --------------------------------------------------------------- CSR5
CSR5_ARRAYS = ("_csr_row_pointer", "_csr_column_index", "_csr_value", "_csr5_tile_pointer",
               "_csr5_tile_descriptor", "_csr5_tile_empty_offset_pointer",
               "_csr5_tile_empty_offset")
CSR5_SCALARS = ("_m", "_n", "_nnz", "_p", "_omega", "_sigma", "_bit_y_offset", "_bit_seg_offset",
                "_tile_desc_len", "_tail_tile_start", "_num_offsets")


def csr5_build(m, n, rst, col, val, omega=64, sigma=16):
    h = getattr(orc, "orc_csr5_build_" + _sfx(val.dtype))(m, n, col.shape[0], _p(rst), _p(col),
                                                         _p(val), omega, sigma)
    o = Obj(h, orc.orc_obj_array, orc.orc_obj_free, _orc_scalars(h, CSR5_SCALARS))
    o.sfx = _sfx(val.dtype)
    return o


def obj_from(arrays: dict, scalars: dict, sfx: str) -> Obj:
    """An oracle structure holding copies of the given arrays and scalars."""
    h = orc.orc_obj_new()
    for k, a in arrays.items():
        a = np.ascontiguousarray(a)
        orc.orc_obj_set_array(h, k.encode(), _p(a), a.shape[0], a.dtype.itemsize)
    for k, v in scalars.items():
        orc.orc_obj_set_scalar(h, k.encode(), float(v))
    o = Obj(h, orc.orc_obj_array, orc.orc_obj_free, dict(scalars))
    o.sfx = sfx
    return o


def vnec_ird(m, n, rst, col, nthreads, nlanes):
    """(IRD, coords[nthreads, 4]) as the reference's IRD_VNEC computes them
    on one thread; coords rows are (start.x, start.y, end.x, end.y)."""
    coords = np.zeros((nthreads, 4), dtype=np.int32)
    ird = orc.orc_vnec_ird(m, n, col.shape[0], _p(rst), _p(col), nthreads, nlanes, _p(coords))
    return ird, coords


def cvr_spmv_ref(arrays: dict, scalars: dict, x, alpha=1.0, beta=0.0, y=None):
    """The reference's own scalar CVR kernel on a CVR layout."""
    vv = arrays["vPack_vec_vals"]
    dt = vv.dtype
    m, n = int(scalars["numRows"]), int(scalars["numCols"])
    y = np.zeros(m, dtype=dt) if y is None else np.array(y, dtype=dt, copy=True)
    x = np.ascontiguousarray(x, dtype=dt)
    a = {k: np.ascontiguousarray(v) for k, v in arrays.items()}
    getattr(ref, "ref_cvr_spmv_" + _sfx(dt))(
        m, n, int(scalars["nItems"]), int(scalars["SIMD_LEN"]), int(scalars["Nthreads"]),
        _p(a["vPack_vec_vals"]), _p(a["vPack_vec_cols"]), a["vPack_vec_record"].shape[0],
        _p(a["vPack_vec_record"]), _p(a["writeback_pos_record"]), _p(x), alpha, beta, _p(y))
    return y


def lav_spmv_ref(arrays: dict, scalars: dict, x, alpha=1.0, beta=0.0, y=None):
    """The reference's own scalar LAV kernel on a LAV layout."""
    vv = arrays["val"]
    dt = vv.dtype
    m, n = int(scalars["m"]), int(scalars["n"])
    y = np.zeros(m, dtype=dt) if y is None else np.array(y, dtype=dt, copy=True)
    x = np.ascontiguousarray(x, dtype=dt)
    a = {k: np.ascontiguousarray(v) for k, v in arrays.items()}
    getattr(ref, "ref_lav_spmv_" + _sfx(dt))(
        m, n, int(scalars["nz"]), int(scalars["SIMD_Lanes"]), int(scalars["segment_count"]),
        int(scalars["chunk_count"]),
        _p(a["chunk_offsets"]), a["chunk_offsets"].shape[0],
        _p(a["chunk_offsets_full_mask"]), a["chunk_offsets_full_mask"].shape[0],
        _p(a["segment_ptr"]), a["segment_ptr"].shape[0],
        _p(a["out_order"]), a["out_order"].shape[0], _p(a["mask"]), a["mask"].shape[0],
        _p(a["val"]), _p(a["col_id"]), a["val"].shape[0], a["sparse_part.Col"].shape[0],
        _p(a["sparse_part.Rst"]), _p(a["sparse_part.Col"]), _p(a["sparse_part.Val"]),
        _p(x), alpha, beta, _p(y))
    return y


def vhcc_spmv(arrays: dict, scalars: dict, x, alpha=1.0, beta=0.0, y=None):
    """The reference vector VHCC multiply restated, on a VHCC layout given
    as its arrays (_thr_info as 10 int32 per partition) and scalars."""
    vv = arrays["_vals_arr"]
    dt = vv.dtype
    m = int(scalars["num_rows"])
    y = np.zeros(m, dtype=dt) if y is None else np.array(y, dtype=dt, copy=True)
    x = np.ascontiguousarray(x, dtype=dt)
    a = {k: np.ascontiguousarray(v) for k, v in arrays.items()}
    getattr(orc, "orc_vhcc_spmv_" + _sfx(dt))(
        m, int(scalars["vec_size"]), int(scalars["num_threads"]), int(scalars["pad_rows"]),
        _p(a["_thr_info"]), _p(a["_veceor_ptr"]), _p(a["_scan_mask"]), _p(a["_row_arr"]),
        _p(a["_col_arr"]), _p(a["_vals_arr"]), _p(x), alpha, beta, _p(y))
    return y


def csr5_spmv(obj, x, alpha=1.0, beta=0.0, y=None):
    """The reference vector CSR5 multiply restated for one thread."""
    dt = np.float64 if obj.sfx == "f64" else np.float32
    m = int(obj.scalars["_m"])
    y = np.zeros(m, dtype=dt) if y is None else np.array(y, dtype=dt, copy=True)
    x = np.ascontiguousarray(x, dtype=dt)
    getattr(orc, "orc_csr5_spmv_" + obj.sfx)(obj._h, _p(x), alpha, beta, _p(y))
    return y

"""Writes the golden fixtures under tests/golden/ from the reference's own
scalar library (oracle/_ref/libref.so, compiled from the reference sources by
oracle/Makefile).  Run in a container where the reference sources exist:

    make -C oracle && python tests/golden/make_golden.py

Fixtures are plain .npz files (inputs and expected outputs only).
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from oracle import oracle as O  # noqa: E402

OUT = Path(__file__).resolve().parent


def rand_csr(m, n, density, seed, empty_every=0):
    rng = np.random.default_rng(seed)
    rows, cols = [], []
    for i in range(m):
        if empty_every and i % empty_every == 0:
            k = 0
        else:
            k = rng.binomial(n, density)
        c = np.sort(rng.choice(n, size=k, replace=False)) if k else np.empty(0, np.int64)
        rows.append(k)
        cols.append(c)
    rst = np.zeros(m + 1, np.int32)
    rst[1:] = np.cumsum(rows)
    col = np.concatenate(cols).astype(np.int32) if cols else np.empty(0, np.int32)
    val = rng.uniform(-1, 1, size=col.shape[0])
    return rst, col, val


def worked_6x6():
    # rows {0:[0,1,4], 1:[2], 2:[5], 3:[0], 4:[3,4], 5:[1,5]}, values 1..10
    cols = [[0, 1, 4], [2], [5], [0], [3, 4], [1, 5]]
    rst = np.zeros(7, np.int32)
    rst[1:] = np.cumsum([len(c) for c in cols])
    col = np.array([c for r in cols for c in r], np.int32)
    val = np.arange(1, 11, dtype=np.float64)
    return rst, col, val


def tie_matrix():
    # 4x4 with R=C=2, p=1: block (0,0) holds exactly rows*p = 2 entries (tie
    # -> COO), block (0,1) 3 entries (CSR), block (1,0) 1, block (1,1) 4.
    dense = np.array([[1, 0, 1, 1], [0, 1, 1, 0], [1, 0, 1, 1], [0, 0, 1, 1]], dtype=float)
    rst = np.zeros(5, np.int32)
    cols, vals = [], []
    for i in range(4):
        nz = np.nonzero(dense[i])[0]
        cols += list(nz)
        vals += [float(10 * i + j + 1) for j in nz]
        rst[i + 1] = len(cols)
    return rst, np.array(cols, np.int32), np.array(vals)


def laplacian(g):
    n = g * g
    rst, col, val = [0], [], []
    for r in range(n):
        gy, gx = divmod(r, g)
        for c, v in ((r - g, -1.0), (r - 1, -1.0), (r, 4.0), (r + 1, -1.0), (r + g, -1.0)):
            if c == r - g and gy == 0: continue
            if c == r + g and gy == g - 1: continue
            if c == r - 1 and gx == 0: continue
            if c == r + 1 and gx == g - 1: continue
            col.append(c); val.append(v)
        rst.append(len(col))
    return np.array(rst, np.int32), np.array(col, np.int32), np.array(val)


HCSR_CASES = [
    # name, (rst, col, val), m, n, R, C, p, dtype
    ("worked6_p1", worked_6x6(), 6, 6, 3, 3, 1.0, "f64"),
    ("worked6_p05", worked_6x6(), 6, 6, 3, 3, 0.5, "f64"),
    ("tie4", tie_matrix(), 4, 4, 2, 2, 1.0, "f64"),
    ("rand37x41", rand_csr(37, 41, 0.2, 1), 37, 41, 5, 7, 2.0, "f64"),
    ("rand_empty", rand_csr(50, 33, 0.3, 2, empty_every=3), 50, 33, 8, 4, 1.5, "f64"),
    ("lap8", laplacian(8), 64, 64, 10, 10, 2.0, "f64"),
    ("lap8_f32", laplacian(8), 64, 64, 16, 12, 2.0, "f32"),
    ("rand200_f32", rand_csr(200, 180, 0.05, 3), 200, 180, 64, 50, 0.7, "f32"),
]


def main():
    assert O.ref is not None, "oracle/_ref/libref.so is not built (make -C oracle)"
    hc = {}
    for name, (rst, col, val), m, n, R, C, p, dt in HCSR_CASES:
        v = val.astype(np.float64 if dt == "f64" else np.float32)
        obj, err = O.hcsr_convert(m, n, rst, col, v, R, C, p, use_ref=True)
        assert err == 0, (name, err)
        hc[f"{name}/rst"] = rst
        hc[f"{name}/col"] = col
        hc[f"{name}/val"] = v
        hc[f"{name}/shape"] = np.array([m, n, R, C], np.int64)
        hc[f"{name}/p"] = np.array([p])
        for f in ("block_ptr", "row_offset", "row_ptr", "col_idx", "values"):
            a = obj.array(f)
            hc[f"{name}/{f}"] = a.view(v.dtype) if f == "values" else a
        x = np.linspace(-1, 1, n).astype(v.dtype)
        hc[f"{name}/x"] = x
        hc[f"{name}/y_hcsr"] = O.hcsr_spmv(obj, x, 1.5, 0.0)
    np.savez(OUT / "hcsr_ref.npz", **hc)

    cs = {}
    for name, (rst, col, val), m, n, seed in (
            ("rand_empty", rand_csr(300, 250, 0.04, 7, empty_every=5), 300, 250, 7),
            ("lap12", laplacian(12), 144, 144, 8)):
        rng = np.random.default_rng(seed)
        x = rng.uniform(-1, 1, n)
        y0 = rng.uniform(-1, 1, m)
        for dt in (np.float64, np.float32):
            v, xx, yy = val.astype(dt), x.astype(dt), y0.astype(dt)
            cs[f"{name}/{np.dtype(dt).name}/y"] = O.csr_spmv(rst, col, v, xx, 0.75, -0.5, yy,
                                                            use_ref=True)
            cs[f"{name}/{np.dtype(dt).name}/val"] = v
            cs[f"{name}/{np.dtype(dt).name}/x"] = xx
            cs[f"{name}/{np.dtype(dt).name}/y0"] = yy
        cs[f"{name}/rst"] = rst
        cs[f"{name}/col"] = col
        cs[f"{name}/shape"] = np.array([m, n], np.int64)
    np.savez(OUT / "csr_ref.npz", **cs)
    print("wrote", OUT / "hcsr_ref.npz", OUT / "csr_ref.npz")


if __name__ == "__main__":
    main()

"""Shared test helpers: small deterministic matrices and the parity check."""
from __future__ import annotations

import numpy as np

from oracle import oracle as O

EPS = {np.dtype(np.float64): 2.0 ** -53, np.dtype(np.float32): 2.0 ** -24}


def rand_csr(m, n, density, seed, empty_every=0, long_rows=()):
    """Random sorted distinct columns per row; rows i % empty_every == 0 are
    empty; rows listed in long_rows get `n // 2` entries."""
    rng = np.random.default_rng(seed)
    cols = []
    for i in range(m):
        if empty_every and i % empty_every == 0:
            k = 0
        elif i in long_rows:
            k = n // 2
        else:
            k = int(rng.binomial(n, density))
        cols.append(np.sort(rng.choice(n, size=k, replace=False)) if k else np.empty(0, np.int64))
    rst = np.zeros(m + 1, np.int32)
    rst[1:] = np.cumsum([len(c) for c in cols])
    col = (np.concatenate(cols) if cols else np.empty(0)).astype(np.int32)
    val = rng.uniform(-1, 1, size=col.shape[0])
    return rst, col, val


def laplacian(g):
    n = g * g
    rst, col, val = [0], [], []
    for r in range(n):
        gy, gx = divmod(r, g)
        if gy > 0: col.append(r - g); val.append(-1.0)
        if gx > 0: col.append(r - 1); val.append(-1.0)
        col.append(r); val.append(4.0)
        if gx < g - 1: col.append(r + 1); val.append(-1.0)
        if gy < g - 1: col.append(r + g); val.append(-1.0)
        rst.append(len(col))
    return np.array(rst, np.int32), np.array(col, np.int32), np.array(val)


def assert_within_bound(y, y_ref, rst, col, val, x, alpha, beta, y0, dtype, what=""):
    """|y - y_ref| <= 4*(len+2)*u*(|alpha|*sum|a||x| + |beta|*|y0|) per row,
    u the unit roundoff of `dtype` (both y and y_ref carry that rounding)."""
    u = EPS[np.dtype(dtype)]
    bound = O.spmv_bound(rst, col, np.asarray(val, np.float64), np.asarray(x, np.float64),
                         alpha, beta, y0, u)
    err = np.abs(np.asarray(y, np.float64) - np.asarray(y_ref, np.float64))
    bad = np.nonzero(err > bound)[0]
    assert bad.size == 0, (f"{what}: {bad.size} rows exceed the bound; first row {bad[0]}: "
                           f"err {err[bad[0]]:.3e} bound {bound[bad[0]]:.3e}")

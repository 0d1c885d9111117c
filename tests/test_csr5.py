"""CSR5: the library's builder against the oracle's independent builder
(bit-exact integer structures), layout invariants, and the reference's
vector CSR5 multiply restated lane by lane (oracle) run on the library's
layout -- the reference RVV kernel itself cannot be built here."""
from __future__ import annotations

import numpy as np
import pytest

import rvvlasparse_amd as R
from helpers import assert_within_bound, laplacian, rand_csr
from oracle import oracle as O

CASES = {
    "rand_empty": (lambda: rand_csr(300, 260, 0.05, 1, empty_every=9, long_rows=(17, 200)), 260),
    "lap20": (lambda: laplacian(20), 400),
    "sparse_empty": (lambda: rand_csr(2000, 3000, 0.002, 3, empty_every=3), 3000),
    "long_rows": (lambda: rand_csr(500, 5000, 0.01, 4, long_rows=(3, 4, 5)), 5000),
    "tiny": (lambda: rand_csr(7, 5, 0.5, 5), 5),
    "all_empty": (lambda: (np.zeros(9, np.int32), np.zeros(0, np.int32), np.zeros(0)), 4),
}


@pytest.mark.parametrize("sigma", [1, 4, 16, 17, 32, 64])
@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_csr5_builder_matches_oracle_and_reference_semantics(name, sigma, dt):
    (rst, col, val), n = CASES[name][0](), CASES[name][1]
    m = len(rst) - 1
    v = val.astype(dt)
    L = R.convert("csr5", R.CSR(m, n, rst, col, val), R.F64 if dt == np.float64 else R.F32,
                  f"sigma={sigma}")
    o = O.csr5_build(m, n, rst, col, v, 64, sigma)
    for f in O.CSR5_ARRAYS:
        assert np.array_equal(L.array(f), o.array(f)), f
    for k in O.CSR5_SCALARS:
        assert L.scalar(k) == o.scalars[k], k
    # invariants: un-transposing the tiles gives back the CSR entry order
    p, W = int(o.scalars["_p"]), 64 * sigma
    c = L.array("_csr_column_index").astype(np.int64)
    if p > 1:
        t = c[:(p - 1) * W].reshape(p - 1, sigma, 64).transpose(0, 2, 1).reshape(-1)
        c = np.concatenate([t, c[(p - 1) * W:]])
    assert np.array_equal(c, col.astype(np.int64))
    # tile_ptr[t] is the row of entry t*W
    tp = L.array("_csr5_tile_pointer") & 0x7FFFFFFF
    rows = np.repeat(np.arange(m), np.diff(rst))
    for t in range(p):
        if t * W < len(col):
            assert tp[t] == rows[t * W]
    # the reference multiply restated, on the library's layout
    arrays = {f: L.array(f) for f in O.CSR5_ARRAYS}
    arrays["_csr_value"] = arrays["_csr_value"].view(dt)
    lib_obj = O.obj_from(arrays, {k: L.scalar(k) for k in O.CSR5_SCALARS},
                         "f64" if dt == np.float64 else "f32")
    rng = np.random.default_rng(sigma)
    x = rng.uniform(-1, 1, n).astype(dt)
    y0 = rng.uniform(-1, 1, m).astype(dt)
    y = O.csr5_spmv(lib_obj, x, 1.5, -0.5, y0)
    y_ref = O.csr_spmv(rst, col, v, x, 1.5, -0.5, y0)
    # The reference algorithm sums segments across lanes as differences of
    # a prefix scan over the whole tile, so a row can pick up rounding of
    # the tile's other partial sums: allow u * (tile size) * (largest row
    # scale) on top of the per-row bound.  A wrong layout errs by O(1).
    u = 2.0 ** -53 if dt == np.float64 else 2.0 ** -24
    scale = 1.5 * O.csr_absdot(rst, col, v, x) + 0.5 * np.abs(y0.astype(np.float64))
    bound = O.spmv_bound(rst, col, v, x, 1.5, -0.5, y0, u) + 4 * 64 * sigma * u * scale.max()
    err = np.abs(y.astype(np.float64) - y_ref)
    assert np.all(err <= bound), f"csr5 {name} s={sigma}: max err {err.max():.3e}"


def test_csr5_rejects_bad_parameters():
    rst, col, val = rand_csr(50, 50, 0.1, 0)
    a = R.CSR(50, 50, rst, col, val)
    with pytest.raises(R.RvvlaError) as e:
        R.convert("csr5", a, R.F64, "omega=32")
    assert e.value.code == -2
    with pytest.raises(R.RvvlaError) as e:
        R.convert("csr5", a, R.F64, "sigma=65")
    assert e.value.code == -3

"""Autotuner (rvvla_autotune): argument checks on the host, and on the GPU
the chosen image of several matrices multiplied against the oracle, the
format list honoured, and the choice cached (in the process and in a file)."""
from __future__ import annotations

import numpy as np
import pytest

import rvvlasparse_amd as R
from helpers import assert_within_bound, laplacian, rand_csr
from oracle import oracle as O

torch = pytest.importorskip("torch")


def _lap(g):
    rst, col, val = laplacian(g)
    return R.CSR(len(rst) - 1, g * g, rst, col, val)


def test_autotune_refuses_bad_arguments():
    a = _lap(8)
    with pytest.raises(R.RvvlaError):  # an unknown format in the list
        R.autotune(a, R.F64, "formats=csr+dense")
    with pytest.raises(R.RvvlaError):  # reps out of range
        R.autotune(a, R.F64, "reps=0")
    bad = R.CSR(a.m, a.n, a.rst, a.col.copy(), a.val)
    bad.col[3] = a.n  # a column past n
    with pytest.raises(R.RvvlaError):
        R.autotune(bad, R.F64)


def _check_choice(a, dtype, fmt, params, seed):
    dev = torch.device("cuda:0")
    dt = np.float64 if dtype == R.F64 else np.float32
    rng = np.random.default_rng(seed)
    x64, y064 = rng.uniform(-1, 1, a.n), rng.uniform(-1, 1, a.m)
    M = R.convert(fmt, a, dtype, params).upload()
    x = torch.from_numpy(x64.astype(dt)).to(dev)
    y = torch.from_numpy(y064.astype(dt)).to(dev)
    M.spmv(x, y, 1.0, 0.5)
    torch.cuda.synchronize()
    v = a.values_for(dtype)
    y_ref = O.csr_spmv(a.rst, a.col, v, x64.astype(dt), 1.0, 0.5, y064.astype(dt))
    assert_within_bound(y.cpu().numpy(), y_ref, a.rst, a.col, v, x64.astype(dt), 1.0, 0.5,
                        y064.astype(dt), dt, f"autotuned {fmt}:{params}")


@pytest.mark.gpu
@pytest.mark.parametrize("dtype", [R.F64, R.F32])
@pytest.mark.parametrize("name", ["lap", "rand", "long"])
def test_autotune_choice_is_correct(name, dtype):
    if name == "lap":
        a = _lap(200)
    elif name == "rand":
        a = R.CSR(30000, 30000, *rand_csr(30000, 30000, 0.0005, 7))
    else:
        a = R.CSR(3000, 60000, *rand_csr(3000, 60000, 0.002, 8, long_rows=(5, 700)))
    print("autotune", name, dtype, a.m, a.nnz, flush=True)
    fmt, params = R.autotune(a, dtype, "reps=5")
    assert fmt in R.FORMATS
    _check_choice(a, dtype, fmt, params, 3)


@pytest.mark.gpu
def test_autotune_format_list_and_cache(tmp_path):
    a = _lap(128)
    fmt, params = R.autotune(a, R.F64, "formats=csr5,reps=3")
    assert fmt == "csr5" and params.startswith("sigma=")
    fmt, params = R.autotune(a, R.F64, "formats=hcsr,reps=3")
    assert fmt == "hcsr"
    _check_choice(a, R.F64, fmt, params, 4)
    path = tmp_path / "tune.txt"
    fmt2, params2 = R.autotune(a, R.F64, f"formats=hcsr,cache={path},reps=3")
    assert (fmt2, params2) == (fmt, params)  # the process cache
    lines = path.read_text().split("\n")
    assert any(l.split(" ")[1:] == ["1", params or "-"] for l in lines if l), lines

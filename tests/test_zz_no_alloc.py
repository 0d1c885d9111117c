"""rvvla_spmv_noalloc (the SpMV entry point that allocates nothing: its
per-product buffers live in the library's 16 MiB device array) against
rvvla_spmv on the same matrix and inputs, through the C-ABI: the same
kernels in the same order, so the results are the same bits.  Every format
and image whose products use such buffers (the LDS- and column-sliced CSR
images, SELL and LAV with split slices / chunks, CSR5, VHCC, VNEC) and two
that use none (HCSR, CVR), each called twice in a row, the second time on a
smaller matrix; and a product needing more than 16 MiB is refused before
anything runs."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import rvvlasparse_amd as R
from helpers import rand_csr

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

ERR_LENGTH = -11


def _dev():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch.device("cuda:0")


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _pair(name):
    """(larger, smaller) matrix of one case."""
    if name == "powerlaw":
        return (R.gen_powerlaw(40000, 1.8, 6000, 600000, 256, seed=3),
                R.gen_powerlaw(9000, 1.8, 3000, 120000, 256, seed=4))
    if name == "wide":
        big = rand_csr(3000, 400000, 0.00005, 21, empty_every=7, long_rows=(5,))
        small = rand_csr(800, 150000, 0.0001, 22, empty_every=5)
        return R.CSR(3000, 400000, *big), R.CSR(800, 150000, *small)
    big = rand_csr(2500, 3000, 0.004, 23, empty_every=13, long_rows=(4, 1500))
    small = rand_csr(700, 900, 0.01, 24, long_rows=(3,))
    return R.CSR(2500, 3000, *big), R.CSR(700, 900, *small)


CASES = [
    ("csr", "xlds=1", "wide"),
    ("csr", "xslices=8", "wide"),
    ("csr", "sell=1", "long"),
    ("sell", "", "long"),
    ("lav", "", "powerlaw"),
    ("csr5", "", "powerlaw"),
    ("vhcc", "", "powerlaw"),
    ("vnec", "", "powerlaw"),
    ("hcsr", "", "long"),
    ("cvr", "", "long"),
]


@pytest.mark.parametrize("dtype", [R.F64, R.F32, R.BF16])
@pytest.mark.parametrize("fmt,params,name", CASES)
def test_noalloc_matches_spmv(fmt, params, name, dtype):
    dev = _dev()
    io = torch.float64 if dtype == R.F64 else torch.float32
    for step, a in enumerate(_pair(name)):
        print(f"noalloc {fmt} {params} {name} dtype={dtype} call {step} m={a.m} nnz={a.nnz}",
              flush=True)
        M = R.convert(fmt, a, dtype, params).upload()
        rng = np.random.default_rng(step)
        x = torch.from_numpy(rng.uniform(-1, 1, a.n)).to(io).to(dev)
        y0 = torch.from_numpy(rng.uniform(-1, 1, a.m)).to(io).to(dev)
        y_ref, y_new = y0.clone(), y0.clone()
        s = torch.cuda.current_stream().cuda_stream
        R.check(R.lib.rvvla_spmv(C.c_void_p(M.handle), 1.25, _ptr(x), -0.5, _ptr(y_ref), s))
        R.check(R.lib.rvvla_spmv_noalloc(C.c_void_p(M.handle), 1.25, _ptr(x), -0.5, _ptr(y_new), s))
        torch.cuda.synchronize()
        assert torch.equal(y_new, y_ref), f"{fmt} {params} {name}: call {step} differs"


def test_noalloc_refuses_large_scratch():
    """An LDS-sliced CSR image of ~2.6 M single-entry pieces needs ~21 MB of
    partials: refused with RVVLA_ERR_LENGTH, y untouched."""
    dev = _dev()
    rng = np.random.default_rng(7)
    m, n, k = 130000, 4_000_000, 20
    col = np.sort(rng.integers(0, n, size=(m, k)), axis=1).astype(np.int32)
    rst = np.arange(m + 1, dtype=np.int32) * k
    a = R.CSR(m, n, rst, col.reshape(-1), rng.uniform(-1, 1, m * k))
    M = R.convert("csr", a, R.F64, "xlds=1").upload()
    x = torch.zeros(n, dtype=torch.float64, device=dev)
    y = torch.full((m,), 3.0, dtype=torch.float64, device=dev)
    rc = R.lib.rvvla_spmv_noalloc(C.c_void_p(M.handle), 1.0, _ptr(x), 0.0, _ptr(y),
                                  torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    assert rc == ERR_LENGTH
    assert torch.all(y == 3.0)

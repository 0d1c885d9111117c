#!/usr/bin/env python3
"""SpMV throughput on one MI355X: GB/s and fraction of the HBM roofline per
format, FP64 and FP32, on the 5-point Laplacian (2048 x 2048 grid,
n = 4,194,304, nnz = 20,963,328), inputs resident in HBM.

A step is one SpMV y = A x (alpha = 1, beta = 0) in one format.  Every
format and precision is timed for `--warmup` + `--steps` multiplies; `value`
is the best FP64 format's throughput under the compulsory byte model
  B = nnz*(s_val + 4) + (m+1)*4 + n*s_x + m*s_y
(the CSR-equivalent bytes, the same for every format).  The average kernel
duration comes from HIP events recorded on the stream the kernels run on.
The CPU baselines are the reference's own scalar OpenMP CSR (and, at the
n = 65,536 configuration, HCSR) SpMVs compiled from its sources into
oracle/_ref (the oracle's restatement where that library is absent), timed
on this host's cores in the same run, with all cores and on one thread, for
the Laplacian (FP64, FP32), the power-law matrix (FP64, FP32) and the
n = 65,536 configuration; the power-law errors of the FP32 and 16-bit forms
are stated against the reference's FP64 CPU product.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--grid 2048]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import statistics
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))
# host CPUs allowed to this process; the CPU baseline runs in a child
# process whose OpenMP threads are bound one per physical core (binding them
# here would pin this process's main thread, and the GPU runtime's threads,
# to one core)
HOST_CPUS = sorted(os.sched_getaffinity(0))


def physical_cores(cpus) -> int:
    """Distinct (package, core) pairs among `cpus` (sysfs topology)."""
    seen = set()
    for c in cpus:
        base = f"/sys/devices/system/cpu/cpu{c}/topology/"
        try:
            seen.add((Path(base + "physical_package_id").read_text().strip(),
                      Path(base + "core_id").read_text().strip()))
        except OSError:
            return len(cpus)
    return max(1, len(seen))

METRIC = "SpMV GB/s & % HBM peak per format; BF16-val vs FP64 speedup @ stated rel err"
PEAK_GBS = 8000.0  # MI355X HBM3E, AMD specification


def algo_bytes(m, n, nnz, s_v, s_x, s_y, beta_nonzero=False):
    return nnz * (s_v + 4) + (m + 1) * 4 + n * s_x + m * s_y * (2 if beta_nonzero else 1)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def time_format(M, x, y, steps, warmup, torch):
    """(event ms per SpMV, wall ms per SpMV) for `steps` back-to-back SpMVs.
    A one-element fill runs first, outside the timed region: in a counter
    run it marks where each leg's products begin in the dispatch order."""
    stream = torch.cuda.current_stream()
    torch.zeros(1, device=x.device)
    for _ in range(warmup):
        M.spmv(x, y, 1.0, 0.0, stream)
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    for _ in range(steps):
        M.spmv(x, y, 1.0, 0.0, stream)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    return ev0.elapsed_time(ev1) / steps, wall


def copy_probe(torch, nbytes=1 << 30, steps=20):
    """Achievable device-to-device copy rate (read + write bytes per second)
    with torch's copy kernel on `nbytes` buffers; also the known-size kernel
    used to scale the FETCH_SIZE / WRITE_SIZE counters."""
    src = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    del src, dst
    return {"bytes_moved": 2 * nbytes, "ms": round(ms, 4), "GBs": round(2 * nbytes / ms / 1e6, 1)}


# CPU baseline jobs: (matrix, format, precision, seconds with all cores,
# seconds on one thread); each is timed for at least 5 (all cores) or 3 (one
# thread) runs
CPU_JOBS = [("laplacian", "csr", "f64", 3.0, 1.5), ("laplacian", "csr", "f32", 2.0, 1.0),
            ("powerlaw", "csr", "f64", 3.0, 2.0), ("powerlaw", "csr", "f32", 2.0, 1.5),
            ("c1", "csr", "f64", 1.0, 0.5), ("c1", "hcsr", "f64", 1.0, 0.5)]


def cpu_baseline(mats):
    """The reference's scalar OpenMP SpMVs (reference src/scalar/spmv_mtx_crs.cpp
    :22-71 and, for HCSR, src/scalar/crs_to_hcsr.cpp + spmv_mtx_hcsr.cpp
    :104-134 at the reference's R = C = 2000, p = 2), compiled from its sources
    into oracle/_ref with the reference's flags (kind "reference"), or where
    that library is absent the oracle's CSR restatement (kind "port"; HCSR
    jobs skipped), on the host cores, in a child process with OMP_NUM_THREADS
    = the physical cores allowed to this process, OMP_PROC_BIND=close,
    OMP_PLACES=cores: per job (CPU_JOBS) the median of the full SpMVs that
    fit its budget after warm-ups, then the same on one thread.  `mats`:
    name -> (CSR, x).  Returns the Laplacian FP64 figure as the headline, the
    rest under "workloads", and the power-law FP64 result vector of the CPU
    run (the reference's own FP64 CSR product) for the error figures."""
    cores = physical_cores(HOST_CPUS)
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        cores = min(cores, int(os.environ["OMP_NUM_THREADS"]))
    tmpdir = "/dev/shm" if os.path.isdir("/dev/shm") else None
    with tempfile.TemporaryDirectory(dir=tmpdir) as td:
        for name, (a, x) in mats.items():
            for k, v in (("rst", a.rst), ("col", a.col), ("val", a.val), ("x", x)):
                np.save(os.path.join(td, f"{name}.{k}.npy"), v)
        env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="close",
                   OMP_PLACES="cores")
        res = subprocess.run([sys.executable, str(Path(__file__).resolve()), "--cpu-child", td],
                             env=env, stdout=subprocess.PIPE, text=True, timeout=900, check=True)
        out = json.loads(res.stdout.strip().splitlines()[-1])
        ypath = os.path.join(td, "powerlaw.y64.npy")
        y_pl = np.load(ypath) if os.path.exists(ypath) else None
    wl = {}
    for job in out["jobs"]:
        name, fmt, prec = job["job"]
        a = mats[name][0]
        sv = 8 if prec == "f64" else 4
        byts = algo_bytes(a.m, a.n, a.nnz, sv, sv, sv)
        wl[f"{name}/{fmt}/{prec}"] = {
            "GBs": round(byts / job["med"] / 1e9, 2), "ms": round(job["med"] * 1e3, 4),
            "runs": job["cnt"], "threads": cores,
            "single_thread": {"GBs": round(byts / job["med1"] / 1e9, 2),
                              "ms": round(job["med1"] * 1e3, 4), "runs": job["cnt1"]}}
    head = wl["laplacian/csr/f64"]  # always present
    lap = mats["laplacian"][0]
    return {"value": head["GBs"], "unit": "GB/s", "cores": cores, "kind": out["kind"],
            "sample": f"full Laplacian {int(np.sqrt(lap.m))}^2 FP64 CSR SpMV, median of "
                      f"{head['runs']} runs after warm-ups ({head['ms']:.3f} ms each, {cores} "
                      f"threads, one per physical core, OMP_PROC_BIND=close); other workloads "
                      f"under 'workloads', each also on one thread",
            "single_thread": {"value": head["single_thread"]["GBs"], "unit": "GB/s", "cores": 1,
                              "sample": f"median of {head['single_thread']['runs']} runs "
                                        f"({head['single_thread']['ms']:.3f} ms each)"},
            "workloads": wl}, y_pl


def cpu_child(td):
    """The timed part of cpu_baseline (runs in the child process; no GPU)."""
    from oracle import oracle as O
    if O.orc is None:
        O.build(ref=False)
        import importlib
        importlib.reload(O)
    kind = "reference" if O.ref is not None else "port"
    results = []
    threads = int(os.environ.get("OMP_NUM_THREADS", "1"))
    loaded = {}
    for name, fmt, prec, secs, secs1 in CPU_JOBS:
        if (fmt == "hcsr" and kind != "reference") or not os.path.exists(
                os.path.join(td, f"{name}.rst.npy")):
            continue
        if name not in loaded:
            loaded = {name: {k: np.load(os.path.join(td, f"{name}.{k}.npy"))
                             for k in ("rst", "col", "val", "x")}}
        d = loaded[name]
        dt = np.float64 if prec == "f64" else np.float32
        rst, col = np.ascontiguousarray(d["rst"]), np.ascontiguousarray(d["col"])
        val, x = np.ascontiguousarray(d["val"].astype(dt)), np.ascontiguousarray(d["x"].astype(dt))
        m = len(rst) - 1
        if kind == "reference":
            if fmt == "hcsr":
                h = O.ref.ref_hcsr_bench_new_f64(m, len(x), len(col), rst.ctypes.data,
                                                 col.ctypes.data, val.ctypes.data, x.ctypes.data,
                                                 2000, 2000, 2.0)
            else:
                new = O.ref.ref_csr_bench_new_f64 if prec == "f64" else O.ref.ref_csr_bench_new_f32
                h = new(m, len(x), len(col), rst.ctypes.data, col.ctypes.data, val.ctypes.data,
                        x.ctypes.data)

            def run():
                O.ref.ref_bench_run(h)
        else:
            y = np.zeros(m, dt)

            def run():
                O.csr_spmv(rst, col, val, x, 1.0, 0.0, y)

        def sample(nthreads, budget, least):
            O.orc.orc_set_threads(nthreads)  # the process's OpenMP thread count
            for _ in range(2):
                run()
            ts = []
            t_end = time.perf_counter() + budget
            while len(ts) < least or time.perf_counter() < t_end:
                t0 = time.perf_counter()
                run()
                ts.append(time.perf_counter() - t0)
            return statistics.median(ts), len(ts)

        med, cnt = sample(threads, secs, 5)
        if kind == "reference" and (name, fmt, prec) == ("powerlaw", "csr", "f64"):
            yv = np.ctypeslib.as_array(C.cast(O.ref.ref_bench_y(h), C.POINTER(C.c_double)), (m,))
            np.save(os.path.join(td, "powerlaw.y64.npy"), yv.copy())
        med1, cnt1 = sample(1, secs1, 3)
        if kind == "reference":
            O.ref.ref_csr_bench_free(h)
        results.append({"job": [name, fmt, prec], "med": med, "cnt": cnt, "med1": med1,
                        "cnt1": cnt1})
    print(json.dumps({"kind": kind, "jobs": results}), flush=True)


def powerlaw_leg(R, torch, steps, warmup, fmts):
    """Power-law matrix (n = 4,194,304, lengths ~ k^-1.8 on [1, 50000] rescaled
    to 64M entries, half within +-1024 of the diagonal, values N(0,1)): the
    FP64 formats, then FP32 and the 16-bit value forms (FP32 x/y and
    accumulation) in the best FP64 format.  Errors of the reduced forms are
    first stated against the FP64 result of the same format on the GPU;
    main() replaces them by errors against the reference's FP64 CSR product
    on the CPU when the CPU baseline runs.  Returns the leg's figures, the
    matrix, x and the GPU results."""
    dev = torch.device("cuda:0")
    log("bench: generating power-law matrix")
    a = R.gen_powerlaw(4194304, 1.8, 50000, 64_000_000, 1024, seed=0)
    x64 = R.gen_uniform(a.n, -1.0, 1.0, seed=0, purpose=3)
    out = {"workload": f"power-law n={a.m}, nnz={a.nnz}, alpha=1.8, max row {int(a.row_lengths().max())}"}
    per = {}
    xs = {R.F64: torch.from_numpy(x64).to(dev), R.F32: torch.from_numpy(x64).float().to(dev)}
    ys = {}
    for fmt in fmts:
        log(f"bench: powerlaw {fmt} f64")
        try:
            M = R.convert(fmt, a, R.F64).upload()
        except R.RvvlaError as e:
            per[fmt] = {"f64": {"error": str(e)}}
            continue
        y = torch.zeros(a.m, dtype=torch.float64, device=dev)
        ev_ms, wall_ms = time_format(M, xs[R.F64], y, steps, warmup, torch)
        byts = algo_bytes(a.m, a.n, a.nnz, 8, 8, 8)
        per[fmt] = {"f64": {"ms": round(ev_ms, 5), "GBs": round(byts / ev_ms / 1e6, 1),
                            "frac": round(byts / ev_ms / 1e6 / PEAK_GBS, 4)}}
        del M
    best = max((f for f in per if "ms" in per[f]["f64"]), key=lambda f: per[f]["f64"]["GBs"])
    out["best_format"] = best
    for dtype, tag, sv in ((R.F64, "f64", 8), (R.F32, "f32", 4), (R.BF16, "bf16", 2), (R.F16, "f16", 2)):
        log(f"bench: powerlaw {best} {tag}")
        M = R.convert(best, a, dtype).upload()
        io = torch.float64 if dtype == R.F64 else torch.float32
        y = torch.zeros(a.m, dtype=io, device=dev)
        ev_ms, _ = time_format(M, xs[R.F64 if dtype == R.F64 else R.F32], y, steps, warmup, torch)
        sx = 8 if dtype == R.F64 else 4
        byts = algo_bytes(a.m, a.n, a.nnz, sv, sx, sx)
        per[best][tag] = {"ms": round(ev_ms, 5), "GBs": round(byts / ev_ms / 1e6, 1),
                          "frac": round(byts / ev_ms / 1e6 / PEAK_GBS, 4)}
        ys[tag] = y.double()
        del M
    for tag in ("f32", "bf16", "f16"):
        per[best][tag]["speedup_vs_f64"] = round(per[best]["f64"]["ms"] / per[best][tag]["ms"], 3)
    out["formats"] = per
    out["bf16_vs_f64_speedup"] = per[best]["bf16"]["speedup_vs_f64"]
    set_errors(out, {t: v.cpu().numpy() for t, v in ys.items()}, ys["f64"].cpu().numpy(),
               "the FP64 result of the same format on the GPU")
    return out, a, x64, {t: v.cpu().numpy() for t, v in ys.items()}


def set_errors(out, ys, y64, against):
    """Normwise ||y - y64||_2 / ||y64||_2 and max-row max_i |y_i - y64_i| /
    ||y64||_inf of the FP32, BF16 and FP16 results against y64."""
    best = out["best_format"]
    for tag in ("f32", "bf16", "f16"):
        d = ys[tag] - y64
        out["formats"][best][tag]["normwise_rel_err"] = float(np.linalg.norm(d) / np.linalg.norm(y64))
        out["formats"][best][tag]["maxrow_rel_err"] = float(np.abs(d).max() / np.abs(y64).max())
    out["bf16_normwise_rel_err"] = out["formats"][best]["bf16"]["normwise_rel_err"]
    out["errors_against"] = against


def config1_leg(R, torch, steps, warmup):
    """The reference's own CPU-path configuration on the GPU: n = 65,536, 16
    entries per row (the diagonal and 15 distinct uniform columns), values
    and x uniform[-1, 1], seed 0, FP64, CSR and HCSR (the reference's
    R = C = 2000, p = 2).  Returns the timings, the matrix and x."""
    dev = torch.device("cuda:0")
    a = R.gen_random_rows(65536, 16, seed=0, include_diag=True)
    x64 = R.gen_uniform(a.n, -1.0, 1.0, seed=0, purpose=3)
    x = torch.from_numpy(x64).to(dev)
    y = torch.zeros(a.m, dtype=torch.float64, device=dev)
    byts = algo_bytes(a.m, a.n, a.nnz, 8, 8, 8)
    gpu = {}
    for fmt in ("csr", "hcsr"):
        M = R.convert(fmt, a, R.F64).upload()
        ev_ms, wall_ms = time_format(M, x, y, steps, warmup, torch)
        gpu[fmt] = {"us": round(ev_ms * 1e3, 2), "wall_us": round(wall_ms * 1e3, 2),
                    "GBs": round(byts / ev_ms / 1e6, 1)}
        del M
    return {"workload": f"n={a.m}, nnz={a.nnz} (diagonal + 15 uniform columns per row), FP64",
            "gpu": gpu}, a, x64


def read_probe(R, torch, nbytes=1 << 31, steps=20):
    """Achievable streaming-read rate: the library's read probe (16 workgroups
    a CU, each streaming its own contiguous share with 16-byte nontemporal
    loads) over a 2 GiB buffer, HIP events on the launch stream."""
    buf = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    sink = torch.zeros(1 << 16, dtype=torch.int32, device="cuda")
    s = torch.cuda.current_stream().cuda_stream

    def go():
        R.check(R.lib.rvvla_read_probe(buf.data_ptr(), nbytes, sink.data_ptr(), s))

    for _ in range(3):
        go()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        go()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    del buf, sink
    return {"bytes_read": nbytes, "ms": round(ms, 4), "GBs": round(nbytes / ms / 1e6, 1),
            "kernel": "k_read_probe"}


def autograd_leg(R, torch, steps, warmup):
    """Differentiable SpMM Y = A(theta) X: A n=2^20 with 32 distinct uniform
    columns per row, theta N(0,1/32), X 2^20 x 64, FP32; loss = sum(W*Y).
    Times the forward and the backward (dX = A^T dY and dtheta) with HIP
    events on the current stream."""
    from rvvlasparse_amd.autograd import SparseStructure, spmm
    n, k = 1 << 20, 64
    log("bench: autograd SpMM setup")
    a = R.gen_random_rows(n, 32, seed=5, include_diag=False, normal=True, stdev=32 ** -0.5)
    A = SparseStructure(a.m, a.n, a.rst, a.col, torch.float32)
    g = torch.Generator(device="cuda").manual_seed(0)
    theta = torch.tensor(a.val, dtype=torch.float32, device="cuda", requires_grad=True)
    X = torch.randn(n, k, device="cuda", generator=g, requires_grad=True)
    W = torch.randn(n, k, device="cuda", generator=g)

    def fwd():
        return spmm(A, theta, X)

    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    for _ in range(warmup):
        Y = fwd()
        Y.backward(W)
    torch.cuda.synchronize()
    f_ms = b_ms = 0.0
    for _ in range(steps):
        ev[0].record()
        Y = fwd()
        ev[1].record()
        Y.backward(W)
        ev[2].record()
        torch.cuda.synchronize()
        f_ms += ev[0].elapsed_time(ev[1])
        b_ms += ev[1].elapsed_time(ev[2])
    f_ms /= steps
    b_ms /= steps
    fb = a.nnz * 8 + (n + 1) * 4 + 2 * n * k * 4  # vals+cols, row ptr, X and Y
    # the columns are uniform over 2^20 rows of X (268 MB, more than the
    # caches hold): every entry gathers its own k-row of X
    gb = a.nnz * (8 + k * 4) + (n + 1) * 4 + n * k * 4
    return {"workload": f"SpMM n={n}, nnz={a.nnz}, k={k}, FP32", "forward_ms": round(f_ms, 4),
            "backward_ms": round(b_ms, 4), "backward_over_forward": round(b_ms / f_ms, 3),
            "forward_GFLOPs": round(2 * a.nnz * k / f_ms / 1e6, 1),
            "forward_compulsory_GBs": round(fb / f_ms / 1e6, 1),
            "forward_gather_GBs": round(gb / f_ms / 1e6, 1)}


def batch_leg(R, torch, steps, warmup, count=512):
    """512 square matrices, n log-uniform in [1e3, 2e5], 8..64 distinct
    uniform columns per row, values N(0,1), k in 1..16 right-hand sides,
    FP32, one persistent launch per step (autotuned item kernels)."""
    rng = np.random.default_rng(11)
    ns = np.exp(rng.uniform(np.log(1e3), np.log(2e5), count)).astype(int)
    ks = [int(k) for k in rng.integers(1, 17, count)]
    log("bench: batch generation")
    mats = [R.gen_random_rows_var(int(n), 8, 64, seed=100 + i) for i, n in enumerate(ns)]
    nnz = sum(a.nnz for a in mats)
    log(f"bench: batch create ({count} matrices, {nnz} entries)")
    t0 = time.perf_counter()
    B = R.Batch(mats, ks, R.F32)
    create_s = time.perf_counter() - t0
    s = torch.cuda.current_stream().cuda_stream
    for _ in range(warmup):
        B.run(1.0, 0.0, s)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        B.run(1.0, 0.0, s)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    B.capture(1.0, 0.0)
    e0.record()
    for _ in range(steps):
        B.replay(s)
    e1.record()
    torch.cuda.synchronize()
    gms = e0.elapsed_time(e1) / steps
    byts = sum(a.nnz * 8 + (a.m + 1) * 4 + a.n * k * 4 + a.m * k * 4 for a, k in zip(mats, ks))
    flops = sum(2 * a.nnz * k for a, k in zip(mats, ks))
    fmts = [B.block(i)[2] for i in range(count)]
    return {"workload": f"{count} matrices, {sum(a.m for a in mats)} rows, {nnz} entries, k 1..16, FP32",
            "ms": round(ms, 4), "graph_ms": round(gms, 4), "GBs": round(byts / ms / 1e6, 1),
            "GFLOPs": round(flops / ms / 1e6, 1), "create_s": round(create_s, 2),
            "item_kernels": {"csr": fmts.count(0), "sell": fmts.count(2)}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--grid", type=int, default=2048)
    ap.add_argument("--formats", default="all")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-powerlaw", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the autotuner, autograd-SpMM and batched legs")
    ap.add_argument("--copy-only", action="store_true",
                    help="run only the 1 GiB copy probe (counter calibration)")
    ap.add_argument("--only", choices=("batch", "autograd"), default=None,
                    help="run only the batched or the autograd-SpMM leg (for profiles)")
    ap.add_argument("--cpu-child", default=None, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_child:
        cpu_child(args.cpu_child)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "row-block sharding is a design note, not a measured path", file=sys.stderr)
        sys.exit(2)

    import torch

    import rvvlasparse_amd as R

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    dev = torch.device("cuda:0")
    if args.copy_only:
        print(json.dumps({"copy_probe": copy_probe(torch, steps=args.steps)}), flush=True)
        return
    if args.only == "batch":
        print(json.dumps({"batched": batch_leg(R, torch, args.steps, args.warmup)}), flush=True)
        return
    if args.only == "autograd":
        print(json.dumps({"autograd_spmm": autograd_leg(R, torch, args.steps, args.warmup)}),
              flush=True)
        return
    g = args.grid
    log(f"bench: generating Laplacian {g}x{g}")
    a = R.gen_laplacian5(g)
    x64 = R.gen_uniform(a.n, -1.0, 1.0, seed=0, purpose=3)
    fmts = [f for f in R.FORMATS if f in READY] if args.formats == "all" else args.formats.split(",")

    traffic_db = {}
    tf = ROOT / "profiles" / "traffic.json"
    if tf.exists():
        traffic_db = json.loads(tf.read_text())

    # each format with the reference's parameters, plus HCSR with blocks of
    # 2048 x 2048 (the grid's row length: a block-row's neighbours fall in
    # whole blocks)
    variants = [(f, f, "") for f in fmts]
    if "hcsr" in fmts:
        variants.append(("hcsr_R2048", "hcsr", "R=2048,C=2048"))
    results = {}
    for dtype, tag, s in ((R.F64, "f64", 8), (R.F32, "f32", 4)):
        io = torch.float64 if dtype == R.F64 else torch.float32
        x = torch.from_numpy(x64).to(io).to(dev)
        y = torch.zeros(a.m, dtype=io, device=dev)
        byts = algo_bytes(a.m, a.n, a.nnz, s, s, s)
        for name, fmt, params in variants:
            log(f"bench: {name} {tag}: convert + upload")
            M = R.convert(fmt, a, dtype, params).upload()
            log(f"bench: {name} {tag}: timing {args.warmup}+{args.steps}")
            ev_ms, wall_ms = time_format(M, x, y, args.steps, args.warmup, torch)
            results.setdefault(name, {})[tag] = {
                "ms": round(ev_ms, 5), "wall_ms": round(wall_ms, 5),
                "GBs": round(byts / (ev_ms * 1e-3) / 1e9, 1),
                "frac": round(byts / (ev_ms * 1e-3) / 1e9 / PEAK_GBS, 4),
                "GFLOPs": round(2 * a.nnz / (ev_ms * 1e-3) / 1e9, 1),
                "format_bytes": M.device_bytes,
            }
            del M
    best = max(results, key=lambda f: results[f]["f64"]["GBs"])
    r = results[best]["f64"]
    byts64 = algo_bytes(a.m, a.n, a.nnz, 8, 8, 8)
    key = f"{best}/f64/laplacian{g}"
    tr = traffic_db.get(key)
    roofline = {"bound": "hbm", "achieved": r["GBs"], "peak": PEAK_GBS, "unit": "GB/s",
                "frac": round(r["GBs"] / PEAK_GBS, 4),
                "traffic": tr["bytes_per_launch"] if tr else None,
                "algorithmic_bytes": byts64, "kernel": best,
                # the same model for every format: a format that stores less
                # than CSR (HCSR's 16-bit columns) can pass the peak on it;
                # traffic_GBs is the rate of the bytes it actually moves
                "bytes_model": "CSR-equivalent: nnz*(s_val+4) + (m+1)*4 + n*s_x + m*s_y"}
    if tr:  # the rate of the bytes the kernel actually moves (HCSR's 16-bit columns move fewer)
        roofline["traffic_GBs"] = round(tr["bytes_per_launch"] / (r["ms"] * 1e-3) / 1e9, 1)
        roofline["traffic_frac"] = round(roofline["traffic_GBs"] / PEAK_GBS, 4)
    line = {
        "metric": METRIC, "value": r["GBs"], "unit": "GB/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": r["wall_ms"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"5-point Laplacian {g}x{g} grid, n={a.m}, nnz={a.nnz}, FP64 "
                               f"SpMV y=Ax, best of {len(fmts)} formats",
                   "best_format": best, "parallelism": "single GPU"},
        "roofline": roofline,
        "formats": results,
    }
    line["copy_probe"] = copy_probe(torch)
    line["read_probe"] = read_probe(R, torch)
    if not args.no_extra:
        log("bench: autotuner on the Laplacian (FP64)")
        t0 = time.perf_counter()
        tf_, tp_ = R.autotune(a, R.F64, "reps=20")
        line["autotune"] = {"laplacian/f64": f"{tf_}:{tp_}", "candidates": "all eight formats",
                            "seconds": round(time.perf_counter() - t0, 1)}
    mats = {"laplacian": (a, x64)}
    log("bench: config-1 (n = 65,536, 16 per row) CSR and HCSR")
    line["config1"], a1, x1 = config1_leg(R, torch, args.steps, args.warmup)
    mats["c1"] = (a1, x1)
    pl_ys = None
    if not args.no_powerlaw:
        line["powerlaw"], apl, xpl, pl_ys = powerlaw_leg(R, torch, args.steps, args.warmup, fmts)
        mats["powerlaw"] = (apl, xpl)
    if not args.no_extra:
        line["autograd_spmm"] = autograd_leg(R, torch, min(args.steps, 20), 3)
        line["batched"] = batch_leg(R, torch, min(args.steps, 20), 3)
    if not args.no_cpu:
        log("bench: CPU baselines")
        line["cpu_baseline"], y_ref = cpu_baseline(mats)
        wl = line["cpu_baseline"]["workloads"]
        line["config1"]["cpu_reference"] = {k.split("/")[1]: v for k, v in wl.items()
                                            if k.startswith("c1/")}
        if pl_ys is not None and y_ref is not None:
            set_errors(line["powerlaw"], pl_ys, y_ref,
                       "the reference's FP64 CSR product on the CPU (oracle/_ref)")
    print(json.dumps(line), flush=True)


READY = ("csr", "hcsr", "sell", "csr5", "lav", "vnec", "vhcc", "cvr")

if __name__ == "__main__":
    main()

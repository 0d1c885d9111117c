"""Matrix Market -> binary CSR converter.

    python -m rvvlasparse_amd.mtx2bin matrix.mtx matrix.bin

Reads a coordinate Matrix Market file (rvvla_mtx_read: real / integer /
pattern, general / symmetric / skew-symmetric, duplicates summed) and writes
the library's binary CSR file (rvvla_bin_write), which rvvla_bin_read and
SparseMatrixLib::read_bin load without parsing.
"""
from __future__ import annotations

import argparse
import sys
import time


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m rvvlasparse_amd.mtx2bin",
                                 description="Matrix Market coordinate file -> binary CSR")
    ap.add_argument("src")
    ap.add_argument("dst")
    args = ap.parse_args(argv)
    import rvvlasparse_amd as R
    t0 = time.perf_counter()
    a, code = R.read_mtx(args.src)
    t1 = time.perf_counter()
    R.write_bin(args.dst, a, code)
    t2 = time.perf_counter()
    print(f"{args.src}: {a.m} x {a.n}, {a.nnz} entries ({code}); read {t1 - t0:.2f} s, "
          f"wrote {args.dst} in {t2 - t1:.2f} s")
    return 0


if __name__ == "__main__":
    sys.exit(main())

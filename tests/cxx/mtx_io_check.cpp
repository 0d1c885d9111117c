// mtx_io_check.cpp -- the C++ file wrappers (mtx_io.h) on the reference's
// CRS container: read a Matrix Market file, write it as binary CSR, read
// that back and compare field by field, including the MM_typecode.
// Usage: mtx_io_check <in.mtx> <tmp.bin>; prints "m n nnz code" and exits 0
// when the round trip is exact.
#include <cstdio>
#include <cstring>

#include "mtx_io.h"

using namespace SparseMatrixLib;

int main(int argc, char** argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s in.mtx tmp.bin\n", argv[0]);
    return 2;
  }
  try {
    spMtxCRS<double> a;
    read_mtx(argv[1], a);
    write_bin(argv[2], a);
    spMtxCRS<float> f;
    read_bin(argv[2], f);
    spMtxCRS<double> b;
    read_bin(argv[2], b);
    bool same = a.m == b.m && a.n == b.n && a.nz == b.nz &&
                std::memcmp(a.matcode, b.matcode, 4) == 0 &&
                std::memcmp(a.Rst, b.Rst, sizeof(int) * (a.m + 1)) == 0 &&
                std::memcmp(a.Col, b.Col, sizeof(int) * a.nz) == 0 &&
                std::memcmp(a.Val, b.Val, sizeof(double) * a.nz) == 0 && f.nz == a.nz;
    for (size_t e = 0; same && e < a.nz; ++e) same = f.Val[e] == (float)a.Val[e];
    std::printf("%zu %zu %zu %.4s\n", a.m, a.n, a.nz, a.matcode);
    if (!same) {
      std::fprintf(stderr, "round trip differs\n");
      return 1;
    }
    try {
      read_mtx("/nonexistent/file.mtx", a);
      std::fprintf(stderr, "missing file not reported\n");
      return 1;
    } catch (const std::runtime_error&) {
    }
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}

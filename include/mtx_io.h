// mtx_io.h -- matrix files into the reference's CRS container.
//
// read_mtx / read_bin fill a spMtxCRS<T> (reference include/CRS.h:10-256)
// from a Matrix Market or binary CSR file, including its MM_typecode;
// write_bin stores one.  They wrap rvvla_mtx_read / rvvla_bin_read /
// rvvla_bin_write (rvvla_hip.h) and throw std::runtime_error with the
// library's message on failure.
#pragma once
#include <cstring>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "CRS.h"
#include "rvvla_hip.h"

namespace SparseMatrixLib {

namespace detail {
inline void io_check(int rc, const std::string& what) {
  if (rc != RVVLA_OK) throw std::runtime_error(what + ": " + rvvla_strerror(rc));
}
template <class T>
void take(rvvla_gen_csr& g, const char code[4], spMtxCRS<T>& out) {
  spMtxCRS<T> a((size_t)g.m, (size_t)g.n, (size_t)g.nnz);
  std::memcpy(a.Rst, g.rst, sizeof(int) * (size_t)(g.m + 1));
  if (g.nnz > 0) std::memcpy(a.Col, g.col, sizeof(int) * (size_t)g.nnz);
  for (int64_t e = 0; e < g.nnz; ++e) a.Val[e] = (T)g.val[e];
  std::memcpy(a.matcode, code, 4);
  rvvla_gen_free(&g);
  out = std::move(a);
}
}  // namespace detail

template <class T>
void read_mtx(const std::string& path, spMtxCRS<T>& out) {
  rvvla_gen_csr g{};
  char code[4];
  detail::io_check(rvvla_mtx_read(path.c_str(), &g, code), "read_mtx " + path);
  detail::take(g, code, out);
}

template <class T>
void read_bin(const std::string& path, spMtxCRS<T>& out) {
  rvvla_gen_csr g{};
  char code[4];
  detail::io_check(rvvla_bin_read(path.c_str(), &g, code), "read_bin " + path);
  detail::take(g, code, out);
}

template <class T>
void write_bin(const std::string& path, const spMtxCRS<T>& a) {
  std::vector<double> v(a.nz);
  for (size_t e = 0; e < a.nz; ++e) v[e] = (double)a.Val[e];
  rvvla_gen_csr g{(int64_t)a.m, (int64_t)a.n, (int64_t)a.nz, a.Rst, a.Col, v.data()};
  detail::io_check(rvvla_bin_write(path.c_str(), &g, a.matcode), "write_bin " + path);
}

}  // namespace SparseMatrixLib

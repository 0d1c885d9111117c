/* CRS.h -- compressed row storage, the input of every converter.
 * Field names and ownership follow reference include/CRS.h:10-256: raw
 * owned arrays Rst[m+1], Col[nz], Val[nz] (int32 indices). */
#pragma once
#include <algorithm>
#include <cstddef>
#include <cstring>
#include <ostream>

#include "device_image.h"

namespace SparseMatrixLib {

typedef char MM_typecode[4];

template <typename ValT>
class spMtxCRS {
 public:
  size_t m = 0;
  size_t n = 0;
  size_t nz = 0;
  size_t capacity = 0;
  MM_typecode matcode = {0, 0, 0, 0};
  int* Rst = nullptr;
  int* Col = nullptr;
  ValT* Val = nullptr;
  /* device image, uploaded on the first multiply (not copied with the matrix) */
  mutable DeviceImage device;

  spMtxCRS() = default;
  spMtxCRS(size_t rows, size_t cols) : m(rows), n(cols) { Rst = new int[m + 1](); }
  spMtxCRS(size_t rows, size_t cols, size_t nnz) : m(rows), n(cols), nz(nnz), capacity(nnz) {
    Rst = new int[m + 1];
    Col = new int[nz];
    Val = new ValT[nz];
  }
  spMtxCRS(size_t rows, size_t cols, size_t nnz, const int* rst, const int* col, const ValT* val)
      : spMtxCRS(rows, cols, nnz) {
    std::memcpy(Rst, rst, (m + 1) * sizeof(int));
    std::memcpy(Col, col, nz * sizeof(int));
    if (val) std::memcpy(Val, val, nz * sizeof(ValT));
  }
  spMtxCRS(const spMtxCRS& o) : spMtxCRS(o.m, o.n, o.nz, o.Rst, o.Col, o.Val) {
    std::memcpy(matcode, o.matcode, sizeof(MM_typecode));
  }
  spMtxCRS(spMtxCRS&& o) noexcept { take(o); }
  ~spMtxCRS() { freeMem(); }

  spMtxCRS& operator=(const spMtxCRS& o) {
    if (this != &o) {
      spMtxCRS tmp(o);
      freeMem();
      take(tmp);
    }
    return *this;
  }
  spMtxCRS& operator=(spMtxCRS&& o) noexcept {
    if (this != &o) {
      freeMem();
      take(o);
    }
    return *this;
  }

  void freeMem() {
    delete[] Rst;
    delete[] Col;
    delete[] Val;
    Rst = nullptr;
    Col = nullptr;
    Val = nullptr;
    m = n = nz = capacity = 0;
  }

  /* Rows [begin, end) as a new matrix with rebased row pointers. */
  spMtxCRS extractRows(size_t begin, size_t end) const {
    const int base = Rst[begin];
    spMtxCRS r(end - begin, n, (size_t)(Rst[end] - base));
    for (size_t i = 0; i <= end - begin; ++i) r.Rst[i] = Rst[begin + i] - base;
    std::copy(Col + base, Col + Rst[end], r.Col);
    std::copy(Val + base, Val + Rst[end], r.Val);
    std::memcpy(r.matcode, matcode, sizeof(MM_typecode));
    return r;
  }

  bool operator==(const spMtxCRS& o) const {
    if (m != o.m || n != o.n || nz != o.nz) return false;
    return std::equal(Rst, Rst + m + 1, o.Rst) && std::equal(Col, Col + nz, o.Col) &&
           std::equal(Val, Val + nz, o.Val);
  }

  friend std::ostream& operator<<(std::ostream& os, const spMtxCRS& a) {
    return os << "spMtxCRS( m  = " << a.m << ", n  = " << a.n << ", nz = " << a.nz << ")";
  }

 private:
  void take(spMtxCRS& o) {
    m = o.m; n = o.n; nz = o.nz; capacity = o.capacity;
    std::memcpy(matcode, o.matcode, sizeof(MM_typecode));
    Rst = o.Rst; Col = o.Col; Val = o.Val;
    o.Rst = nullptr; o.Col = nullptr; o.Val = nullptr;
    o.m = o.n = o.nz = o.capacity = 0;
  }
};

}  // namespace SparseMatrixLib

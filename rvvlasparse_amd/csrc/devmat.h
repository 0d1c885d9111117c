// devmat.h -- device-resident matrix images and kernel launchers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "rvvla_hip.h"

namespace rvvla {

// CSR: adaptive row blocks.  blocks[b] .. blocks[b+1] is the row range of
// workgroup b: either one row longer than kLongRow entries, or up to
// kStreamRows rows holding at most kStreamCap entries in total.
constexpr int kWG = 256;
constexpr int kStreamCap = 2048;
constexpr int kStreamRows = 256;
constexpr int kLongRow = 1024;

struct DevCSR {
  const int32_t* rst = nullptr;
  const int32_t* col = nullptr;
  const void* val = nullptr;
  const int32_t* blocks = nullptr;
  int64_t nblocks = 0;
};

// HCSR: non-empty blocks only, grouped by block-row.
struct DevHCSR {
  int64_t R = 0, C = 0, b_m = 0, b_n = 0;
  const int32_t* brow_ptr = nullptr;  // b_m+1 offsets into the non-empty block list
  const int32_t* nb_col = nullptr;    // block column of each non-empty block
  const int32_t* nb_ptr = nullptr;    // first entry of each non-empty block (+ sentinel nnz)
  const uint32_t* nb_roff = nullptr;  // kind << 30 | offset into rp32 (CSR kind)
  const uint32_t* rp32 = nullptr;     // local row offsets of CSR-kind blocks
  const uint16_t* coo16 = nullptr;    // local row of each COO-kind entry (nnz-sized)
  const uint16_t* col16 = nullptr;
  const void* val = nullptr;
};

// SELL-C-sigma with C = 64: one wavefront per slice.
struct DevSELL {
  int64_t nslices = 0;
  const int32_t* cl = nullptr;
  const int32_t* cs = nullptr;
  const int32_t* bs = nullptr;
  const int32_t* sbs = nullptr;
  const uint32_t* col = nullptr;
  const uint32_t* perm = nullptr;
  const void* val = nullptr;
};

}  // namespace rvvla

struct rvvla_mat_s {
  int format = -1;
  int dtype = -1;
  int64_t m = 0, n = 0, nnz = 0;
  int device = 0;
  int64_t device_bytes = 0;
  std::vector<void*> allocs;
  rvvla::DevCSR csr;
  rvvla::DevHCSR hcsr;
  rvvla::DevSELL sell;
  // staging for rvvla_spmv_host
  void* xstage = nullptr;
  void* ystage = nullptr;

  ~rvvla_mat_s();
  // Allocate `bytes` of device memory, copy `host` into it (if non-null)
  // and keep it for the lifetime of the matrix.
  void* put(const void* host, size_t bytes);
  template <class T>
  const T* put_vec(const std::vector<T>& v) {
    return (const T*)put(v.data(), v.size() * sizeof(T));
  }
};

namespace rvvla {

// Status of a HIP call as an RVVLA code (records the error text).
int hip_check(hipError_t e, const char* what);

int launch_csr(const rvvla_mat_s& A, double alpha, const void* x, double beta, void* y,
               hipStream_t s);
int launch_csr_spmm(const rvvla_mat_s& A, int k, double alpha, const void* X, int64_t ldx,
                    double beta, void* Y, int64_t ldy, hipStream_t s);
int launch_hcsr(const rvvla_mat_s& A, double alpha, const void* x, double beta, void* y,
                hipStream_t s);
int launch_sell(const rvvla_mat_s& A, double alpha, const void* x, double beta, void* y,
                hipStream_t s);

}  // namespace rvvla

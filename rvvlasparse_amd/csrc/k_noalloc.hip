// k_noalloc.hip -- the SpMV entry point that allocates nothing.
//
// rvvla_spmv_noalloc computes what rvvla_spmv computes, with the same
// kernels, but the partial-sum and hand-over buffers those kernels use
// (CSR's sliced images, the pieces of wide SELL slices and LAV chunks, the
// CSR5 tile sums, the VHCC and VNEC partition hand-overs) live in one
// file-scope device array of kNoAllocScratchBytes instead of the matrix's
// own buffers.  Each device holds its own copy of the array until the
// process ends; the first call on a device looks up that device's address
// of it (hipGetSymbolAddress, with the device current) and keeps it.  The
// array's name is a host-side stand-in on the host, so only that address
// is ever handed to a kernel.  Every word of the array that a product reads
// was written earlier in the same product, so nothing is cleared between
// calls.  All calls on one device share the array: calls that overlap on
// one device (other streams, other threads) race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "internal.h"
#include "kcommon.h"

namespace rvvla {
namespace {

constexpr int64_t kNoAllocScratchBytes = int64_t(16) << 20;  // 16 MiB a device
constexpr int kNoAllocMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_scratch_mu;
void* g_scratch_ptr[kNoAllocMaxDevices] = {};

// The device address of `scratch` on `dev`, looked up once per device.
int scratch_on(int dev, void** out) {
  if (dev < 0 || dev >= kNoAllocMaxDevices) {
    set_last_error("rvvla_spmv_noalloc: device index beyond the supported range");
    return RVVLA_ERR_INVALID_ARG;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int cur = 0;
    if (int rc = hip_check(hipGetDevice(&cur), "hipGetDevice")) return rc;
    if (cur != dev)
      if (int rc = hip_check(hipSetDevice(dev), "hipSetDevice")) return rc;
    void* p = nullptr;
    const int rc = hip_check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)),
                             "hipGetSymbolAddress(scratch)");
    if (cur != dev) (void)hipSetDevice(cur);
    if (rc) return rc;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return RVVLA_OK;
}

}  // namespace
}  // namespace rvvla

extern "C" int rvvla_spmv_noalloc(rvvla_mat_t A, double alpha, const void* x, double beta, void* y,
                                  void* stream) {
  if (int rc = rvvla::spmv_check_args(A, x, y)) return rc;
  const int64_t need = rvvla::spmv_scratch_bytes(*A);
  if (need > rvvla::kNoAllocScratchBytes) {
    rvvla::set_last_error("rvvla_spmv_noalloc: the product needs " + std::to_string(need) +
                          " bytes of scratch, more than the " +
                          std::to_string(rvvla::kNoAllocScratchBytes) + " the library holds");
    return RVVLA_ERR_LENGTH;
  }
  void* ptr = nullptr;
  if (need > 0)
    if (int rc = rvvla::scratch_on(A->device, &ptr)) return rc;
  return rvvla::spmv_dispatch(*A, alpha, x, beta, y, (hipStream_t)stream, ptr);
}

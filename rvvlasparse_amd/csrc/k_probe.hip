// k_probe.hip -- the streaming-read rate of the device (no reference
// counterpart): the ceiling a read-mostly SpMV can approach, measured by
// the benchmark on the same device as the products.
//
// rvvla_read_probe reads `bytes` of device memory once: a grid of 16
// workgroups per CU, each streaming its own contiguous share of the buffer
// with 16-byte nontemporal loads, eight in flight per thread (measured
// faster than a grid-strided walk: 7.0 against 6.3 TB/s on 2 GiB).  The
// loaded words are folded into one value per thread that is stored only if
// it equals a constant the fold of a buffer filled by the caller cannot
// produce in practice; the store keeps the loads live.
#include <hip/hip_runtime.h>

#include "internal.h"
#include "kcommon.h"

namespace rvvla {

__global__ __launch_bounds__(256) void k_read_probe(const uint4* __restrict__ p, int64_t n16,
                                                    uint32_t key, uint32_t* __restrict__ sink) {
  typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
  constexpr int U = 8;
  const int64_t per = (n16 + gridDim.x - 1) / gridDim.x;
  const int64_t b0 = (int64_t)blockIdx.x * per, b1 = min(n16, b0 + per);
  uint32_t h = 0;
  int64_t i = b0 + threadIdx.x;
  for (; i + (U - 1) * 256 < b1; i += U * 256) {
    u32x4 a[U];
#pragma unroll
    for (int u = 0; u < U; ++u)
      a[u] = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p + i + u * 256));
#pragma unroll
    for (int u = 0; u < U; ++u) h ^= a[u].x ^ a[u].y ^ a[u].z ^ a[u].w;
  }
  for (; i < b1; i += 256) {
    const u32x4 a = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(p + i));
    h ^= a.x ^ a.y ^ a.z ^ a.w;
  }
  if (h == key) sink[blockIdx.x] = h;
}

}  // namespace rvvla

extern "C" int rvvla_read_probe(const void* buf, int64_t bytes, void* sink, void* stream) {
  using namespace rvvla;
  if (!buf || !sink || bytes < 16 || bytes % 16 != 0 || (reinterpret_cast<uintptr_t>(buf) & 15) != 0)
    return RVVLA_ERR_INVALID_ARG;
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    cus = 256;
  const int64_t n16 = bytes / 16;
  const int64_t grid = std::max<int64_t>(1, std::min<int64_t>(16 * (int64_t)cus, (n16 + 255) / 256));
  hipLaunchKernelGGL(k_read_probe, dim3((unsigned)grid), dim3(256), 0, (hipStream_t)stream,
                     (const uint4*)buf, n16, 0x9E3779B9u, (uint32_t*)sink);
  return hip_check(hipGetLastError(), "k_read_probe launch");
}

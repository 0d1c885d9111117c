/* device_image.h -- shared handle to the device-resident copy of a format.
 * Containers hold it so a const matrix can be multiplied repeatedly after a
 * single upload; copies of a container share the (read-only) device image. */
#pragma once
#include <memory>
#include "rvvla_hip.h"

namespace SparseMatrixLib {

struct DeviceImage {
  std::shared_ptr<rvvla_mat_s> handle;
  rvvla_mat_t get() const { return handle.get(); }
  explicit operator bool() const { return (bool)handle; }
  void reset(rvvla_mat_t h) {
    if (h) handle.reset(h, [](rvvla_mat_t p) { rvvla_mat_free(p); });
    else handle.reset();
  }
  /* drop the device copy (after changing the host fields) */
  void clear() { handle.reset(); }
};

}  // namespace SparseMatrixLib

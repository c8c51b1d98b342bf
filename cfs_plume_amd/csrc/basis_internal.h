// basis_internal.h -- library-internal access to a loaded basis and the
// error reporting of the C ABI (shared by basis.cpp and batch.cpp).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "cfs_device.h"

struct cfs_basis;

namespace cfs {
// Sets the calling thread's cfs_last_error() message; returns `code`.
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

int basis_device(const cfs_basis* b);
const DevBasis& basis_dev(const cfs_basis* b);
const float* basis_damp(const cfs_basis* b);
const void* basis_tables(const cfs_basis* b);
int basis_num_blocks(const cfs_basis* b);
int64_t basis_total_tokens(const cfs_basis* b);

// Makes `dev` the calling thread's current device for the guard's lifetime.
struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) == hipSuccess && prev != dev) (void)hipSetDevice(dev);
    else prev = -1;
  }
  ~DeviceGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
};
}  // namespace cfs

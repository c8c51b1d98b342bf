// noalloc.hip -- the exact reduced advection without caller workspace.
//
// cfs_reduced_advect needs two full-grid FP32 fields of scratch, which its
// caller allocates and passes in.  Callers that cannot allocate device
// memory (code under HIP graph capture, serving loops, harnesses that forbid
// hipMalloc) call cfs_reduced_advect_noalloc instead: the same launches, with
// both fields carved from one file-scope device array of fixed size that
// every GPU the library runs on holds until the process ends.
//
// The array's name is a host-side stand-in in host code; its device address
// comes from hipGetSymbolAddress, looked up once per device (with that device
// current) and cached.  Calls on one GPU share its array: the first field is
// cleared by every call (unproject writes interior cells only), the second is
// written whole by the advection, and calls that overlap on one GPU race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/cfs_plume.h"
#include "basis_internal.h"

namespace cfs {
hipError_t launch_zero_words(void* p, long nwords, hipStream_t s);
hipError_t launch_axpb(double* q, int r, double alpha, const double* forcing, hipStream_t s);
}  // namespace cfs

namespace {

// Two fields of at most CFS_NOALLOC_FIELD_FLOATS floats each (32 MiB in all):
// full grids of up to 1,398,101 cells, e.g. 96 x 128 x 96.
constexpr long NOALLOC_FIELD_FLOATS = CFS_NOALLOC_FIELD_FLOATS;
static __device__ __align__(256) float scratch[2 * NOALLOC_FIELD_FLOATS];

constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
float* g_scratch_ptr[MAX_DEVICES] = {};

// The device address of `scratch` on device `dev` (current), looked up once.
int scratch_for_device(int dev, float** out) {
  if (dev < 0 || dev >= MAX_DEVICES) return cfs::fail(CFS_ERR_ARG, "device ordinal out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return cfs::hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!p) return cfs::fail(CFS_ERR_HIP, "hipGetSymbolAddress(scratch) returned null");
    g_scratch_ptr[dev] = static_cast<float*>(p);
  }
  *out = g_scratch_ptr[dev];
  return CFS_OK;
}

}  // namespace

extern "C" {

int cfs_reduced_advect_noalloc(cfs_basis* U, const double* q_in, double* q_out, float dt0, double alpha,
                               const double* forcing, void* stream) {
  if (!U || !q_in || !q_out) return cfs::fail(CFS_ERR_ARG, "null argument");
  cfs_basis_info info;
  int st = cfs_basis_get_info(U, &info);
  if (st) return st;
  const long X = info.dims[0] + 2, Y = info.dims[1] + 2, Z = info.dims[2] + 2;
  const long n = 3 * X * Y * Z;
  if (n > NOALLOC_FIELD_FLOATS)
    return cfs::fail(CFS_ERR_ARG, "full grid of " + std::to_string(3 * X * Y * Z / 3) +
                                      " cells exceeds the no-allocation scratch (" +
                                      std::to_string(NOALLOC_FIELD_FLOATS / 3) + " cells per field)");
  cfs::DeviceGuard g(cfs::basis_device(U));
  float* base = nullptr;
  if ((st = scratch_for_device(cfs::basis_device(U), &base))) return st;
  float* field_a = base;
  float* field_b = base + NOALLOC_FIELD_FLOATS;
  const hipStream_t s = (hipStream_t)stream;
  hipError_t e = cfs::launch_zero_words(field_a, n, s);
  if (e != hipSuccess) return cfs::hip_fail(e, "zero launch");
  if ((st = cfs_unproject(U, q_in, field_a, stream))) return st;
  if ((st = cfs_advect_field(field_a, field_b, (int)X, (int)Y, (int)Z, dt0, stream))) return st;
  if ((st = cfs_project(U, field_b, q_out, stream))) return st;
  e = cfs::launch_axpb(q_out, info.num_cols, alpha, forcing, s);
  if (e != hipSuccess) return cfs::hip_fail(e, "axpb launch");
  return CFS_OK;
}

}  // extern "C"

// cubature.hip -- cubature reduced advection over a compressed basis.
//
// reducedAdvectCompressionFriendly (reference SUBSPACE_FLUID_3D_COMPRESSED_EIGEN.cpp:
// 1781-1854) advects in the subspace at K key cells x_i with weights w_i
// (readAdvectionCubature :1188-1327 reads them from the `cubature` file and
// builds After_i = w_i U_after(x_i)^T, an r x 3 matrix per cell):
//
//   v_i   = U_pre(x_i) q                              velocity at the key cell
//   p_i   = x_i - dt0 v_i, clamped to [1.5, res - 2.5] per axis (full-grid units)
//   a_i   = sum over the 8 corners c of p_i's cell of w_c U_pre(c) q   (trilinear)
//   q'    = sum_i After_i a_i                          (accumAdvectRequests :1859-1993)
//
// The reference decodes the 3 x r rows U_pre(c) of every corner from the
// compressed streams (GetSubmatrixNoSVDSparse, one block cached) and
// multiplies each by q.  Every U_pre(c) q is a value of the field U_pre q,
// so here the step first unprojects q once (the fused step's decoder, every
// block decoded once) and the key-cell kernel interpolates that field:
//
//   launch 1   cfs_unproject(U_pre, q) -> field (FP32, interior cells)
//   launch 2   workgroup of 16 key cells x 64 outputs: the backtrace and
//              trilinear a_i in FP64 from the field, then lane j sums
//              After_i[j] . a_i over the workgroup's cells (FP64) -> partials
//   launch 3   q'_j = alpha * sum of the partials (workgroup order, FP64) + f_j
//
// After_i (FP64, [K][r][3]) is built once from U_after on the device with the
// point decoder (GetSubmatrixNoSVDSparse) and scaled by w_i.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/cfs_plume.h"
#include "basis_internal.h"

namespace {

// A workgroup of one wave takes CUB_CELLS key cells and 64 of the r outputs:
// lanes 0..CUB_CELLS-1 compute the cells' interpolated velocities (FP64) into
// LDS, then lane j sums After_i[j] . a_i over the cells with every load of
// the (unrolled) loop in flight.  Many small workgroups keep enough loads in
// flight: the after-matrices are small (K x r x 24 bytes) and the kernel is
// bound by load latency, not bandwidth.
constexpr int CUB_CELLS = 16;
constexpr int CUB_TPB = 64;

__global__ void __launch_bounds__(CUB_TPB)
cubature_kernel(const float* __restrict__ field, int X, int Y, int Z, const int32_t* __restrict__ cells, int K,
                const double* __restrict__ after, int r, double dt0, double* __restrict__ partials) {
  __shared__ double adv[CUB_CELLS][3];
  const int t = threadIdx.x;
  const int i0 = blockIdx.x * CUB_CELLS;
  if (t < CUB_CELLS) {
    const int i = i0 + t;
    double a[3] = {0.0, 0.0, 0.0};
    if (i < K) {
      // peeled cell -> full-grid coordinates
      const int slab = X * Y;
      const int idx = cells[i];
      const int z = idx / slab + 1, y = (idx % slab) / X + 1, x = (idx % slab) % X + 1;
      const int GX = X + 2, GY = Y + 2;
      const size_t gslab = (size_t)GX * GY;
      const float* f = field + 3 * ((size_t)x + (size_t)GX * ((size_t)y + (size_t)GY * z));
      double xt = x - dt0 * (double)f[0];
      double yt = y - dt0 * (double)f[1];
      double zt = z - dt0 * (double)f[2];
      // clamp to [1.5, res - 2.5], res the full-grid resolution (peeled + 2)
      xt = xt < 1.5 ? 1.5 : xt;
      xt = xt > GX - 2.5 ? GX - 2.5 : xt;
      yt = yt < 1.5 ? 1.5 : yt;
      yt = yt > GY - 2.5 ? GY - 2.5 : yt;
      zt = zt < 1.5 ? 1.5 : zt;
      zt = zt > (Z + 2) - 2.5 ? (Z + 2) - 2.5 : zt;
      // peeled corner x0 = (int)xt - 1 is full-grid cell (int)xt
      const int cx = (int)xt, cy = (int)yt, cz = (int)zt;
      const double s1 = xt - cx, s0 = 1.0 - s1;
      const double t1 = yt - cy, t0 = 1.0 - t1;
      const double u1 = zt - cz, u0 = 1.0 - u1;
      const float* c000 = field + 3 * ((size_t)cx + (size_t)GX * ((size_t)cy + (size_t)GY * cz));
      const float* c010 = c000 + 3 * GX;
      const float* c100 = c000 + 3;
      const float* c110 = c010 + 3;
      const float* c001 = c000 + 3 * gslab;
      const float* c011 = c010 + 3 * gslab;
      const float* c101 = c100 + 3 * gslab;
      const float* c111 = c110 + 3 * gslab;
      const double w000 = u0 * s0 * t0, w010 = u0 * s0 * t1, w100 = u0 * s1 * t0, w110 = u0 * s1 * t1;
      const double w001 = u1 * s0 * t0, w011 = u1 * s0 * t1, w101 = u1 * s1 * t0, w111 = u1 * s1 * t1;
#pragma unroll
      for (int c = 0; c < 3; c++)
        a[c] = w000 * c000[c] + w010 * c010[c] + w100 * c100[c] + w110 * c110[c] + w001 * c001[c] +
               w011 * c011[c] + w101 * c101[c] + w111 * c111[c];
    }
    adv[t][0] = a[0];
    adv[t][1] = a[1];
    adv[t][2] = a[2];
  }
  __syncthreads();
  const int j = blockIdx.y * CUB_TPB + t;
  if (j >= r) return;
  const int n = min(CUB_CELLS, K - i0);
  const double* A = after + ((size_t)i0 * r + j) * 3;
  double s = 0.0;
  if (n == CUB_CELLS) {
    double m[CUB_CELLS][3];
#pragma unroll
    for (int k = 0; k < CUB_CELLS; k++)
#pragma unroll
      for (int c = 0; c < 3; c++) m[k][c] = A[(size_t)k * r * 3 + c];
#pragma unroll
    for (int k = 0; k < CUB_CELLS; k++) s += m[k][0] * adv[k][0] + m[k][1] * adv[k][1] + m[k][2] * adv[k][2];
  } else {
    for (int k = 0; k < n; k++) {
      const double* mk = A + (size_t)k * r * 3;
      s += mk[0] * adv[k][0] + mk[1] * adv[k][1] + mk[2] * adv[k][2];
    }
  }
  partials[(size_t)blockIdx.x * r + j] = s;
}

// q_out[j] = alpha * sum_w partials[w][j] + forcing[j] (forcing may be null)
__global__ void __launch_bounds__(256) cubature_reduce_kernel(const double* __restrict__ partials, int nwg, int r,
                                                              double alpha, const double* __restrict__ forcing,
                                                              double* __restrict__ q_out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= r) return;
  // four running sums so that four loads are in flight, combined in a fixed order
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int w = 0;
  for (; w + 4 <= nwg; w += 4) {
    s0 += partials[(size_t)w * r + j];
    s1 += partials[(size_t)(w + 1) * r + j];
    s2 += partials[(size_t)(w + 2) * r + j];
    s3 += partials[(size_t)(w + 3) * r + j];
  }
  for (; w < nwg; w++) s0 += partials[(size_t)w * r + j];
  q_out[j] = alpha * ((s0 + s1) + (s2 + s3)) + (forcing ? forcing[j] : 0.0);
}

// after[i][j][c] *= weights[i]
__global__ void __launch_bounds__(256) scale_rows_kernel(double* __restrict__ after, const double* __restrict__ w,
                                                         int K, int per) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)K * per) return;
  after[e] *= w[e / per];
}

}  // namespace

struct cfs_cubature {
  int device = 0;
  int K = 0, r = 0, X = 0, Y = 0, Z = 0;
  int nwg = 0;
  std::vector<int32_t> cells;
  std::vector<double> weights;
  int32_t* d_cells = nullptr;
  double* d_after = nullptr;
  double* d_partials = nullptr;
  ~cfs_cubature() {
    (void)hipFree(d_cells);
    (void)hipFree(d_after);
    (void)hipFree(d_partials);
  }
};

namespace {

#define CUB_HIP(call, what)                         \
  do {                                              \
    const hipError_t e_ = (call);                   \
    if (e_ != hipSuccess) return cfs::hip_fail(e_, what); \
  } while (0)

}  // namespace

extern "C" {

}  // extern "C"

namespace {

// A cubature over the peeled grid dims[3] with `r` outputs; `fill` writes the
// unscaled after-rows ([K][r][3], entry (c, j) of cell i at i 3r + c + 3j)
// into the device array it is given; they are then scaled by the weights.
template <typename Fill>
int make_cubature(int K, const int32_t* cells, const double* weights, const int dims[3], int r, int device,
                  Fill fill, cfs_cubature** out) {
  if (!out) return cfs::fail(CFS_ERR_ARG, "null output handle");
  *out = nullptr;
  if (!cells || !weights) return cfs::fail(CFS_ERR_ARG, "null argument");
  if (K < 1) return cfs::fail(CFS_ERR_ARG, "a cubature needs at least one key cell");
  if (r < 1) return cfs::fail(CFS_ERR_ARG, "the after-basis needs at least one column");
  const long ncell = (long)dims[0] * dims[1] * dims[2];
  for (int i = 0; i < K; i++) {
    if (cells[i] < 0 || cells[i] >= ncell)
      return cfs::fail(CFS_ERR_ARG, "key cell " + std::to_string(i) + " (" + std::to_string(cells[i]) +
                                        ") outside the peeled grid");
    if (!std::isfinite(weights[i])) return cfs::fail(CFS_ERR_ARG, "key cell weight is not finite");
  }
  std::unique_ptr<cfs_cubature> C(new cfs_cubature);
  C->device = device;
  cfs::DeviceGuard g(C->device);
  C->K = K;
  C->r = r;
  C->X = dims[0];
  C->Y = dims[1];
  C->Z = dims[2];
  C->nwg = (K + CUB_CELLS - 1) / CUB_CELLS;
  C->cells.assign(cells, cells + K);
  C->weights.assign(weights, weights + K);
  const size_t per = (size_t)3 * C->r;
  CUB_HIP(hipMalloc(&C->d_cells, (size_t)K * 4), "hipMalloc");
  CUB_HIP(hipMalloc(&C->d_after, (size_t)K * per * 8), "hipMalloc");
  CUB_HIP(hipMalloc(&C->d_partials, (size_t)C->nwg * C->r * 8), "hipMalloc");
  CUB_HIP(hipMemcpy(C->d_cells, cells, (size_t)K * 4, hipMemcpyHostToDevice), "hipMemcpy");
  int st = fill(C->d_after);
  if (st) return st;
  double* d_w = nullptr;
  CUB_HIP(hipMalloc(&d_w, (size_t)K * 8), "hipMalloc");
  hipError_t e = hipMemcpy(d_w, weights, (size_t)K * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    const long n = (long)K * (long)per;
    hipLaunchKernelGGL(scale_rows_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, nullptr, C->d_after,
                       d_w, K, (int)per);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  (void)hipFree(d_w);
  if (e != hipSuccess) return cfs::hip_fail(e, "building the after-matrices");
  *out = C.release();
  return CFS_OK;
}

}  // namespace

extern "C" {

int cfs_cubature_create(int K, const int32_t* cells, const double* weights, cfs_basis* after, cfs_cubature** out) {
  if (!after) return cfs::fail(CFS_ERR_ARG, "null argument");
  cfs_basis_info info;
  int st = cfs_basis_get_info(after, &info);
  if (st) return st;
  // After_i = w_i U_after(x_i)^T: the 3 x r rows of each key cell
  // (GetSubmatrixNoSVDSparse), column-major, entry (c, j) at c + 3 j
  auto fill = [&](double* d_after) {
    const size_t per = (size_t)3 * info.num_cols;
    for (int i = 0; i < K; i++) {
      const int s2 = cfs_get_submatrix(after, 3 * cells[i], d_after + (size_t)i * per, nullptr);
      if (s2) return s2;
    }
    return (int)CFS_OK;
  };
  return make_cubature(K, cells, weights, info.dims, info.num_cols, cfs::basis_device(after), fill, out);
}

int cfs_cubature_create_rows(int K, const int32_t* cells, const double* weights, int X, int Y, int Z, int r,
                             const double* rows, cfs_cubature** out) {
  if (!rows) return cfs::fail(CFS_ERR_ARG, "null rows");
  if (X < 1 || Y < 1 || Z < 1) return cfs::fail(CFS_ERR_ARG, "bad grid");
  int device = 0;
  CUB_HIP(hipGetDevice(&device), "hipGetDevice");
  const int dims[3] = {X, Y, Z};
  auto fill = [&](double* d_after) {
    const hipError_t e = hipMemcpy(d_after, rows, (size_t)K * 3 * r * 8, hipMemcpyHostToDevice);
    return e == hipSuccess ? (int)CFS_OK : cfs::hip_fail(e, "hipMemcpy");
  };
  return make_cubature(K, cells, weights, dims, r, device, fill, out);
}

int cfs_cubature_open_file(const char* path, cfs_basis* after, cfs_cubature** out) {
  if (!out) return cfs::fail(CFS_ERR_ARG, "null output handle");
  *out = nullptr;
  if (!path) return cfs::fail(CFS_ERR_ARG, "null path");
  FILE* f = std::fopen(path, "rb");
  if (!f) return cfs::fail(CFS_ERR_IO, std::string("cannot open ") + path);
  std::fseek(f, 0, SEEK_END);
  const long n = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  int32_t K = 0;
  if (n < 4 || std::fread(&K, 4, 1, f) != 1) {
    std::fclose(f);
    return cfs::fail(CFS_ERR_FORMAT, std::string(path) + ": truncated cubature header");
  }
  // int K; int cells[K]; double weights[K] (CUBATURE_GENERATOR_EIGEN::writeCubature)
  if (K < 1 || (long)4 + 12L * K != n) {
    std::fclose(f);
    return cfs::fail(CFS_ERR_FORMAT, std::string(path) + ": size does not match " + std::to_string(K) +
                                         " key cells");
  }
  std::vector<int32_t> cells(K);
  std::vector<double> w(K);
  const bool ok = std::fread(cells.data(), 4, K, f) == (size_t)K && std::fread(w.data(), 8, K, f) == (size_t)K;
  std::fclose(f);
  if (!ok) return cfs::fail(CFS_ERR_IO, std::string("short read on ") + path);
  return cfs_cubature_create(K, cells.data(), w.data(), after, out);
}

int cfs_cubature_write_file(const char* path, int K, const int32_t* cells, const double* weights) {
  if (!path || !cells || !weights || K < 1) return cfs::fail(CFS_ERR_ARG, "bad argument");
  FILE* f = std::fopen(path, "wb");
  if (!f) return cfs::fail(CFS_ERR_IO, std::string("cannot open ") + path);
  const int32_t k = K;
  bool ok = std::fwrite(&k, 4, 1, f) == 1;
  ok = ok && std::fwrite(cells, 4, K, f) == (size_t)K;
  ok = ok && std::fwrite(weights, 8, K, f) == (size_t)K;
  ok = (std::fclose(f) == 0) && ok;
  return ok ? CFS_OK : cfs::fail(CFS_ERR_IO, std::string("short write on ") + path);
}

int cfs_cubature_get_info(const cfs_cubature* c, int* K, int* r) {
  if (!c) return cfs::fail(CFS_ERR_ARG, "null cubature");
  if (K) *K = c->K;
  if (r) *r = c->r;
  return CFS_OK;
}

int cfs_cubature_get_cells(const cfs_cubature* c, int32_t* cells, double* weights) {
  if (!c) return cfs::fail(CFS_ERR_ARG, "null cubature");
  if (cells) std::memcpy(cells, c->cells.data(), (size_t)c->K * 4);
  if (weights) std::memcpy(weights, c->weights.data(), (size_t)c->K * 8);
  return CFS_OK;
}

int cfs_cubature_advect_field(cfs_cubature* c, const float* field, double* q_out, float dt0, double alpha,
                              const double* forcing, void* stream) {
  if (!c || !field || !q_out) return cfs::fail(CFS_ERR_ARG, "null argument");
  cfs::DeviceGuard g(c->device);
  const hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(cubature_kernel, dim3(c->nwg, (c->r + CUB_TPB - 1) / CUB_TPB), dim3(CUB_TPB), 0, s, field, c->X,
                     c->Y, c->Z, c->d_cells, c->K, c->d_after, c->r, (double)dt0, c->d_partials);
  CUB_HIP(hipGetLastError(), "cubature launch");
  hipLaunchKernelGGL(cubature_reduce_kernel, dim3((c->r + 255) / 256), dim3(256), 0, s, c->d_partials, c->nwg, c->r,
                     alpha, forcing, q_out);
  CUB_HIP(hipGetLastError(), "cubature reduce launch");
  return CFS_OK;
}

int cfs_cubature_advect(cfs_cubature* c, cfs_basis* pre, const double* q_in, double* q_out, float* field, float dt0,
                        double alpha, const double* forcing, void* stream) {
  if (!c || !pre || !q_in || !q_out || !field) return cfs::fail(CFS_ERR_ARG, "null argument");
  cfs_basis_info info;
  int st = cfs_basis_get_info(pre, &info);
  if (st) return st;
  if (info.dims[0] != c->X || info.dims[1] != c->Y || info.dims[2] != c->Z)
    return cfs::fail(CFS_ERR_ARG, "the basis grid differs from the cubature's");
  if ((st = cfs_unproject(pre, q_in, field, stream))) return st;
  return cfs_cubature_advect_field(c, field, q_out, dt0, alpha, forcing, stream);
}

int cfs_cubature_destroy(cfs_cubature* c) {
  if (!c) return cfs::fail(CFS_ERR_ARG, "null cubature");
  cfs::DeviceGuard g(c->device);
  delete c;
  return CFS_OK;
}

}  // extern "C"

// cfs_plume.hpp -- C++ host layer over the C ABI (cfs_plume.h) with the
// reference's call surface for the compressed-subspace step
// (aadjones/CFS_PLUME, src/compression/COMPRESSION.h:256-325).
//
// The reference passes Eigen vectors and VECTOR3_FIELD_3D objects; this layer
// uses plain std::vector containers with the same layouts:
//   a vector field is 3 * xRes * yRes * zRes doubles, cell (x, y, z) at
//   3 * (x + xRes * (y + yRes * z)), components interleaved
//   (VECTOR3_FIELD_3D.cpp:69-85); the step's fields are full grids
//   (X+2)(Y+2)(Z+2) and DecodeVectorField's output is the peeled X*Y*Z grid;
//   q is an r-vector of doubles; a submatrix is 3 x r, column-major.
// Fields travel host <-> device as FP32 per call (the GPU path computes in
// FP32 with FP64 accumulation of q).  Where the reference asserts, prints and
// returns NULL or calls exit(), these functions throw cfs_plume::Error
// carrying the C ABI status and message.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <stdexcept>
#include <string>
#include <vector>

#include "cfs_plume.h"

namespace cfs_plume {

class Error : public std::runtime_error {
 public:
  Error(int status, const std::string& what) : std::runtime_error(what), status_(status) {}
  int status() const { return status_; }

 private:
  int status_;
};

inline void check(int status, const char* what) {
  if (status != CFS_OK) throw Error(status, std::string(what) + ": " + cfs_last_error());
}
inline void check_hip(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error(CFS_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Device scratch owned by the basis wrapper (one FP32 full-grid field, q).
template <class T>
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  explicit DeviceBuffer(size_t n) : n_(n) { check_hip(hipMalloc(&p_, n * sizeof(T) + 4), "hipMalloc"); }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; }
  DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
    if (this != &o) {
      if (p_) (void)hipFree(p_);
      p_ = o.p_;
      n_ = o.n_;
      o.p_ = nullptr;
    }
    return *this;
  }
  ~DeviceBuffer() {
    if (p_) (void)hipFree(p_);
  }
  T* get() const { return reinterpret_cast<T*>(p_); }
  size_t size() const { return n_; }
  void upload(const T* h) { check_hip(hipMemcpy(p_, h, n_ * sizeof(T), hipMemcpyHostToDevice), "hipMemcpy H2D"); }
  void download(T* h) const { check_hip(hipMemcpy(h, p_, n_ * sizeof(T), hipMemcpyDeviceToHost), "hipMemcpy D2H"); }

 private:
  void* p_ = nullptr;
  size_t n_ = 0;
};

// MATRIX_COMPRESSION_DATA (MATRIX_COMPRESSION_DATA.h:12-141) resident on the
// GPU: the three component files U.<name>.component0..2.
class MatrixCompressionData {
 public:
  // ReadBinaryFileToMemory x3 + constructor + set_dampingArrayLists + dct_setup
  // + init_cache (COMPRESSION.cpp:1456-1594; MATRIX_COMPRESSION_DATA.h:59-122).
  MatrixCompressionData(const std::string& file0, const std::string& file1, const std::string& file2) {
    check(cfs_basis_open_files(file0.c_str(), file1.c_str(), file2.c_str(), &basis_), "ReadBinaryFileToMemory");
    check(cfs_basis_get_info(basis_, &info_), "cfs_basis_get_info");
    const size_t cells = (size_t)(info_.dims[0] + 2) * (info_.dims[1] + 2) * (info_.dims[2] + 2);
    field_ = DeviceBuffer<float>(3 * cells);
    q_ = DeviceBuffer<double>((size_t)info_.num_cols);
  }
  MatrixCompressionData(const MatrixCompressionData&) = delete;
  MatrixCompressionData& operator=(const MatrixCompressionData&) = delete;
  ~MatrixCompressionData() {
    if (basis_) cfs_basis_close(basis_);
  }

  cfs_basis* handle() const { return basis_; }
  int get_numCols() const { return info_.num_cols; }
  int get_numBlocks() const { return info_.num_blocks; }
  int get_nBits() const { return info_.n_bits; }
  // peeled dims (X, Y, Z) and padded dims
  int xRes() const { return info_.dims[0]; }
  int yRes() const { return info_.dims[1]; }
  int zRes() const { return info_.dims[2]; }
  size_t full_cells() const { return (size_t)(xRes() + 2) * (yRes() + 2) * (zRes() + 2); }
  size_t peeled_cells() const { return (size_t)xRes() * yRes() * zRes(); }

  DeviceBuffer<float>& field_scratch() { return field_; }
  DeviceBuffer<double>& q_scratch() { return q_; }

 private:
  cfs_basis* basis_ = nullptr;
  cfs_basis_info info_{};
  DeviceBuffer<float> field_;
  DeviceBuffer<double> q_;
};

namespace detail {
inline std::vector<float> to_float(const std::vector<double>& v) { return std::vector<float>(v.begin(), v.end()); }
}  // namespace detail

// PeeledCompressedProjectTransformNoSVD (COMPRESSION.cpp:2978-3016):
// q = U^T V over the peeled cells; V is the full grid; q is resized.
inline void PeeledCompressedProjectTransformNoSVD(const std::vector<double>& V, MatrixCompressionData* U,
                                                  std::vector<double>* q) {
  if (!U || !q) throw Error(CFS_ERR_ARG, "null argument");
  if (V.size() != 3 * U->full_cells()) throw Error(CFS_ERR_ARG, "V must be the full grid");
  std::vector<float> f = detail::to_float(V);
  U->field_scratch().upload(f.data());
  check(cfs_project(U->handle(), U->field_scratch().get(), U->q_scratch().get(), nullptr),
        "PeeledCompressedProjectTransformNoSVD");
  q->resize((size_t)U->get_numCols());
  U->q_scratch().download(q->data());
}

// PeeledCompressedUnprojectTransform (COMPRESSION.cpp:2739-2900): interior
// cells of the full-grid V = U q; boundary cells keep their values.
inline void PeeledCompressedUnprojectTransform(MatrixCompressionData* U, const std::vector<double>& q,
                                               std::vector<double>* V) {
  if (!U || !V) throw Error(CFS_ERR_ARG, "null argument");
  if (q.size() != (size_t)U->get_numCols()) throw Error(CFS_ERR_ARG, "q must have numCols entries");
  if (V->size() != 3 * U->full_cells()) throw Error(CFS_ERR_ARG, "V must be the full grid");
  std::vector<float> f = detail::to_float(*V);
  U->field_scratch().upload(f.data());
  U->q_scratch().upload(q.data());
  check(cfs_unproject(U->handle(), U->q_scratch().get(), U->field_scratch().get(), nullptr),
        "PeeledCompressedUnprojectTransform");
  U->field_scratch().download(f.data());
  V->assign(f.begin(), f.end());
}

// DecodeVectorField (COMPRESSION.cpp:2640-2670): column `col` of U as a
// peeled X*Y*Z vector field.
inline void DecodeVectorField(MatrixCompressionData* U, int col, std::vector<double>* out) {
  if (!U || !out) throw Error(CFS_ERR_ARG, "null argument");
  DeviceBuffer<float> d(3 * U->peeled_cells());
  check(cfs_decode_columns(U->handle(), col, 1, d.get(), nullptr), "DecodeVectorField");
  std::vector<float> f(d.size());
  d.download(f.data());
  out->assign(f.begin(), f.end());
}

// GetSubmatrixNoSVDSparse (COMPRESSION.cpp:1949-2158): rows
// [startRow, startRow+3) of U as a 3 x r column-major matrix.
inline void GetSubmatrixNoSVDSparse(int startRow, MatrixCompressionData* U, std::vector<double>* submatrix) {
  if (!U || !submatrix) throw Error(CFS_ERR_ARG, "null argument");
  DeviceBuffer<double> d(3 * (size_t)U->get_numCols());
  check(cfs_get_submatrix(U->handle(), startRow, d.get(), nullptr), "GetSubmatrixNoSVDSparse");
  submatrix->resize(d.size());
  d.download(submatrix->data());
}

// CompressAndWriteMatrixComponents (COMPRESSION.cpp:2271-2359), SVD off:
// U column-major, column c = 3*X*Y*Z doubles; writes <prefix>0..2.
inline void CompressAndWriteMatrixComponents(const std::string& prefix, const std::vector<double>& U, int X, int Y,
                                             int Z, int numCols, int nBits = 32, double percent = 0.9995,
                                             int maxIterations = 50) {
  if (U.size() != (size_t)3 * X * Y * Z * numCols) throw Error(CFS_ERR_ARG, "U must be 3*X*Y*Z x numCols");
  cfs_encoded* enc = nullptr;
  check(cfs_encode_matrix(U.data(), X, Y, Z, numCols, nBits, percent, maxIterations, 0, &enc),
        "CompressAndWriteMatrixComponents");
  const int st = cfs_encoded_write(enc, prefix.c_str());
  cfs_encoded_free(enc);
  check(st, "write component files");
}

}  // namespace cfs_plume

// linalg.hip -- the dense products and sparse operators of the basis builder.
//
// The reduced bases come from an SVD of the snapshot matrices
// (BIG_MATRIX::outOfCoreSVD, reference src/linearalgebra/BIG_MATRIX.cpp:
// 276-420) and the reduced operators from projecting sparse full-grid
// operators onto them (SUBSPACE_FLUID_3D_EIGEN::buildOutOfCoreMatrices,
// src/integrators/SUBSPACE_FLUID_3D_EIGEN.cpp:1355-1489).  Both reduce to
// two tall-skinny FP64 products on the device, the snapshot/basis matrices
// staying in HBM (column-major, one column per snapshot or mode):
//
//   gemm_tn   C = A^T B (m x n, rows -> split over workgroups; partial
//             products reduced in a fixed order, so results are repeatable)
//   gemm_nn   C = A M   (M small, k x n, staged in LDS per 16-column tile)
//
// and the four sparse operators of the reduced pressure, divergence,
// gradient and damping stages applied to many columns at once, on the
// peeled grid (X-2)(Y-2)(Z-2):
//
//   POISSON         buildFlatA (:2166-2207): the pressure Poisson matrix with
//                   the wall obstacles
//   VEL_TO_DIV      computeVelocityToDivergence (:1989-2100)
//   PRESSURE_TO_VEL computePressureToVelocity (:627-695)
//   DAMPING         buildPeeledDampingMatrixFlat (:2107-2160)
#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "../../include/cfs_plume.h"
#include "basis_internal.h"

namespace {

constexpr int TT = 16;          // output tile edge of gemm_tn
constexpr int KROWS = 32;       // rows staged per LDS round
constexpr int CHUNK = 8192;     // rows per split of gemm_tn
constexpr int NN_TPB = 256;
constexpr int NN_COLS = 16;     // output columns per gemm_nn workgroup
constexpr int NN_MAXK = 512;

#define LA_HIP(call, what)                                  \
  do {                                                      \
    const hipError_t e_ = (call);                           \
    if (e_ != hipSuccess) return cfs::hip_fail(e_, what);   \
  } while (0)

// partial[chunk][i][j] = sum over the chunk's rows of A[row, i] B[row, j]
__global__ void __launch_bounds__(TT* TT) gemm_tn_kernel(const double* __restrict__ A, long rows, int m, long lda,
                                                         const double* __restrict__ B, int n, long ldb,
                                                         double* __restrict__ partial) {
  __shared__ double As[KROWS][TT + 1], Bs[KROWS][TT + 1];
  const int ti = threadIdx.x % TT, tj = threadIdx.x / TT;
  const int i0 = blockIdx.x * TT, j0 = blockIdx.y * TT;
  const long r0 = (long)blockIdx.z * CHUNK, r1 = min(rows, r0 + CHUNK);
  double acc = 0.0;
  for (long k0 = r0; k0 < r1; k0 += KROWS) {
    // 2 x 32 x 16 loads by 256 threads: row-contiguous (coalesced) per column
    for (int e = threadIdx.x; e < KROWS * TT; e += TT * TT) {
      const int rr = e % KROWS, cc = e / KROWS;
      const long row = k0 + rr;
      As[rr][cc] = (row < r1 && i0 + cc < m) ? A[row + (long)(i0 + cc) * lda] : 0.0;
      Bs[rr][cc] = (row < r1 && j0 + cc < n) ? B[row + (long)(j0 + cc) * ldb] : 0.0;
    }
    __syncthreads();
#pragma unroll 8
    for (int rr = 0; rr < KROWS; rr++) acc += As[rr][ti] * Bs[rr][tj];
    __syncthreads();
  }
  const int i = i0 + ti, j = j0 + tj;
  if (i < m && j < n) partial[((size_t)blockIdx.z * m + i) * n + j] = acc;
}

__global__ void __launch_bounds__(256) gemm_tn_reduce_kernel(const double* __restrict__ partial, int chunks, int m,
                                                             int n, double* __restrict__ C) {
  const int e = blockIdx.x * 256 + threadIdx.x;
  if (e >= m * n) return;
  double s = 0.0;
  for (int c = 0; c < chunks; c++) s += partial[(size_t)c * m * n + e];
  C[e] = s;
}

// C[row, j] = sum_p A[row, p] M[p, j] for the workgroup's 16 columns j
__global__ void __launch_bounds__(NN_TPB) gemm_nn_kernel(const double* __restrict__ A, long rows, int k, long lda,
                                                         const double* __restrict__ M, int n, double* __restrict__ C,
                                                         long ldc) {
  __shared__ double Ms[NN_MAXK][NN_COLS];
  const int j0 = blockIdx.y * NN_COLS;
  for (int e = threadIdx.x; e < k * NN_COLS; e += NN_TPB) {
    const int p = e / NN_COLS, jj = e % NN_COLS;
    Ms[p][jj] = j0 + jj < n ? M[(size_t)p * n + j0 + jj] : 0.0;
  }
  __syncthreads();
  const long row = (long)blockIdx.x * NN_TPB + threadIdx.x;
  if (row >= rows) return;
  double acc[NN_COLS];
#pragma unroll
  for (int jj = 0; jj < NN_COLS; jj++) acc[jj] = 0.0;
  for (int p = 0; p < k; p++) {
    const double a = A[row + (long)p * lda];
#pragma unroll
    for (int jj = 0; jj < NN_COLS; jj++) acc[jj] += a * Ms[p][jj];
  }
#pragma unroll
  for (int jj = 0; jj < NN_COLS; jj++)
    if (j0 + jj < n) C[row + (long)(j0 + jj) * ldc] = acc[jj];
}

struct Grid {
  int X, Y, Z;     // full grid
  int Xp, Yp, Zp;  // peeled
  int bc[6];       // front, back, left, right, top, bottom (1 wall)
  double dx, w;
};

__device__ __forceinline__ bool is_wall(const Grid& g, int x, int y, int z) {
  // the obstacle slabs of the FLUID_3D_MIC constructor, full-grid coordinates
  return (z == 0 && g.bc[0] == 1) || (z == g.Z - 1 && g.bc[1] == 1) || (x == 0 && g.bc[2] == 1) ||
         (x == g.X - 1 && g.bc[3] == 1) || (y == g.Y - 1 && g.bc[4] == 1) || (y == 0 && g.bc[5] == 1);
}

// out column c = Op (in column c), one thread per (peeled cell, column)
__global__ void __launch_bounds__(256) operator_kernel(int op, Grid g, const double* __restrict__ in, long ld_in,
                                                       double* __restrict__ out, long ld_out, int ncols) {
  const long np = (long)g.Xp * g.Yp * g.Zp;
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  if (t >= np * ncols) return;
  const int col = (int)(t / np);
  const int j = (int)(t % np);
  const int px = j % g.Xp, py = (j / g.Xp) % g.Yp, pz = j / (g.Xp * g.Yp);
  const int slab = g.Xp * g.Yp;
  const double* v = in + (long)col * ld_in;
  double* o = out + (long)col * ld_out;
  if (op == CFS_OP_POISSON) {
    // full-grid cell (px+1, py+1, pz+1); interior cells are never walls
    const int x = px + 1, y = py + 1, z = pz + 1;
    double diag = 0.0;
    diag += is_wall(g, x + 1, y, z) ? 0.0 : 1.0;
    diag += is_wall(g, x - 1, y, z) ? 0.0 : 1.0;
    diag += is_wall(g, x, y + 1, z) ? 0.0 : 1.0;
    diag += is_wall(g, x, y - 1, z) ? 0.0 : 1.0;
    diag += is_wall(g, x, y, z + 1) ? 0.0 : 1.0;
    diag += is_wall(g, x, y, z - 1) ? 0.0 : 1.0;
    double s = diag * v[j];
    if (x != 1 && !is_wall(g, x - 1, y, z)) s -= v[j - 1];
    if (x != g.X - 2 && !is_wall(g, x + 1, y, z)) s -= v[j + 1];
    if (y != 1 && !is_wall(g, x, y - 1, z)) s -= v[j - g.Xp];
    if (y != g.Y - 2 && !is_wall(g, x, y + 1, z)) s -= v[j + g.Xp];
    if (z != 1 && !is_wall(g, x, y, z - 1)) s -= v[j - slab];
    if (z != g.Z - 2 && !is_wall(g, x, y, z + 1)) s -= v[j + slab];
    o[j] = s;
  } else if (op == CFS_OP_VEL_TO_DIV) {
    // divergence of a peeled velocity; a missing neighbour across a wall is
    // the cell's own value negated, across an open side (Neumann) the value
    // beyond the cell
    const double coeff = -g.dx * 0.5;
    const int right = j + 1, left = j - 1, up = j + g.Xp, down = j - g.Xp, far = j + slab, near = j - slab;
    double s = 0.0;
    if (px != g.Xp - 1) s += coeff * v[3 * right];
    else if (g.bc[3] != 0) s += -coeff * v[3 * j];
    else s += coeff * v[3 * left];
    if (px != 0) s += -coeff * v[3 * left];
    else if (g.bc[2] != 0) s += coeff * v[3 * j];
    else s += -coeff * v[3 * right];
    if (py != g.Yp - 1) s += coeff * v[3 * up + 1];
    else if (g.bc[4] != 0) s += -coeff * v[3 * j + 1];
    else s += coeff * v[3 * down + 1];
    if (py != 0) s += -coeff * v[3 * down + 1];
    else if (g.bc[5] != 0) s += coeff * v[3 * j + 1];
    else s += -coeff * v[3 * up + 1];
    if (pz != g.Zp - 1) s += coeff * v[3 * far + 2];
    else if (g.bc[1] != 0) s += -coeff * v[3 * j + 2];
    else s += coeff * v[3 * near + 2];
    if (pz != 0) s += -coeff * v[3 * near + 2];
    else if (g.bc[0] != 0) s += coeff * v[3 * j + 2];
    else s += -coeff * v[3 * far + 2];
    o[j] = s;
  } else if (op == CFS_OP_PRESSURE_TO_VEL) {
    // minus the pressure gradient, one-sided on the first and last layers
    const double inv = 1.0 / g.dx, half = 0.5 / g.dx;
    const int strides[3] = {1, g.Xp, slab};
    const int pos[3] = {px, py, pz}, lim[3] = {g.Xp, g.Yp, g.Zp};
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const int s = strides[c];
      double val;
      if (pos[c] == 0)
        val = -(v[j + s] - v[j]) * inv;
      else if (pos[c] == lim[c] - 1)
        val = -(v[j] - v[j - s]) * inv;
      else
        val = -(v[j + s] - v[j - s]) * half;
      o[3 * j + c] = val;
    }
  } else {  // CFS_OP_DAMPING: (1 - w) v + (w / 6) sum of the peeled neighbours
    const double sixth = (1.0 / 6.0) * g.w, keep = 1.0 - g.w;
#pragma unroll
    for (int c = 0; c < 3; c++) {
      double acc = 0.0;
      if (pz > 0) acc += v[3 * (j - slab) + c] * sixth;
      if (py > 0) acc += v[3 * (j - g.Xp) + c] * sixth;
      if (px > 0) acc += v[3 * (j - 1) + c] * sixth;
      acc += v[3 * j + c] * keep;
      if (px < g.Xp - 1) acc += v[3 * (j + 1) + c] * sixth;
      if (py < g.Yp - 1) acc += v[3 * (j + g.Xp) + c] * sixth;
      if (pz < g.Zp - 1) acc += v[3 * (j + slab) + c] * sixth;
      o[3 * j + c] = acc;
    }
  }
}

}  // namespace

extern "C" {

int cfs_gemm_tn(const double* A, int64_t rows, int m, int64_t lda, const double* B, int n, int64_t ldb, double* C,
                void* stream) {
  if (!A || !B || !C) return cfs::fail(CFS_ERR_ARG, "null argument");
  if (rows < 0 || m < 1 || n < 1 || lda < rows || ldb < rows) return cfs::fail(CFS_ERR_ARG, "bad dimensions");
  const hipStream_t s = (hipStream_t)stream;
  const int chunks = (int)((rows + CHUNK - 1) / CHUNK);
  if (chunks == 0) {
    hipLaunchKernelGGL(gemm_tn_reduce_kernel, dim3((m * n + 255) / 256), dim3(256), 0, s, C, 0, m, n, C);
    LA_HIP(hipGetLastError(), "gemm_tn launch");
    return CFS_OK;
  }
  if (chunks > 65535) return cfs::fail(CFS_ERR_ARG, "too many rows");
  double* partial = nullptr;
  LA_HIP(hipMallocAsync((void**)&partial, (size_t)chunks * m * n * 8, s), "hipMallocAsync");
  hipLaunchKernelGGL(gemm_tn_kernel, dim3((m + TT - 1) / TT, (n + TT - 1) / TT, chunks), dim3(TT * TT), 0, s, A, rows,
                     m, (long)lda, B, n, (long)ldb, partial);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) {
    hipLaunchKernelGGL(gemm_tn_reduce_kernel, dim3((m * n + 255) / 256), dim3(256), 0, s, partial, chunks, m, n, C);
    e = hipGetLastError();
  }
  const hipError_t f = hipFreeAsync(partial, s);
  if (e != hipSuccess) return cfs::hip_fail(e, "gemm_tn launch");
  LA_HIP(f, "hipFreeAsync");
  return CFS_OK;
}

int cfs_gemm_nn(const double* A, int64_t rows, int k, int64_t lda, const double* M, int n, double* C, int64_t ldc,
                void* stream) {
  if (!A || !M || !C) return cfs::fail(CFS_ERR_ARG, "null argument");
  if (rows < 1 || k < 1 || n < 1 || lda < rows || ldc < rows) return cfs::fail(CFS_ERR_ARG, "bad dimensions");
  if (k > NN_MAXK) return cfs::fail(CFS_ERR_ARG, "inner dimension above 512");
  hipLaunchKernelGGL(gemm_nn_kernel, dim3((unsigned)((rows + NN_TPB - 1) / NN_TPB), (n + NN_COLS - 1) / NN_COLS),
                     dim3(NN_TPB), 0, (hipStream_t)stream, A, (long)rows, k, (long)lda, M, n, C, (long)ldc);
  LA_HIP(hipGetLastError(), "gemm_nn launch");
  return CFS_OK;
}

int cfs_apply_operator(int op, int X, int Y, int Z, const int* boundaries, double dx, double w, const double* in,
                       int64_t ld_in, int ncols, double* out, int64_t ld_out, void* stream) {
  if (!in || !out || !boundaries) return cfs::fail(CFS_ERR_ARG, "null argument");
  if (op < CFS_OP_POISSON || op > CFS_OP_DAMPING) return cfs::fail(CFS_ERR_ARG, "unknown operator");
  if (X < 4 || Y < 4 || Z < 4 || ncols < 1) return cfs::fail(CFS_ERR_ARG, "bad dimensions");
  Grid g;
  g.X = X;
  g.Y = Y;
  g.Z = Z;
  g.Xp = X - 2;
  g.Yp = Y - 2;
  g.Zp = Z - 2;
  for (int k = 0; k < 6; k++) g.bc[k] = boundaries[k];
  g.dx = dx;
  g.w = w;
  const long np = (long)g.Xp * g.Yp * g.Zp;
  const long need_in = (op == CFS_OP_POISSON || op == CFS_OP_PRESSURE_TO_VEL) ? np : 3 * np;
  const long need_out = (op == CFS_OP_POISSON || op == CFS_OP_VEL_TO_DIV) ? np : 3 * np;
  if (ld_in < need_in || ld_out < need_out) return cfs::fail(CFS_ERR_ARG, "leading dimension below the column size");
  const long total = np * ncols;
  hipLaunchKernelGGL(operator_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, (hipStream_t)stream, op, g,
                     in, (long)ld_in, out, (long)ld_out, ncols);
  LA_HIP(hipGetLastError(), "operator launch");
  return CFS_OK;
}

}  // extern "C"

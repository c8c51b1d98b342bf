// obstacle.hip -- a moving box obstacle: the box itself, its full-space
// Neumann stomp, and the reduced stomp of the subspace solver.
//
//   box             BOX (reference src/geometry/BOX.cpp:21-131): spins about z
//                   with `period` (spin :105-114, rotation matrix
//                   MATRIX3::rotation(z, -theta), MATRIX3.h:96-117) and slides
//                   its centre along x as 0.25 sin(phi) (translate_center
//                   :116-126); inside() :58-70, update_r :72-81 and
//                   update_linearVelocity (BOX.h:61-65): omega x r + the slide
//   full stomp      setPeeledSparseMovingIOP and neumannIOP [v; 1]
//                   (FLUID_3D_MIC.cpp:1645-1702, used by stepWithMovingObstacle
//                   :279-282) and setInitialVelocity (:744-763): interior cells
//                   whose centre (cellCenter :1704-1733) lies in the box take
//                   the box's velocity there
//   reduced stomp   reducedSetMovingBox (SUBSPACE_FLUID_3D_COMPRESSED_EIGEN.cpp:
//                   779-790): q = R q + n with, per step, R = U_iop^T U_proj -
//                   U_iop^T C U_proj and n = U_iop^T nvec (C the diagonal of the
//                   box's peeled components, nvec their box velocities;
//                   buildObstacleMatrices, SUBSPACE_FLUID_3D_EIGEN.cpp:1699-1711).
//                   The reference precomputes one (R, n) file pair per step;
//                   here the step evaluates R q + n directly as
//                   M q + sum over box cells of U_iop[c]^T (v_box(c) - U_proj[c] q)
//                   with M = U_iop^T U_proj fixed: only the box's cells are read.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "../../include/cfs_plume.h"
#include "basis_internal.h"
#include "obstacle_internal.h"

namespace cfs {
namespace {

constexpr int TPB = 256;
constexpr double TWO_PI = 6.283185307179586476925286766559;

struct CellTest {
  bool inside;
  double v[3];
};

// cell (x, y, z) of the full grid against the box; the box velocity when
// inside.  Every product is rounded on its own (no fused multiply-adds), as
// in the reference's arithmetic, so that cells on the box's faces fall on the
// same side.
__device__ __forceinline__ CellTest test_cell(const BoxState& s, int x, int y, int z) {
#pragma clang fp contract(off)
  const double dx = 1.0 / s.X, dy = 1.0 / s.Y, dz = 1.0 / s.Z;
  const double p0 = x * dx + dx * 0.5, p1 = y * dy + dy * 0.5, p2 = z * dz + dz * 0.5;
  const double u0 = p0 - s.c[0], u1 = p1 - s.c[1], u2 = p2 - s.c[2];
  const double t0 = s.R[0] * u0 + s.R[1] * u1 + s.R[2] * u2;
  const double t1 = s.R[3] * u0 + s.R[4] * u1 + s.R[5] * u2;
  const double t2 = s.R[6] * u0 + s.R[7] * u1 + s.R[8] * u2;
  CellTest r;
  r.inside = fabs(t0) <= s.h[0] && fabs(t1) <= s.h[1] && fabs(t2) <= s.h[2];
  // r = u without its component along the spin axis z; omega x r + slide
  r.v[0] = -s.omega * u1 + s.slide;
  r.v[1] = s.omega * u0;
  r.v[2] = 0.0;
  return r;
}

__global__ void __launch_bounds__(TPB) box_stomp_kernel(double* __restrict__ vel, BoxState s) {
  const long i = (long)blockIdx.x * TPB + threadIdx.x;
  const long n = box_span_cells(s);
  if (i >= n) return;
  const int x = s.x0 + (int)(i % s.nx), y = s.y0 + (int)((i / s.nx) % s.ny), z = s.z0 + (int)(i / ((long)s.nx * s.ny));
  const CellTest t = test_cell(s, x, y, z);
  if (!t.inside) return;
  double* v = vel + 3 * ((long)x + (long)s.X * (y + (long)s.Y * z));
  v[0] = t.v[0];
  v[1] = t.v[1];
  v[2] = t.v[2];
}

// per cell of the box's span: flags = inside, w = v_box - U_proj[c] q
__global__ void __launch_bounds__(TPB) iop_rows_kernel(const double* __restrict__ U_proj, int r_proj,
                                                       const double* __restrict__ q, BoxState s,
                                                       double* __restrict__ w, uint8_t* __restrict__ flags) {
  const long i = (long)blockIdx.x * TPB + threadIdx.x;
  if (i >= box_span_cells(s)) return;
  const int x = s.x0 + (int)(i % s.nx), y = s.y0 + (int)((i / s.nx) % s.ny), z = s.z0 + (int)(i / ((long)s.nx * s.ny));
  const CellTest t = test_cell(s, x, y, z);
  flags[i] = t.inside ? 1 : 0;
  if (!t.inside) return;
  const long Xp = s.X - 2, Yp = s.Y - 2;
  const long n3 = 3 * Xp * Yp * (s.Z - 2);
  const long row = 3 * ((x - 1) + Xp * ((y - 1) + Yp * (z - 1)));
  double u0 = 0.0, u1 = 0.0, u2 = 0.0;
  for (int j = 0; j < r_proj; j++) {
    const double* col = U_proj + (long)j * n3 + row;
    const double qj = q[j];
    u0 += col[0] * qj;
    u1 += col[1] * qj;
    u2 += col[2] * qj;
  }
  w[3 * i + 0] = t.v[0] - u0;
  w[3 * i + 1] = t.v[1] - u1;
  w[3 * i + 2] = t.v[2] - u2;
}

// one workgroup per output j: q_out[j] = sum_k M[j][k] q[k] + sum over box
// cells of U_iop[c, j] . w[c]; partial sums meet in LDS in a fixed order
__global__ void __launch_bounds__(TPB) iop_reduce_kernel(const double* __restrict__ U_iop, const double* __restrict__ M,
                                                         int r_proj, const double* __restrict__ q, BoxState s,
                                                         const double* __restrict__ w,
                                                         const uint8_t* __restrict__ flags, double* __restrict__ q_out) {
  __shared__ double red[TPB];
  const int j = blockIdx.x, tid = threadIdx.x;
  const long Xp = s.X - 2, Yp = s.Y - 2;
  const long n3 = 3 * Xp * Yp * (s.Z - 2);
  const long nb = box_span_cells(s);
  const double* col = U_iop + (long)j * n3;
  double acc = 0.0;
  for (long i = tid; i < nb; i += TPB) {
    if (!flags[i]) continue;
    const int x = s.x0 + (int)(i % s.nx), y = s.y0 + (int)((i / s.nx) % s.ny),
              z = s.z0 + (int)(i / ((long)s.nx * s.ny));
    const long row = 3 * ((x - 1) + Xp * ((y - 1) + Yp * (z - 1)));
    acc += col[row] * w[3 * i] + col[row + 1] * w[3 * i + 1] + col[row + 2] * w[3 * i + 2];
  }
  for (int k = tid; k < r_proj; k += TPB) acc += M[(long)j * r_proj + k] * q[k];
  red[tid] = acc;
  __syncthreads();
  for (int off = TPB / 2; off > 0; off >>= 1) {
    if (tid < off) red[tid] += red[tid + off];
    __syncthreads();
  }
  if (tid == 0) q_out[j] = red[0];
}

unsigned blocks_for(long n) { return (unsigned)((n + TPB - 1) / TPB); }

}  // namespace

BoxState box_state(const cfs_box& b, int X, int Y, int Z) {
  BoxState s{};
  for (int i = 0; i < 3; i++) {
    s.c[i] = b.center[i];
    s.h[i] = b.half_lengths[i];
  }
  // MATRIX3::rotation about the unit z axis by -theta
  const double th = -b.theta;
  const double ax[3] = {0.0, 0.0, 1.0};
  const double c = std::cos(th), sn = std::sin(th), v = 1.0 - c;
  s.R[0] = ax[0] * ax[0] * v + c;
  s.R[1] = ax[0] * ax[1] * v - ax[2] * sn;
  s.R[2] = ax[0] * ax[2] * v + ax[1] * sn;
  s.R[3] = ax[1] * ax[0] * v + ax[2] * sn;
  s.R[4] = ax[1] * ax[1] * v + c;
  s.R[5] = ax[1] * ax[2] * v - ax[0] * sn;
  s.R[6] = ax[2] * ax[0] * v - ax[1] * sn;
  s.R[7] = ax[2] * ax[1] * v + ax[0] * sn;
  s.R[8] = ax[2] * ax[2] * v + c;
  s.omega = TWO_PI / b.period;
  s.slide = 0.25 * TWO_PI / b.translation_period * std::cos(b.phi);
  s.X = X;
  s.Y = Y;
  s.Z = Z;
  // the box's extent in the unit cube (corners c + R^T (+-h)), then the cells
  // whose centres can lie in it, one cell of margin, interior only
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int k = 0; k < 8; k++) {
    const double l[3] = {(k & 1) ? s.h[0] : -s.h[0], (k & 2) ? s.h[1] : -s.h[1], (k & 4) ? s.h[2] : -s.h[2]};
    for (int i = 0; i < 3; i++) {
      const double p = s.c[i] + s.R[0 + i] * l[0] + s.R[3 + i] * l[1] + s.R[6 + i] * l[2];
      lo[i] = std::min(lo[i], p);
      hi[i] = std::max(hi[i], p);
    }
  }
  const int res[3] = {X, Y, Z};
  int a[3], e[3];
  for (int i = 0; i < 3; i++) {
    const double fa = std::floor(lo[i] * res[i] - 0.5) - 1.0, fe = std::ceil(hi[i] * res[i] - 0.5) + 1.0;
    a[i] = (int)std::max(1.0, std::min(fa, (double)res[i]));
    e[i] = (int)std::min((double)(res[i] - 2), std::max(fe, 0.0));
  }
  s.x0 = a[0];
  s.y0 = a[1];
  s.z0 = a[2];
  s.nx = std::max(0, e[0] - a[0] + 1);
  s.ny = std::max(0, e[1] - a[1] + 1);
  s.nz = std::max(0, e[2] - a[2] + 1);
  if (s.nx == 0 || s.ny == 0 || s.nz == 0) s.nx = s.ny = s.nz = 0;
  return s;
}

hipError_t launch_box_stomp(double* vel, const BoxState& s, hipStream_t st) {
  const long n = box_span_cells(s);
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(box_stomp_kernel, dim3(blocks_for(n)), dim3(TPB), 0, st, vel, s);
  return hipGetLastError();
}

hipError_t launch_reduced_iop(const double* U_iop, int r_iop, const double* U_proj, int r_proj, const double* M,
                              const BoxState& s, const double* q, double* q_out, double* w, uint8_t* flags,
                              hipStream_t st) {
  const long n = box_span_cells(s);
  if (n > 0) {
    hipLaunchKernelGGL(iop_rows_kernel, dim3(blocks_for(n)), dim3(TPB), 0, st, U_proj, r_proj, q, s, w, flags);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(iop_reduce_kernel, dim3(r_iop), dim3(TPB), 0, st, U_iop, M, r_proj, q, s, w, flags, q_out);
  return hipGetLastError();
}

}  // namespace cfs

extern "C" {

int cfs_box_init(const double center[3], const double lengths[3], double period, double translation_period,
                 double dt, cfs_box* box) {
  if (!center || !lengths || !box) return cfs::fail(CFS_ERR_ARG, "null argument");
  if (!(period > 0.0) || !(translation_period > 0.0) || !(dt >= 0.0))
    return cfs::fail(CFS_ERR_ARG, "the periods must be positive and dt non-negative");
  for (int i = 0; i < 3; i++) {
    if (!(lengths[i] >= 0.0)) return cfs::fail(CFS_ERR_ARG, "box lengths must be non-negative");
    box->center[i] = center[i];
    box->original_center[i] = center[i];
    box->half_lengths[i] = 0.5 * lengths[i];
  }
  box->period = period;
  box->translation_period = translation_period;
  box->theta = 0.0;
  box->phi = 0.0;
  box->time = 0.0;
  box->dt = dt;
  return CFS_OK;
}

int cfs_box_advance(cfs_box* box) {
  if (!box) return cfs::fail(CFS_ERR_ARG, "null box");
  // translate_center
  box->phi += cfs::TWO_PI * box->dt / box->translation_period;
  if (box->phi >= cfs::TWO_PI) box->phi = 0.0;
  box->center[0] = box->original_center[0] + 0.25 * std::sin(box->phi);
  // spin
  box->theta += cfs::TWO_PI * box->dt / box->period;
  if (box->theta >= cfs::TWO_PI) box->theta = 0.0;
  // update_time
  box->time += box->dt;
  return CFS_OK;
}

int64_t cfs_reduced_iop_scratch_bytes(int X, int Y, int Z) {
  if (X < 1 || Y < 1 || Z < 1) return -1;
  const int64_t n = (int64_t)X * Y * Z;
  return 3 * n * (int64_t)sizeof(double) + ((n + 255) / 256) * 256;
}

int cfs_reduced_iop(const double* U_iop, int r_iop, const double* U_proj, int r_proj, const double* iop_T_proj,
                    int X, int Y, int Z, const cfs_box* box, const double* q, double* q_out, void* scratch,
                    void* stream) {
  if (!U_iop || !U_proj || !iop_T_proj || !box || !q || !q_out) return cfs::fail(CFS_ERR_ARG, "null argument");
  if (r_iop < 1 || r_proj < 1) return cfs::fail(CFS_ERR_ARG, "column counts must be positive");
  if (X < 1 || Y < 1 || Z < 1) return cfs::fail(CFS_ERR_ARG, "bad peeled grid");
  const hipStream_t s = (hipStream_t)stream;
  const cfs::BoxState bs = cfs::box_state(*box, X + 2, Y + 2, Z + 2);
  const int64_t n = (int64_t)X * Y * Z;
  void* buf = scratch;
  if (!buf) {
    const hipError_t e = hipMallocAsync(&buf, (size_t)cfs_reduced_iop_scratch_bytes(X, Y, Z), s);
    if (e != hipSuccess) return cfs::hip_fail(e, "hipMallocAsync");
  }
  double* w = static_cast<double*>(buf);
  uint8_t* flags = reinterpret_cast<uint8_t*>(w + 3 * n);
  const hipError_t e = cfs::launch_reduced_iop(U_iop, r_iop, U_proj, r_proj, iop_T_proj, bs, q, q_out, w, flags, s);
  if (!scratch) (void)hipFreeAsync(buf, s);
  return e == hipSuccess ? CFS_OK : cfs::hip_fail(e, "reduced IOP launch");
}

}  // extern "C"

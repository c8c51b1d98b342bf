// plume.hip -- the whole compressed plume timestep on the device.
//
// stepPlume (reference src/integrators/SUBSPACE_FLUID_3D_COMPRESSED_EIGEN.cpp:
// 453-501), preceded each frame by addSmokeColumn (FLUID_3D_MIC.cpp:655-703,
// called from reducedCompressed3D.cpp:385), as launches on one stream with
// no host round trip and no allocation, so a caller can capture it in a HIP
// graph:
//
//   forces      velocity.setZeroBorder(); velocity += dt * buoyancy * heat (y)
//               (addBuoyancy, FLUID_3D_MIC.cpp:491-502); vorticity confinement
//               (addVorticity :524-593): the curl, its magnitude, then
//               velocity += dt * eps * dx * (N x w) with N the normalised
//               gradient of |w| -- two stencil kernels
//   project     q = U_pre^T velocity                       (cfs_project)
//   scalars     velocity border per axis (setZeroX/copyBorderY/setZeroZ for the
//               plume's walls, VECTOR3_FIELD_3D.cpp:2872-3057); density and heat
//               advected by velocity with dt0 = dt / dx, then their borders
//               zeroed (advectHeatAndDensityStam :796-824)
//   advection   q = Adv(q): the cubature (reducedAdvectCompressionFriendly),
//               the learned closure, or the exact U^T A(U q)
//   reduced     q = D q; q = P q + inverseProduct q  (reducedPeeledDiffusion
//               :848-854, reducedStagedProject :942-946), one workgroup, FP64
//   unproject   velocity interior = U_final q             (cfs_unproject)
//
// Fields are full grids X x Y x Z (the peeled basis dims + 2), x fastest,
// FP32; velocity is interleaved xyz.
#include <hip/hip_runtime.h>

#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "../../include/cfs_plume.h"
#include "basis_internal.h"
#include "grid_ops.h"
#include "obstacle_internal.h"

using namespace cfs;  // the shared grid stencils

namespace cfs {
hipError_t launch_advect_scalar(const float* vel, const float* old, float* out, int X, int Y, int Z, float dt0,
                                hipStream_t s);
}  // namespace cfs

namespace {

constexpr int TPB = GRID_TPB;
constexpr int MAX_R = 256;

__global__ void __launch_bounds__(TPB) copy2_kernel(const float* __restrict__ a, float* __restrict__ a_old,
                                                    const float* __restrict__ b, float* __restrict__ b_old, int n) {
  const int i = blockIdx.x * TPB + threadIdx.x;
  if (i >= n) return;
  a_old[i] = a[i];
  b_old[i] = b[i];
}

// the peeled interior of an FP32 full-grid vector field as an FP64 column
__global__ void __launch_bounds__(TPB) peel_to_column_kernel(const float* __restrict__ f, double* __restrict__ col, int X,
                                                             int Y, int Z) {
  const int Xp = X - 2, Yp = Y - 2, Zp = Z - 2;
  const int j = blockIdx.x * TPB + threadIdx.x;
  if (j >= Xp * Yp * Zp) return;
  const int x = j % Xp + 1, y = (j / Xp) % Yp + 1, z = j / (Xp * Yp) + 1;
  const size_t i = (size_t)x + (size_t)X * (y + (size_t)Y * z);
  for (int c = 0; c < 3; c++) col[3 * (size_t)j + c] = (double)f[3 * i + c];
}

// setWithPeeled: an FP64 peeled column into the interior of an FP32 field
__global__ void __launch_bounds__(TPB) column_to_interior_kernel(const double* __restrict__ col, float* __restrict__ f,
                                                                 int X, int Y, int Z) {
  const int Xp = X - 2, Yp = Y - 2, Zp = Z - 2;
  const int j = blockIdx.x * TPB + threadIdx.x;
  if (j >= Xp * Yp * Zp) return;
  const int x = j % Xp + 1, y = (j / Xp) % Yp + 1, z = j / (Xp * Yp) + 1;
  const size_t i = (size_t)x + (size_t)X * (y + (size_t)Y * z);
  for (int c = 0; c < 3; c++) f[3 * i + c] = (float)col[3 * (size_t)j + c];
}

__global__ void copy_q_kernel(const double* __restrict__ a, double* __restrict__ b, int r) {
  if ((int)threadIdx.x < r) b[threadIdx.x] = a[threadIdx.x];
}

// q <- D q (r_proj x r_adv; D null: the identity, r_proj = r_adv), then
// q <- P q + I q (r_fin x r_proj), FP64, every dimension <= 256, one
// workgroup.  q_in and q_out may alias.
__global__ void __launch_bounds__(MAX_R) reduced_ops_kernel(const double* __restrict__ D, const double* __restrict__ P,
                                                            const double* __restrict__ I, const double* q_in,
                                                            double* q_out, int r_adv, int r_proj, int r_fin) {
  __shared__ double a[MAX_R], t[MAX_R];
  const int j = threadIdx.x;
  if (j < r_adv) a[j] = q_in[j];
  __syncthreads();
  if (j < r_proj) {
    double s = 0.0;
    if (D) {
      for (int k = 0; k < r_adv; k++) s += D[(size_t)j * r_adv + k] * a[k];
    } else {
      s = a[j];
    }
    t[j] = s;
  }
  __syncthreads();
  if (j < r_fin) {
    double p = 0.0, c = 0.0;
    for (int k = 0; k < r_proj; k++) {
      p += P[(size_t)j * r_proj + k] * t[k];
      c += I[(size_t)j * r_proj + k] * t[k];
    }
    q_out[j] = p + c;
  }
}

// y = A x (rows x cols, row-major), one workgroup, dimensions <= 256
__global__ void __launch_bounds__(MAX_R) matvec_kernel(const double* __restrict__ A, const double* __restrict__ x,
                                                       double* __restrict__ y, int rows, int cols) {
  const int j = threadIdx.x;
  if (j >= rows) return;
  double s = 0.0;
  for (int k = 0; k < cols; k++) s += A[(size_t)j * cols + k] * x[k];
  y[j] = s;
}

}  // namespace

struct cfs_plume_solver {
  int device = 0;
  cfs_basis* pre = nullptr;
  cfs_basis* fin = nullptr;
  cfs_plume_params prm{};
  int X = 0, Y = 0, Z = 0;                      // full grid
  int r_pre = 0, r_adv = 0, r_proj = 0, r_fin = 0;  // columns of U_preadvect, the advected, preprojected
                                                    // and final coordinates
  int mode = CFS_PLUME_ADVECT_EXACT;
  cfs_cubature* cub = nullptr;
  cfs_closure* closure = nullptr;
  double *d_D = nullptr, *d_P = nullptr, *d_I = nullptr, *d_q = nullptr;
  double* d_qpre = nullptr;  // the advection's input: q after the projection, zero-padded to MAX_R
  // dense bases (the uncompressed solver): device FP64, 3 Npeel x r column-major,
  // owned by the caller; null for compressed bases
  const double* U_pre_dense = nullptr;
  const double* U_fin_dense = nullptr;
  const double* U_adv_dense = nullptr;  // dense exact advection: project onto this after-basis
  double* d_col = nullptr;  // an FP64 peeled vector field
  // moving obstacle (cfs_plume_set_obstacle): dense U_iop and U_preproject
  // (device, owned by the caller), U_iop^T U_preproject, and the projection
  // matrices of the IOP basis (r_fin x r_iop)
  int r_iop = 0;
  const double* U_iop = nullptr;
  const double* U_proj = nullptr;
  double *d_M = nullptr, *d_Piop = nullptr, *d_Iiop = nullptr, *d_t = nullptr, *d_u = nullptr;
  void* d_iop_scratch = nullptr;
  float *d_w = nullptr, *d_wmag = nullptr, *d_dold = nullptr, *d_hold = nullptr, *d_fa = nullptr, *d_fb = nullptr;
  std::vector<void*> allocs;
  ~cfs_plume_solver() {
    for (void* p : allocs) (void)hipFree(p);
  }
};

namespace {

template <class T>
int dev_alloc(cfs_plume_solver* p, T** out, size_t count, const T* host) {
  const hipError_t e = hipMalloc(reinterpret_cast<void**>(out), std::max<size_t>(count, 1) * sizeof(T));
  if (e != hipSuccess) return cfs::hip_fail(e, "hipMalloc");
  p->allocs.push_back(*out);
  const hipError_t e2 = host ? hipMemcpy(*out, host, count * sizeof(T), hipMemcpyHostToDevice)
                             : hipMemset(*out, 0, count * sizeof(T));
  if (e2 != hipSuccess) return cfs::hip_fail(e2, host ? "hipMemcpy" : "hipMemset");
  return CFS_OK;
}

#define PL_HIP(call, what)                                  \
  do {                                                      \
    const hipError_t e_ = (call);                           \
    if (e_ != hipSuccess) return cfs::hip_fail(e_, what);   \
  } while (0)

long peeled(const cfs_plume_solver* p) { return (long)(p->X - 2) * (p->Y - 2) * (p->Z - 2); }

// q = U^T v with U_preadvect, compressed or dense
int project_pre(cfs_plume_solver* p, const float* field, double* q, void* stream) {
  if (!p->U_pre_dense) return cfs_project(p->pre, field, q, stream);
  const hipStream_t s = (hipStream_t)stream;
  const long np = peeled(p);
  hipLaunchKernelGGL(peel_to_column_kernel, dim3((unsigned)((np + TPB - 1) / TPB)), dim3(TPB), 0, s, field, p->d_col,
                     p->X, p->Y, p->Z);
  PL_HIP(hipGetLastError(), "peel launch");
  return cfs_gemm_tn(p->U_pre_dense, 3 * np, p->r_pre, 3 * np, p->d_col, 1, 3 * np, q, stream);
}

// field interior = U q, with U_preadvect (which = 0) or U_final (which = 1)
int unproject_with(cfs_plume_solver* p, int which, const double* q, float* field, void* stream) {
  const double* U = which ? p->U_fin_dense : p->U_pre_dense;
  if (!U) return cfs_unproject(which ? p->fin : p->pre, q, field, stream);
  const hipStream_t s = (hipStream_t)stream;
  const long np = peeled(p);
  const int r = which ? p->r_fin : p->r_pre;
  int st = cfs_gemm_nn(U, 3 * np, r, 3 * np, q, 1, p->d_col, 3 * np, stream);
  if (st) return st;
  hipLaunchKernelGGL(column_to_interior_kernel, dim3((unsigned)((np + TPB - 1) / TPB)), dim3(TPB), 0, s, p->d_col,
                     field, p->X, p->Y, p->Z);
  PL_HIP(hipGetLastError(), "interior launch");
  return CFS_OK;
}

}  // namespace

extern "C" {

void cfs_plume_default_params(int X, int Y, int Z, cfs_plume_params* prm) {
  if (!prm) return;
  // FLUID_3D_MIC constructor (FLUID_3D_MIC.cpp:30-56): dt 0.1, buoyancy 0.1,
  // vorticity 2 / max(1, 64 / maxRes), dx = 1 / maxRes; the plume driver's
  // walls front/back/left/right, open top/bottom (reducedCompressed3D.cpp:325-331)
  const int maxres = std::max(X, std::max(Y, Z));
  prm->dt = 0.1;
  prm->dx = 1.0 / (double)(float)maxres;
  prm->buoyancy = 0.1;
  const double scaling = std::max(1.0, (double)(64.0f / (float)maxres));
  prm->vorticity_eps = 2.0 / scaling;
  prm->smoke_offset_x = 0.0;
  prm->smoke_offset_z = 0.0;
  prm->open_x = 0;
  prm->open_y = 1;
  prm->open_z = 0;
}

}  // extern "C"

namespace {

// the reduced matrices and the scratch of a plume whose grid and dimensions are set
int finish_plume(cfs_plume_solver* P, const double* damping, const double* ptof, const double* inverse_product,
                 const cfs_plume_params* prm) {
  if (P->r_pre < 1 || P->r_pre > MAX_R || P->r_fin < 1 || P->r_fin > MAX_R || P->r_adv < 1 || P->r_adv > MAX_R ||
      P->r_proj < 1 || P->r_proj > MAX_R)
    return cfs::fail(CFS_ERR_ARG, "every reduced dimension must lie in [1, 256]");
  if (!(prm->dt > 0.0) || !(prm->dx > 0.0)) return cfs::fail(CFS_ERR_ARG, "dt and dx must be positive");
  P->prm = *prm;
  cfs::DeviceGuard g(P->device);
  const size_t n = (size_t)P->X * P->Y * P->Z;
  int st;
  if ((st = dev_alloc(P, &P->d_D, (size_t)P->r_proj * P->r_adv, damping))) return st;
  if ((st = dev_alloc(P, &P->d_P, (size_t)P->r_fin * P->r_proj, ptof))) return st;
  if ((st = dev_alloc(P, &P->d_I, (size_t)P->r_fin * P->r_proj, inverse_product))) return st;
  if ((st = dev_alloc<double>(P, &P->d_q, (size_t)MAX_R, nullptr))) return st;
  if ((st = dev_alloc<double>(P, &P->d_qpre, (size_t)MAX_R, nullptr))) return st;
  if ((st = dev_alloc<float>(P, &P->d_w, 3 * n, nullptr))) return st;
  if ((st = dev_alloc<float>(P, &P->d_wmag, n, nullptr))) return st;
  if ((st = dev_alloc<float>(P, &P->d_dold, n, nullptr))) return st;
  if ((st = dev_alloc<float>(P, &P->d_hold, n, nullptr))) return st;
  if ((st = dev_alloc<float>(P, &P->d_fa, 3 * n, nullptr))) return st;
  if ((st = dev_alloc<float>(P, &P->d_fb, 3 * n, nullptr))) return st;
  if (P->U_pre_dense && (st = dev_alloc<double>(P, &P->d_col, 3 * (size_t)peeled(P), nullptr))) return st;
  return CFS_OK;
}

}  // namespace

extern "C" {

int cfs_plume_create(cfs_basis* pre, cfs_basis* fin, int r_adv, int r_proj, const double* damping,
                     const double* ptof, const double* inverse_product, const cfs_plume_params* prm,
                     cfs_plume_solver** out) {
  if (!out) return cfs::fail(CFS_ERR_ARG, "null output handle");
  *out = nullptr;
  if (!pre || !fin || !damping || !ptof || !inverse_product || !prm) return cfs::fail(CFS_ERR_ARG, "null argument");
  cfs_basis_info a, b;
  int st = cfs_basis_get_info(pre, &a);
  if (st || (st = cfs_basis_get_info(fin, &b))) return st;
  if (a.dims[0] != b.dims[0] || a.dims[1] != b.dims[1] || a.dims[2] != b.dims[2])
    return cfs::fail(CFS_ERR_ARG, "U_preadvect and U_final must share the grid");
  if (cfs::basis_device(pre) != cfs::basis_device(fin)) return cfs::fail(CFS_ERR_ARG, "bases on different devices");
  std::unique_ptr<cfs_plume_solver> P(new cfs_plume_solver);
  P->device = cfs::basis_device(pre);
  P->pre = pre;
  P->fin = fin;
  P->X = a.dims[0] + 2;
  P->Y = a.dims[1] + 2;
  P->Z = a.dims[2] + 2;
  P->r_pre = a.num_cols;
  P->r_adv = r_adv;
  P->r_proj = r_proj;
  P->r_fin = b.num_cols;
  if ((st = finish_plume(P.get(), damping, ptof, inverse_product, prm))) return st;
  *out = P.release();
  return CFS_OK;
}

int cfs_plume_create_dense(const double* U_pre, int r_pre, const double* U_fin, int r_fin, int X, int Y, int Z,
                           int r_adv, int r_proj, const double* damping, const double* ptof,
                           const double* inverse_product, const cfs_plume_params* prm, cfs_plume_solver** out) {
  if (!out) return cfs::fail(CFS_ERR_ARG, "null output handle");
  *out = nullptr;
  if (!U_pre || !U_fin || !damping || !ptof || !inverse_product || !prm)
    return cfs::fail(CFS_ERR_ARG, "null argument");
  if (X < 2 || Y < 2 || Z < 2) return cfs::fail(CFS_ERR_ARG, "bad peeled grid");
  std::unique_ptr<cfs_plume_solver> P(new cfs_plume_solver);
  PL_HIP(hipGetDevice(&P->device), "hipGetDevice");
  P->U_pre_dense = U_pre;
  P->U_fin_dense = U_fin;
  P->X = X + 2;
  P->Y = Y + 2;
  P->Z = Z + 2;
  P->r_pre = r_pre;
  P->r_adv = r_adv;
  P->r_proj = r_proj;
  P->r_fin = r_fin;
  int st = finish_plume(P.get(), damping, ptof, inverse_product, prm);
  if (st) return st;
  *out = P.release();
  return CFS_OK;
}

int cfs_plume_set_exact_after(cfs_plume_solver* p, const double* U_adv) {
  if (!p || !U_adv) return cfs::fail(CFS_ERR_ARG, "null argument");
  if (!p->U_pre_dense) return cfs::fail(CFS_ERR_ARG, "the exact after-basis needs a dense plume");
  p->U_adv_dense = U_adv;
  return CFS_OK;
}

int cfs_plume_set_advection(cfs_plume_solver* p, int mode, cfs_cubature* cub, cfs_closure* closure) {
  if (!p) return cfs::fail(CFS_ERR_ARG, "null plume");
  if (mode == CFS_PLUME_ADVECT_CUBATURE) {
    int K = 0, r = 0;
    if (!cub) return cfs::fail(CFS_ERR_ARG, "cubature mode needs a cubature");
    int st = cfs_cubature_get_info(cub, &K, &r);
    if (st) return st;
    if (r != p->r_adv) return cfs::fail(CFS_ERR_ARG, "the cubature's after-basis has another column count");
  } else if (mode == CFS_PLUME_ADVECT_CLOSURE) {
    int r = 0, rp = 0, h = 0;
    int64_t np = 0;
    if (!closure) return cfs::fail(CFS_ERR_ARG, "closure mode needs a closure");
    int st = cfs_closure_get_dims(closure, &r, &rp, &h, &np);
    if (st) return st;
    if (r != p->r_adv || r < p->r_pre)
      return cfs::fail(CFS_ERR_ARG, "the closure's r must equal r_adv and be at least U_preadvect's r");
  } else if (mode == CFS_PLUME_ADVECT_EXACT) {
    if (p->r_adv != p->r_pre && !p->U_adv_dense)
      return cfs::fail(CFS_ERR_ARG, "the exact advection stays in U_preadvect coordinates: r_adv must equal its r");
  } else {
    return cfs::fail(CFS_ERR_ARG, "unknown advection mode");
  }
  p->mode = mode;
  p->cub = cub;
  p->closure = closure;
  return CFS_OK;
}

int cfs_plume_add_smoke(cfs_plume_solver* p, float* density, float* heat, void* stream) {
  if (!p || !density || !heat) return cfs::fail(CFS_ERR_ARG, "null argument");
  cfs::DeviceGuard g(p->device);
  const int n = p->X * p->Y * p->Z;
  const int y0 = (int)(0.05 * p->Y), y1 = (int)(0.10 * p->Y);
  hipLaunchKernelGGL(smoke_kernel<float>, dim3((n + TPB - 1) / TPB), dim3(TPB), 0, (hipStream_t)stream, density, heat,
                     p->X, p->Y, p->Z, y0, y1, p->prm.smoke_offset_x, p->prm.smoke_offset_z);
  PL_HIP(hipGetLastError(), "smoke launch");
  return CFS_OK;
}

int cfs_plume_set_obstacle(cfs_plume_solver* p, int r_iop, const double* U_iop, const double* U_proj,
                           const double* iop_T_proj, const double* ptof_iop, const double* inverse_iop) {
  if (!p || !U_iop || !U_proj || !iop_T_proj || !ptof_iop || !inverse_iop)
    return cfs::fail(CFS_ERR_ARG, "null argument");
  if (r_iop < 1 || r_iop > MAX_R) return cfs::fail(CFS_ERR_ARG, "r_iop must lie in [1, 256]");
  if (p->r_iop) return cfs::fail(CFS_ERR_ARG, "the plume already has an obstacle");
  cfs::DeviceGuard g(p->device);
  int st;
  if ((st = dev_alloc(p, &p->d_M, (size_t)r_iop * p->r_proj, iop_T_proj))) return st;
  if ((st = dev_alloc(p, &p->d_Piop, (size_t)p->r_fin * r_iop, ptof_iop))) return st;
  if ((st = dev_alloc(p, &p->d_Iiop, (size_t)p->r_fin * r_iop, inverse_iop))) return st;
  if ((st = dev_alloc<double>(p, &p->d_t, (size_t)MAX_R, nullptr))) return st;
  if ((st = dev_alloc<double>(p, &p->d_u, (size_t)MAX_R, nullptr))) return st;
  const int64_t bytes = cfs_reduced_iop_scratch_bytes(p->X - 2, p->Y - 2, p->Z - 2);
  uint8_t* scratch = nullptr;
  if ((st = dev_alloc<uint8_t>(p, &scratch, (size_t)bytes, nullptr))) return st;
  p->d_iop_scratch = scratch;
  p->U_iop = U_iop;
  p->U_proj = U_proj;
  p->r_iop = r_iop;
  return CFS_OK;
}

}  // extern "C"

namespace {

// stepPlume, or with a box stepMovingObstacle (SUBSPACE_FLUID_3D_COMPRESSED_
// EIGEN.cpp:508-566): no buoyancy, and between the diffusion and the pressure
// projection q = R q + n (reducedSetMovingBox :779-790), the projection then
// with the IOP basis' matrices.
int plume_step(cfs_plume_solver* p, const cfs_box* box, float* velocity, float* density, float* heat, double* q,
               void* stream) {
  if (p->mode == CFS_PLUME_ADVECT_EXACT && p->r_adv != p->r_pre && !p->U_adv_dense)
    return cfs::fail(CFS_ERR_ARG, "no advection set: the exact advection needs r_adv equal to U_preadvect's r");
  cfs::DeviceGuard g(p->device);
  const hipStream_t s = (hipStream_t)stream;
  const int X = p->X, Y = p->Y, Z = p->Z, n = X * Y * Z;
  const dim3 grid((n + TPB - 1) / TPB), block(TPB);
  const float dt = (float)p->prm.dt, inv_dx = (float)(1.0 / p->prm.dx);
  const float dt0 = (float)(p->prm.dt / p->prm.dx);
  int st;
  // forces
  hipLaunchKernelGGL(buoyancy_kernel<float>, grid, block, 0, s, velocity, heat, X, Y, Z, dt,
                     box ? 0.0f : (float)p->prm.buoyancy);
  PL_HIP(hipGetLastError(), "buoyancy launch");
  if (p->prm.vorticity_eps > 0.0) {
    hipLaunchKernelGGL(curl_kernel<float>, grid, block, 0, s, velocity, p->d_w, p->d_wmag, X, Y, Z, inv_dx);
    PL_HIP(hipGetLastError(), "curl launch");
    hipLaunchKernelGGL(confinement_kernel<float>, grid, block, 0, s, velocity, p->d_w, p->d_wmag, X, Y, Z, inv_dx, dt,
                       (float)(p->prm.vorticity_eps * p->prm.dx));
    PL_HIP(hipGetLastError(), "confinement launch");
  }
  // projection into the subspace; the advection reads its copy, zero-padded
  if ((st = project_pre(p, velocity, q, stream))) return st;
  hipLaunchKernelGGL(copy_q_kernel, dim3(1), dim3(MAX_R), 0, s, q, p->d_qpre, p->r_pre);
  PL_HIP(hipGetLastError(), "copy launch");
  // density and heat
  hipLaunchKernelGGL(velocity_border_kernel<float>, grid, block, 0, s, velocity, X, Y, Z, p->prm.open_x,
                     p->prm.open_y, p->prm.open_z, 1);
  PL_HIP(hipGetLastError(), "border launch");
  hipLaunchKernelGGL(copy2_kernel, grid, block, 0, s, density, p->d_dold, heat, p->d_hold, n);
  PL_HIP(hipGetLastError(), "copy launch");
  PL_HIP(cfs::launch_advect_scalar(velocity, p->d_dold, density, X, Y, Z, dt0, s), "advect launch");
  PL_HIP(cfs::launch_advect_scalar(velocity, p->d_hold, heat, X, Y, Z, dt0, s), "advect launch");
  hipLaunchKernelGGL(zero_border2_kernel<float>, grid, block, 0, s, density, heat, X, Y, Z);
  PL_HIP(hipGetLastError(), "border launch");
  // advection in the subspace
  if (p->mode == CFS_PLUME_ADVECT_CLOSURE) {
    st = cfs_closure_step(p->closure, p->d_qpre, p->d_q, 1, 1.0, nullptr, stream);
  } else if (!p->U_pre_dense) {
    st = p->mode == CFS_PLUME_ADVECT_CUBATURE
             ? cfs_cubature_advect(p->cub, p->pre, p->d_qpre, p->d_q, p->d_fa, dt0, 1.0, nullptr, stream)
             : cfs_reduced_advect(p->pre, p->d_qpre, p->d_q, p->d_fa, p->d_fb, dt0, 1.0, nullptr, stream);
  } else {
    // dense U_preadvect: unproject once (d_fa keeps a zero border), then the
    // cubature from the field, or the exact advection and the projection
    if ((st = unproject_with(p, 0, p->d_qpre, p->d_fa, stream))) return st;
    if (p->mode == CFS_PLUME_ADVECT_CUBATURE) {
      st = cfs_cubature_advect_field(p->cub, p->d_fa, p->d_q, dt0, 1.0, nullptr, stream);
    } else {
      if ((st = cfs_advect_field(p->d_fa, p->d_fb, X, Y, Z, dt0, stream))) return st;
      if (p->U_adv_dense) {
        const long np = peeled(p);
        hipLaunchKernelGGL(peel_to_column_kernel, dim3((unsigned)((np + TPB - 1) / TPB)), dim3(TPB), 0, s, p->d_fb,
                           p->d_col, X, Y, Z);
        PL_HIP(hipGetLastError(), "peel launch");
        st = cfs_gemm_tn(p->U_adv_dense, 3 * np, p->r_adv, 3 * np, p->d_col, 1, 3 * np, p->d_q, stream);
      } else {
        st = project_pre(p, p->d_fb, p->d_q, stream);
      }
    }
  }
  if (st) return st;
  if (box) {
    // diffusion, the box, the pressure projection with the IOP basis
    hipLaunchKernelGGL(matvec_kernel, dim3(1), dim3(MAX_R), 0, s, p->d_D, p->d_q, p->d_t, p->r_proj, p->r_adv);
    PL_HIP(hipGetLastError(), "diffusion launch");
    const int64_t n = (int64_t)(X - 2) * (Y - 2) * (Z - 2);
    double* w = static_cast<double*>(p->d_iop_scratch);
    uint8_t* flags = reinterpret_cast<uint8_t*>(w + 3 * n);
    PL_HIP(cfs::launch_reduced_iop(p->U_iop, p->r_iop, p->U_proj, p->r_proj, p->d_M, cfs::box_state(*box, X, Y, Z),
                                   p->d_t, p->d_u, w, flags, s),
           "reduced IOP launch");
    hipLaunchKernelGGL(reduced_ops_kernel, dim3(1), dim3(MAX_R), 0, s, (const double*)nullptr, p->d_Piop, p->d_Iiop,
                       p->d_u, q, p->r_iop, p->r_iop, p->r_fin);
  } else {
    // diffusion and pressure projection
    hipLaunchKernelGGL(reduced_ops_kernel, dim3(1), dim3(MAX_R), 0, s, p->d_D, p->d_P, p->d_I, p->d_q, q, p->r_adv,
                       p->r_proj, p->r_fin);
  }
  PL_HIP(hipGetLastError(), "reduced operators launch");
  // unprojection
  return unproject_with(p, 1, q, velocity, stream);
}

}  // namespace

extern "C" {

int cfs_plume_step(cfs_plume_solver* p, float* velocity, float* density, float* heat, double* q, void* stream) {
  if (!p || !velocity || !density || !heat || !q) return cfs::fail(CFS_ERR_ARG, "null argument");
  return plume_step(p, nullptr, velocity, density, heat, q, stream);
}

int cfs_plume_step_obstacle(cfs_plume_solver* p, const cfs_box* box, float* velocity, float* density, float* heat,
                            double* q, void* stream) {
  if (!p || !box || !velocity || !density || !heat || !q) return cfs::fail(CFS_ERR_ARG, "null argument");
  if (!p->r_iop) return cfs::fail(CFS_ERR_ARG, "no obstacle set (cfs_plume_set_obstacle)");
  return plume_step(p, box, velocity, density, heat, q, stream);
}

int cfs_plume_set_params(cfs_plume_solver* p, const cfs_plume_params* prm) {
  if (!p || !prm) return cfs::fail(CFS_ERR_ARG, "null argument");
  if (!(prm->dt > 0.0) || !(prm->dx > 0.0)) return cfs::fail(CFS_ERR_ARG, "dt and dx must be positive");
  p->prm = *prm;
  return CFS_OK;
}

int cfs_plume_advection_state(const cfs_plume_solver* p, const double** q_pre, const double** q_adv) {
  if (!p) return cfs::fail(CFS_ERR_ARG, "null plume");
  if (q_pre) *q_pre = p->d_qpre;
  if (q_adv) *q_adv = p->d_q;
  return CFS_OK;
}

int cfs_plume_destroy(cfs_plume_solver* p) {
  if (!p) return cfs::fail(CFS_ERR_ARG, "null plume");
  cfs::DeviceGuard g(p->device);
  delete p;
  return CFS_OK;
}

}  // extern "C"

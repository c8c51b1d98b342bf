// fluid.hip -- the full-rank Stam smoke plume on the device, in FP64.
//
// Restates FLUID_3D_MIC (reference src/integrators/FLUID_3D_MIC.cpp), the
// solver whose snapshots the reduced bases are built from (projects/fluid3D:
// addSmokeColumn, step, appendStreams per frame):
//
//   step :179-233   velocity.setZeroBorder, density.setZeroBorder; buoyancy;
//                   vorticity confinement; advectStam :595-653 (border rules,
//                   semi-Lagrangian advection of density, heat and velocity,
//                   zero borders, border rules); the peeled damping matrix
//                   (buildPeeledDampingMatrixFull :1251-1307, w = 0.9);
//                   project :314-402 (Neumann/zero border rules, divergence,
//                   MIC-preconditioned CG for the pressure, gradient subtract)
//   snapshots       appendStreams :1329-1480: peeled velocity before
//                   advection, before diffusion, before projection and
//                   final; peeled pressure and divergence
//
// Fields are the reference's `Real` = double.  The Poisson matrix and the
// MIC(0) preconditioner depend only on the obstacles (the domain walls), so
// they are built once on the host with the reference's loops (buildA
// :772-801, buildMICPreconditioner :825-862).  Per CG iteration the device
// runs the stencil product, the two triangular sweeps of the preconditioner
// (applyMICPreconditioner :864-899) and the vector updates; the sweeps are
// level-scheduled: interior cells grouped by x + y + z, whose dependencies
// (x-1, y-1, z-1 forward; x+1, y+1, z+1 backward) lie in the previous level,
// one workgroup walking the levels with a barrier between them.  Dot products
// and the residual maximum are deterministic two-stage reductions; their
// results come back to the host, which runs the CG control flow
// (solvePoisson :918-1017: stop when the residual maximum is below eps, after
// `iterations`, or when it changes by less than 1e-7 relative).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "../../include/cfs_plume.h"
#include "basis_internal.h"
#include "grid_ops.h"
#include "obstacle_internal.h"

using namespace cfs;  // the shared grid stencils

namespace {

constexpr int TPB = GRID_TPB;
constexpr int RED_BLOCKS = 256;   // first-stage blocks of a reduction
constexpr int SWEEP_TPB = 1024;   // the preconditioner's single workgroup

#define FL_HIP(call, what)                                  \
  do {                                                      \
    const hipError_t e_ = (call);                           \
    if (e_ != hipSuccess) return cfs::hip_fail(e_, what);   \
  } while (0)

__device__ __forceinline__ bool interior(int x, int y, int z, int X, int Y, int Z) {
  return x > 0 && y > 0 && z > 0 && x < X - 1 && y < Y - 1 && z < Z - 1;
}

__global__ void __launch_bounds__(TPB) copy_kernel(const double* __restrict__ a, double* __restrict__ b, long n) {
  const long i = (long)blockIdx.x * TPB + threadIdx.x;
  if (i < n) b[i] = a[i];
}

__global__ void __launch_bounds__(TPB) fill_kernel(double* __restrict__ a, long n, double v) {
  const long i = (long)blockIdx.x * TPB + threadIdx.x;
  if (i < n) a[i] = v;
}

// The peeled damping matrix applied to the velocity interior
// (buildPeeledDampingMatrixFull, FLUID_3D_MIC.cpp:1251-1307, then
// setWithPeeled): (1 - w) v + (w / 6) sum of the neighbours inside the peeled
// domain, summed in the order of the sparse row's columns (z-1, y-1, x-1,
// self, x+1, y+1, z+1), as the map-ordered product SPARSE_MATRIX * VectorXd
// (SPARSE_MATRIX.cpp:83-105) does.  in -> out, interior cells only.
__global__ void __launch_bounds__(TPB) damping_kernel(const double* __restrict__ in, double* __restrict__ out, int X,
                                                      int Y, int Z, double w) {
  const int i = blockIdx.x * TPB + threadIdx.x;
  if (i >= X * Y * Z) return;
  const int x = i % X, y = (i / X) % Y, z = i / (X * Y);
  if (!interior(x, y, z, X, Y, Z)) return;
  const int slab = X * Y;
  const double sixth = (1.0 / 6.0) * w, keep = 1.0 - w;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const double* v = in + c;
    double acc = 0.0;
    if (z > 1) acc += v[3 * (size_t)(i - slab)] * sixth;
    if (y > 1) acc += v[3 * (size_t)(i - X)] * sixth;
    if (x > 1) acc += v[3 * (size_t)(i - 1)] * sixth;
    acc += v[3 * (size_t)i] * keep;
    if (x < X - 2) acc += v[3 * (size_t)(i + 1)] * sixth;
    if (y < Y - 2) acc += v[3 * (size_t)(i + X)] * sixth;
    if (z < Z - 2) acc += v[3 * (size_t)(i + slab)] * sixth;
    out[3 * (size_t)i + c] = acc;
  }
}

// copies the interior of a vector field
__global__ void __launch_bounds__(TPB) copy_interior_kernel(const double* __restrict__ a, double* __restrict__ b, int X,
                                                            int Y, int Z) {
  const int i = blockIdx.x * TPB + threadIdx.x;
  if (i >= X * Y * Z) return;
  const int x = i % X, y = (i / X) % Y, z = i / (X * Y);
  if (!interior(x, y, z, X, Y, Z)) return;
  b[3 * (size_t)i] = a[3 * (size_t)i];
  b[3 * (size_t)i + 1] = a[3 * (size_t)i + 1];
  b[3 * (size_t)i + 2] = a[3 * (size_t)i + 2];
}

// peelBoundary().flattened() of a field with `comps` components into a
// snapshot column (peeled cells x fastest, components interleaved)
__global__ void __launch_bounds__(TPB) peel_kernel(const double* __restrict__ f, double* __restrict__ col, int X, int Y,
                                                   int Z, int comps) {
  const int Xp = X - 2, Yp = Y - 2, Zp = Z - 2;
  const int j = blockIdx.x * TPB + threadIdx.x;
  if (j >= Xp * Yp * Zp) return;
  const int x = j % Xp + 1, y = (j / Xp) % Yp + 1, z = j / (Xp * Yp) + 1;
  const size_t i = (size_t)x + (size_t)X * (y + (size_t)Y * z);
  for (int c = 0; c < comps; c++) col[(size_t)comps * j + c] = f[comps * i + c];
}

// The divergence of project() (FLUID_3D_MIC.cpp:336-362): at interior cells,
// -dx/2 times the central differences, a neighbour inside an obstacle
// replaced by the cell's own value negated.
__global__ void __launch_bounds__(TPB) divergence_kernel(const double* __restrict__ vel, const uint8_t* __restrict__ obst,
                                                         double* __restrict__ div, int X, int Y, int Z, double dx) {
  const int i = blockIdx.x * TPB + threadIdx.x;
  if (i >= X * Y * Z) return;
  const int x = i % X, y = (i / X) % Y, z = i / (X * Y);
  if (!interior(x, y, z, X, Y, Z)) return;
  const int slab = X * Y;
  const double* v = vel;
  double xright = v[3 * (size_t)(i + 1)], xleft = v[3 * (size_t)(i - 1)];
  double yup = v[3 * (size_t)(i + X) + 1], ydown = v[3 * (size_t)(i - X) + 1];
  double ztop = v[3 * (size_t)(i + slab) + 2], zbottom = v[3 * (size_t)(i - slab) + 2];
  if (obst[i + 1]) xright = -v[3 * (size_t)i];
  if (obst[i - 1]) xleft = -v[3 * (size_t)i];
  if (obst[i + X]) yup = -v[3 * (size_t)i + 1];
  if (obst[i - X]) ydown = -v[3 * (size_t)i + 1];
  if (obst[i + slab]) ztop = -v[3 * (size_t)i + 2];
  if (obst[i - slab]) zbottom = -v[3 * (size_t)i + 2];
  div[i] = -dx * 0.5 * (xright - xleft + yup - ydown + ztop - zbottom);
}

// bf = A xf at interior cells (multA, FLUID_3D_MIC.cpp:803-823)
__global__ void __launch_bounds__(TPB) mult_a_kernel(const double* __restrict__ xf, double* __restrict__ bf,
                                                     const double* __restrict__ Ad, const double* __restrict__ Api,
                                                     const double* __restrict__ Apj, const double* __restrict__ Apk,
                                                     int X, int Y, int Z) {
  const int i = blockIdx.x * TPB + threadIdx.x;
  if (i >= X * Y * Z) return;
  const int x = i % X, y = (i / X) % Y, z = i / (X * Y);
  if (!interior(x, y, z, X, Y, Z)) return;
  const int slab = X * Y;
  bf[i] = Ad[i] * xf[i] + xf[i - 1] * Api[i - 1] + xf[i + 1] * Api[i] + xf[i - X] * Apj[i - X] + xf[i + X] * Apj[i] +
          xf[i - slab] * Apk[i - slab] + xf[i + slab] * Apk[i];
}

// r = (b - r) at interior cells, then zeroBoundary(r, obstacles)
__global__ void __launch_bounds__(TPB) residual_kernel(const double* __restrict__ b, double* __restrict__ r,
                                                       const uint8_t* __restrict__ obst, int X, int Y, int Z) {
  const int i = blockIdx.x * TPB + threadIdx.x;
  if (i >= X * Y * Z) return;
  const int x = i % X, y = (i / X) % Y, z = i / (X * Y);
  if (interior(x, y, z, X, Y, Z)) r[i] = b[i] - r[i];
  if (obst[i]) r[i] = 0.0;
}

// x += alpha d; r -= alpha z (whole fields, as FIELD_3D::axpy)
__global__ void __launch_bounds__(TPB) cg_update_kernel(double* __restrict__ x, double* __restrict__ r,
                                                        const double* __restrict__ d, const double* __restrict__ z,
                                                        long n, double alpha) {
  const long i = (long)blockIdx.x * TPB + threadIdx.x;
  if (i >= n) return;
  x[i] += alpha * d[i];
  r[i] += -alpha * z[i];
}

// d = beta d + z, then zeroBoundary(d, obstacles)
__global__ void __launch_bounds__(TPB) direction_kernel(double* __restrict__ d, const double* __restrict__ z,
                                                        const uint8_t* __restrict__ obst, long n, double beta) {
  const long i = (long)blockIdx.x * TPB + threadIdx.x;
  if (i >= n) return;
  double v = d[i] * beta;
  v += z[i];
  d[i] = obst[i] ? 0.0 : v;
}

// applyMICPreconditioner (FLUID_3D_MIC.cpp:864-899): L q = r forward, then
// L^T zf = q backward, over the interior cells listed level by level
// (level = x + y + z; cells [off[l], off[l+1]) of `cells`).  One workgroup;
// cells of one level are independent, and every value a level reads was
// written by an earlier level or is a border zero.
__global__ void __launch_bounds__(SWEEP_TPB) mic_sweep_kernel(
    const double* __restrict__ r, double* __restrict__ q, double* __restrict__ zf, const double* __restrict__ Api,
    const double* __restrict__ Apj, const double* __restrict__ Apk, const double* __restrict__ pc,
    const int* __restrict__ cells, const int* __restrict__ off, int nlev, int X, int Y) {
  const int slab = X * Y;
  for (int l = 0; l < nlev; l++) {
    for (int k = off[l] + threadIdx.x; k < off[l + 1]; k += SWEEP_TPB) {
      const int i = cells[k];
      const int xm = i - 1, ym = i - X, zm = i - slab;
      const double t = r[i] - Api[xm] * pc[xm] * q[xm] - Apj[ym] * pc[ym] * q[ym] - Apk[zm] * pc[zm] * q[zm];
      q[i] = t * pc[i];
    }
    __syncthreads();
  }
  for (int l = nlev - 1; l >= 0; l--) {
    for (int k = off[l] + threadIdx.x; k < off[l + 1]; k += SWEEP_TPB) {
      const int i = cells[k];
      const double p = pc[i];
      const double t = q[i] - Api[i] * p * zf[i + 1] - Apj[i] * p * zf[i + X] - Apk[i] * p * zf[i + slab];
      zf[i] = t * p;
    }
    __syncthreads();
  }
}

// First stage of the reductions: block b sums a[i] b[i] (dot) and takes the
// maximum of m[i] (max) over its fixed slice, in a fixed tree order.
__global__ void __launch_bounds__(TPB) reduce1_kernel(const double* __restrict__ a, const double* __restrict__ b,
                                                      const double* __restrict__ m, long n, double* __restrict__ part) {
  __shared__ double s_dot[TPB], s_max[TPB];
  const long per = (n + RED_BLOCKS - 1) / RED_BLOCKS;
  const long lo = per * blockIdx.x, hi = std::min(n, lo + per);
  double d = 0.0, mx = -INFINITY;
  for (long i = lo + threadIdx.x; i < hi; i += TPB) {
    if (a) d += a[i] * b[i];
    if (m) mx = fmax(mx, m[i]);
  }
  s_dot[threadIdx.x] = d;
  s_max[threadIdx.x] = mx;
  __syncthreads();
  for (int h = TPB / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      s_dot[threadIdx.x] += s_dot[threadIdx.x + h];
      s_max[threadIdx.x] = fmax(s_max[threadIdx.x], s_max[threadIdx.x + h]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[blockIdx.x] = s_dot[0];
    part[RED_BLOCKS + blockIdx.x] = s_max[0];
  }
}

// second stage: out[0] = sum of the dots, out[1] = max of the maxima
__global__ void __launch_bounds__(RED_BLOCKS) reduce2_kernel(const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double s_dot[RED_BLOCKS], s_max[RED_BLOCKS];
  s_dot[threadIdx.x] = part[threadIdx.x];
  s_max[threadIdx.x] = part[RED_BLOCKS + threadIdx.x];
  __syncthreads();
  for (int h = RED_BLOCKS / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      s_dot[threadIdx.x] += s_dot[threadIdx.x + h];
      s_max[threadIdx.x] = fmax(s_max[threadIdx.x], s_max[threadIdx.x + h]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = s_dot[0];
    out[1] = s_max[0];
  }
}

// velocity -= grad p at interior cells, one-sided next to the border
// (project, FLUID_3D_MIC.cpp:373-401)
__global__ void __launch_bounds__(TPB) gradient_kernel(double* __restrict__ vel, const double* __restrict__ p, int X,
                                                       int Y, int Z, double inv_dx) {
  const int i = blockIdx.x * TPB + threadIdx.x;
  if (i >= X * Y * Z) return;
  const int x = i % X, y = (i / X) % Y, z = i / (X * Y);
  if (!interior(x, y, z, X, Y, Z)) return;
  const int slab = X * Y;
  double* v = vel + 3 * (size_t)i;
  if (x == 1)
    v[0] -= (p[i + 1] - p[i]) * inv_dx;
  else if (x == X - 2)
    v[0] -= (p[i] - p[i - 1]) * inv_dx;
  else
    v[0] -= 0.5 * (p[i + 1] - p[i - 1]) * inv_dx;
  if (y == 1)
    v[1] -= (p[i + X] - p[i]) * inv_dx;
  else if (y == Y - 2)
    v[1] -= (p[i] - p[i - X]) * inv_dx;
  else
    v[1] -= 0.5 * (p[i + X] - p[i - X]) * inv_dx;
  if (z == 1)
    v[2] -= (p[i + slab] - p[i]) * inv_dx;
  else if (z == Z - 2)
    v[2] -= (p[i] - p[i - slab]) * inv_dx;
  else
    v[2] -= 0.5 * (p[i + slab] - p[i - slab]) * inv_dx;
}

}  // namespace

struct cfs_fluid {
  int device = 0;
  int X = 0, Y = 0, Z = 0;
  long n = 0;
  cfs_fluid_params prm{};
  // fields (full grid)
  double *vel = nullptr, *vel_old = nullptr, *dens = nullptr, *dens_old = nullptr, *heat = nullptr,
         *heat_old = nullptr;
  double *pres = nullptr, *div = nullptr, *res = nullptr, *dir = nullptr, *zf = nullptr, *qf = nullptr;
  double *w = nullptr, *wmag = nullptr;
  // Poisson matrix and preconditioner
  double *Ad = nullptr, *Api = nullptr, *Apj = nullptr, *Apk = nullptr, *pc = nullptr;
  uint8_t* obst = nullptr;
  int *cells = nullptr, *off = nullptr;
  int nlev = 0;
  double* red = nullptr;  // [2 * RED_BLOCKS] partials + [2] results
  int last_iters = 0;
  double last_max = 0.0;
  std::vector<void*> allocs;
  ~cfs_fluid() {
    cfs::DeviceGuard g(device);
    for (void* p : allocs) (void)hipFree(p);
  }
};

namespace {

template <typename T>
int fl_alloc(cfs_fluid* F, T** p, size_t count) {
  void* q = nullptr;
  const hipError_t e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
  if (e != hipSuccess) return cfs::hip_fail(e, "hipMalloc (fluid)");
  F->allocs.push_back(q);
  *p = static_cast<T*>(q);
  return (int)hipMemset(q, 0, std::max<size_t>(count, 1) * sizeof(T)) == 0 ? CFS_OK
                                                                            : cfs::fail(CFS_ERR_HIP, "hipMemset");
}

// dot / max reduction to the host (synchronous on stream s)
int reduce_to_host(cfs_fluid* F, const double* a, const double* b, const double* m, hipStream_t s, double* dot,
                   double* mx) {
  hipLaunchKernelGGL(reduce1_kernel, dim3(RED_BLOCKS), dim3(TPB), 0, s, a, b, m, F->n, F->red);
  FL_HIP(hipGetLastError(), "reduction launch");
  hipLaunchKernelGGL(reduce2_kernel, dim3(1), dim3(RED_BLOCKS), 0, s, F->red, F->red + 2 * RED_BLOCKS);
  FL_HIP(hipGetLastError(), "reduction launch");
  double out[2];
  FL_HIP(hipMemcpyAsync(out, F->red + 2 * RED_BLOCKS, sizeof(out), hipMemcpyDeviceToHost, s), "reduction copy");
  FL_HIP(hipStreamSynchronize(s), "reduction sync");
  if (dot) *dot = out[0];
  if (mx) *mx = out[1];
  return CFS_OK;
}

// solvePoisson (FLUID_3D_MIC.cpp:918-1017) for the pressure: x = pres, b = div
int solve_poisson(cfs_fluid* F, hipStream_t s) {
  const int X = F->X, Y = F->Y, Z = F->Z;
  const dim3 grid(grid_blocks(F->n)), block(TPB);
  int st;
  hipLaunchKernelGGL(mult_a_kernel, grid, block, 0, s, F->pres, F->res, F->Ad, F->Api, F->Apj, F->Apk, X, Y, Z);
  hipLaunchKernelGGL(residual_kernel, grid, block, 0, s, F->div, F->res, F->obst, X, Y, Z);
  auto precondition = [&]() {
    hipLaunchKernelGGL(mic_sweep_kernel, dim3(1), dim3(SWEEP_TPB), 0, s, F->res, F->qf, F->zf, F->Api, F->Apj, F->Apk,
                       F->pc, F->cells, F->off, F->nlev, X, Y);
  };
  precondition();
  hipLaunchKernelGGL(copy_kernel, grid, block, 0, s, F->zf, F->dir, F->n);
  FL_HIP(hipGetLastError(), "solver launch");
  double delta_new = 0.0;
  if ((st = reduce_to_host(F, F->res, F->zf, nullptr, s, &delta_new, nullptr))) return st;
  const double eps = F->prm.solver_eps;
  double max_r = 2.0 * eps;
  int i = 0;
  while (i < F->prm.iterations && max_r > eps) {
    hipLaunchKernelGGL(mult_a_kernel, grid, block, 0, s, F->dir, F->zf, F->Ad, F->Api, F->Apj, F->Apk, X, Y, Z);
    FL_HIP(hipGetLastError(), "solver launch");
    double alpha = 0.0;
    if ((st = reduce_to_host(F, F->dir, F->zf, nullptr, s, &alpha, nullptr))) return st;
    if (std::fabs(alpha) > 0.0)
      alpha = delta_new / alpha;
    else
      break;  // the reference bails when alpha breaks down
    hipLaunchKernelGGL(cg_update_kernel, grid, block, 0, s, F->pres, F->res, F->dir, F->zf, F->n, alpha);
    precondition();
    FL_HIP(hipGetLastError(), "solver launch");
    // The residual's border is zero already (every update above writes the
    // interior or adds border zeros), so setZeroBorder is the identity here,
    // and the maximum runs over the whole field as FIELD_3D::max does.
    const double old_max = max_r;
    const double delta_old = delta_new;
    if ((st = reduce_to_host(F, F->res, F->zf, F->res, s, &delta_new, &max_r))) return st;
    const double beta = delta_new / delta_old;
    hipLaunchKernelGGL(direction_kernel, grid, block, 0, s, F->dir, F->zf, F->obst, F->n, beta);
    FL_HIP(hipGetLastError(), "solver launch");
    i++;
    if (std::fabs((max_r - old_max) / old_max) < 1e-7) break;
  }
  F->last_iters = i;
  F->last_max = max_r;
  return CFS_OK;
}

}  // namespace

extern "C" {

void cfs_fluid_default_params(int X, int Y, int Z, cfs_fluid_params* prm) {
  if (!prm) return;
  // FLUID_3D_MIC constructor (FLUID_3D_MIC.cpp:29-56, boundaries :100-118)
  const int maxres = std::max(X, std::max(Y, Z));
  prm->dt = 0.1;
  prm->dx = 1.0 / (double)maxres;
  prm->buoyancy = (double)0.1f;
  const double scaling = std::max(1.0, (double)(64.0f / (float)maxres));
  prm->vorticity_eps = 2.0 / scaling;
  prm->damping_w = 0.9;
  prm->solver_eps = 1e-10;
  prm->iterations = 1000;
  const int walls[6] = {1, 1, 1, 1, 0, 0};
  for (int k = 0; k < 6; k++) prm->boundaries[k] = walls[k];
}

int cfs_fluid_create(int X, int Y, int Z, const cfs_fluid_params* prm, cfs_fluid** out) {
  if (!out) return cfs::fail(CFS_ERR_ARG, "null output handle");
  *out = nullptr;
  if (!prm) return cfs::fail(CFS_ERR_ARG, "null parameters");
  if (X < 4 || Y < 4 || Z < 4) return cfs::fail(CFS_ERR_ARG, "the full grid must be at least 4 cells per axis");
  if ((long)X * Y * Z > (1L << 28)) return cfs::fail(CFS_ERR_ARG, "grid too large");
  if (!(prm->dt > 0.0) || !(prm->dx > 0.0) || prm->iterations < 0)
    return cfs::fail(CFS_ERR_ARG, "dt, dx must be positive and iterations non-negative");
  std::unique_ptr<cfs_fluid> F(new cfs_fluid);
  FL_HIP(hipGetDevice(&F->device), "hipGetDevice");
  F->X = X;
  F->Y = Y;
  F->Z = Z;
  F->n = (long)X * Y * Z;
  F->prm = *prm;
  const size_t n = (size_t)F->n;
  int st;
  double** vec[] = {&F->vel, &F->vel_old, &F->w};
  for (double** p : vec)
    if ((st = fl_alloc(F.get(), p, 3 * n))) return st;
  double** sca[] = {&F->dens, &F->dens_old, &F->heat, &F->heat_old, &F->pres, &F->div, &F->res, &F->dir,
                    &F->zf,   &F->qf,       &F->wmag, &F->Ad,       &F->Api,  &F->Apj, &F->Apk, &F->pc};
  for (double** p : sca)
    if ((st = fl_alloc(F.get(), p, n))) return st;
  if ((st = fl_alloc(F.get(), &F->red, 2 * RED_BLOCKS + 2))) return st;

  // obstacles: the domain walls (FLUID_3D_MIC.cpp:120-160); boundaries =
  // front, back (z), left, right (x), top, bottom (y)
  const int slab = X * Y;
  std::vector<uint8_t> obst(n, 0);
  const int* bc = prm->boundaries;
  for (int y = 0; y < Y; y++)
    for (int x = 0; x < X; x++) {
      if (bc[0] == 1) obst[x + y * X] = 1;
      if (bc[1] == 1) obst[x + y * X + n - slab] = 1;
    }
  for (int z = 0; z < Z; z++)
    for (int x = 0; x < X; x++) {
      if (bc[5] == 1) obst[x + z * slab] = 1;
      if (bc[4] == 1) obst[x + z * slab + slab - X] = 1;
    }
  for (int z = 0; z < Z; z++)
    for (int y = 0; y < Y; y++) {
      if (bc[2] == 1) obst[y * X + z * slab] = 1;
      if (bc[3] == 1) obst[y * X + z * slab + X - 1] = 1;
    }

  // buildA (:772-801) with skip = obstacles, heat = false (Aoff = 1)
  std::vector<double> Ad(n, 0.0), Api(n, 0.0), Apj(n, 0.0), Apk(n, 0.0), pc(n, 0.0);
  for (int z = 1; z < Z - 1; z++)
    for (int y = 1; y < Y - 1; y++)
      for (int x = 1; x < X - 1; x++) {
        const size_t i = (size_t)x + (size_t)X * (y + (size_t)Y * z);
        if (obst[i]) continue;
        Ad[i] = 0.0;
        const size_t nb[6] = {i + 1, i - 1, i + X, i - X, i + slab, i - slab};
        for (size_t j : nb)
          if (!obst[j]) Ad[i] += 1.0;
        Api[i - 1] = obst[i - 1] ? 0.0 : -1.0;
        Api[i] = obst[i + 1] ? 0.0 : -1.0;
        Apj[i - X] = obst[i - X] ? 0.0 : -1.0;
        Apj[i] = obst[i + X] ? 0.0 : -1.0;
        Apk[i - slab] = obst[i - slab] ? 0.0 : -1.0;
        Apk[i] = obst[i + slab] ? 0.0 : -1.0;
      }
  // buildMICPreconditioner (:825-862), tau 0.97, sigma 0.25
  const double tau = 0.97, sig = 0.25;
  for (int z = 1; z < Z - 1; z++)
    for (int y = 1; y < Y - 1; y++)
      for (int x = 1; x < X - 1; x++) {
        const size_t i = (size_t)x + (size_t)X * (y + (size_t)Y * z);
        const size_t xm = i - 1, ym = i - X, zm = i - slab;
        const double pi2 = pc[xm] * pc[xm], pj2 = pc[ym] * pc[ym], pk2 = pc[zm] * pc[zm];
        const double e0 = Api[xm] * Api[xm] * pi2 + Apj[ym] * Apj[ym] * pj2 + Apk[zm] * Apk[zm] * pk2;
        const double e1 = tau * (Api[xm] * (Apj[xm] + Apk[xm]) * pi2 + Apj[ym] * (Api[ym] + Apk[ym]) * pj2 +
                                 Apk[zm] * (Api[zm] + Apj[zm]) * pk2);
        double e = Ad[i] - e0 - e1;
        if (e < sig * Ad[i]) e = Ad[i];
        pc[i] = std::fabs(Ad[i]) < 0.000001 ? 0.0 : 1.0 / std::sqrt(e);
      }
  // interior cells by level x + y + z
  const int nlev = (X - 2) + (Y - 2) + (Z - 2) - 2;
  std::vector<int> off(nlev + 1, 0), cells;
  cells.reserve((size_t)(X - 2) * (Y - 2) * (Z - 2));
  for (int l = 0; l < nlev; l++) {
    off[l] = (int)cells.size();
    const int s = l + 3;
    for (int z = 1; z < Z - 1; z++)
      for (int y = 1; y < Y - 1; y++) {
        const int x = s - y - z;
        if (x >= 1 && x < X - 1) cells.push_back(x + X * (y + Y * z));
      }
  }
  off[nlev] = (int)cells.size();
  F->nlev = nlev;
  if ((st = fl_alloc(F.get(), &F->obst, n))) return st;
  if ((st = fl_alloc(F.get(), &F->cells, cells.size()))) return st;
  if ((st = fl_alloc(F.get(), &F->off, off.size()))) return st;
  FL_HIP(hipMemcpy(F->obst, obst.data(), n, hipMemcpyHostToDevice), "upload");
  FL_HIP(hipMemcpy(F->cells, cells.data(), cells.size() * sizeof(int), hipMemcpyHostToDevice), "upload");
  FL_HIP(hipMemcpy(F->off, off.data(), off.size() * sizeof(int), hipMemcpyHostToDevice), "upload");
  FL_HIP(hipMemcpy(F->Ad, Ad.data(), n * 8, hipMemcpyHostToDevice), "upload");
  FL_HIP(hipMemcpy(F->Api, Api.data(), n * 8, hipMemcpyHostToDevice), "upload");
  FL_HIP(hipMemcpy(F->Apj, Apj.data(), n * 8, hipMemcpyHostToDevice), "upload");
  FL_HIP(hipMemcpy(F->Apk, Apk.data(), n * 8, hipMemcpyHostToDevice), "upload");
  FL_HIP(hipMemcpy(F->pc, pc.data(), n * 8, hipMemcpyHostToDevice), "upload");
  *out = F.release();
  return CFS_OK;
}

int cfs_fluid_fields(cfs_fluid* f, double** velocity, double** density, double** heat, double** pressure) {
  if (!f) return cfs::fail(CFS_ERR_ARG, "null fluid");
  if (velocity) *velocity = f->vel;
  if (density) *density = f->dens;
  if (heat) *heat = f->heat;
  if (pressure) *pressure = f->pres;
  return CFS_OK;
}

int cfs_fluid_add_smoke(cfs_fluid* f, void* stream) {
  if (!f) return cfs::fail(CFS_ERR_ARG, "null fluid");
  cfs::DeviceGuard g(f->device);
  const int y0 = (int)(0.05 * f->Y), y1 = (int)(0.10 * f->Y);
  hipLaunchKernelGGL(smoke_kernel<double>, dim3(grid_blocks(f->n)), dim3(TPB), 0, (hipStream_t)stream, f->dens,
                     f->heat, f->X, f->Y, f->Z, y0, y1, 0.0, 0.0);
  FL_HIP(hipGetLastError(), "smoke launch");
  return CFS_OK;
}

}  // extern "C"

namespace {

// FLUID_3D_MIC::step, or with a box stepWithMovingObstacle (:236-310): no
// buoyancy, and the box's cells stomped between the damping and the
// projection (the post-stomp velocity is the IOP snapshot).
int fluid_step(cfs_fluid* f, const cfs_fluid_snapshots* snap, const cfs_box* box, double* iop_snapshot,
               void* stream) {
  cfs::DeviceGuard g(f->device);
  const hipStream_t s = (hipStream_t)stream;
  const int X = f->X, Y = f->Y, Z = f->Z;
  const long n = f->n;
  const dim3 grid(grid_blocks(n)), grid3(grid_blocks(3 * n)), block(TPB);
  const long np = (long)(X - 2) * (Y - 2) * (Z - 2);
  const dim3 gridp(grid_blocks(np));
  const cfs_fluid_params& P = f->prm;
  const int open_x = P.boundaries[2] == 0, open_y = P.boundaries[4] == 0, open_z = P.boundaries[0] == 0;
  const double dt = P.dt, inv_dx = 1.0 / P.dx;
  auto peel = [&](const double* src, double* dst, int comps) {
    if (dst) hipLaunchKernelGGL(peel_kernel, gridp, block, 0, s, src, dst, X, Y, Z, comps);
  };
  // forces: velocity and density borders zeroed, buoyancy (the force is
  // added to every cell), vorticity confinement
  hipLaunchKernelGGL(buoyancy_kernel<double>, grid, block, 0, s, f->vel, f->heat, X, Y, Z, dt,
                     box ? 0.0 : P.buoyancy);
  hipLaunchKernelGGL(zero_border2_kernel<double>, grid, block, 0, s, f->dens, (double*)nullptr, X, Y, Z);
  if (P.vorticity_eps != 0.0) {
    hipLaunchKernelGGL(curl_kernel<double>, grid, block, 0, s, f->vel, f->w, f->wmag, X, Y, Z, inv_dx);
    hipLaunchKernelGGL(confinement_kernel<double>, grid, block, 0, s, f->vel, f->w, f->wmag, X, Y, Z, inv_dx, dt,
                       P.vorticity_eps * P.dx);
  }
  FL_HIP(hipGetLastError(), "force launch");
  // advectStam
  hipLaunchKernelGGL(velocity_border_kernel<double>, grid, block, 0, s, f->vel, X, Y, Z, open_x, open_y, open_z, 1);
  if (snap) peel(f->vel, snap->preadvect, 3);
  hipLaunchKernelGGL(copy_kernel, grid3, block, 0, s, f->vel, f->vel_old, 3 * n);
  hipLaunchKernelGGL(copy_kernel, grid, block, 0, s, f->dens, f->dens_old, n);
  hipLaunchKernelGGL(copy_kernel, grid, block, 0, s, f->heat, f->heat_old, n);
  const double dt0 = dt / P.dx;
  hipLaunchKernelGGL(advect_scalar_field_kernel<double>, grid, block, 0, s, f->vel_old, f->dens_old, f->dens, X, Y, Z,
                     dt0);
  hipLaunchKernelGGL(advect_scalar_field_kernel<double>, grid, block, 0, s, f->vel_old, f->heat_old, f->heat, X, Y, Z,
                     dt0);
  hipLaunchKernelGGL(advect_vector_kernel<double>, grid, block, 0, s, f->vel_old, f->vel, X, Y, Z, dt0);
  hipLaunchKernelGGL(zero_border_vector_kernel<double>, grid, block, 0, s, f->vel, X, Y, Z);
  hipLaunchKernelGGL(velocity_border_kernel<double>, grid, block, 0, s, f->vel, X, Y, Z, open_x, open_y, open_z, 1);
  hipLaunchKernelGGL(zero_border2_kernel<double>, grid, block, 0, s, f->dens, f->heat, X, Y, Z);
  FL_HIP(hipGetLastError(), "advection launch");
  // diffusion: the peeled damping matrix
  if (snap) peel(f->vel, snap->prediffuse, 3);
  hipLaunchKernelGGL(damping_kernel, grid, block, 0, s, f->vel, f->vel_old, X, Y, Z, P.damping_w);
  hipLaunchKernelGGL(copy_interior_kernel, grid, block, 0, s, f->vel_old, f->vel, X, Y, Z);
  FL_HIP(hipGetLastError(), "damping launch");
  if (snap) peel(f->vel, snap->preproject, 3);
  if (box) {
    FL_HIP(cfs::launch_box_stomp(f->vel, cfs::box_state(*box, X, Y, Z), s), "box stomp launch");
    peel(f->vel, iop_snapshot, 3);
  }
  // project
  hipLaunchKernelGGL(velocity_border_kernel<double>, grid, block, 0, s, f->vel, X, Y, Z, open_x, open_y, open_z, 2);
  hipLaunchKernelGGL(divergence_kernel, grid, block, 0, s, f->vel, f->obst, f->div, X, Y, Z, P.dx);
  // the pressure's interior is reset to 0 and copyBorderAll copies zeros
  // onto every border cell the solve reads
  hipLaunchKernelGGL(fill_kernel, grid, block, 0, s, f->pres, n, 0.0);
  FL_HIP(hipGetLastError(), "projection launch");
  int st = solve_poisson(f, s);
  if (st) return st;
  hipLaunchKernelGGL(gradient_kernel, grid, block, 0, s, f->vel, f->pres, X, Y, Z, inv_dx);
  FL_HIP(hipGetLastError(), "projection launch");
  if (snap) {
    peel(f->vel, snap->final_velocity, 3);
    peel(f->pres, snap->pressure, 1);
    peel(f->div, snap->divergence, 1);
    FL_HIP(hipGetLastError(), "snapshot launch");
  }
  return CFS_OK;
}

}  // namespace

extern "C" {

int cfs_fluid_step(cfs_fluid* f, const cfs_fluid_snapshots* snap, void* stream) {
  if (!f) return cfs::fail(CFS_ERR_ARG, "null fluid");
  return fluid_step(f, snap, nullptr, nullptr, stream);
}

int cfs_fluid_step_obstacle(cfs_fluid* f, const cfs_box* box, const cfs_fluid_snapshots* snap, double* iop_snapshot,
                            void* stream) {
  if (!f || !box) return cfs::fail(CFS_ERR_ARG, "null argument");
  return fluid_step(f, snap, box, iop_snapshot, stream);
}

int cfs_fluid_set_initial_velocity(cfs_fluid* f, const cfs_box* box, void* stream) {
  if (!f || !box) return cfs::fail(CFS_ERR_ARG, "null argument");
  cfs::DeviceGuard g(f->device);
  FL_HIP(cfs::launch_box_stomp(f->vel, cfs::box_state(*box, f->X, f->Y, f->Z), (hipStream_t)stream),
         "box stomp launch");
  return CFS_OK;
}

int cfs_fluid_last_solve(const cfs_fluid* f, int* iterations, double* residual_max) {
  if (!f) return cfs::fail(CFS_ERR_ARG, "null fluid");
  if (iterations) *iterations = f->last_iters;
  if (residual_max) *residual_max = f->last_max;
  return CFS_OK;
}

int cfs_fluid_destroy(cfs_fluid* f) {
  if (!f) return cfs::fail(CFS_ERR_ARG, "null fluid");
  delete f;
  return CFS_OK;
}

}  // extern "C"

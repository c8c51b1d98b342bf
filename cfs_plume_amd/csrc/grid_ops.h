// grid_ops.h -- per-cell stencils of the Stam plume on full grids, for any
// scalar type: FP32 in the reduced step (plume.hip), FP64 in the full-rank
// solver (fluid.hip), whose fields are the reference's `Real` (double,
// SETTINGS.h).
//
// Full grid X x Y x Z, x fastest; vector fields interleaved xyz.  Every
// kernel takes one thread per cell.  The kernels live in an anonymous
// namespace: each translation unit instantiates its own.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

namespace cfs {
namespace {

constexpr int GRID_TPB = 256;

__device__ __forceinline__ float tmin(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ double tmin(double a, double b) { return fmin(a, b); }
__device__ __forceinline__ float tmax(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ double tmax(double a, double b) { return fmax(a, b); }
__device__ __forceinline__ float tsqrt(float a) { return sqrtf(a); }
__device__ __forceinline__ double tsqrt(double a) { return sqrt(a); }

__device__ __forceinline__ bool on_border(int x, int y, int z, int X, int Y, int Z) {
  return x == 0 || y == 0 || z == 0 || x == X - 1 || y == Y - 1 || z == Z - 1;
}

// setZeroBorder of the velocity, then velocity += dt * (buoyancy * heat) e_y on
// every cell (addBuoyancy, FLUID_3D_MIC.cpp:491-502, adds to the force of
// every cell; the step adds dt * force).
template <typename T>
__global__ void __launch_bounds__(GRID_TPB) buoyancy_kernel(T* __restrict__ vel, const T* __restrict__ heat, int X,
                                                            int Y, int Z, T dt, T beta) {
  const int i = blockIdx.x * GRID_TPB + threadIdx.x;
  if (i >= X * Y * Z) return;
  const int x = i % X, y = (i / X) % Y, z = i / (X * Y);
  T* v = vel + 3 * (size_t)i;
  if (on_border(x, y, z, X, Y, Z)) {
    v[0] = T(0);
    v[1] = T(0);
    v[2] = T(0);
  }
  v[1] += dt * (beta * heat[i]);
}

// Neighbours of an interior cell for the vorticity stencils (addVorticity,
// FLUID_3D_MIC.cpp:524-593): the cell itself in place of a neighbour on the
// border layer, and the matching inverse spacings (one-sided 1/dx, central
// 0.5/dx).
template <typename T>
struct Stencil {
  int up, down, out, in, right, left;
  T dy, dz, dx;
};
template <typename T>
__device__ __forceinline__ Stencil<T> stencil(int i, int x, int y, int z, int X, int Y, int Z, T inv_dx) {
  Stencil<T> s;
  const int slab = X * Y;
  const T grid = T(0.5) * inv_dx;
  s.up = y == Y - 2 ? i : i + X;
  s.down = y == 1 ? i : i - X;
  s.dy = (s.up == i || s.down == i) ? inv_dx : grid;
  s.out = z == Z - 2 ? i : i + slab;
  s.in = z == 1 ? i : i - slab;
  s.dz = (s.out == i || s.in == i) ? inv_dx : grid;
  s.right = x == X - 2 ? i : i + 1;
  s.left = x == 1 ? i : i - 1;
  s.dx = (s.left == i || s.right == i) ? inv_dx : grid;
  return s;
}

// the curl w of the velocity and |w| at interior cells (border entries stay 0)
template <typename T>
__global__ void __launch_bounds__(GRID_TPB) curl_kernel(const T* __restrict__ vel, T* __restrict__ w,
                                                        T* __restrict__ wmag, int X, int Y, int Z, T inv_dx) {
  const int i = blockIdx.x * GRID_TPB + threadIdx.x;
  if (i >= X * Y * Z) return;
  const int x = i % X, y = (i / X) % Y, z = i / (X * Y);
  if (on_border(x, y, z, X, Y, Z)) return;
  const Stencil<T> s = stencil<T>(i, x, y, z, X, Y, Z, inv_dx);
  const T* v = vel;
  const T w0 = (v[3 * s.up + 2] - v[3 * s.down + 2]) * s.dy + (-v[3 * s.out + 1] + v[3 * s.in + 1]) * s.dz;
  const T w1 = (v[3 * s.out + 0] - v[3 * s.in + 0]) * s.dz + (-v[3 * s.right + 2] + v[3 * s.left + 2]) * s.dx;
  const T w2 = (v[3 * s.right + 1] - v[3 * s.left + 1]) * s.dx + (-v[3 * s.up + 0] + v[3 * s.down + 0]) * s.dy;
  w[3 * (size_t)i] = w0;
  w[3 * (size_t)i + 1] = w1;
  w[3 * (size_t)i + 2] = w2;
  wmag[i] = tsqrt(w0 * w0 + w1 * w1 + w2 * w2);
}

// velocity += dt * eps * dx * (N x w) at interior cells, N = grad|w| / |grad|w||
template <typename T>
__global__ void __launch_bounds__(GRID_TPB) confinement_kernel(T* __restrict__ vel, const T* __restrict__ w,
                                                               const T* __restrict__ wmag, int X, int Y, int Z,
                                                               T inv_dx, T dt, T eps_dx) {
  const int i = blockIdx.x * GRID_TPB + threadIdx.x;
  if (i >= X * Y * Z) return;
  const int x = i % X, y = (i / X) % Y, z = i / (X * Y);
  if (on_border(x, y, z, X, Y, Z)) return;
  const Stencil<T> s = stencil<T>(i, x, y, z, X, Y, Z, inv_dx);
  T n0 = (wmag[s.right] - wmag[s.left]) * s.dx;
  T n1 = (wmag[s.up] - wmag[s.down]) * s.dy;
  T n2 = (wmag[s.out] - wmag[s.in]) * s.dz;
  T m = tsqrt(n0 * n0 + n1 * n1 + n2 * n2);
  if (m > T(0)) {
    m = T(1) / m;
    n0 *= m;
    n1 *= m;
    n2 *= m;
    const T w0 = w[3 * (size_t)i], w1 = w[3 * (size_t)i + 1], w2 = w[3 * (size_t)i + 2];
    T* v = vel + 3 * (size_t)i;
    v[0] += dt * ((n1 * w2 - n2 * w1) * eps_dx);
    v[1] += dt * (-(n0 * w2 - n2 * w0) * eps_dx);
    v[2] += dt * ((n0 * w1 - n1 * w0) * eps_dx);
  }
}

// The velocity border per axis: a wall zeroes the normal component on its
// two faces (setZeroX/Y/Z), an open side copies it from the layer `reach`
// cells inside: reach 1 is copyBorderX/Y/Z (the advection's border,
// VECTOR3_FIELD_3D.cpp:2872-2927), reach 2 is setNeumannX/Y/Z (the
// projection's, :2929-3005).  Each axis's rule writes and reads only its
// own component (x: component 0 ...), and reads it on a layer the rule does
// not write, so one launch gives what the reference's sequential per-axis
// loops give.
template <typename T>
__global__ void __launch_bounds__(GRID_TPB) velocity_border_kernel(T* __restrict__ vel, int X, int Y, int Z,
                                                                   int open_x, int open_y, int open_z, int reach) {
  const int i = blockIdx.x * GRID_TPB + threadIdx.x;
  if (i >= X * Y * Z) return;
  const int x = i % X, y = (i / X) % Y, z = i / (X * Y);
  T* v = vel + 3 * (size_t)i;
  if (x == 0 || x == X - 1) v[0] = open_x ? vel[3 * (size_t)(x == 0 ? i + reach : i - reach)] : T(0);
  if (y == 0 || y == Y - 1) v[1] = open_y ? vel[3 * (size_t)(y == 0 ? i + reach * X : i - reach * X) + 1] : T(0);
  if (z == 0 || z == Z - 1) {
    const int slab = X * Y;
    v[2] = open_z ? vel[3 * (size_t)(z == 0 ? i + reach * slab : i - reach * slab) + 2] : T(0);
  }
}

// VECTOR3_FIELD_3D::setZeroBorder
template <typename T>
__global__ void __launch_bounds__(GRID_TPB) zero_border_vector_kernel(T* __restrict__ vel, int X, int Y, int Z) {
  const int i = blockIdx.x * GRID_TPB + threadIdx.x;
  if (i >= X * Y * Z) return;
  const int x = i % X, y = (i / X) % Y, z = i / (X * Y);
  if (on_border(x, y, z, X, Y, Z)) {
    vel[3 * (size_t)i] = T(0);
    vel[3 * (size_t)i + 1] = T(0);
    vel[3 * (size_t)i + 2] = T(0);
  }
}

// FIELD_3D::setZeroBorder of two scalar fields (b may be null)
template <typename T>
__global__ void __launch_bounds__(GRID_TPB) zero_border2_kernel(T* __restrict__ a, T* __restrict__ b, int X, int Y,
                                                                int Z) {
  const int i = blockIdx.x * GRID_TPB + threadIdx.x;
  if (i >= X * Y * Z) return;
  const int x = i % X, y = (i / X) % Y, z = i / (X * Y);
  if (on_border(x, y, z, X, Y, Z)) {
    a[i] = T(0);
    if (b) b[i] = T(0);
  }
}

// addSmokeTestCase (FLUID_3D_MIC.cpp:674-703) on density and heat: 1 inside
// the column of radius 0.075 xTotal around (0.4 xTotal, 0.5 zTotal) for
// y in [(int)(0.05 Y), (int)(0.10 Y)], lengths in units of dx = 1 / maxRes.
// (ox, oz): a shift of the column's axis in cells (0 for the reference's
// source; the seeded training runs of the learned closure move it).
template <typename T>
__global__ void __launch_bounds__(GRID_TPB) smoke_kernel(T* __restrict__ dens, T* __restrict__ heat, int X, int Y,
                                                         int Z, int y0, int y1, double ox, double oz) {
  const int i = blockIdx.x * GRID_TPB + threadIdx.x;
  if (i >= X * Y * Z) return;
  const int x = i % X, y = (i / X) % Y, z = i / (X * Y);
  if (y < y0 || y > y1) return;
  const int maxres = max(X, max(Y, Z));
  const double dx = 1.0 / (double)(float)maxres;
  const double xtot = dx * X, ztot = dx * Z;
  const double xl = x * dx - xtot * (double)0.4f - ox * dx;
  const double zl = z * dx - ztot * (double)0.5f - oz * dx;
  const float radius = sqrtf((float)(xl * xl + zl * zl));
  if ((double)radius < (double)0.075f * xtot) {
    dens[i] = T(1);
    heat[i] = T(1);
  }
}

// One cell of VECTOR3_FIELD_3D::advect (VECTOR3_FIELD_3D.cpp:2001-2160):
// backtrace by dt0 * v, clamp to [1.5, res - 2.5] per axis, and the trilinear
// weights of the backtraced point; returns the corner index i000.
template <typename T>
struct Backtrace {
  int i000;
  T s0, s1, t0, t1, u0, u1;
};
template <typename T>
__device__ __forceinline__ Backtrace<T> backtrace(const T* __restrict__ vel, int idx, int X, int Y, int Z, T dt0) {
  const int x = idx % X, y = (idx / X) % Y, z = idx / (X * Y);
  T xt = (T)x - dt0 * vel[3 * (size_t)idx];
  T yt = (T)y - dt0 * vel[3 * (size_t)idx + 1];
  T zt = (T)z - dt0 * vel[3 * (size_t)idx + 2];
  xt = tmin(tmax(xt, T(1.5)), (T)X - T(2.5));
  yt = tmin(tmax(yt, T(1.5)), (T)Y - T(2.5));
  zt = tmin(tmax(zt, T(1.5)), (T)Z - T(2.5));
  const int x0 = (int)xt, y0 = (int)yt, z0 = (int)zt;
  Backtrace<T> b;
  b.s1 = xt - (T)x0;
  b.s0 = T(1) - b.s1;
  b.t1 = yt - (T)y0;
  b.t0 = T(1) - b.t1;
  b.u1 = zt - (T)z0;
  b.u0 = T(1) - b.u1;
  b.i000 = x0 + y0 * X + z0 * X * Y;
  return b;
}

// trilinear sample of the field f (element stride `st`) at a backtrace
template <typename T>
__device__ __forceinline__ T trilinear(const T* __restrict__ f, const Backtrace<T>& b, int X, int slab, int st) {
  const size_t i000 = (size_t)b.i000;
  const T* p = f;
  return b.u0 * (b.s0 * (b.t0 * p[st * i000] + b.t1 * p[st * (i000 + X)]) +
                 b.s1 * (b.t0 * p[st * (i000 + 1)] + b.t1 * p[st * (i000 + X + 1)])) +
         b.u1 * (b.s0 * (b.t0 * p[st * (i000 + slab)] + b.t1 * p[st * (i000 + slab + X)]) +
                 b.s1 * (b.t0 * p[st * (i000 + slab + 1)] + b.t1 * p[st * (i000 + slab + X + 1)]));
}

// out = advect(vel, vel): the semi-Lagrangian self-advection of a vector field
template <typename T>
__global__ void __launch_bounds__(GRID_TPB) advect_vector_kernel(const T* __restrict__ vel, T* __restrict__ out,
                                                                 int X, int Y, int Z, T dt0) {
  const int idx = blockIdx.x * GRID_TPB + threadIdx.x;
  if (idx >= X * Y * Z) return;
  const Backtrace<T> b = backtrace<T>(vel, idx, X, Y, Z, dt0);
  const int slab = X * Y;
#pragma unroll
  for (int c = 0; c < 3; c++) out[3 * (size_t)idx + c] = trilinear<T>(vel + c, b, X, slab, 3);
}

// out = advect(vel, old) of a scalar field (density, heat)
template <typename T>
__global__ void __launch_bounds__(GRID_TPB) advect_scalar_field_kernel(const T* __restrict__ vel,
                                                                       const T* __restrict__ old, T* __restrict__ out,
                                                                       int X, int Y, int Z, T dt0) {
  const int idx = blockIdx.x * GRID_TPB + threadIdx.x;
  if (idx >= X * Y * Z) return;
  const Backtrace<T> b = backtrace<T>(vel, idx, X, Y, Z, dt0);
  out[idx] = trilinear<T>(old, b, X, X * Y, 1);
}

inline int grid_blocks(long n) { return (int)((n + GRID_TPB - 1) / GRID_TPB); }

}  // namespace
}  // namespace cfs

// obstacle_internal.h -- the moving box obstacle shared by the full-rank
// solver (fluid.hip) and the reduced plume (plume.hip); defined in
// obstacle.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/cfs_plume.h"

namespace cfs {

// A box at one instant, as the kernels test cells against it: centre c,
// rotation R (row-major, point' = R (point - c)), half edge lengths h, the
// spin rate omega about z and the sliding velocity along x; the full grid
// X x Y x Z whose cell centres ((x + 0.5) / X, ...) are tested.
struct BoxState {
  double c[3];
  double R[9];
  double h[3];
  double omega, slide;
  int X, Y, Z;
  // full-grid cells [x0, x0 + nx) x ... that can hold a box cell (interior only)
  int x0, y0, z0, nx, ny, nz;
};

BoxState box_state(const cfs_box& b, int X, int Y, int Z);

// Interior cells of a full-grid FP64 velocity field inside the box take the
// box's velocity there.
hipError_t launch_box_stomp(double* vel, const BoxState& s, hipStream_t st);

// q_out = M q + sum over box cells c of U_iop[c]^T (v_box(c) - U_proj[c] q):
// the reduced IOP with U (3 Npeel x r, column-major FP64) and M = U_iop^T U_proj
// (r_iop x r_proj, row-major).  w: 3 nb doubles, flags: nb bytes, nb the
// number of cells of the box's span.
hipError_t launch_reduced_iop(const double* U_iop, int r_iop, const double* U_proj, int r_proj, const double* M,
                              const BoxState& s, const double* q, double* q_out, double* w, uint8_t* flags,
                              hipStream_t st);

__host__ __device__ inline int64_t box_span_cells(const BoxState& s) { return (int64_t)s.nx * s.ny * s.nz; }

}  // namespace cfs

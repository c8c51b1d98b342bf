// block_dct.h -- shared tables and the forward 8x8x8 block DCT (VALU FP32);
// the inverse DCT runs on MFMA (mfma_idct.h).
//
// Orthonormal DCT-II / DCT-III along x, y, z, the closed form of the
// reference's FFTW REDFT10/REDFT01 plus unitary normalisation
// (COMPRESSION.cpp:173-272, 511-555).
#pragma once
#include <hip/hip_runtime.h>

#include "cfs_device.h"
#include "wave_ops.h"

namespace cfs {

// Per-lane constant operands of the MFMA inverse DCT (decode_kernels.hip),
// as BF16 bit patterns split hi + lo (value = hi + lo to ~2^-17):
//   s1b[h][lane][i]: B operand of the x pass, blockdiag(C, C) (16 x 16),
//                    element (k = 4*(lane/16) + i, n = lane % 16);
//   s2a[h][t][ks][lane][i]: A operand of the joint (y, z) pass, rows
//                    (z, y) = 16 t + lane % 16, K index 32 ks + 8*(lane/16) + i.
struct MfmaFrags {
  uint16_t s1b[2][64][4];
  uint16_t s2a[2][4][2][64][8];
};

// Small read-only tables, shared by every basis on the device.
struct Tables {
  float dct[64];        // C[k][n] = a_k cos(pi (2n+1) k / 16), row k
  uint16_t zig[BC];     // zig[x + 8y + 64z] = zigzag position (COMPRESSION_DATA.cpp:90-120)
  uint16_t rev[BC];     // rev[j] = x + 8y + 64z of zigzag position j
  uint8_t dz[BC];       // dz[j] = u+v+w of the cell at zigzag position j
  MfmaFrags mf;
};

// Orthonormal 8-point DCT constants c_m = a cos(m pi / 16), rounded to FP32
// (c4 also equals the k = 0 normalisation sqrt(1/8)).  Compile-time
// constants become instruction literals: no scalar registers are held.
constexpr float DC1 = 0.490392625f, DC2 = 0.461939752f, DC3 = 0.415734798f, DC4 = 0.353553385f,
                DC5 = 0.277785122f, DC6 = 0.191341713f, DC7 = 0.0975451618f;

// X = C a (DCT-II) by even/odd butterflies: C[k][7-n] = (-1)^k C[k][n].
__device__ __forceinline__ void dct8_fwd(const float a[BS], float X[BS]) {
  const float s0 = a[0] + a[7], s1 = a[1] + a[6], s2 = a[2] + a[5], s3 = a[3] + a[4];
  const float d0 = a[0] - a[7], d1 = a[1] - a[6], d2 = a[2] - a[5], d3 = a[3] - a[4];
  const float ss0 = s0 + s3, ss1 = s1 + s2, sd0 = s0 - s3, sd1 = s1 - s2;
  X[0] = DC4 * (ss0 + ss1);
  X[4] = DC4 * (ss0 - ss1);
  X[2] = DC2 * sd0 + DC6 * sd1;
  X[6] = DC6 * sd0 - DC2 * sd1;
  X[1] = DC1 * d0 + DC3 * d1 + DC5 * d2 + DC7 * d3;
  X[3] = DC3 * d0 - DC7 * d1 - DC1 * d2 - DC5 * d3;
  X[5] = DC5 * d0 - DC1 * d1 + DC7 * d2 + DC3 * d3;
  X[7] = DC7 * d0 - DC5 * d1 + DC3 * d2 - DC1 * d3;
}

// a = C^T X (DCT-III, the inverse of dct8_fwd) by the same butterflies.
__device__ __forceinline__ void dct8_inv(const float X[BS], float a[BS]) {
  const float E0 = DC4 * (X[0] + X[4]), E1 = DC4 * (X[0] - X[4]);
  const float P = DC2 * X[2] + DC6 * X[6], Q = DC6 * X[2] - DC2 * X[6];
  const float e0 = E0 + P, e3 = E0 - P, e1 = E1 + Q, e2 = E1 - Q;
  const float o0 = DC1 * X[1] + DC3 * X[3] + DC5 * X[5] + DC7 * X[7];
  const float o1 = DC3 * X[1] - DC7 * X[3] - DC1 * X[5] - DC5 * X[7];
  const float o2 = DC5 * X[1] - DC1 * X[3] + DC7 * X[5] + DC3 * X[7];
  const float o3 = DC7 * X[1] - DC5 * X[3] + DC3 * X[5] - DC1 * X[7];
  a[0] = e0 + o0;
  a[7] = e0 - o0;
  a[1] = e1 + o1;
  a[6] = e1 - o1;
  a[2] = e2 + o2;
  a[5] = e2 - o2;
  a[3] = e3 + o3;
  a[4] = e3 - o3;
}

// Offset (in floats) of component k of full-grid cell (gx+1, gy+1, gz+1)
// of the interleaved field (VECTOR3_FIELD_3D.cpp:69-85).  Full grids here
// hold fewer than 2^31 floats (the loader checks).
__device__ __forceinline__ int field_offset(const DevBasis& B, int gx, int gy, int gz, int k) {
  return 3 * ((gx + 1) + (B.X + 2) * ((gy + 1) + (B.Y + 2) * (gz + 1))) + k;
}

// Lane (y, z) gathers row (y, z) of the zero-padded peeled block b of
// component k from the full-grid interleaved field v (GetBlocksEigen,
// COMPRESSION.cpp:410-451).
__device__ __forceinline__ void gather_block_rows(const float* __restrict__ v, const DevBasis& B, int k, int b,
                                                  float a[BS]) {
  const int lane = lane_id();
  const int bx = b % B.nbx, by = (b / B.nbx) % B.nby, bz = b / (B.nbx * B.nby);
  const int gy = by * BS + (lane & 7), gz = bz * BS + (lane >> 3);
  const bool row_in = gy < B.Y && gz < B.Z;
  const float* row = v + field_offset(B, bx * BS, gy, gz, k);
#pragma unroll
  for (int x = 0; x < BS; x++) a[x] = (row_in && bx * BS + x < B.X) ? row[3 * x] : 0.0f;
}

// Forward orthonormal 3D DCT (UnitaryBlockDCTEigen(1), COMPRESSION.cpp:536-555)
// of a block whose rows were gathered by gather_block_rows, written to
// vz[STRIDE * j] for zigzag position j.  `scr` is a 512-float LDS scratch.
template <int STRIDE = 1>
__device__ void forward_dct_rows(const float a0[BS], const uint16_t* zig, float* scr, float* vz) {
  const int lane = lane_id();
  float X[BS], a[BS];
  // pass along x: lane = (y, z)
  dct8_fwd(a0, X);
  {
    const int y = lane & 7, z = lane >> 3;
#pragma unroll
    for (int u = 0; u < BS; u++) scr[z * 64 + y * 8 + u] = X[u];
  }
  wave_lds_sync();
  // pass along y: lane = (u, z)
  {
    const int u = lane & 7, z = lane >> 3;
#pragma unroll
    for (int y = 0; y < BS; y++) a[y] = scr[z * 64 + y * 8 + u];
    wave_lds_sync();
    dct8_fwd(a, X);
#pragma unroll
    for (int vv = 0; vv < BS; vv++) scr[z * 64 + vv * 8 + u] = X[vv];
  }
  wave_lds_sync();
  // pass along z: lane = (u, v); scatter to zigzag order
  {
    const int u = lane & 7, vv = lane >> 3;
#pragma unroll
    for (int z = 0; z < BS; z++) a[z] = scr[z * 64 + vv * 8 + u];
    dct8_fwd(a, X);
#pragma unroll
    for (int w = 0; w < BS; w++) vz[STRIDE * zig[u + 8 * vv + 64 * w]] = X[w];
  }
  wave_lds_sync();
}

// Gather + forward DCT of block b of component k, written to vz in zigzag order.
template <int STRIDE = 1>
__device__ void forward_dct_block(const float* __restrict__ v, const DevBasis& B, int k, int b, const uint16_t* zig,
                                  float* scr, float* vz) {
  float a[BS];
  gather_block_rows(v, B, k, b, a);
  forward_dct_rows<STRIDE>(a, zig, scr, vz);
}

}  // namespace cfs

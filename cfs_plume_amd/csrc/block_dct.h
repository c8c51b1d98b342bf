// block_dct.h -- shared tables and the 8x8x8 block DCT helpers (VALU FP32).
//
// Orthonormal DCT-II / DCT-III along x, y, z, the closed form of the
// reference's FFTW REDFT10/REDFT01 plus unitary normalisation
// (COMPRESSION.cpp:173-272, 511-555).
#pragma once
#include <hip/hip_runtime.h>

#include "cfs_device.h"
#include "wave_ops.h"

namespace cfs {

// Small read-only tables, shared by every basis on the device.
struct Tables {
  float dct[64];        // C[k][n] = a_k cos(pi (2n+1) k / 16), row k
  uint16_t zig[BC];     // zig[x + 8y + 64z] = zigzag position (COMPRESSION_DATA.cpp:90-120)
  uint16_t rev[BC];     // rev[j] = x + 8y + 64z of zigzag position j
  uint8_t dz[BC];       // dz[j] = u+v+w of the cell at zigzag position j
};

// Lane (y, z) gathers row (y, z) of the zero-padded peeled block b of
// component k from the full-grid interleaved field v (GetBlocksEigen,
// COMPRESSION.cpp:410-451).
__device__ __forceinline__ void gather_block_rows(const float* __restrict__ v, const DevBasis& B, int k, int b,
                                                  float a[BS]) {
  const int lane = lane_id();
  const int bx = b % B.nbx, by = (b / B.nbx) % B.nby, bz = b / (B.nbx * B.nby);
  const int GX = B.X + 2, GY = B.Y + 2;
  const int y = lane & 7, z = lane >> 3;
  const int gy = by * BS + y, gz = bz * BS + z;
#pragma unroll
  for (int x = 0; x < BS; x++) {
    const int gx = bx * BS + x;
    a[x] = (gx < B.X && gy < B.Y && gz < B.Z)
               ? v[3 * ((size_t)(gx + 1) + (size_t)GX * ((gy + 1) + (size_t)GY * (gz + 1))) + k]
               : 0.0f;
  }
}

// Forward orthonormal 3D DCT (UnitaryBlockDCTEigen(1), COMPRESSION.cpp:536-555)
// of a block whose rows were gathered by gather_block_rows, written to
// vz[STRIDE * j] for zigzag position j.  `scr` is a 512-float LDS scratch.
template <int STRIDE = 1>
__device__ void forward_dct_rows(const float a0[BS], const float* __restrict__ C, const uint16_t* zig, float* scr,
                                 float* vz) {
  const int lane = lane_id();
  // pass along x: lane = (y, z)
  {
    const int y = lane & 7, z = lane >> 3;
#pragma unroll
    for (int u = 0; u < BS; u++) {
      float s = 0.0f;
#pragma unroll
      for (int x = 0; x < BS; x++) s += C[u * BS + x] * a0[x];
      scr[z * 64 + y * 8 + u] = s;
    }
  }
  wave_lds_sync();
  // pass along y: lane = (u, z)
  {
    const int u = lane & 7, z = lane >> 3;
    float a[BS];
#pragma unroll
    for (int y = 0; y < BS; y++) a[y] = scr[z * 64 + y * 8 + u];
    wave_lds_sync();
#pragma unroll
    for (int vv = 0; vv < BS; vv++) {
      float s = 0.0f;
#pragma unroll
      for (int y = 0; y < BS; y++) s += C[vv * BS + y] * a[y];
      scr[z * 64 + vv * 8 + u] = s;
    }
  }
  wave_lds_sync();
  // pass along z: lane = (u, v); scatter to zigzag order
  {
    const int u = lane & 7, vv = lane >> 3;
    float a[BS];
#pragma unroll
    for (int z = 0; z < BS; z++) a[z] = scr[z * 64 + vv * 8 + u];
#pragma unroll
    for (int w = 0; w < BS; w++) {
      float s = 0.0f;
#pragma unroll
      for (int z = 0; z < BS; z++) s += C[w * BS + z] * a[z];
      vz[STRIDE * zig[u + 8 * vv + 64 * w]] = s;
    }
  }
  wave_lds_sync();
}

// Gather + forward DCT of block b of component k, written to vz in zigzag order.
template <int STRIDE = 1>
__device__ void forward_dct_block(const float* __restrict__ v, const DevBasis& B, int k, int b,
                                  const float* __restrict__ C, const uint16_t* zig, float* scr, float* vz) {
  float a[BS];
  gather_block_rows(v, B, k, b, a);
  forward_dct_rows<STRIDE>(a, C, zig, scr, vz);
}

// Inverse orthonormal 3D DCT of the zigzag-ordered fixed-point accumulator
// `az` (value = az * unit) and
// scatter of the interior cells of component k into the full-grid field
// (UnitaryBlockDCTEigen(-1) + AssimilateBlocksEigen + subfield + setWithPeeled,
// COMPRESSION.cpp:2880-2898, VECTOR3_FIELD_3D.cpp:3415-3430).  Leaves az zero.
__device__ void inverse_dct_block(float* __restrict__ field, const DevBasis& B, int k, int b,
                                  const float* __restrict__ C, const uint16_t* zig, int* az, float unit, float* scr) {
  const int lane = lane_id();
  // gather + pass along z: lane = (u, v)
  {
    const int u = lane & 7, vv = lane >> 3;
    float a[BS];
#pragma unroll
    for (int w = 0; w < BS; w++) a[w] = (float)az[zig[u + 8 * vv + 64 * w]] * unit;
#pragma unroll
    for (int z = 0; z < BS; z++) {
      float s = 0.0f;
#pragma unroll
      for (int w = 0; w < BS; w++) s += C[w * BS + z] * a[w];
      scr[z * 64 + vv * 8 + u] = s;
    }
  }
  wave_lds_sync();
  for (int i = lane; i < BC; i += WAVE) az[i] = 0;
  // pass along y: lane = (u, z)
  {
    const int u = lane & 7, z = lane >> 3;
    float a[BS];
#pragma unroll
    for (int vv = 0; vv < BS; vv++) a[vv] = scr[z * 64 + vv * 8 + u];
    wave_lds_sync();
#pragma unroll
    for (int y = 0; y < BS; y++) {
      float s = 0.0f;
#pragma unroll
      for (int vv = 0; vv < BS; vv++) s += C[vv * BS + y] * a[vv];
      scr[z * 64 + y * 8 + u] = s;
    }
  }
  wave_lds_sync();
  // pass along x: lane = (y, z); write interior cells
  {
    const int y = lane & 7, z = lane >> 3;
    const int bx = b % B.nbx, by = (b / B.nbx) % B.nby, bz = b / (B.nbx * B.nby);
    const int gy = by * BS + y, gz = bz * BS + z;
    const int GX = B.X + 2, GY = B.Y + 2;
    float a[BS];
#pragma unroll
    for (int u = 0; u < BS; u++) a[u] = scr[z * 64 + y * 8 + u];
#pragma unroll
    for (int x = 0; x < BS; x++) {
      float s = 0.0f;
#pragma unroll
      for (int u = 0; u < BS; u++) s += C[u * BS + x] * a[u];
      const int gx = bx * BS + x;
      if (gx < B.X && gy < B.Y && gz < B.Z)
        field[3 * ((size_t)(gx + 1) + (size_t)GX * ((gy + 1) + (size_t)GY * (gz + 1))) + k] = s;
    }
  }
  wave_lds_sync();
}

}  // namespace cfs

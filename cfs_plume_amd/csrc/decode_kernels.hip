// decode_kernels.hip -- random-access decoders over a repacked basis.
//
// One wave decodes one segment (component k, block b, column c) with the
// same wavefront RLE parser as the fused step, then:
//   DEC_INTS    ints in u+8v+64w order      RunLengthDecodeBinary + ZigzagUnflatten
//                                           (reference COMPRESSION.cpp:1198-1283, 1427-1445)
//   DEC_COEFFS  dequantised FP32 coefficients DecodeBlockWithCompressionData (:1003-1057)
//   DEC_SPATIAL inverse DCT, written to the peeled unpadded interleaved column
//                                           DecodeScalarField / DecodeVectorField (:2389-2445, :2640-2670)
//   DEC_POINT   the spatial value at one in-block cell, i.e. one entry of
//               the 3 x r submatrix          GetSubmatrixNoSVDSparse (:1949-2158)
#include <hip/hip_runtime.h>

#include "block_dct.h"
#include "cfs_device.h"
#include "stream_decode.h"
#include "wave_ops.h"

namespace cfs {

enum : int { DEC_INTS = 0, DEC_COEFFS = 1, DEC_SPATIAL = 2, DEC_POINT = 3 };
constexpr int DEC_WAVES = 4;

template <int OUT>
__global__ void __launch_bounds__(DEC_WAVES * WAVE)
decode_kernel(DevBasis B, DecodeArgs a, const float* __restrict__ g_damp, const Tables* __restrict__ g_tab) {
  __shared__ __align__(16) float buf[DEC_WAVES][BC];
  __shared__ __align__(16) float scr[DEC_WAVES][BC];
  const int wave = threadIdx.x / WAVE;
  const int lane = lane_id();
  const long item = (long)blockIdx.x * DEC_WAVES + wave;
  const long total = (long)a.ncol * a.nk3 * a.nblk;
  if (item >= total) return;  // no workgroup barriers below
  const int b = a.b0 + (int)(item % a.nblk);
  const int k = a.k0 + (int)((item / a.nblk) % a.nk3);
  const int col = a.col0 + (int)(item / ((long)a.nblk * a.nk3));
  const size_t kb = (size_t)k * B.nB + b;
  const size_t si = kb * B.r + col;
  const uint32_t lk = B.seg_lk[si];
  const int len = (int)(lk & 0xffffu);
  const int row = (int)(lk >> 16) * NDIAG;
  const float sinv = B.seg_sinv[si];
  const uint64_t off = B.blk_base[kb] + B.seg_off[si];
  const uint64_t aligned = off & ~(uint64_t)3;
  const int lo = (int)(off - aligned);

  float* bf = buf[wave];
  int* bi = reinterpret_cast<int*>(bf);
  for (int i = lane; i < BC; i += WAVE) bi[i] = 0;
  wave_lds_sync();

  // one segment: nothing follows it inside [lo, lo+len), so no resets
  StreamParser<DEC_TPL> parser(B.tokens + aligned, lo, lo + len, nullptr, 0);
  for (int base = 0; base < lo + len; base += 64 * DEC_TPL) {
    Chunk<DEC_TPL> ch;
    parser.template parse<false>(base, ch);
#pragma unroll
    for (int i = 0; i < DEC_TPL; i++) {
      if (ch.em[i] > 0 && ch.v[i] != 0) {
        for (int m = 0; m < ch.em[i]; m++) {
          const int j = (ch.P[i] + m) & (BC - 1);
          const int idx = g_tab->rev[j];
          if (OUT == DEC_INTS)
            bi[idx] = ch.v[i];
          else
            bf[idx] = (float)ch.v[i] * (g_damp[row + g_tab->dz[j]] * sinv);
        }
      }
    }
  }
  if (!parser.complete(1) && lane == 0) atomicOr(B.err, (uint32_t)ERR_INCOMPLETE);
  wave_lds_sync();

  if (OUT == DEC_INTS || OUT == DEC_COEFFS) {
    int* o = reinterpret_cast<int*>(a.out) + item * BC;
    for (int i = lane; i < BC; i += WAVE) o[i] = bi[i];
    return;
  }
  const float* C = g_tab->dct;
  if (OUT == DEC_POINT) {
    // value at cell p = sum_idx coeff[idx] * C[u][x] C[v][y] C[w][z]
    const int px = a.point & 7, py = (a.point >> 3) & 7, pz = a.point >> 6;
    float s = 0.0f;
    for (int idx = lane; idx < BC; idx += WAVE) {
      const int u = idx & 7, vv = (idx >> 3) & 7, w = idx >> 6;
      s += bf[idx] * (C[u * BS + px] * C[vv * BS + py] * C[w * BS + pz]);
    }
    s = wave_sum(s);
    if (lane == 0) reinterpret_cast<double*>(a.out)[k + 3 * (col - a.col0)] = (double)s;
    return;
  }
  // DEC_SPATIAL: inverse DCT along z, y, x; write the peeled unpadded cells.
  float* sc = scr[wave];
  {
    const int u = lane & 7, vv = lane >> 3;
    float x8[BS];
#pragma unroll
    for (int w = 0; w < BS; w++) x8[w] = bf[u + 8 * vv + 64 * w];
#pragma unroll
    for (int z = 0; z < BS; z++) {
      float s = 0.0f;
#pragma unroll
      for (int w = 0; w < BS; w++) s += C[w * BS + z] * x8[w];
      sc[z * 64 + vv * 8 + u] = s;
    }
  }
  wave_lds_sync();
  {
    const int u = lane & 7, z = lane >> 3;
    float x8[BS];
#pragma unroll
    for (int vv = 0; vv < BS; vv++) x8[vv] = sc[z * 64 + vv * 8 + u];
#pragma unroll
    for (int y = 0; y < BS; y++) {
      float s = 0.0f;
#pragma unroll
      for (int vv = 0; vv < BS; vv++) s += C[vv * BS + y] * x8[vv];
      bf[z * 64 + y * 8 + u] = s;
    }
  }
  wave_lds_sync();
  {
    const int y = lane & 7, z = lane >> 3;
    const int bx = b % B.nbx, by = (b / B.nbx) % B.nby, bz = b / (B.nbx * B.nby);
    const int gy = by * BS + y, gz = bz * BS + z;
    float x8[BS];
#pragma unroll
    for (int u = 0; u < BS; u++) x8[u] = bf[z * 64 + y * 8 + u];
    float* o = reinterpret_cast<float*>(a.out) + (size_t)(col - a.col0) * 3 * ((size_t)B.X * B.Y * B.Z);
#pragma unroll
    for (int x = 0; x < BS; x++) {
      float s = 0.0f;
#pragma unroll
      for (int u = 0; u < BS; u++) s += C[u * BS + x] * x8[u];
      const int gx = bx * BS + x;
      if (gx < B.X && gy < B.Y && gz < B.Z) o[3 * ((size_t)gx + (size_t)B.X * (gy + (size_t)B.Y * gz)) + k] = s;
    }
  }
}

hipError_t launch_decode(const DevBasis& B, int mode, const DecodeArgs& a, const float* d_damp, const void* d_tab,
                         hipStream_t s) {
  const long total = (long)a.ncol * a.nk3 * a.nblk;
  if (total <= 0) return hipSuccess;
  const dim3 grid((unsigned)((total + DEC_WAVES - 1) / DEC_WAVES));
  const Tables* tab = reinterpret_cast<const Tables*>(d_tab);
  switch (mode) {
    case DEC_INTS:
      hipLaunchKernelGGL(decode_kernel<DEC_INTS>, grid, dim3(DEC_WAVES * WAVE), 0, s, B, a, d_damp, tab);
      break;
    case DEC_COEFFS:
      hipLaunchKernelGGL(decode_kernel<DEC_COEFFS>, grid, dim3(DEC_WAVES * WAVE), 0, s, B, a, d_damp, tab);
      break;
    case DEC_SPATIAL:
      hipLaunchKernelGGL(decode_kernel<DEC_SPATIAL>, grid, dim3(DEC_WAVES * WAVE), 0, s, B, a, d_damp, tab);
      break;
    default:
      hipLaunchKernelGGL(decode_kernel<DEC_POINT>, grid, dim3(DEC_WAVES * WAVE), 0, s, B, a, d_damp, tab);
      break;
  }
  return hipGetLastError();
}

}  // namespace cfs

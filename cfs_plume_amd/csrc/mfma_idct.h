// mfma_idct.h -- inverse 8x8x8 DCT of two blocks on MFMA (gfx950), BF16
// hi + lo split operands, FP32 accumulation.
//
//   x pass   T = X C           rows (w, v), K = (block, u) = 16: 4 tiles of
//                              v_mfma_f32_16x16x16_bf16 against blockdiag(C, C)
//   (y, z)   Out = K T         K[(z,y),(w,v)] = C[w][z] C[v][y] (64 x 64): 4 tiles x
//                              2 K-steps of v_mfma_f32_16x16x32_bf16.
// Lane l = 16 g + n.  Input (per tile t): the lane's 4 values of row
// (w, v) = 16 t + n of block g >> 1, u = 4 (g & 1) .. +3.  Output (per tile
// t): Out[(z, y) = 16 t + 4 g + i][x = n % 8] of block n / 8, i = 0..3.
// The x pass's accumulator layout (rows 4g..4g+3 of column n) is the (y, z)
// pass's B operand as it stands: K index 32 ks + 8 g + i maps to row
// (w, v) = 32 ks + 4 g + i (i < 4) or 32 ks + 16 + 4 g + i - 4, and the host
// (fill_mfma_frags) orders the columns of K the same way.  Each hi + lo
// product keeps hi*hi + hi*lo + lo*hi.  Lane layouts of both instructions
// were checked on the device against a host matrix product.
#pragma once
#include <hip/hip_runtime.h>

#include "block_dct.h"

namespace cfs {

typedef short bf16x4_t __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short bf16_bits(__bf16 h) {
  short v;
  __builtin_memcpy(&v, &h, 2);
  return v;
}
// x = hi + lo, both BF16 (round to nearest even)
__device__ __forceinline__ void split_bf16(float x, short& hi, short& lo) {
  const __bf16 h = (__bf16)x;
  hi = bf16_bits(h);
  lo = bf16_bits((__bf16)(x - (float)h));
}
__device__ __forceinline__ bf16x8_t as_bf16x8(const int4& v) {
  bf16x8_t r;
  __builtin_memcpy(&r, &v, 16);
  return r;
}

// The lane's constant operands (from Tables::mf).
struct MfmaConsts {
  bf16x4_t b1h, b1l;
  bf16x8_t a2h[4][2], a2l[4][2];
};

__device__ __forceinline__ void load_mfma_consts(const MfmaFrags& mf, int lane, MfmaConsts& c) {
#pragma unroll
  for (int i = 0; i < 4; i++) {
    c.b1h[i] = (short)mf.s1b[0][lane][i];
    c.b1l[i] = (short)mf.s1b[1][lane][i];
  }
#pragma unroll
  for (int t = 0; t < 4; t++)
#pragma unroll
    for (int ks = 0; ks < 2; ks++) {
      c.a2h[t][ks] = as_bf16x8(*reinterpret_cast<const int4*>(mf.s2a[0][t][ks][lane]));
      c.a2l[t][ks] = as_bf16x8(*reinterpret_cast<const int4*>(mf.s2a[1][t][ks][lane]));
    }
}

// The x pass and the conversion of its result into the (y, z) pass's B operand.
__device__ __forceinline__ void idct_x_pass(const float4 in[4], const bf16x4_t& b1h, const bf16x4_t& b1l,
                                            bf16x8_t b2h[2], bf16x8_t b2l[2]) {
  f32x4_t d1[4];
#pragma unroll
  for (int t = 0; t < 4; t++) {
    short h0, h1, h2, h3, l0, l1, l2, l3;
    split_bf16(in[t].x, h0, l0);
    split_bf16(in[t].y, h1, l1);
    split_bf16(in[t].z, h2, l2);
    split_bf16(in[t].w, h3, l3);
    const bf16x4_t ah = {h0, h1, h2, h3}, al = {l0, l1, l2, l3};
    f32x4_t c = {0.0f, 0.0f, 0.0f, 0.0f};
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, b1h, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, b1l, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, b1h, c, 0, 0, 0);
    d1[t] = c;
  }
#pragma unroll
  for (int ks = 0; ks < 2; ks++)
#pragma unroll
    for (int i = 0; i < 8; i++) {
      short h, l;
      split_bf16(i < 4 ? d1[2 * ks][i] : d1[2 * ks + 1][i - 4], h, l);
      __bf16 hb, lb;
      __builtin_memcpy(&hb, &h, 2);
      __builtin_memcpy(&lb, &l, 2);
      b2h[ks][i] = hb;
      b2l[ks][i] = lb;
    }
}

// The (y, z) pass of output tile t.
__device__ __forceinline__ f32x4_t idct_yz_tile(const bf16x8_t a2h[2], const bf16x8_t a2l[2], const bf16x8_t b2h[2],
                                                const bf16x8_t b2l[2]) {
  f32x4_t c = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int ks = 0; ks < 2; ks++) {
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2h[ks], b2h[ks], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2h[ks], b2l[ks], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2l[ks], b2h[ks], c, 0, 0, 0);
  }
  return c;
}

// Both passes with the constant operands resident in registers.
__device__ __forceinline__ void idct_pair_mfma(const float4 in[4], const MfmaConsts& mc, f32x4_t out[4]) {
  bf16x8_t b2h[2], b2l[2];
  idct_x_pass(in, mc.b1h, mc.b1l, b2h, b2l);
#pragma unroll
  for (int t = 0; t < 4; t++) out[t] = idct_yz_tile(mc.a2h[t], mc.a2l[t], b2h, b2l);
}

// Both passes reading the (y, z) operands from memory tile by tile (fewer
// live registers, for kernels that are register-bound).
__device__ __forceinline__ void idct_pair_mfma_streamed(const float4 in[4], const MfmaFrags& mf, int lane,
                                                        f32x4_t out[4]) {
  bf16x4_t b1h, b1l;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    b1h[i] = (short)mf.s1b[0][lane][i];
    b1l[i] = (short)mf.s1b[1][lane][i];
  }
  bf16x8_t b2h[2], b2l[2];
  idct_x_pass(in, b1h, b1l, b2h, b2l);
#pragma unroll
  for (int t = 0; t < 4; t++) {
    bf16x8_t a2h[2], a2l[2];
#pragma unroll
    for (int ks = 0; ks < 2; ks++) {
      a2h[ks] = as_bf16x8(*reinterpret_cast<const int4*>(mf.s2a[0][t][ks][lane]));
      a2l[ks] = as_bf16x8(*reinterpret_cast<const int4*>(mf.s2a[1][t][ks][lane]));
    }
    out[t] = idct_yz_tile(a2h, a2l, b2h, b2l);
  }
}

}  // namespace cfs

// wave_ops.h -- wave64 primitives for gfx950 built on DPP and v_perm_b32.
//
// The RLE token grammar (reference COMPRESSION.cpp:1139-1169, decoder
// :1241-1282) assigns each token one of three roles -- VALUE, DUP, LEN -- and
// the role of a token depends on every token before it.  Roles are recovered
// in parallel by scanning the composition of per-token transition functions
// of a 3-state machine.  A function f on states {0=VALUE, 1=DUP, 2=LEN} is
// stored as four bytes, byte s = f(s) (byte 3 = 3, unused); composition
// (g o f) is then a single v_perm_b32 whose selector is f.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cfs {

constexpr int WAVE = 64;

// DPP control words (GFX9 encoding; gfx950 keeps the wave-wide forms).
constexpr int DPP_ROW_SHR1 = 0x111;
constexpr int DPP_ROW_SHR2 = 0x112;
constexpr int DPP_ROW_SHR4 = 0x114;
constexpr int DPP_ROW_SHR8 = 0x118;
constexpr int DPP_WAVE_SHL1 = 0x130;
constexpr int DPP_WAVE_SHR1 = 0x138;
constexpr int DPP_ROW_BCAST15 = 0x142;
constexpr int DPP_ROW_BCAST31 = 0x143;

// Token transition functions.
constexpr uint32_t FSM_ID = 0x03020100u;     // identity
constexpr uint32_t FSM_NE = 0x03000200u;     // VALUE->VALUE, DUP->LEN, LEN->VALUE
constexpr uint32_t FSM_EQ = 0x03000201u;     // VALUE->DUP,   DUP->LEN, LEN->VALUE
// A state s is carried as the constant function s*0x01010101: composing any
// function after it yields the constant function of the next state.
constexpr uint32_t FSM_CONST_LEN = 0x02020202u;

__device__ __forceinline__ int lane_id() { return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0)); }

template <int CTRL, int ROW_MASK = 0xf, int BANK_MASK = 0xf>
__device__ __forceinline__ uint32_t dpp(uint32_t old, uint32_t src) {
  return (uint32_t)__builtin_amdgcn_update_dpp((int)old, (int)src, CTRL, ROW_MASK, BANK_MASK, false);
}
template <int CTRL, int ROW_MASK = 0xf, int BANK_MASK = 0xf>
__device__ __forceinline__ float dppf(float old, float src) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(src), CTRL, ROW_MASK, BANK_MASK, false));
}

// (later o earlier): byte s of the result is later[earlier[s]].
__device__ __forceinline__ uint32_t fsm_compose(uint32_t later, uint32_t earlier) {
  return __builtin_amdgcn_perm(0u, later, earlier);
}
__device__ __forceinline__ uint32_t fsm_apply(uint32_t f, uint32_t s) { return (f >> (8u * s)) & 0xffu; }

// Inclusive scan of transition functions across the wave (lane 0 first).
__device__ __forceinline__ uint32_t wave_scan_fsm(uint32_t f) {
  f = fsm_compose(f, dpp<DPP_ROW_SHR1>(FSM_ID, f));
  f = fsm_compose(f, dpp<DPP_ROW_SHR2>(FSM_ID, f));
  f = fsm_compose(f, dpp<DPP_ROW_SHR4>(FSM_ID, f));
  f = fsm_compose(f, dpp<DPP_ROW_SHR8>(FSM_ID, f));
  f = fsm_compose(f, dpp<DPP_ROW_BCAST15, 0xa>(FSM_ID, f));
  f = fsm_compose(f, dpp<DPP_ROW_BCAST31, 0xc>(FSM_ID, f));
  return f;
}

// DPP move whose out-of-range / masked lanes read 0 (bound_ctrl), the form
// the compiler folds into the consuming VALU instruction.
template <int CTRL, int ROW_MASK = 0xf, int BANK_MASK = 0xf>
__device__ __forceinline__ int dpp0(int src) {
  return __builtin_amdgcn_update_dpp(0, src, CTRL, ROW_MASK, BANK_MASK, true);
}

// Inclusive integer prefix sum across the wave.
__device__ __forceinline__ int wave_scan_add(int x) {
  x += dpp0<DPP_ROW_SHR1>(x);
  x += dpp0<DPP_ROW_SHR2>(x);
  x += dpp0<DPP_ROW_SHR4>(x);
  x += dpp0<DPP_ROW_SHR8>(x);
  x += dpp0<DPP_ROW_BCAST15, 0xa>(x);
  x += dpp0<DPP_ROW_BCAST31, 0xc>(x);
  return x;
}

// Inclusive float prefix sum across the wave.
__device__ __forceinline__ float wave_scan_addf(float x) {
  x += __int_as_float(dpp0<DPP_ROW_SHR1>(__float_as_int(x)));
  x += __int_as_float(dpp0<DPP_ROW_SHR2>(__float_as_int(x)));
  x += __int_as_float(dpp0<DPP_ROW_SHR4>(__float_as_int(x)));
  x += __int_as_float(dpp0<DPP_ROW_SHR8>(__float_as_int(x)));
  x += __int_as_float(dpp0<DPP_ROW_BCAST15, 0xa>(__float_as_int(x)));
  x += __int_as_float(dpp0<DPP_ROW_BCAST31, 0xc>(__float_as_int(x)));
  return x;
}

// Value held by lane+1 (lane 63 receives `fill`).
__device__ __forceinline__ uint32_t wave_shl1(uint32_t x, uint32_t fill) { return dpp<DPP_WAVE_SHL1>(fill, x); }
// Value held by lane-1 (lane 0 receives `fill`).
__device__ __forceinline__ uint32_t wave_shr1(uint32_t x, uint32_t fill) { return dpp<DPP_WAVE_SHR1>(fill, x); }

__device__ __forceinline__ int readlane(int x, int lane) { return __builtin_amdgcn_readlane(x, lane); }
__device__ __forceinline__ uint32_t readlane(uint32_t x, int lane) { return (uint32_t)__builtin_amdgcn_readlane((int)x, lane); }
__device__ __forceinline__ float readlanef(float x, int lane) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), lane)); }
__device__ __forceinline__ int uniform(int x) { return __builtin_amdgcn_readfirstlane(x); }

// Sum across the wave; every lane receives the total.
__device__ __forceinline__ float wave_sum(float x) {
  x += __shfl_xor(x, 32);
  x += __shfl_xor(x, 16);
  x += __shfl_xor(x, 8);
  x += __shfl_xor(x, 4);
  x += __shfl_xor(x, 2);
  x += __shfl_xor(x, 1);
  return x;
}
__device__ __forceinline__ float wave_max(float x) {
  x = fmaxf(x, __shfl_xor(x, 32));
  x = fmaxf(x, __shfl_xor(x, 16));
  x = fmaxf(x, __shfl_xor(x, 8));
  x = fmaxf(x, __shfl_xor(x, 4));
  x = fmaxf(x, __shfl_xor(x, 2));
  x = fmaxf(x, __shfl_xor(x, 1));
  return x;
}
__device__ __forceinline__ double wave_sum(double x) {
  x += __shfl_xor(x, 32);
  x += __shfl_xor(x, 16);
  x += __shfl_xor(x, 8);
  x += __shfl_xor(x, 4);
  x += __shfl_xor(x, 2);
  x += __shfl_xor(x, 1);
  return x;
}

// (a & mask) | (b & ~mask) as one v_bfi_b32.  Written as an instruction so
// the optimiser does not turn a 0/-1 mask back into a compare and a select.
__device__ __forceinline__ int bit_select(int mask, int a, int b) {
  int r;
  asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(r) : "v"(mask), "v"(a), "v"(b));
  return r;
}

// Orders this wave's LDS traffic (no cross-wave synchronisation).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace cfs

// step_kernels.hip -- fused decode + project + unproject for one compressed
// basis on gfx950.
//
// Restates, for the GPU, the per-step pair
//   PeeledCompressedProjectTransformNoSVD  (reference COMPRESSION.cpp:2978-3016)
//   PeeledCompressedUnprojectTransform     (reference COMPRESSION.cpp:2739-2900)
// Both stay in the DCT domain: project is q_c = sum_{k,b} <D_{k,c,b}, DCT(v)_{k,b}>,
// unproject accumulates A_{k,b} = sum_c q_c D_{k,c,b} and inverse-transforms
// each block once.  Each (component k, block b) "item" is handled by one wave,
// which walks the block stream of all r columns once: every decoded value
// feeds both the project dot and the unproject accumulator.
//
// Precision: tokens and positions are exact integers; dequantisation is
// int * (damp^gamma * (1/s)) in FP32 (the reference evaluates it in FP64,
// COMPRESSION.cpp:1029-1056); per-block dots are FP32 and are accumulated
// across blocks in FP64.
#include <hip/hip_runtime.h>

#include "block_dct.h"
#include "cfs_device.h"
#include "stream_decode.h"
#include "wave_ops.h"

namespace cfs {

enum : int { MODE_PROJECT = 1, MODE_UNPROJECT = 2, MODE_BOTH = 3 };

constexpr int WAVES_PER_WG = 4;
constexpr int NT = STEP_TPL;
constexpr int CH = 64 * NT;
constexpr int QREG = 4;  // segment parameters per lane: columns lane + 64*m, r <= 256
constexpr int MAX_COLS = 64 * QREG;
constexpr int NO_END = 0x7fffffff;

struct WaveLds {
  int4* prm;     // [r+1] {damp row byte offset, bits(2^64/s), bits(fixed-point weight of q_c), end offset};
                 // entry r is {0, 0, 0, NO_END} (tokens past the stream's end index it)
  double* qacc;  // [r] this wave's FP64 project accumulator
  float* vz;     // [512] DCT(v) block in zigzag order
  int* az;       // [512] unproject accumulator in zigzag order, fixed point
  float* pfx;    // [512] project prefix sums of the current chunk; DCT scratch
};

__host__ __device__ inline size_t step_wave_lds_bytes(int r) {
  return ((size_t)(r + 1) * 16 + (size_t)r * 8 + 3 * BC * 4 + 15) / 16 * 16;
}
__host__ __device__ inline size_t step_shared_lds_bytes(int nk) {
  return ((size_t)nk * NDIAG * 4 + 15) / 16 * 16 + BC;
}
__host__ inline size_t step_lds_bytes(int r, int nk) {
  return step_shared_lds_bytes(nk) + WAVES_PER_WG * step_wave_lds_bytes(r);
}

__device__ __forceinline__ WaveLds carve_wave(char* base, int r) {
  WaveLds w;
  w.prm = reinterpret_cast<int4*>(base);
  char* p = base + (size_t)(r + 1) * 16;
  w.vz = reinterpret_cast<float*>(p); p += BC * 4;
  w.az = reinterpret_cast<int*>(p); p += BC * 4;
  w.pfx = reinterpret_cast<float*>(p); p += BC * 4;
  w.qacc = reinterpret_cast<double*>(p);
  return w;
}

// One value of the stream at output position pos: dequantise (int * damp *
// (1/s), in the scaled units of the damping table) and feed both operators.
template <bool PROJ, bool UNPROJ>
__device__ __forceinline__ float step_value(const char* lds, const uint8_t* dz4, const WaveLds& w, const int4& pr,
                                            int val, int j) {
  const float dm = *reinterpret_cast<const float*>(lds + (pr.x + (int)dz4[j]));
  const float vf = (float)val * dm;
  if (UNPROJ) atomicAdd(&w.az[j], (int)(vf * __int_as_float(pr.z)));
  return PROJ ? vf * w.vz[j] * __int_as_float(pr.y) : 0.0f;
}

// Unproject accumulates A[j] = sum_c q_c D_c[j] with integer LDS atomics
// (gfx950's LDS float atomics are far slower).  |D_c[j]| <= vmax_c / s_c
// (vmax_c from the loader), so |A[j]| <= max|q| * blk_bound with
// blk_bound = sum_c vmax_c / s_c; with S = 2^30 / (max|q| * blk_bound) every
// partial sum of rounded terms stays below 2^31 and the rounding error is
// below r units of 1/S per entry.
//
// Project keeps a running FP64 sum over the block stream (FP32 within a
// chunk); the sum at the last token of segment c minus the sum at the last
// token of segment c-1 is segment c's dot product, added to q_c.
template <int MODE>
__global__ void __launch_bounds__(WAVES_PER_WG * WAVE, 3)  // <= 168 VGPRs: three waves per SIMD
step_kernel(DevBasis B, const float* __restrict__ v_in, const double* __restrict__ q_in,
            float* __restrict__ field_out, const float* __restrict__ g_damp, const Tables* __restrict__ g_tab) {
  extern __shared__ __align__(16) char lds[];
  constexpr bool PROJ = (MODE & MODE_PROJECT) != 0;
  constexpr bool UNPROJ = (MODE & MODE_UNPROJECT) != 0;

  const int nk = B.nk;
  float* damp = reinterpret_cast<float*>(lds);
  const size_t dbytes = ((size_t)nk * NDIAG * 4 + 15) / 16 * 16;
  uint8_t* dz4 = reinterpret_cast<uint8_t*>(lds + dbytes);
  const uint16_t* zig = g_tab->zig;  // used only by the per-block DCTs
  const int wave = uniform((int)(threadIdx.x / WAVE));
  const int lane = lane_id();
  const WaveLds w = carve_wave(lds + step_shared_lds_bytes(nk) + (size_t)wave * step_wave_lds_bytes(B.r), B.r);
  const int r = B.r;

  for (int i = threadIdx.x; i < nk * NDIAG; i += blockDim.x) damp[i] = g_damp[i];
  for (int i = threadIdx.x; i < BC; i += blockDim.x) dz4[i] = (uint8_t)(4 * g_tab->dz[i]);
  for (int i = lane; i < BC; i += WAVE) w.az[i] = 0;
  for (int c = lane; c < r; c += WAVE) w.qacc[c] = 0.0;
  if (lane == 0) w.prm[r] = make_int4(0, 0, 0, NO_END);
  // q is the same for every block: keep this lane's columns in registers
  float qf[QREG];
  float qmax = 0.0f;
#pragma unroll
  for (int m = 0; m < QREG; m++) {
    const int c = lane + WAVE * m;
    qf[m] = (UNPROJ && c < r) ? (float)q_in[c] : 0.0f;
    qmax = fmaxf(qmax, fabsf(qf[m]));
  }
  if (UNPROJ) qmax = wave_max(qmax);
  __syncthreads();
  const float* C = g_tab->dct;

  bool complete = true;
  const int total = 3 * B.nB;
  const int nwaves = gridDim.x * WAVES_PER_WG;
  // Static round-robin over items sorted by decreasing work: the three
  // components of a block go to neighbouring waves and run together.
  for (int item = blockIdx.x * WAVES_PER_WG + wave; item < total; item += nwaves) {
    const int b = B.order[item / 3];
    const int k = item % 3;
    const size_t kb = (size_t)k * B.nB + b;
    const size_t i0 = kb * r;
    const int len = (int)B.blk_len[kb];
    const uint64_t tbase = B.blk_base[kb];
    const uint32_t h0 = B.hz_start[kb], h1 = B.hz_start[kb + 1];
    const uint32_t l0 = B.lr_start[kb], l1 = B.lr_start[kb + 1];
    // per-segment parameters: all loads first, one round trip
    uint32_t lk[QREG];
    float sv[QREG];
#pragma unroll
    for (int m = 0; m < QREG; m++) {
      const int c = lane + WAVE * m;
      lk[m] = c < r ? B.seg_lk[i0 + c] : 0u;
      sv[m] = c < r ? B.seg_sinv[i0 + c] : 0.0f;
    }
    float inv_scale = 0.0f, S = 0.0f;
    if (UNPROJ) {
      const float bound = B.blk_bound[kb] * qmax * 1.0001f;
      S = bound > 0.0f ? 1073741824.0f / bound : 0.0f;
      inv_scale = S > 0.0f ? 1.0f / S : 0.0f;
    }
    // segment end offsets: inclusive prefix sum of the lengths (segments are
    // back to back in the block stream)
    int carry = 0;
#pragma unroll
    for (int m = 0; m < QREG; m++) {
      const int c = lane + WAVE * m;
      const int L = (int)(lk[m] & 0xffffu);
      const int incl = wave_scan_add(L) + carry;
      carry = readlane(incl, WAVE - 1);
      if (c < r)
        w.prm[c] = make_int4((int)(lk[m] >> 16) * (NDIAG * 4), __float_as_int(sv[m]), __float_as_int(qf[m] * sv[m] * S),
                             incl);
    }
    if (PROJ) forward_dct_block(v_in, B, k, b, C, zig, w.pfx, w.vz);
    wave_lds_sync();

    StreamParser<NT> parser(B.tokens + tbase, 0, len, B.hz + h0, (int)(h1 - h0));
    int next_lr = l0 < l1 ? (int)B.lr[l0] : NO_END;
    uint32_t li = l0;
    double run = 0.0;       // project sum of the chunks before this one (wave-uniform)
    double last_end = 0.0;  // running sum at the end of the last closed segment
    int cur = 0;            // first segment not yet closed
    for (int base = 0; base < len; base += CH) {
      Slots<NT> sl;
      if (parser.plain(base))
        parser.template parse_slots<true>(base, sl);
      else
        parser.template parse_slots<false>(base, sl);
      float acc = 0.0f, pre[NT];
#pragma unroll
      for (int i = 0; i < NT; i++) {
        const int4 pr = w.prm[sl.pos[i] >> 9];
        acc += step_value<PROJ, UNPROJ>(lds, dz4, w, pr, sl.val[i], sl.pos[i] & (BC - 1));
        pre[i] = acc;
      }
      // values P+3 .. P+L-1 of nonzero runs with L >= 4 (listed at load)
      while (next_lr < base + CH) {
        const int o = next_lr - base;
        const int ol = o / NT, os = o % NT;
        int pv = 0;
#pragma unroll
        for (int i = 0; i < NT; i++)
          if (i == os) pv = sl.pos[i];
        pv = readlane(pv, ol);                               // position of the value at P+2
        const int Lv = B.tokens[tbase + next_lr];            // run length
        const int vv = B.tokens[tbase + next_lr - 1];        // run value (the DUP token)
        const int4 pr = w.prm[pv >> 9];
        float ex = 0.0f;
        for (int m = lane; m < Lv - 3; m += WAVE)
          ex += step_value<PROJ, UNPROJ>(lds, dz4, w, pr, vv, (pv + 1 + m) & (BC - 1));
        if (PROJ) {
          ex = wave_sum(ex);
          if (lane == ol) {
            acc += ex;
#pragma unroll
            for (int i = 0; i < NT; i++)
              if (i >= os) pre[i] += ex;
          }
        }
        li++;
        next_lr = li < l1 ? (int)B.lr[li] : NO_END;
      }
      if (PROJ) {
        const float incl = wave_scan_addf(acc);
        const float excl = incl - acc;
        float4* pf = reinterpret_cast<float4*>(w.pfx + NT * lane);
#pragma unroll
        for (int q = 0; q < NT / 4; q++)
          pf[q] = make_float4(excl + pre[4 * q], excl + pre[4 * q + 1], excl + pre[4 * q + 2], excl + pre[4 * q + 3]);
        wave_lds_sync();
        // segments cur, cur+1, ... whose last token lies in this chunk
        for (;;) {
          const int c = cur + lane;
          const int e = w.prm[min(c, r)].w;
          const bool closes = e <= base + CH;
          const int pi = closes ? e - 1 - base : 0;
          const double endv = closes ? run + (double)w.pfx[pi] : 0.0;
          const double prevv = __shfl_up(endv, 1);
          const double d = endv - (lane == 0 ? last_end : prevv);
          if (closes) w.qacc[c] += d;
          const uint64_t mask = __ballot(closes);
          const int n = __popcll(mask);
          if (n > 0) last_end = __shfl(endv, n - 1);
          cur += n;
          if (n < WAVE) break;
        }
        run += (double)readlanef(incl, WAVE - 1);
        wave_lds_sync();
      }
    }
    complete = complete && parser.complete(r);
    wave_lds_sync();

    if (UNPROJ) inverse_dct_block(field_out, B, k, b, C, zig, w.az, inv_scale, w.pfx);
    wave_lds_sync();
  }
  if (!complete && lane == 0) atomicOr(B.err, (uint32_t)ERR_INCOMPLETE);

  if (PROJ) {
    // combine the waves' FP64 accumulators in a fixed order
    __syncthreads();
    for (int c = threadIdx.x; c < r; c += blockDim.x) {
      double s = 0.0;
      for (int wv = 0; wv < WAVES_PER_WG; wv++)
        s += carve_wave(lds + step_shared_lds_bytes(nk) + (size_t)wv * step_wave_lds_bytes(r), r).qacc[c];
      B.partials[(size_t)blockIdx.x * r + c] = s;
    }
  }
}

// q[c] = sum over workgroups of partials[wg][c], in workgroup order (FP64).
__global__ void __launch_bounds__(256) reduce_q_kernel(const double* __restrict__ partials, int nrows, int r,
                                                       double* __restrict__ q) {
  __shared__ double red[256];
  const int c = blockIdx.x;
  double s = 0.0;
  for (int i = threadIdx.x; i < nrows; i += blockDim.x) s += partials[(size_t)i * r + c];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) q[c] = red[0];
}

}  // namespace cfs

// ---------------------------------------------------------------------------
// Launch helpers (called from basis.cpp)
// ---------------------------------------------------------------------------
namespace cfs {

// largest r the step kernel supports (q accumulators live in registers)
int step_max_cols() { return MAX_COLS; }

hipError_t launch_step(const DevBasis& B, int mode, const float* v_in, const double* q_in, double* q_out,
                       float* field_out, const float* d_damp, const void* d_tab, int grid, hipStream_t s) {
  const size_t lds = step_lds_bytes(B.r, B.nk);
  const Tables* tab = reinterpret_cast<const Tables*>(d_tab);
  switch (mode) {
    case MODE_PROJECT:
      hipLaunchKernelGGL(step_kernel<MODE_PROJECT>, dim3(grid), dim3(WAVES_PER_WG * WAVE), lds, s, B, v_in, q_in,
                         field_out, d_damp, tab);
      break;
    case MODE_UNPROJECT:
      hipLaunchKernelGGL(step_kernel<MODE_UNPROJECT>, dim3(grid), dim3(WAVES_PER_WG * WAVE), lds, s, B, v_in, q_in,
                         field_out, d_damp, tab);
      break;
    default:
      hipLaunchKernelGGL(step_kernel<MODE_BOTH>, dim3(grid), dim3(WAVES_PER_WG * WAVE), lds, s, B, v_in, q_in,
                         field_out, d_damp, tab);
      break;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (mode & MODE_PROJECT) {
    hipLaunchKernelGGL(reduce_q_kernel, dim3(B.r), dim3(256), 0, s, B.partials, grid, B.r, q_out);
    e = hipGetLastError();
  }
  return e;
}

hipError_t step_occupancy(int mode, int r, int nk, int* wg_per_cu) {
  const size_t lds = step_lds_bytes(r, nk);
  hipError_t e;
  switch (mode) {
    case MODE_PROJECT:
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(wg_per_cu, step_kernel<MODE_PROJECT>, WAVES_PER_WG * WAVE, lds);
      break;
    case MODE_UNPROJECT:
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(wg_per_cu, step_kernel<MODE_UNPROJECT>, WAVES_PER_WG * WAVE, lds);
      break;
    default:
      e = hipOccupancyMaxActiveBlocksPerMultiprocessor(wg_per_cu, step_kernel<MODE_BOTH>, WAVES_PER_WG * WAVE, lds);
      break;
  }
  return e;
}

size_t tables_bytes() { return sizeof(Tables); }

void fill_tables(void* host_tab, const float* dct, const uint16_t* zig, const uint16_t* rev, const uint8_t* dz) {
  Tables* t = reinterpret_cast<Tables*>(host_tab);
  for (int i = 0; i < 64; i++) t->dct[i] = dct[i];
  for (int i = 0; i < BC; i++) {
    t->zig[i] = zig[i];
    t->rev[i] = rev[i];
    t->dz[i] = dz[i];
  }
}

}  // namespace cfs

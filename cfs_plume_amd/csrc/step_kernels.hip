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
struct WaveLds {
  float* vz;     // [APAD] DCT(v) block in zigzag order (padded index)
  float* az;     // [APAD] unproject accumulator, zigzag order (padded index)
  int* ends;     // [r] inclusive end offset of each segment in the block stream
  int* row;      // [r] damping-table row offset 22*k4
  float* sinv;   // [r] 1/s
  float* qs;     // [r] q_c/s (unproject weight)
  float* q32;    // [r] this block's FP32 project partials
  double* qacc;  // [r] this wave's FP64 project accumulator
};

__host__ __device__ inline size_t step_wave_lds_bytes(int r) {
  const size_t rp = (size_t)((r + 1) & ~1);
  return 2 * APAD * 4 + rp * (4 * 5 + 8);
}
__host__ __device__ inline size_t step_shared_lds_bytes(int nk) {
  return ((size_t)nk * NDIAG * 4 + 15) / 16 * 16 + BC * 2 + BC;
}
__host__ inline size_t step_lds_bytes(int r, int nk) {
  return step_shared_lds_bytes(nk) + WAVES_PER_WG * step_wave_lds_bytes(r);
}

__device__ __forceinline__ WaveLds carve_wave(char* base, int r) {
  const int rp = (r + 1) & ~1;
  WaveLds w;
  w.qacc = reinterpret_cast<double*>(base);
  char* p = base + (size_t)rp * 8;
  w.vz = reinterpret_cast<float*>(p); p += APAD * 4;
  w.az = reinterpret_cast<float*>(p); p += APAD * 4;
  w.ends = reinterpret_cast<int*>(p); p += (size_t)rp * 4;
  w.row = reinterpret_cast<int*>(p); p += (size_t)rp * 4;
  w.sinv = reinterpret_cast<float*>(p); p += (size_t)rp * 4;
  w.qs = reinterpret_cast<float*>(p); p += (size_t)rp * 4;
  w.q32 = reinterpret_cast<float*>(p);
  return w;
}

// Loads the per-segment parameters of block stream (k, b) into LDS.
__device__ void load_block_params(const DevBasis& B, int k, int b, const double* __restrict__ q_in,
                                  const WaveLds& w, bool unproject) {
  const int lane = lane_id();
  const size_t i0 = ((size_t)k * B.nB + b) * B.r;
  int carry = 0;
  for (int c0 = 0; c0 < B.r; c0 += WAVE) {
    const int c = c0 + lane;
    uint32_t lk = 0;
    float sv = 0.0f;
    if (c < B.r) {
      lk = B.seg_lk[i0 + c];
      sv = B.seg_sinv[i0 + c];
    }
    const int len = (int)(lk & 0xffffu);
    const int incl = wave_scan_add(len);
    if (c < B.r) {
      w.ends[c] = carry + incl - 1;
      w.row[c] = (int)(lk >> 16) * NDIAG;
      w.sinv[c] = sv;
      w.qs[c] = unproject ? (float)(q_in[c] * (double)sv) : 0.0f;
      w.q32[c] = 0.0f;
    }
    carry += readlane(incl, WAVE - 1);
  }
}

template <int MODE>
__global__ void __launch_bounds__(WAVES_PER_WG * WAVE)
step_kernel(DevBasis B, const float* __restrict__ v_in, const double* __restrict__ q_in,
            float* __restrict__ field_out, const float* __restrict__ g_damp, const Tables* __restrict__ g_tab) {
  extern __shared__ __align__(16) char lds[];
  constexpr bool PROJ = (MODE & MODE_PROJECT) != 0;
  constexpr bool UNPROJ = (MODE & MODE_UNPROJECT) != 0;

  const int nk = B.nk;
  float* damp = reinterpret_cast<float*>(lds);
  const size_t dbytes = ((size_t)nk * NDIAG * 4 + 15) / 16 * 16;
  uint16_t* zig = reinterpret_cast<uint16_t*>(lds + dbytes);
  uint8_t* dz = reinterpret_cast<uint8_t*>(lds + dbytes + BC * 2);
  const int wave = threadIdx.x / WAVE;
  const int lane = lane_id();
  char* wbase = lds + step_shared_lds_bytes(nk) + (size_t)wave * step_wave_lds_bytes(B.r);
  const WaveLds w = carve_wave(wbase, B.r);

  for (int i = threadIdx.x; i < nk * NDIAG; i += blockDim.x) damp[i] = g_damp[i];
  for (int i = threadIdx.x; i < BC; i += blockDim.x) {
    zig[i] = g_tab->zig[i];
    dz[i] = g_tab->dz[i];
  }
  for (int c = lane; c < B.r; c += WAVE) w.qacc[c] = 0.0;
  for (int i = lane; i < APAD; i += WAVE) w.az[i] = 0.0f;
  __syncthreads();
  const float* C = g_tab->dct;

  uint32_t err = 0;
  const int total = 3 * B.nB;
  for (;;) {
    int item = 0;
    if (lane == 0) item = (int)atomicAdd(B.counter, 1u);
    item = uniform(item);
    if (item >= total) break;
    const int b = B.order[item / 3];
    const int k = item % 3;
    const size_t kb = (size_t)k * B.nB + b;

    load_block_params(B, k, b, q_in, w, UNPROJ);
    if (PROJ) forward_dct_block(v_in, B, k, b, C, zig, w.az, w.vz);
    if (PROJ && UNPROJ)
      for (int i = lane; i < APAD; i += WAVE) w.az[i] = 0.0f;  // az served as DCT scratch
    wave_lds_sync();

    const int len = (int)B.blk_len[kb];
    StreamParser parser(B.tokens + B.blk_base[kb], 0, len, w.ends, B.r);
    for (int base = 0; base < len; base += CHUNK) {
      Chunk ch;
      parser.parse(base, ch);
      // this lane's head and tail segments
      const int last_seg = B.r - 1;
      int sh = min(ch.P[0] >> 9, last_seg);
      int st = sh;
#pragma unroll
      for (int i = 1; i < TPL; i++)
        if (ch.em[i] > 0) st = min(ch.P[i] >> 9, last_seg);
      const int row_h = w.row[sh], row_t = w.row[st];
      float qs_h = 0.0f, qs_t = 0.0f;
      if (UNPROJ) {
        qs_h = w.qs[sh];
        qs_t = w.qs[st];
      }
      float dot_h = 0.0f, dot_t = 0.0f;
#pragma unroll
      for (int i = 0; i < TPL; i++) {
        if (ch.em[i] > 0 && ch.v[i] != 0) {
          const bool hd = (ch.P[i] >> 9) <= sh;
          const int row = hd ? row_h : row_t;
          const float qs = hd ? qs_h : qs_t;
          const float vf0 = (float)ch.v[i];
          for (int m = 0; m < ch.em[i]; m++) {
            const int j = (ch.P[i] + m) & (BC - 1);
            const int jj = padj(j);
            const float vf = vf0 * damp[row + dz[j]];
            if (PROJ) {
              const float p = vf * w.vz[jj];
              if (hd) dot_h += p; else dot_t += p;
            }
            if (UNPROJ) atomicAdd(&w.az[jj], vf * qs);
          }
        }
      }
      if (PROJ) {
        if (dot_h != 0.0f) atomicAdd(&w.q32[sh], dot_h * w.sinv[sh]);
        if (st != sh && dot_t != 0.0f) atomicAdd(&w.q32[st], dot_t * w.sinv[st]);
      }
    }
    err |= parser.finish();
    wave_lds_sync();

    if (PROJ)
      for (int c = lane; c < B.r; c += WAVE) w.qacc[c] += (double)w.q32[c];
    if (UNPROJ) inverse_dct_block(field_out, B, k, b, C, zig, w.az, w.vz);
    wave_lds_sync();
  }
  if (err && lane == 0) atomicOr(B.err, err);

  if (PROJ) {
    __syncthreads();
    // combine the waves' FP64 accumulators in a fixed order
    for (int c = threadIdx.x; c < B.r; c += blockDim.x) {
      double s = 0.0;
      for (int wv = 0; wv < WAVES_PER_WG; wv++) {
        const WaveLds o = carve_wave(lds + step_shared_lds_bytes(nk) + (size_t)wv * step_wave_lds_bytes(B.r), B.r);
        s += o.qacc[c];
      }
      B.partials[(size_t)blockIdx.x * B.r + c] = s;
    }
  }
  // The last workgroup to finish resets the work counter for the next launch.
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    const unsigned done = atomicAdd(B.counter + 1, 1u);
    if (done == gridDim.x - 1) {
      B.counter[0] = 0u;
      B.counter[1] = 0u;
      __threadfence();
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

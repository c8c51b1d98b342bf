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

#include <algorithm>
#include <cmath>
#include <cstring>

#include "block_dct.h"
#include "cfs_device.h"
#include "mfma_idct.h"
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

// ---------------------------------------------------------------------------
// Spatial decode of whole columns (DecodeVectorField, COMPRESSION.cpp:2640-2670)
// with the inverse DCT on MFMA, BF16 hi + lo split products, FP32 accumulation.
//
// Work item: (column group of SP_G columns, block b), one workgroup of three
// waves, wave k decoding component k.  A group's segments are contiguous
// inside block stream (k, b), so each wave parses them as one stream and
// scatters the dequantised coefficients into LDS (natural u + 8v + 64w
// order, one 512-float buffer per column).  Then, per pair of columns, each
// wave transforms its component:
//   x pass   T = X C           rows (w, v), K = (block, u) = 16: 4 tiles of
//                              v_mfma_f32_16x16x16_bf16 against blockdiag(C, C)
//   (y, z)   Out = K T         K[(z,y),(w,v)] = C[w][z] C[v][y] (64 x 64): 4 tiles x
//                              2 K-steps of v_mfma_f32_16x16x32_bf16.
// The x pass's accumulator layout (lane: rows 4g..4g+3 of column n) is used
// as the (y, z) pass's B operand as it stands: K index 32 ks + 8 g + i maps
// to row (w, v) = 32 ks + 4 g + i (i < 4) or 32 ks + 16 + 4 g + i - 4, and the
// host orders the columns of K the same way.  Each product of a hi + lo pair
// keeps hi*hi + hi*lo + lo*hi.  Each wave writes its transformed pair over
// the pair's (already read) coefficient buffers; the store phase gathers the
// three components from the three waves' buffers and writes the column as
// contiguous 96-byte rows of interleaved cells.
// ---------------------------------------------------------------------------
constexpr int SP_G = 4;    // columns per work item
constexpr int SP_NT = 4;   // tokens per lane per chunk of the parser

// LDS: damping rows | jinfo[512] | per component: X[SP_G][512] + prm[SP_G+1]
__host__ __device__ inline size_t spatial_shared_lds_bytes(int nk) { return ((size_t)nk * NDIAG * 4 + 15) / 16 * 16 + BC * 4; }
__host__ __device__ inline size_t spatial_comp_lds_bytes() { return (size_t)SP_G * BC * 4 + (SP_G + 2) * 8; }
__host__ inline size_t spatial_lds_bytes(int nk) {
  return spatial_shared_lds_bytes(nk) + 3 * spatial_comp_lds_bytes();
}

// One work item's column group and this wave's segment parameters (lane < G).
struct SpatialItem {
  int b, c0, G;
  uint32_t lk, off;
  float sv;
};

__device__ __forceinline__ void spatial_load(const DevBasis& B, const DecodeArgs& a, int k, long item,
                                             SpatialItem& it) {
  const int grp = (int)(item / B.nB);
  it.b = (int)(item % B.nB);
  it.c0 = a.col0 + grp * SP_G;
  it.G = min(SP_G, a.col0 + a.ncol - it.c0);
  const size_t si0 = ((size_t)k * B.nB + it.b) * B.r + it.c0;
  const int lane = lane_id();
  it.lk = 0;
  it.off = 0;
  it.sv = 0.0f;
  if (lane < it.G) {
    it.lk = B.seg_lk[si0 + lane];
    it.off = B.seg_off[si0 + lane];
    it.sv = B.seg_sinv[si0 + lane];
  }
}
// token range [lo, hi) of the group inside its block stream
__device__ __forceinline__ int spatial_lo(const SpatialItem& it) { return readlane((int)it.off, 0); }
__device__ __forceinline__ int spatial_hi(const SpatialItem& it) {
  return readlane((int)it.off, it.G - 1) + (int)(readlane(it.lk, it.G - 1) & 0xffffu);
}
__device__ __forceinline__ StreamParser<SP_NT> spatial_parser(const DevBasis& B, int k, const SpatialItem& it) {
  const size_t kb = (size_t)k * B.nB + it.b;
  const int lo = spatial_lo(it), hi = spatial_hi(it);
  const int lo_al = lo & ~3;
  const uint32_t h0 = B.hz_start[kb], h1 = B.hz_start[kb + 1];
  return StreamParser<SP_NT>(B.tokens + B.blk_base[kb] + lo_al, lo - lo_al, hi - lo_al, B.hz + h0, (int)(h1 - h0),
                             lo_al);
}

// Persistent: workgroups stride over the work items; the next item's
// parameters and first token chunks are loaded while this one transforms.
__global__ void __launch_bounds__(3 * WAVE) __attribute__((amdgpu_waves_per_eu(3)))  // <= 168 registers
spatial_kernel(DevBasis B, DecodeArgs a, const float* __restrict__ g_damp, const Tables* __restrict__ g_tab) {
  extern __shared__ __align__(16) char lds[];
  const int nk = B.nk;
  float* damp = reinterpret_cast<float*>(lds);
  // per zigzag position j: natural index rev[j] | 4*(u+v+w) << 16
  uint32_t* jinfo = reinterpret_cast<uint32_t*>(lds + ((size_t)nk * NDIAG * 4 + 15) / 16 * 16);
  const int k = uniform((int)(threadIdx.x / WAVE));  // this wave's component
  const int lane = lane_id();
  char* cbase = lds + spatial_shared_lds_bytes(nk) + (size_t)k * spatial_comp_lds_bytes();
  float* X = reinterpret_cast<float*>(cbase);                           // [SP_G][512]
  int2* prm = reinterpret_cast<int2*>(cbase + (size_t)SP_G * BC * 4);   // [SP_G + 1] {row byte offset, bits(1/s)}

  // Each workgroup takes a contiguous range of items: neighbouring blocks
  // along x, whose 96-byte rows share 128-byte lines of the output, are
  // written one after the other by the same workgroup (same L2).
  const long all = (long)((a.ncol + SP_G - 1) / SP_G) * B.nB;
  const long per = (all + gridDim.x - 1) / gridDim.x;
  long item = (long)blockIdx.x * per;
  const long total = min(all, item + per);
  if (item >= total) return;  // workgroup-uniform: no items (no barrier reached)
  SpatialItem cur;
  spatial_load(B, a, k, item, cur);
  StreamParser<SP_NT> parser = spatial_parser(B, k, cur);

  for (int i = threadIdx.x; i < nk * NDIAG; i += blockDim.x) damp[i] = g_damp[i];
  for (int i = threadIdx.x; i < BC; i += blockDim.x)
    jinfo[i] = (uint32_t)g_tab->rev[i] | ((uint32_t)(4 * g_tab->dz[i]) << 16);
  // constant MFMA operands of this lane
  MfmaConsts mc;
  load_mfma_consts(g_tab->mf, lane, mc);
  const int g = lane >> 4, n = lane & 15;
  const size_t col_stride = (size_t)3 * B.X * B.Y * B.Z;
  bool complete = true;

  for (; item < total; item++) {
    const int b = cur.b, c0 = cur.c0, G = cur.G;
    const size_t kb = (size_t)k * B.nB + b;
    const size_t si0 = kb * B.r + c0;
    const int lo = spatial_lo(cur), hi = spatial_hi(cur);
    const int lo_al = lo & ~3;
    if (lane <= G)
      prm[lane] = lane < G ? make_int2((int)(cur.lk >> 16) * (NDIAG * 4), __float_as_int(cur.sv)) : make_int2(0, 0);
    {
      float4* x4 = reinterpret_cast<float4*>(X);
      for (int i = lane; i < SP_G * BC / 4; i += WAVE) x4[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    }
    __syncthreads();  // tables (first item) and this wave's buffers
    const long nitem = item + 1;
    SpatialItem nx = cur;
    if (nitem < total) spatial_load(B, a, k, nitem, nx);

    // parse the group's tokens as one stream; tokens outside [lo, hi) are ignored
    const int32_t* tok = B.tokens + B.blk_base[kb] + lo_al;
    // long runs of columns c0 .. c0+G-1: [seg_lr[si0], seg_lr[si0 + G]) (seg_lr has one extra entry)
    uint32_t li = B.seg_lr[si0];
    const uint32_t l1 = B.seg_lr[si0 + G];
    int next_lr = li < l1 ? (int)B.lr[li] - lo_al : 0x7fffffff;
    for (int base = 0; base < hi - lo_al; base += 64 * SP_NT) {
      Slots<SP_NT> sl;
      if (parser.plain(base))
        parser.template parse_slots<true>(base, sl);
      else
        parser.template parse_slots<false>(base, sl);
#pragma unroll
      for (int i = 0; i < SP_NT; i++) {
        if (sl.val[i] != 0) {
          const int2 pr = prm[sl.pos[i] >> 9];
          const uint32_t ji = jinfo[sl.pos[i] & (BC - 1)];
          const float dm = *reinterpret_cast<const float*>(lds + (pr.x + (int)(ji >> 16)));
          X[(sl.pos[i] >> 9) * BC + (int)(ji & 0xffffu)] = (float)sl.val[i] * (dm * __int_as_float(pr.y));
        }
      }
      // values P+3 .. P+L-1 of nonzero runs with L >= 4
      while (next_lr < min(base + 64 * SP_NT, hi - lo_al)) {
        const int o = next_lr - base;
        const int ol = o / SP_NT, os = o % SP_NT;
        int pv = 0;
#pragma unroll
        for (int i = 0; i < SP_NT; i++)
          if (i == os) pv = sl.pos[i];
        pv = readlane(pv, ol);
        const int Lv = tok[next_lr];
        const int vv = tok[next_lr - 1];
        const int2 pr = prm[pv >> 9];
        for (int m = lane; m < Lv - 3; m += WAVE) {
          const int p = pv + 1 + m;
          const uint32_t ji = jinfo[p & (BC - 1)];
          const float dm = *reinterpret_cast<const float*>(lds + (pr.x + (int)(ji >> 16)));
          X[(p >> 9) * BC + (int)(ji & 0xffffu)] = (float)vv * (dm * __int_as_float(pr.y));
        }
        li++;
        next_lr = li < l1 ? (int)B.lr[li] - lo_al : 0x7fffffff;
      }
    }
    complete = complete && parser.complete(G);
    // the next item's first token chunks, in flight during this item's transforms
    if (nitem < total) parser = spatial_parser(B, k, nx);
    wave_lds_sync();

    const int bx = b % B.nbx, by = (b / B.nbx) % B.nby, bz = b / (B.nbx * B.nby);
    for (int s0 = 0; s0 < G; s0 += 2) {
      // lane (n, g) reads row 16 t + n of block s0 + (g >> 1), u = 4 (g & 1) .. +3
      const float* xs = X + (s0 + (g >> 1)) * BC + 4 * (g & 1);
      float4 in[4];
#pragma unroll
      for (int t = 0; t < 4; t++) in[t] = *reinterpret_cast<const float4*>(xs + (16 * t + n) * BS);
      f32x4_t d2[4];
      idct_pair_mfma(in, mc, d2);
      // lane column n = (block s0 + n / 8, x = n % 8), rows (z, y) = 16 t + 4 g + i,
      // written over this wave's buffers of the pair: X[s0 + block][x + 8y + 64z]
      // (the MFMAs above consumed every lane's loads of them)
      float* tp = X + (size_t)(s0 + (n >> 3)) * BC + (n & 7);
#pragma unroll
      for (int t = 0; t < 4; t++)
#pragma unroll
        for (int i = 0; i < 4; i++) tp[(16 * t + 4 * g + i) * BS] = d2[t][i];
      __syncthreads();  // the three components of the pair's cells are in LDS
      // rows of 8 cells x 3 components = 24 contiguous floats per (block, z, y):
      // thread (y, f) = (tid / 24, tid % 24) writes float f of row (z, y) for every z
      {
        const int f24 = threadIdx.x % 24, y = threadIdx.x / 24;
        const int comp = f24 % 3, xx = f24 / 3;
        const float* pl = reinterpret_cast<const float*>(lds + spatial_shared_lds_bytes(nk) +
                                                         (size_t)comp * spatial_comp_lds_bytes()) + (size_t)s0 * BC;
        const int gx = bx * BS + xx, gy = by * BS + y;
        const int zmax = min(BS, B.Z - bz * BS);
        if (gx < B.X && gy < B.Y) {
          const int cell0 = 3 * (gx + B.X * (gy + B.Y * bz * BS)) + f24 % 3;
          const int zstep = 3 * B.X * B.Y;
#pragma unroll
          for (int sb = 0; sb < 2; sb++) {
            if (s0 + sb < G) {
              float* o = reinterpret_cast<float*>(a.out) + (size_t)(c0 + s0 + sb - a.col0) * col_stride + cell0;
              const float* tr = pl + sb * BC + y * BS + xx;
#pragma unroll
              for (int z = 0; z < BS; z++)
                if (z < zmax) o[z * zstep] = tr[z * 64];
            }
          }
        }
      }
      __syncthreads();  // the pair's buffers are free again
    }
    cur = nx;
  }
  if (!complete && lane == 0) atomicOr(B.err, (uint32_t)ERR_INCOMPLETE);
}

// UntransformVectorFieldSVD (reference COMPRESSION.cpp:713-727) on decoded
// columns: every cell's xyz triple of column c becomes V_c u' (V_c row-major
// 3 x 3, FP64; the reference multiplies the interleaved column by the block
// diagonal of V_c, SparseBlockDiagonal :566-598).  One thread per cell of one
// column; the FP32 triple is read once and written once.
__global__ void __launch_bounds__(256) untransform_svd_kernel(float* __restrict__ cols, long cells, int ncols,
                                                              const double* __restrict__ V) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int c = (int)blockIdx.y;
  if (i >= cells || c >= ncols) return;
  const double* m = V + 9 * (size_t)c;
  float* p = cols + (size_t)c * 3 * (size_t)cells + 3 * (size_t)i;
  const double a = p[0], b = p[1], d = p[2];
  p[0] = (float)(m[0] * a + m[1] * b + m[2] * d);
  p[1] = (float)(m[3] * a + m[4] * b + m[5] * d);
  p[2] = (float)(m[6] * a + m[7] * b + m[8] * d);
}

hipError_t launch_untransform_svd(float* cols, long cells, int ncols, const double* V, hipStream_t s) {
  if (cells <= 0 || ncols <= 0) return hipSuccess;
  const long blocks = (cells + 255) / 256;
  if (blocks > 0x7fffffffL || ncols > 65535) return hipErrorInvalidValue;
  hipLaunchKernelGGL(untransform_svd_kernel, dim3((unsigned)blocks, (unsigned)ncols), dim3(256), 0, s, cols, cells,
                     ncols, V);
  return hipGetLastError();
}

// PeeledCompressedUnproject's accumulation (COMPRESSION.cpp:2677-2711) for a
// group of decoded columns: the interior of the full-grid field gets
// sum_k q[k] cols[k] (FP64 sums, added to the field's value unless `first`).
__global__ void __launch_bounds__(256) accumulate_columns_kernel(const float* __restrict__ cols, int ncols, long n3,
                                                                 const double* __restrict__ q,
                                                                 float* __restrict__ field, int X, int Y, int Z,
                                                                 int first) {
  const long j = (long)blockIdx.x * 256 + threadIdx.x;
  if (j >= n3) return;
  const long cell = j / 3;
  const int comp = (int)(j % 3);
  const int x = (int)(cell % X), y = (int)((cell / X) % Y), z = (int)(cell / ((long)X * Y));
  const long f = 3 * ((long)(x + 1) + (long)(X + 2) * ((y + 1) + (long)(Y + 2) * (z + 1))) + comp;
  double acc = 0.0;
  for (int k = 0; k < ncols; k++) acc += q[k] * (double)cols[(long)k * n3 + j];
  field[f] = first ? (float)acc : (float)((double)field[f] + acc);
}

hipError_t launch_accumulate_columns(const float* cols, int ncols, long n3, const double* q, float* field, int X,
                                     int Y, int Z, int first, hipStream_t s) {
  if (n3 <= 0 || ncols <= 0) return hipSuccess;
  const long blocks = (n3 + 255) / 256;
  if (blocks > 0x7fffffffL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(accumulate_columns_kernel, dim3((unsigned)blocks), dim3(256), 0, s, cols, ncols, n3, q, field,
                     X, Y, Z, first);
  return hipGetLastError();
}

// Host: the MFMA operands of the inverse DCT in lane order (FP64 closed form
// split into BF16 hi + lo, round to nearest even).
static uint16_t host_bf16_bits(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}
static float host_bf16_value(uint16_t h) {
  const uint32_t u = (uint32_t)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
static void host_split(double x, uint16_t* hi, uint16_t* lo) {
  *hi = host_bf16_bits((float)x);
  *lo = host_bf16_bits((float)(x - (double)host_bf16_value(*hi)));
}

void fill_mfma_frags(void* host_tab) {
  MfmaFrags& mf = reinterpret_cast<Tables*>(host_tab)->mf;
  double C[BS][BS];
  for (int kk = 0; kk < BS; kk++)
    for (int nn = 0; nn < BS; nn++)
      C[kk][nn] = (kk == 0 ? std::sqrt(0.125) : 0.5) * std::cos(M_PI * (double)((2 * nn + 1) * kk) / 16.0);
  for (int l = 0; l < WAVE; l++) {
    const int g = l / 16, nn = l % 16;
    for (int i = 0; i < 4; i++) {
      const int kk = 4 * g + i;  // blockdiag(C, C)[kk][nn]
      const double v = ((kk < 8) == (nn < 8)) ? C[kk % 8][nn % 8] : 0.0;
      host_split(v, &mf.s1b[0][l][i], &mf.s1b[1][l][i]);
    }
    for (int t = 0; t < 4; t++)
      for (int ks = 0; ks < 2; ks++)
        for (int i = 0; i < 8; i++) {
          const int q = 16 * t + nn;                                         // output row (z, y)
          const int rw = i < 4 ? 32 * ks + 4 * g + i : 32 * ks + 16 + 4 * g + (i - 4);  // input row (w, v)
          const double v = C[rw >> 3][q >> 3] * C[rw & 7][q & 7];             // C[w][z] C[v][y]
          host_split(v, &mf.s2a[0][t][ks][l][i], &mf.s2a[1][t][ks][l][i]);
        }
  }
}

hipError_t launch_decode(const DevBasis& B, int mode, const DecodeArgs& a, const float* d_damp, const void* d_tab,
                         hipStream_t s) {
  if (mode == DEC_SPATIAL) {
    const long items = (long)((a.ncol + SP_G - 1) / SP_G) * B.nB;
    if (a.ncol == 0) return hipSuccess;  // an empty column range, as the other decoders
    if (items <= 0 || a.k0 != 0 || a.nk3 != 3 || a.b0 != 0 || a.nblk != B.nB) return hipErrorInvalidValue;
    const size_t lds = spatial_lds_bytes(B.nk);
    int dev = 0, cus = 0, per = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (e == hipSuccess) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, spatial_kernel, 3 * WAVE, lds);
    if (e != hipSuccess) return e;
    if (per < 1) return hipErrorInvalidConfiguration;
    // two workgroups per CU: fewer items in flight per XCD between a block's
    // partial output lines and its x-neighbour's completing them (measured
    // 3 % faster than the full occupancy of 4)
    const long grid = std::min<long>(items, (long)cus * std::min(per, 2));
    hipLaunchKernelGGL(spatial_kernel, dim3((unsigned)grid), dim3(3 * WAVE), lds, s, B, a, d_damp,
                       reinterpret_cast<const Tables*>(d_tab));
    return hipGetLastError();
  }
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

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
// COMPRESSION.cpp:1029-1056); per-segment dots are FP32, summed in FP64 over
// the blocks one wave handles and across waves.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "block_dct.h"
#include "cfs_device.h"
#include "mfma_idct.h"
#include "stream_decode.h"
#include "wave_ops.h"

namespace cfs {

enum : int { MODE_PROJECT = 1, MODE_UNPROJECT = 2, MODE_BOTH = 3 };

// The single-basis step runs one workgroup of STEP_WAVES waves per CU (three
// per SIMD); the batched step keeps the autotuned 2- or 4-wave shapes.
constexpr int STEP_WAVES = 12;
constexpr int NT = STEP_TPL;
constexpr int CH = 64 * NT;
constexpr int QREG = 4;  // segment parameters per lane: columns lane + 64*m, r <= 256
constexpr int MAX_COLS = 64 * QREG;
constexpr int NO_END = 0x7fffffff;

struct WaveLds {
  int4* prm;     // [r+1] {damp row byte offset, bits(fixed-point weight of q_c), bits(2^64/s), end offset};
                 // entry r is {0, 0, 0, NO_END} (tokens past the stream's end index it)
  double* qacc;  // [r] this wave's project accumulator (the dots of its items)
  float* vz;     // [512] DCT(v) block in zigzag order
  int* az;       // [512] unproject accumulator in zigzag order, fixed point
  float* pfx;    // [512] project prefix sums of the current chunk; DCT scratch
};

__host__ __device__ inline size_t step_wave_lds_bytes(int r) {
  return (size_t)(r + 1) * 16 + 3 * (size_t)BC * 4 + ((size_t)r * 8 + 15) / 16 * 16;
}
// Shared part of the workgroup's dynamic LDS: the damping rows (the tables
// 4*(u+v+w) and zigzag by index are static LDS, 3*512 bytes).
__host__ __device__ inline size_t step_shared_lds_bytes(int nk) {
  return ((size_t)nk * NDIAG * 4 + 15) / 16 * 16;
}
__host__ inline size_t step_lds_bytes(int r, int nk, int wpw) {
  return step_shared_lds_bytes(nk) + (size_t)wpw * step_wave_lds_bytes(r);
}
// gfx950: 160 KB of LDS per CU, all of it available to one workgroup; the
// step's static LDS (the 4*(u+v+w) and zigzag tables, the slot counter)
// takes less than 2 KB.
constexpr size_t LDS_PER_CU = 160 * 1024;
constexpr size_t STEP_STATIC_LDS = 2048;
// The single-basis step copies the inverse DCT's constant operand table
// (17 KB) into its LDS, after the waves' buffers, when the table fits there.
__host__ inline bool step_mf_in_lds(int mode, int r, int nk) {
  return (mode & 2) != 0 && step_lds_bytes(r, nk, STEP_WAVES) + sizeof(MfmaFrags) + STEP_STATIC_LDS <= LDS_PER_CU;
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

// Unproject accumulates A[j] = sum_c q_c D_c[j] with integer LDS atomics
// (gfx950's LDS float atomics are far slower).  |D_c[j]| <= vmax_c / s_c
// (vmax_c from the loader), so |A[j]| <= max|q| * blk_bound with
// blk_bound = sum_c vmax_c / s_c; with S = 2^30 / (max|q| * blk_bound) every
// partial sum of rounded terms stays below 2^31 and the rounding error is
// below r units of 1/S per entry.
//
// Project keeps, per chunk, the prefix sum of int * damp * DCT(v) over the
// chunk's tokens (FP32, starting at zero).  The prefix at the last token of
// segment c minus the prefix at the last token of segment c-1 (or, for the
// first segment closed in a chunk, plus its sum over earlier chunks), times
// 1/s_c, is segment c's dot product.  The encoder scales every segment to the
// same peak (2^(nBits-1)-1), so the unscaled terms of neighbouring segments
// have comparable magnitudes.
// Everything an item needs from global memory before its chunk loop; loaded
// one item ahead so the latency overlaps the previous item's work.
struct ItemPrefetch {
  uint4 d0, d1, d2;     // the item's descriptor (wave-uniform)
  uint32_t lk[QREG];    // seg_lk of columns lane + 64*m
  float sv[QREG];       // seg_sinv of the same columns
  float rows[BS];       // row (y, z) = lane of the block of v
};

template <bool PROJ>
__device__ __forceinline__ void prefetch_item(const DevBasis& B, const float* __restrict__ v_in, int item,
                                              ItemPrefetch& p) {
  const uint4* d = B.items + (size_t)3 * item;
  p.d0 = d[0];
  p.d1 = d[1];
  p.d2 = d[2];
  const size_t i0 = (size_t)p.d0.w * B.r;
  const int lane = lane_id();
#pragma unroll
  for (int m = 0; m < QREG; m++) {
    const int c = lane + WAVE * m;
    p.lk[m] = c < B.r ? B.seg_lk[i0 + c] : 0u;
    p.sv[m] = c < B.r ? B.seg_sinv[i0 + c] : 0.0f;
  }
  if (PROJ) gather_block_rows(v_in, B, (int)p.d2.z, (int)p.d2.y, p.rows);
}

// Inverse DCT of the zigzag-ordered fixed-point accumulator `az` (value =
// az * unit) on MFMA (mfma_idct.h: one block in the first half of a pair) and
// the store of the interior cells of component k into the full-grid field
// (UnitaryBlockDCTEigen(-1) + AssimilateBlocksEigen + subfield +
// setWithPeeled, COMPRESSION.cpp:2880-2898, VECTOR3_FIELD_3D.cpp:3415-3430).
// Leaves az zero.  `mf` is the constant operand table, in LDS for the
// single-basis step (so that its reads do not queue behind the next item's
// prefetch in the in-order vector-memory counter) or in global memory.
__device__ __forceinline__ void inverse_dct_block_mfma(float* __restrict__ field, const DevBasis& B, int k, int b,
                                                       const uint16_t* zig, int* az, float unit,
                                                       const MfmaFrags& mf) {
  const int lane = lane_id();
  const int g = lane >> 4, n = lane & 15;
  // lane (n, g < 2) holds row (w, v) = 16 t + n, u = 4 g .. 4 g + 3; lanes g >= 2 (second block) hold zeros
  float4 in[4];
#pragma unroll
  for (int t = 0; t < 4; t++) {
    float v[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (g < 2) {
      const int nat = (16 * t + n) * BS + 4 * g;
#pragma unroll
      for (int i = 0; i < 4; i++) v[i] = (float)az[zig[nat + i]] * unit;
    }
    in[t] = make_float4(v[0], v[1], v[2], v[3]);
  }
  wave_lds_sync();
  for (int i = lane; i < BC; i += WAVE) az[i] = 0;
  bf16x4_t b1h, b1l;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    b1h[i] = (short)mf.s1b[0][lane][i];
    b1l[i] = (short)mf.s1b[1][lane][i];
  }
  bf16x8_t b2h[2], b2l[2];
  idct_x_pass(in, b1h, b1l, b2h, b2l);
  // (y, z) pass tile by tile (one tile's operands live at a time); lane
  // (n < 8, g) stores cells x = n, (z, y) = 16 t + 4 g + i
  const int bx = b % B.nbx, by = (b / B.nbx) % B.nby, bz = b / (B.nbx * B.nby);
  const int gx = bx * BS + n;
  const bool store = n < BS && gx < B.X;
#pragma unroll 1
  for (int t = 0; t < 4; t++) {
    bf16x8_t a2h[2], a2l[2];
#pragma unroll
    for (int ks = 0; ks < 2; ks++) {
      a2h[ks] = as_bf16x8(*reinterpret_cast<const int4*>(mf.s2a[0][t][ks][lane]));
      a2l[ks] = as_bf16x8(*reinterpret_cast<const int4*>(mf.s2a[1][t][ks][lane]));
    }
    const f32x4_t d = idct_yz_tile(a2h, a2l, b2h, b2l);
    if (store) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int q = 16 * t + 4 * g + i;
        const int gy = by * BS + (q & 7), gz = bz * BS + (q >> 3);
        if (gy < B.Y && gz < B.Z) field[field_offset(B, gx, gy, gz, k)] = d[i];
      }
    }
  }
}

// Workgroups are dispatched to the chip's NXCD dies in turn: workgroup i
// runs on die i % NXCD, whose L2 it shares with the die's other workgroups.
constexpr int NXCD = 8;

// Position of a wave among the nwaves items of each round of the schedule
// (item = position + round * nwaves; items are blocks in decreasing-work
// order, three consecutive items per block).  DIE_GROUPED (WPW a multiple of
// three): a block's three component items go to three consecutive waves of
// one workgroup, so they read its v rows and write its interleaved cells
// through one L2; consecutive blocks go to different dies, and a die's
// blocks go to its workgroups in turn (the workgroup's first triple of
// waves, then its second, ...), so every die and every workgroup gets an
// equal share of every round, also of a last, partial one.  Otherwise the
// position is the wave's global index.
template <int WPW, bool DIE_GROUPED>
__device__ __forceinline__ int round_position(int wg, int nwg, int wave) {
  if (DIE_GROUPED && WPW % 3 == 0 && nwg % NXCD == 0) {
    const int triple = (wave / 3) * (nwg / NXCD) + wg / NXCD;  // block index on the die within the round
    return (triple * NXCD + wg % NXCD) * 3 + wave % 3;
  }
  return wg * WPW + wave;
}

// The fused step of one basis, run by workgroup `wg` of `nwg` that share it;
// WPW waves per workgroup.  Item slots of the workgroup: slot s is round
// s / WPW at position round_position(wg, nwg, s % WPW), i.e. the items a
// static round-robin would give its waves, in round order.  DYNAMIC: each
// wave starts with slot `wave` and then takes the workgroup's next slot from
// an LDS counter, so the waves of a CU that the SIMD arbiter runs faster (the
// older ones) take more items; otherwise wave w takes slots w, w + WPW, ...
template <int MODE, int WPW, bool DIE_GROUPED = false, bool DYNAMIC = false, bool MF_LDS = false>
__device__ __forceinline__ void step_body(const DevBasis& B, const float* __restrict__ v_in,
                                          const double* __restrict__ q_in, float* __restrict__ field_out,
                                          const float* __restrict__ g_damp, const Tables* __restrict__ g_tab,
                                          double* __restrict__ partials, int wg, int nwg) {
  extern __shared__ __align__(16) char lds[];
  constexpr bool PROJ = (MODE & MODE_PROJECT) != 0;
  constexpr bool UNPROJ = (MODE & MODE_UNPROJECT) != 0;

  const int nk = B.nk;
  // static LDS: its addresses are compile-time constants, so dz4[j] reads
  // address j + constant with no address arithmetic
  __shared__ uint8_t dz4[BC];
  __shared__ __align__(16) uint16_t zig[BC];  // used by the per-block DCTs
  __shared__ uint32_t next_slot;               // DYNAMIC: the workgroup's next unassigned slot
  float* damp = reinterpret_cast<float*>(lds);
  const int wave = uniform((int)(threadIdx.x / WAVE));
  const int lane = lane_id();
  const WaveLds w = carve_wave(lds + step_shared_lds_bytes(nk) + (size_t)wave * step_wave_lds_bytes(B.r), B.r);
  const int r = B.r;
  const int total = 3 * B.nB;
  const int nwaves = nwg * WPW;
  int item = round_position<WPW, DIE_GROUPED>(wg, nwg, wave);

  ItemPrefetch nx;
  if (item < total) prefetch_item<PROJ>(B, v_in, item, nx);
  for (int i = threadIdx.x; i < nk * NDIAG; i += blockDim.x) damp[i] = g_damp[i];
  const MfmaFrags* mf = &g_tab->mf;
  if (MF_LDS) {
    MfmaFrags* m = reinterpret_cast<MfmaFrags*>(lds + step_shared_lds_bytes(nk) + (size_t)WPW * step_wave_lds_bytes(r));
    const int4* src = reinterpret_cast<const int4*>(&g_tab->mf);
    int4* dst = reinterpret_cast<int4*>(m);
    for (int i = threadIdx.x; i < (int)(sizeof(MfmaFrags) / 16); i += blockDim.x) dst[i] = src[i];
    mf = m;
  }
  for (int i = threadIdx.x; i < BC; i += blockDim.x) {
    dz4[i] = (uint8_t)(4 * g_tab->dz[i]);
    zig[i] = g_tab->zig[i];
  }
  for (int i = lane; i < BC; i += WAVE) w.az[i] = 0;
  for (int c = lane; c < r; c += WAVE) w.qacc[c] = 0.0;
  if (lane == 0) w.prm[r] = make_int4(0, 0, 0, NO_END);
  if (DYNAMIC && threadIdx.x == 0) next_slot = WPW;
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
  bool complete = true;
  // Items sorted by decreasing work: the three components of a block go to
  // neighbouring waves and run together.
  while (item < total) {
    const ItemPrefetch it = nx;
    const int b = (int)it.d2.y;
    const int k = (int)it.d2.z;
    const int len = (int)it.d0.z;
    const uint64_t tbase = (uint64_t)it.d0.x | ((uint64_t)it.d0.y << 32);
    const uint32_t l0 = it.d1.z, l1 = it.d1.w;
    // the item's first token chunks and long run, in flight during the
    // segment set-up and the forward DCT below
    MarkedStream<NT> parser(B.tokens + tbase, B.len_bits + (size_t)it.d2.w * (CH / 8), len);
    int next_lr = l0 < l1 ? (int)B.lr[l0] : NO_END;
    float inv_scale = 0.0f, S = 0.0f;
    if (UNPROJ) {
      const float bound = __uint_as_float(it.d2.x) * qmax * 1.0001f;
      S = bound > 0.0f ? 1073741824.0f / bound : 0.0f;
      inv_scale = S > 0.0f ? 1.0f / S : 0.0f;
    }
    // segment end offsets: inclusive prefix sum of the lengths (segments are
    // back to back in the block stream)
    int carry = 0;
#pragma unroll
    for (int m = 0; m < QREG; m++) {
      const int c = lane + WAVE * m;
      const int L = (int)(it.lk[m] & 0xffffu);
      const int incl = wave_scan_add(L) + carry;
      carry = readlane(incl, WAVE - 1);
      if (c < r)
        w.prm[c] = make_int4((int)(it.lk[m] >> 16) * (NDIAG * 4), __float_as_int(qf[m] * it.sv[m] * S),
                             __float_as_int(it.sv[m]), incl);
    }
    if (PROJ) forward_dct_rows(it.rows, zig, w.pfx, w.vz);
    wave_lds_sync();

    uint32_t li = l0;
    float open_sum = 0.0f;  // project sum of segment cur's tokens in earlier chunks (wave-uniform)
    int cur = 0;            // first segment not yet closed
    // Each chunk runs in three stages: parse (tokens -> values and
    // positions), compute (table reads, products, unproject atomics, long
    // runs, the project prefix sums) and close (the project dots of the
    // segments that end in the chunk).  The next chunk is parsed between a
    // chunk's compute and its close, so the parse's arithmetic covers the
    // latency of the atomics and prefix writes that the close waits for.
    // Chunks alternate between the parser's two token buffers, so the loads
    // of chunk c+2 are not waited for until chunk c+2.
    auto parse = [&](const int base, auto par, Slots<NT>& sl) {
      constexpr int PAR = decltype(par)::value;
      if (parser.plain(base))
        parser.template parse_slots<true, PAR>(base, sl);
      else
        parser.template parse_slots<false, PAR>(base, sl);
    };
    // {bits(1/s), end} of the segments cur, cur+1, ... (lane j: cur + j)
    int2 se = make_int2(0, NO_END);
    // Returns the chunk's project sum (wave-uniform).
    auto compute = [&](const int base, const Slots<NT>& sl) -> float {
      // Phases: all table reads first, then the arithmetic, then the atomics,
      // so that the LDS round trips of the NT slots overlap.
      float acc = 0.0f, pre[NT];
      {
        int2 pr[NT];
        int jj[NT], d4[NT];
        float vzv[NT];
#pragma unroll
        for (int i = 0; i < NT; i++) {
          jj[i] = sl.pos[i] & (BC - 1);
          pr[i] = *reinterpret_cast<const int2*>(w.prm + (sl.pos[i] >> 9));
          d4[i] = dz4[jj[i]];
          vzv[i] = PROJ ? w.vz[jj[i]] : 0.0f;
        }
        float dm[NT];
#pragma unroll
        for (int i = 0; i < NT; i++) dm[i] = *reinterpret_cast<const float*>(lds + (pr[i].x + d4[i]));
        int qv[NT];
#pragma unroll
        for (int i = 0; i < NT; i++) {
          const float vf = (float)sl.val[i] * dm[i];
          if (PROJ) acc += vf * vzv[i];
          pre[i] = acc;
          qv[i] = (int)(vf * __int_as_float(pr[i].y));
        }
        if (UNPROJ) {
#pragma unroll
          // an atomic of zero changes nothing, so a zero term (the tokens of
          // zero runs) goes to the lane's own entry instead of its position:
          // in one slot the 64 lanes' positions lie about 8 apart and share
          // few banks, entry `lane` shares none
          for (int i = 0; i < NT; i++) atomicAdd(&w.az[qv[i] != 0 ? jj[i] : lane], qv[i]);
        }
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
        const int2 pr = *reinterpret_cast<const int2*>(w.prm + (pv >> 9));
        float ex = 0.0f;
        for (int m = lane; m < Lv - 3; m += WAVE) {
          const int j = (pv + 1 + m) & (BC - 1);
          const float vf = (float)vv * *reinterpret_cast<const float*>(lds + (pr.x + (int)dz4[j]));
          if (PROJ) ex += vf * w.vz[j];
          if (UNPROJ) atomicAdd(&w.az[j], (int)(vf * __int_as_float(pr.y)));
        }
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
      if (!PROJ) return 0.0f;
      const float incl = wave_scan_addf(acc);
      const float excl = incl - acc;
      float4* pf = reinterpret_cast<float4*>(w.pfx + NT * lane);
#pragma unroll
      for (int q = 0; q < NT / 4; q++)
        pf[q] = make_float4(excl + pre[4 * q], excl + pre[4 * q + 1], excl + pre[4 * q + 2], excl + pre[4 * q + 3]);
      se = *reinterpret_cast<const int2*>(&w.prm[min(cur + lane, r)].z);
      return readlanef(incl, WAVE - 1);
    };
    auto close = [&](const int base, const float chunk_sum) {
      if (!PROJ) return;
      wave_lds_sync();
      // segments cur, cur+1, ... whose last token lies in this chunk: the
      // dot of segment cur + j is the chunk-local prefix at its last token
      // minus the one at the last token of segment cur + j - 1 (lane j - 1,
      // or bv: the last segment closed before, or -open_sum for segment cur)
      float bv = -open_sum;
      for (;;) {
        const int c = cur + lane;
        const bool closes = se.y <= base + CH;
        const int pi = closes ? se.y - 1 - base : 0;
        const float pv = w.pfx[pi];
        const double q0 = closes ? w.qacc[c] : 0.0;
        const float prevv = __uint_as_float(wave_shr1(__float_as_uint(pv), __float_as_uint(bv)));
        if (closes) w.qacc[c] = q0 + (double)((pv - prevv) * __int_as_float(se.x));
        const int n = __popcll(__ballot(closes));
        if (n > 0) bv = readlanef(pv, n - 1);
        cur += n;
        if (n < WAVE) break;
        se = *reinterpret_cast<const int2*>(&w.prm[min(cur + lane, r)].z);
      }
      open_sum = chunk_sum - bv;
      wave_lds_sync();
    };
    // Pairs of chunks in the loop while a third chunk follows them, the last
    // one or two after it.  Every path into the loop's head has issued the
    // loads of both buffers in the same order (second buffer, then first),
    // so each parse waits for its own loads only: a conditional parse inside
    // the loop would make the wait-count pass wait for every load there.
    Slots<NT> sa, sb;
    parse(0, std::integral_constant<int, 0>(), sa);
    int base = 0;
    for (; base + 2 * CH < len; base += 2 * CH) {
      float cs = compute(base, sa);
      parser.template parse_slots<true, 1>(base + CH, sb);  // base + 2 CH < len: a whole chunk
      close(base, cs);
      cs = compute(base + CH, sb);
      parse(base + 2 * CH, std::integral_constant<int, 0>(), sa);
      close(base + CH, cs);
    }
    {
      float cs = compute(base, sa);
      const bool two = base + CH < len;
      if (two) parse(base + CH, std::integral_constant<int, 1>(), sb);
      close(base, cs);
      if (two) {
        cs = compute(base + CH, sb);
        close(base + CH, cs);
      }
    }
    complete = complete && parser.complete(r);
    // the next item (slots map to items in increasing order, so the first
    // slot past the end ends the wave); its descriptor, segment parameters
    // and block of v are in flight during this item's inverse DCT
    int nxt = item + nwaves;
    if (DYNAMIC) {
      uint32_t sl = 0;
      if (lane == 0) sl = atomicAdd(&next_slot, 1u);
      const int slot = (int)readlane(sl, 0);
      nxt = slot / WPW * nwaves + round_position<WPW, DIE_GROUPED>(wg, nwg, slot % WPW);
    }
    if (nxt < total) prefetch_item<PROJ>(B, v_in, nxt, nx);
    wave_lds_sync();

    if (UNPROJ) inverse_dct_block_mfma(field_out, B, k, b, zig, w.az, inv_scale, *mf);
    wave_lds_sync();
    item = nxt;
  }
  if (!complete && lane == 0) atomicOr(B.err, (uint32_t)ERR_INCOMPLETE);

  if (PROJ) {
    // combine the waves' FP64 accumulators in a fixed order; column-major
    // partials, so the reduction reads each column's row contiguously
    __syncthreads();
    for (int c = threadIdx.x; c < r; c += blockDim.x) {
      double s = 0.0;
      for (int wv = 0; wv < WPW; wv++)
        s += carve_wave(lds + step_shared_lds_bytes(nk) + (size_t)wv * step_wave_lds_bytes(r), r).qacc[c];
      partials[(size_t)c * nwg + wg] = s;
    }
  }
}

template <int MODE, bool MF_LDS>
__global__ void __launch_bounds__(STEP_WAVES * WAVE, 3)  // <= 168 VGPRs: three waves per SIMD
step_kernel(DevBasis B, const float* __restrict__ v_in, const double* __restrict__ q_in,
            float* __restrict__ field_out, const float* __restrict__ g_damp, const Tables* __restrict__ g_tab) {
  step_body<MODE, STEP_WAVES, true, true, MF_LDS>(B, v_in, q_in, field_out, g_damp, g_tab, B.partials, blockIdx.x,
                                                  gridDim.x);
}

// Several independent bases in one persistent launch: workgroups
// [wg_begin, wg_begin + wg_count) of the grid step entry e (entries sorted by
// wg_begin, contiguous).
// WPW waves per workgroup (a tile shape of the batch autotuner).  The
// second launch bound is HIP's minimum number of waves per SIMD: three,
// the register budget of the step body (168 VGPRs), for either shape.
template <int MODE, int WPW>
__global__ void __launch_bounds__(WPW * WAVE, 3)
batch_step_kernel(const BatchEntry* __restrict__ E, int n) {
  int e = 0;
  while (e + 1 < n && (int)blockIdx.x >= E[e + 1].wg_begin) e++;
  const BatchEntry& be = E[e];
  const DevBasis B = be.B;
  step_body<MODE, WPW>(B, be.v_in, be.q_in, be.field, be.damp, be.tab, be.partials, (int)blockIdx.x - be.wg_begin,
                       be.wg_count);
}

// Sum of row[0..n) in a fixed order (FP64): per thread a strided sum, then
// per wave a butterfly, then the four waves in order.  256 threads.
__device__ __forceinline__ double block_row_sum(const double* __restrict__ row, int n) {
  __shared__ double red[4];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += row[i];
  s = wave_sum(s);
  if (lane_id() == 0) red[threadIdx.x / WAVE] = s;
  __syncthreads();
  return (red[0] + red[1]) + (red[2] + red[3]);
}

// q of every entry: block (e, c) sums row c of entry e's partials.
__global__ void __launch_bounds__(256) batch_reduce_kernel(const BatchEntry* __restrict__ E, int n) {
  int e = 0;
  while (e + 1 < n && (int)blockIdx.x >= E[e + 1].q_begin) e++;
  const int c = (int)blockIdx.x - E[e].q_begin;
  const double s = block_row_sum(E[e].partials + (size_t)c * E[e].wg_count, E[e].wg_count);
  if (threadIdx.x == 0) E[e].q_out[c] = s;
}

// q[c] = sum over workgroups of partials[c][wg], in a fixed order (FP64).
__global__ void __launch_bounds__(256) reduce_q_kernel(const double* __restrict__ partials, int nrows,
                                                       double* __restrict__ q) {
  const double s = block_row_sum(partials + (size_t)blockIdx.x * nrows, nrows);
  if (threadIdx.x == 0) q[blockIdx.x] = s;
}

}  // namespace cfs

// ---------------------------------------------------------------------------
// Launch helpers (called from basis.cpp)
// ---------------------------------------------------------------------------
namespace cfs {

// largest r the step kernel supports (q accumulators live in registers)
int step_max_cols() { return MAX_COLS; }
// rows of the FP64 partial-q buffer a launch of `grid` workgroups needs
int step_partial_rows(int grid) { return grid; }

hipError_t launch_step(const DevBasis& B, int mode, const float* v_in, const double* q_in, double* q_out,
                       float* field_out, const float* d_damp, const void* d_tab, int grid, hipStream_t s) {
  const bool mfl = step_mf_in_lds(mode, B.r, B.nk);
  const size_t lds = step_lds_bytes(B.r, B.nk, STEP_WAVES) + (mfl ? sizeof(MfmaFrags) : 0);
  const Tables* tab = reinterpret_cast<const Tables*>(d_tab);
  const dim3 block(STEP_WAVES * WAVE);
  switch (mode) {
    case MODE_PROJECT:
      hipLaunchKernelGGL((step_kernel<MODE_PROJECT, false>), dim3(grid), block, lds, s, B, v_in, q_in, field_out,
                         d_damp, tab);
      break;
    case MODE_UNPROJECT:
      if (mfl)
        hipLaunchKernelGGL((step_kernel<MODE_UNPROJECT, true>), dim3(grid), block, lds, s, B, v_in, q_in, field_out,
                           d_damp, tab);
      else
        hipLaunchKernelGGL((step_kernel<MODE_UNPROJECT, false>), dim3(grid), block, lds, s, B, v_in, q_in, field_out,
                           d_damp, tab);
      break;
    default:
      if (mfl)
        hipLaunchKernelGGL((step_kernel<MODE_BOTH, true>), dim3(grid), block, lds, s, B, v_in, q_in, field_out, d_damp,
                           tab);
      else
        hipLaunchKernelGGL((step_kernel<MODE_BOTH, false>), dim3(grid), block, lds, s, B, v_in, q_in, field_out,
                           d_damp, tab);
      break;
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (mode & MODE_PROJECT) {
    hipLaunchKernelGGL(reduce_q_kernel, dim3(B.r), dim3(256), 0, s, B.partials, grid, q_out);
    e = hipGetLastError();
  }
  return e;
}

hipError_t step_occupancy(int mode, int r, int nk, int* wg_per_cu) {
  const bool mfl = step_mf_in_lds(mode, r, nk);
  const size_t lds = step_lds_bytes(r, nk, STEP_WAVES) + (mfl ? sizeof(MfmaFrags) : 0);
  const int threads = STEP_WAVES * WAVE;
  switch (mode) {
    case MODE_PROJECT:
      return hipOccupancyMaxActiveBlocksPerMultiprocessor(wg_per_cu, step_kernel<MODE_PROJECT, false>, threads, lds);
    case MODE_UNPROJECT:
      return mfl ? hipOccupancyMaxActiveBlocksPerMultiprocessor(wg_per_cu, step_kernel<MODE_UNPROJECT, true>, threads, lds)
                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(wg_per_cu, step_kernel<MODE_UNPROJECT, false>, threads, lds);
    default:
      return mfl ? hipOccupancyMaxActiveBlocksPerMultiprocessor(wg_per_cu, step_kernel<MODE_BOTH, true>, threads, lds)
                 : hipOccupancyMaxActiveBlocksPerMultiprocessor(wg_per_cu, step_kernel<MODE_BOTH, false>, threads, lds);
  }
}

template <int WPW>
hipError_t launch_batch_step_w(const BatchEntry* d_entries, int n, int mode, int grid, size_t lds, hipStream_t s) {
  switch (mode) {
    case MODE_PROJECT:
      hipLaunchKernelGGL((batch_step_kernel<MODE_PROJECT, WPW>), dim3(grid), dim3(WPW * WAVE), lds, s, d_entries, n);
      break;
    case MODE_UNPROJECT:
      hipLaunchKernelGGL((batch_step_kernel<MODE_UNPROJECT, WPW>), dim3(grid), dim3(WPW * WAVE), lds, s, d_entries, n);
      break;
    default:
      hipLaunchKernelGGL((batch_step_kernel<MODE_BOTH, WPW>), dim3(grid), dim3(WPW * WAVE), lds, s, d_entries, n);
      break;
  }
  return hipGetLastError();
}

hipError_t launch_batch_step(const BatchEntry* d_entries, int n, int mode, int grid, int qrows, size_t lds, int wpw,
                             hipStream_t s) {
  hipError_t e = wpw == 2 ? launch_batch_step_w<2>(d_entries, n, mode, grid, lds, s)
                          : launch_batch_step_w<4>(d_entries, n, mode, grid, lds, s);
  if (e != hipSuccess || !(mode & MODE_PROJECT)) return e;
  hipLaunchKernelGGL(batch_reduce_kernel, dim3(qrows), dim3(256), 0, s, d_entries, n);
  return hipGetLastError();
}

template <int WPW>
hipError_t batch_occupancy_w(int mode, size_t lds, int* wg_per_cu) {
  switch (mode) {
    case MODE_PROJECT:
      return hipOccupancyMaxActiveBlocksPerMultiprocessor(wg_per_cu, batch_step_kernel<MODE_PROJECT, WPW>, WPW * WAVE,
                                                          lds);
    case MODE_UNPROJECT:
      return hipOccupancyMaxActiveBlocksPerMultiprocessor(wg_per_cu, batch_step_kernel<MODE_UNPROJECT, WPW>,
                                                          WPW * WAVE, lds);
    default:
      return hipOccupancyMaxActiveBlocksPerMultiprocessor(wg_per_cu, batch_step_kernel<MODE_BOTH, WPW>, WPW * WAVE,
                                                          lds);
  }
}

hipError_t batch_occupancy(int mode, size_t lds, int wpw, int* wg_per_cu) {
  return wpw == 2 ? batch_occupancy_w<2>(mode, lds, wg_per_cu) : batch_occupancy_w<4>(mode, lds, wg_per_cu);
}

size_t step_lds_bytes_for(int r, int nk, int wpw) { return step_lds_bytes(r, nk, wpw); }

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

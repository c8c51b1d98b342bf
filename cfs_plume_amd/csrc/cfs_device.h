// cfs_device.h -- device-side layout of a compressed basis on the GPU.
//
// A basis (three component files in the reference format v1,
// COMPRESSION.cpp:2165-2253) is repacked once at load time:
//
//   * Tokens are regrouped "block-major": for component k and block b, the r
//     segments (k, b, c=0..r-1) are stored back to back, forming one "block
//     stream".  The file stores them column-major (all blocks of column 0,
//     then column 1, ...), which scatters a block's segments across the file.
//     Each block stream starts on a 16-byte boundary; the array ends with a
//     guard of zero tokens so full-width loads past a stream stay in bounds.
//   * Per segment, in the same block-major order: seg_lk = length | (4*gamma)
//     << 16 and seg_sinv = (float)(1/s); seg_off = offset inside the stream.
//   * The damping table (1+d)^(k/4) for d = u+v+w in [0,21] and k = 4*gamma in
//     [kmin, kmax], the range the basis uses (COMPRESSION_DATA.cpp:25-43
//     restricted to the 22 distinct bases), is rounded once from FP64 to FP32.  Table entries are stored times
//     2^-FACTOR_SHIFT and 1/s times 2^FACTOR_SHIFT: 22^32 = 9.4e42 exceeds the
//     FP32 range, the scaled pair does not, and their product is unchanged.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cfs {

constexpr int BS = 8;
constexpr int BC = 512;
constexpr int NDIAG = 22;      // distinct values of u+v+w
constexpr int STEP_TPL = 8;    // tokens per lane per chunk in the fused step
constexpr int DEC_TPL = 4;     // tokens per lane per chunk in the random-access decoders
constexpr int GUARD_TOKENS = 3 * 64 * STEP_TPL;  // covers the two-chunk prefetch
constexpr int MAX_SEG_TOKENS = 2 * BC;  // reference assert COMPRESSION.cpp:1206
constexpr int FACTOR_SHIFT = 64;

// error bit raised by the device decoders; the loader checks every segment
// before upload, so it signals corrupted device memory, not a bad file
enum : uint32_t {
  ERR_INCOMPLETE = 1u << 1,   // a stream did not decode to 512 values per segment
};

struct DevBasis {
  const int32_t* tokens;      // block streams + guard
  const uint64_t* blk_base;   // [3*nB] first token of block stream (k, b)
  const uint32_t* blk_len;    // [3*nB] tokens in block stream (k, b)
  const uint32_t* seg_lk;     // [3*nB*r] length | (k4 - kmin) << 16, block-major
  const uint32_t* seg_off;    // [3*nB*r] offset of segment inside its block stream
  const float* seg_sinv;      // [3*nB*r] 2^FACTOR_SHIFT / s
  const float* seg_vmax;      // [3*nB*r] max |int| * damp^gamma * 2^-FACTOR_SHIFT of the segment
  const float* blk_bound;     // [3*nB] sum over the block stream's segments of vmax/s
  const float* damp;          // [nk * 22] (1+d)^((kmin+k)/4) * 2^-FACTOR_SHIFT in FP32
  const uint32_t* hz_start;   // [3*nB+1] first hazard of block stream (k, b) in hz
  const uint32_t* hz;         // hazard offsets (segment-final singleton equal to the next token)
  const uint32_t* lr_start;   // [3*nB+1] first long run of block stream (k, b) in lr
  const uint32_t* lr;         // offsets of the length tokens of nonzero runs with L >= 4
  const uint32_t* seg_lr;     // [3*nB*r + 1] index in lr of the first long run of each segment (block-major)
  const uint8_t* len_bits;    // LEN-token bitmap of every block stream, each padded to whole
                              // 512-token chunks (64 bytes), plus two guard chunks
  const int32_t* order;       // [nB] blocks in descending work order
  const uint4* items;         // [3*nB][3] fused-step items in schedule order:
                              //   {base lo, base hi, len, k*nB+b}, {hz0, hz1, lr0, lr1}, {bits(blk_bound), b, k, first 64-byte chunk of len_bits}
  uint32_t* err;              // device error bits
  uint32_t* counter;          // [2] spare device words (kept zero)
  double* partials;           // [r][grid] per-workgroup FP64 partial q (column-major)
  int X, Y, Z;                // peeled dims
  int nbx, nby, nbz, nB;      // blocks per axis, total
  int r, nk;                  // columns, rows of the damping table (kmax - kmin + 1)
};

struct Tables;

// One basis of a batched step (batch.cpp): its device data, bound fields and
// its share of the persistent grid.
struct BatchEntry {
  DevBasis B;
  const float* v_in;     // full-grid field projected (FP32)
  const double* q_in;    // q unprojected
  float* field;          // full-grid field written (interior cells)
  double* q_out;         // q projected
  const float* damp;     // the basis's damping table
  const Tables* tab;     // the basis's small tables
  double* partials;      // [r][wg_count] per-workgroup partial q (column-major)
  int wg_begin, wg_count;
  int q_begin;           // first block of this entry in the batched reduce
  int pad;
};

// arguments of the random-access decode kernel (decode_kernels.hip)
struct DecodeArgs {
  int k0, nk3;      // components [k0, k0+nk3)
  int col0, ncol;   // columns [col0, col0+ncol)
  int b0, nblk;     // blocks [b0, b0+nblk)
  int point;        // DEC_POINT: in-block cell x + 8y + 64z
  void* out;
};

}  // namespace cfs

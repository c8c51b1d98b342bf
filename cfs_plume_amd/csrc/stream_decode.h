// stream_decode.h -- wavefront-parallel RLE parser for concatenated segments.
//
// Restates RunLengthDecodeBinary (reference COMPRESSION.cpp:1198-1283) for a
// whole "stream": segments laid back to back, each decoding to exactly 512
// values (the loader checks every segment before upload).  One wave consumes
// the stream in chunks of 64*NT tokens; lane l owns tokens
// [base + NT*l, base + NT*l + NT).
//
// Per chunk:
//   1. every token gets a transition function of the VALUE/DUP/LEN machine;
//   2. a wave scan of the composed functions gives every token's role;
//   3. VALUE tokens emit 1 value (singleton) or L values (run [v, v, L]);
//      a wave prefix sum of the emission counts gives the position P of each
//      VALUE token in the concatenated output, so segment = P >> 9 and
//      zigzag position = P & 511.
// Segment boundaries need no bookkeeping: a segment that ends in a run's
// length token leaves the machine in LEN, whose successor is VALUE anyway.
// Only a segment ending in a singleton equal to the next segment's first token
// would be misread as a run start; the loader lists those few boundaries
// ("hazards") and the parser reads them as singletons.
#pragma once
#include "cfs_device.h"
#include "wave_ops.h"

namespace cfs {

template <int NT>
struct Chunk {
  int v[NT];     // token value
  int P[NT];     // output position of a VALUE token (global over the stream)
  int em[NT];    // values emitted by the token (0 for DUP/LEN/out of range)
};

// Per-token view used by the fused step: every token contributes at most one
// value inline.  A run [v, v, L] at output position P is split over its three
// tokens: the VALUE token takes P, the DUP token P+1 and the LEN token P+2
// (when L > 2) -- so runs of length 2 and 3, the common ones, need no loop.
// The LEN token's count is L-2; values P+3 .. P+L-1 of nonzero runs with
// L >= 4 (listed by the loader) are added by the caller.
template <int NT>
struct Slots {
  int val[NT];   // value contributed at pos (0 if none)
  int pos[NT];   // output position of that value, global over the stream
};

// NT tokens per lane (a multiple of 4), chunks of 64*NT tokens.
template <int NT>
class StreamParser {
  static_assert(NT % 4 == 0, "tokens are loaded 16 bytes at a time");
  static constexpr int NV = NT / 4;   // int4 loads per lane per chunk
  static constexpr int CH = 64 * NT;  // tokens per chunk

 public:
  // tok must be 16-byte aligned; tokens outside [lo, hi) are ignored.
  // hz/nhz: sorted offsets (from tok) of segment-final singletons that equal
  // the next token; the parser treats them as singletons.
  // hz_bias is subtracted from every hazard offset (lists kept relative to
  // an earlier origin than tok).
  __device__ StreamParser(const int32_t* __restrict__ tok, int lo, int hi, const uint32_t* __restrict__ hz, int nhz,
                          int hz_bias = 0)
      : tok_(tok), lo_(lo), len_(hi), hz_(hz), nhz_(nhz), hz_bias_(hz_bias) {
    lane_ = lane_id();
#pragma unroll
    for (int q = 0; q < NV; q++) {
      cur_[q] = *reinterpret_cast<const int4*>(tok_ + NT * lane_ + 4 * q);
      next_[q] = *reinterpret_cast<const int4*>(tok_ + CH + NT * lane_ + 4 * q);
    }
    next_hz_ = nhz_ > 0 ? (int)hz_[0] - hz_bias_ : 0x7fffffff;
  }

  // True if the chunk at `base` lies inside [lo, hi - 1) and holds no hazard,
  // so parse<true> may skip the range and hazard tests (wave-uniform).
  __device__ __forceinline__ bool plain(int base) const {
    return base >= lo_ && base + CH < len_ && next_hz_ >= base + CH;
  }

  // Parses the chunk starting at token `base` (a multiple of 64*NT).
  template <bool PLAIN>
  __device__ __forceinline__ void parse(int base, Chunk<NT>& ch) {
    // Two chunks are in flight: this one and the next (whose first two tokens
    // complete lane 63's lookahead).  The token array ends with a guard of
    // zero tokens, so the prefetch of chunk base+2*CH stays in bounds.
    int t[NT + 2];
#pragma unroll
    for (int q = 0; q < NV; q++) {
      t[4 * q] = cur_[q].x;
      t[4 * q + 1] = cur_[q].y;
      t[4 * q + 2] = cur_[q].z;
      t[4 * q + 3] = cur_[q].w;
    }
    const uint32_t nx0 = readlane((uint32_t)next_[0].x, 0);
    const uint32_t nx1 = readlane((uint32_t)next_[0].y, 0);
    t[NT] = (int)wave_shl1((uint32_t)t[0], nx0);
    t[NT + 1] = (int)wave_shl1((uint32_t)t[1], nx1);
#pragma unroll
    for (int q = 0; q < NV; q++) {
      cur_[q] = next_[q];
      next_[q] = *reinterpret_cast<const int4*>(tok_ + base + 2 * CH + NT * lane_ + 4 * q);
    }
    const int tid0 = base + NT * lane_;

    // segment-final singletons equal to the next token (rare)
    uint32_t hzm = 0;
    while (!PLAIN && next_hz_ < base + CH) {
      const int o = next_hz_ - base;
      if (o >= 0 && lane_ == o / NT) hzm |= 1u << (o % NT);  // o < 0: before this range
      hz_i_++;
      next_hz_ = hz_i_ < nhz_ ? (int)hz_[hz_i_] - hz_bias_ : 0x7fffffff;
    }

    // transition functions and the lane's composed function
    uint32_t f[NT];
    bool eq[NT];
#pragma unroll
    for (int i = 0; i < NT; i++) {
      if (PLAIN) {
        eq[i] = t[i] == t[i + 1];
        f[i] = eq[i] ? FSM_EQ : FSM_NE;
      } else {
        const bool valid = tid0 + i >= lo_ && tid0 + i < len_;
        eq[i] = !((hzm >> i) & 1u) && (tid0 + i + 1 < len_) && t[i] == t[i + 1];
        f[i] = !valid ? FSM_ID : (eq[i] ? FSM_EQ : FSM_NE);
      }
    }
    uint32_t F = f[0];
#pragma unroll
    for (int i = 1; i < NT; i++) F = fsm_compose(f[i], F);

    // roles: the state is carried as the constant function that maps every
    // state to it, so one v_perm advances it (const_s o f = const_{f(s)}).
    const uint32_t Fin = wave_scan_fsm(F);
    const uint32_t Fex = wave_shr1(Fin, FSM_ID);
    uint32_t st = fsm_compose(Fex, state_);
    state_ = fsm_compose(readlane(Fin, WAVE - 1), state_);

    // emissions and positions
    int lane_sum = 0, pre[NT];
#pragma unroll
    for (int i = 0; i < NT; i++) {
      const bool isv = (st == 0u) && (PLAIN || f[i] != FSM_ID);
      ch.em[i] = isv ? (eq[i] ? t[i + 2] : 1) : 0;
      ch.v[i] = t[i];
      pre[i] = lane_sum;
      lane_sum += ch.em[i];
      st = fsm_compose(f[i], st);
    }
    const int incl = wave_scan_add(lane_sum);
    const int p0 = pos_ + incl - lane_sum;
    pos_ += readlane(incl, WAVE - 1);
#pragma unroll
    for (int i = 0; i < NT; i++) ch.P[i] = p0 + pre[i];
  }

  // Parses the chunk at `base` into per-token slots (see Slots).
  template <bool PLAIN>
  __device__ __forceinline__ void parse_slots(int base, Slots<NT>& s) {
    int t[NT + 2];
#pragma unroll
    for (int q = 0; q < NV; q++) {
      t[4 * q] = cur_[q].x;
      t[4 * q + 1] = cur_[q].y;
      t[4 * q + 2] = cur_[q].z;
      t[4 * q + 3] = cur_[q].w;
    }
    const uint32_t nx0 = readlane((uint32_t)next_[0].x, 0);
    const uint32_t nx1 = readlane((uint32_t)next_[0].y, 0);
    t[NT] = (int)wave_shl1((uint32_t)t[0], nx0);
    t[NT + 1] = (int)wave_shl1((uint32_t)t[1], nx1);
    // the token before this lane's first (lane 0: the previous chunk's last)
    const int tm1 = (int)wave_shr1((uint32_t)t[NT - 1], (uint32_t)prev_last_);
    prev_last_ = (int)readlane((uint32_t)t[NT - 1], WAVE - 1);
#pragma unroll
    for (int q = 0; q < NV; q++) {
      cur_[q] = next_[q];
      next_[q] = *reinterpret_cast<const int4*>(tok_ + base + 2 * CH + NT * lane_ + 4 * q);
    }
    const int tid0 = base + NT * lane_;

    uint32_t hzm = 0;
    while (!PLAIN && next_hz_ < base + CH) {
      const int o = next_hz_ - base;
      if (o >= 0 && lane_ == o / NT) hzm |= 1u << (o % NT);  // o < 0: before this range
      hz_i_++;
      next_hz_ = hz_i_ < nhz_ ? (int)hz_[hz_i_] - hz_bias_ : 0x7fffffff;
    }

    uint32_t f[NT];
    bool valid[NT];
#pragma unroll
    for (int i = 0; i < NT; i++) {
      if (PLAIN) {
        valid[i] = true;
        f[i] = t[i] == t[i + 1] ? FSM_EQ : FSM_NE;
      } else {
        valid[i] = tid0 + i >= lo_ && tid0 + i < len_;
        const bool eq = !((hzm >> i) & 1u) && (tid0 + i + 1 < len_) && t[i] == t[i + 1];
        f[i] = !valid[i] ? FSM_ID : (eq ? FSM_EQ : FSM_NE);
      }
    }
    uint32_t F = f[0];
#pragma unroll
    for (int i = 1; i < NT; i++) F = fsm_compose(f[i], F);
    // If every lane but lane 0 maps all three entry states to one state (the
    // machine resynchronises within NT tokens: ~98% of chunks of coded
    // bases), the inclusive scan equals F itself and is skipped.
    uint32_t Fin = F;
    const bool constant = fsm_compose(F, 0x03000000u) == F;
    if ((__ballot(!constant) & ~1ull) != 0) Fin = wave_scan_fsm(F);
    const uint32_t Fex = wave_shr1(Fin, FSM_ID);
    uint32_t st = fsm_compose(Fex, state_);
    state_ = fsm_compose(readlane(Fin, WAVE - 1), state_);

    // Counts are carried biased by +2 (VALUE/DUP tokens 3, LEN tokens L,
    // invalid tokens 2), which saves a subtraction per token; the bias is
    // removed once per lane and folded into the per-slot constant.
    int lane_sum = 0, pre[NT];
#pragma unroll
    for (int i = 0; i < NT; i++) {
      const int prev = i == 0 ? tm1 : t[i - 1];
      const bool len_tok = st == FSM_CONST_LEN;
      int cnt2 = len_tok ? t[i] : 3;
      int val = len_tok ? (t[i] > 2 ? prev : 0) : t[i];
      if (!PLAIN && !valid[i]) {
        cnt2 = 2;
        val = 0;
      }
      s.val[i] = val;
      pre[i] = lane_sum;
      lane_sum += cnt2;
      st = fsm_compose(f[i], st);
    }
    const int incl = wave_scan_add(lane_sum);
    const int p0 = pos_ + incl - lane_sum - 2 * NT * lane_;
    pos_ += readlane(incl, WAVE - 1) - 2 * NT * WAVE;
#pragma unroll
    for (int i = 0; i < NT; i++) s.pos[i] = p0 + pre[i] - 2 * i;
  }

  // Called after the last chunk: true if the stream produced nseg*512 values
  // and ended outside a run (always so for streams the loader accepted).
  __device__ bool complete(int nseg) const { return pos_ == nseg * BC && state_ == 0u; }

 private:
  const int32_t* __restrict__ tok_;
  int lo_, len_;
  const uint32_t* __restrict__ hz_;
  int nhz_;
  int hz_bias_;
  int lane_;
  int4 cur_[NV], next_[NV];
  int hz_i_ = 0, next_hz_;
  uint32_t state_ = 0u;  // constant function of the state entering the next chunk (0 = VALUE)
  int prev_last_ = 0;    // last token of the previous chunk (parse_slots)
  int pos_ = 0;          // output position entering the next chunk
};

// The fused step's parser for one block stream whose roles were marked at
// load: bit (o & 7) of byte len_bits[o >> 3] is set iff token o is the length
// token of a run [v, v, L] (the loader's canonical parse of every segment,
// RunLengthDecodeBinary, COMPRESSION.cpp:1198-1283).  Tokens and their marks
// are still read every step; the marks replace the per-token role machine
// and the hazard list, so a token's role costs one bit test.  Lane l owns
// tokens [base + 8l, base + 8l + 8) and reads their marks as one byte; chunks
// are 64-byte aligned in the bitmap.  Same Slots output as
// StreamParser::parse_slots.
template <int NT>
class MarkedStream {
  static_assert(NT == 8, "one mark byte per lane");
  static constexpr int NV = NT / 4;
  static constexpr int CH = 64 * NT;

 public:
  // tok: 16-byte aligned block stream of `len` tokens followed by at least
  // one chunk of readable tokens; marks: its LEN bitmap, likewise padded.
  __device__ MarkedStream(const int32_t* __restrict__ tok, const uint8_t* __restrict__ marks, int len)
      : tok_(tok), marks_(marks), len_(len) {
    lane_ = lane_id();
    mark_shift_ = 8u * ((uint32_t)lane_ & 3u);
    last_ = len > 0 ? (len - 1) / CH * CH : 0;
    // first chunk's marks and tokens, then the second's: the first chunk
    // waits only for its own loads
    m_[0] = mark_word(marks_);
#pragma unroll
    for (int q = 0; q < NV; q++) buf_[0][q] = *reinterpret_cast<const int4*>(tok_ + NT * lane_ + 4 * q);
    __builtin_amdgcn_sched_barrier(0);
    m_[1] = mark_word(marks_ + CH / 8);
#pragma unroll
    for (int q = 0; q < NV; q++) buf_[1][q] = *reinterpret_cast<const int4*>(tok_ + CH + NT * lane_ + 4 * q);
  }
  // True if every token of the chunk at `base` belongs to the stream.
  __device__ __forceinline__ bool plain(int base) const { return base + CH <= len_; }

  // The 32-bit word of a chunk's marks (64 bytes, one per lane) that holds
  // this lane's byte.  Loaded as a word, the marks need no zero-extension:
  // extending a byte load let the compiler widen the marks at the end of
  // the previous chunk and wait for their load there.
  __device__ __forceinline__ uint32_t mark_word(const uint8_t* __restrict__ chunk_marks) const {
    return reinterpret_cast<const uint32_t*>(chunk_marks)[(uint32_t)lane_ >> 2];
  }

  // PAR: the buffer holding this chunk (chunks alternate, starting with 0).
  template <bool PLAIN, int PAR>
  __device__ __forceinline__ void parse_slots(int base, Slots<NT>& s) {
    int t[NT];
#pragma unroll
    for (int q = 0; q < NV; q++) {
      t[4 * q] = buf_[PAR][q].x;
      t[4 * q + 1] = buf_[PAR][q].y;
      t[4 * q + 2] = buf_[PAR][q].z;
      t[4 * q + 3] = buf_[PAR][q].w;
    }
    // this lane's byte of the marks word (the bits above it are not read)
    const uint32_t m = m_[PAR] >> mark_shift_;
    // the token before this lane's first (lane 0: the previous chunk's last)
    const int tm1 = (int)wave_shr1((uint32_t)t[NT - 1], (uint32_t)prev_last_);
    prev_last_ = (int)readlane((uint32_t)t[NT - 1], WAVE - 1);
    const int tid0 = base + NT * lane_;

    // Counts are selected biased by +2 (VALUE/DUP tokens 3, LEN tokens L,
    // tokens past the stream 2), which saves a subtraction per token; the
    // bias is removed once per lane and folded into the per-slot constant.
    int cnt[NT];
#pragma unroll
    for (int i = 0; i < NT; i++) {
      const int prev = i == 0 ? tm1 : t[i - 1];
      // all ones for a length token: the two selects below are bitfield
      // inserts with this mask, with no compare to make a lane mask
      const int msk = __builtin_amdgcn_sbfe((int)m, i, 1);
      int cnt2 = bit_select(msk, t[i], 3);
      int val = bit_select(msk, t[i] > 2 ? prev : 0, t[i]);
      if (!PLAIN && tid0 + i >= len_) {
        cnt2 = 2;
        val = 0;
      }
      s.val[i] = val;
      cnt[i] = cnt2;
    }
    // the lane's total as a tree of three-input adds (the positions follow
    // the scan as a chain, one add per slot)
    const int lane_sum = ((cnt[0] + cnt[1] + cnt[2]) + (cnt[3] + cnt[4] + cnt[5])) + (cnt[6] + cnt[7]);
    // chunk base + 2*CH into this buffer, issued once its tokens are no
    // longer read so that the loads may reuse their registers (a wave-uniform
    // base address plus 32-bit lane offsets)
    // The marks' load goes first: the compiler widens them at the end of the
    // previous chunk, and a wait for them there then need not also wait for
    // the tokens (the memory counter is in order).
    // chunk base + 2*CH, or the stream's last chunk again when that lies
    // past the end (a cache hit instead of a read of tokens nobody uses)
    const int nb = min(base + 2 * CH, last_);
    m_[PAR] = mark_word(marks_ + nb / 8);
    const int32_t* tk = tok_ + nb;
#pragma unroll
    for (int q = 0; q < NV; q++) buf_[PAR][q] = *reinterpret_cast<const int4*>(tk + (uint32_t)(NT * lane_ + 4 * q));
    const int incl = wave_scan_add(lane_sum);
    const int p0 = pos_ + incl - lane_sum - 2 * NT * lane_;
    pos_ += readlane(incl, WAVE - 1) - 2 * NT * WAVE;
    s.pos[0] = p0;
#pragma unroll
    for (int i = 1; i < NT; i++) s.pos[i] = s.pos[i - 1] + cnt[i - 1] - 2;
  }

  // Called after the last chunk: true if the stream produced nseg*512 values.
  __device__ bool complete(int nseg) const { return pos_ == nseg * BC; }

 private:
  const int32_t* __restrict__ tok_;
  const uint8_t* __restrict__ marks_;
  int len_;
  int last_;  // first token of the stream's last chunk
  int lane_;
  int4 buf_[2][NV];  // tokens of the chunks at base and base + CH
  uint32_t m_[2];         // marks words of the chunks in the two buffers
  uint32_t mark_shift_;   // 8 * (lane % 4): this lane's byte in its marks word
  int prev_last_ = 0;
  int pos_ = 0;
};

}  // namespace cfs

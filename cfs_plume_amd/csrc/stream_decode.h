// stream_decode.h -- wavefront-parallel RLE parser for concatenated segments.
//
// Restates RunLengthDecodeBinary (reference COMPRESSION.cpp:1198-1283) for a
// whole "stream": segments laid back to back, each decoding to exactly 512
// values (the loader checks every segment before upload).  One wave consumes
// the stream in chunks of 64*TPL tokens; lane l owns tokens
// [base + TPL*l, base + TPL*l + TPL).
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
// ("hazards") and the parser resets the machine there.
#pragma once
#include "cfs_device.h"
#include "wave_ops.h"

namespace cfs {

struct Chunk {
  int v[TPL];     // token value
  int P[TPL];     // output position of a VALUE token (global over the stream)
  int em[TPL];    // values emitted by the token (0 for DUP/LEN/out of range)
};

class StreamParser {
 public:
  // tok must be 16-byte aligned; tokens outside [lo, hi) are ignored.
  // hz/nhz: sorted offsets (from tok) of segment-final tokens where the
  // machine must be reset explicitly (see the header comment).
  __device__ StreamParser(const int32_t* __restrict__ tok, int lo, int hi, const uint32_t* __restrict__ hz, int nhz)
      : tok_(tok), lo_(lo), len_(hi), hz_(hz), nhz_(nhz) {
    lane_ = lane_id();
    cur_ = *reinterpret_cast<const int4*>(tok_ + TPL * lane_);
    next_ = *reinterpret_cast<const int4*>(tok_ + CHUNK + TPL * lane_);
    next_hz_ = nhz_ > 0 ? (int)hz_[0] : 0x7fffffff;
  }

  // Parses the chunk starting at token `base` (a multiple of CHUNK).
  __device__ __forceinline__ void parse(int base, Chunk& ch) {
    // Two chunks are in flight: this one and the next (whose first two tokens
    // complete lane 63's lookahead).  The token array ends with GUARD_TOKENS
    // zeros, so the prefetch of chunk base+2*CHUNK stays in bounds.
    const int4 cur = cur_;
    const uint32_t nx0 = readlane((uint32_t)next_.x, 0);
    const uint32_t nx1 = readlane((uint32_t)next_.y, 0);
    cur_ = next_;
    next_ = *reinterpret_cast<const int4*>(tok_ + base + 2 * CHUNK + TPL * lane_);
    int t[TPL + 2];
    t[0] = cur.x;
    t[1] = cur.y;
    t[2] = cur.z;
    t[3] = cur.w;
    t[4] = (int)wave_shl1((uint32_t)cur.x, nx0);
    t[5] = (int)wave_shl1((uint32_t)cur.y, nx1);
    const int tid0 = base + TPL * lane_;

    // explicit resets (rare): segment-final singletons equal to the next token
    uint32_t resetm = 0;
    while (next_hz_ < base + CHUNK) {
      const int o = next_hz_ - base;
      if (lane_ == (o >> 2)) resetm |= 1u << (o & 3);
      hz_i_++;
      next_hz_ = hz_i_ < nhz_ ? (int)hz_[hz_i_] : 0x7fffffff;
    }

    // transition functions and the lane's composed function
    uint32_t f[TPL];
    bool eq[TPL], valid[TPL];
#pragma unroll
    for (int i = 0; i < TPL; i++) {
      valid[i] = tid0 + i >= lo_ && tid0 + i < len_;
      const bool reset = (resetm >> i) & 1u;
      eq[i] = !reset && (tid0 + i + 1 < len_) && t[i] == t[i + 1];
      f[i] = !valid[i] ? FSM_ID : (reset ? FSM_RESET : (eq[i] ? FSM_EQ : FSM_NE));
    }
    uint32_t F = f[0];
#pragma unroll
    for (int i = 1; i < TPL; i++) F = fsm_compose(f[i], F);

    // roles
    const uint32_t Fin = wave_scan_fsm(F);
    const uint32_t Fex = wave_shr1(Fin, FSM_ID);
    uint32_t s[TPL];
    s[0] = fsm_apply(Fex, state_);
#pragma unroll
    for (int i = 1; i < TPL; i++) s[i] = fsm_apply(f[i - 1], s[i - 1]);
    state_ = fsm_apply(readlane(Fin, WAVE - 1), state_);

    // emissions and positions
    int lane_sum = 0, pre[TPL];
#pragma unroll
    for (int i = 0; i < TPL; i++) {
      const bool isv = valid[i] && s[i] == 0;
      ch.em[i] = isv ? (eq[i] ? t[i + 2] : 1) : 0;
      ch.v[i] = t[i];
      pre[i] = lane_sum;
      lane_sum += ch.em[i];
    }
    const int incl = wave_scan_add(lane_sum);
    const int p0 = pos_ + incl - lane_sum;
    pos_ += readlane(incl, WAVE - 1);
#pragma unroll
    for (int i = 0; i < TPL; i++) ch.P[i] = p0 + pre[i];
  }

  // Called after the last chunk: true if the stream produced nseg*512 values
  // and ended outside a run (always so for streams the loader accepted).
  __device__ bool complete(int nseg) const { return pos_ == nseg * BC && state_ == 0; }

 private:
  const int32_t* __restrict__ tok_;
  int lo_, len_;
  const uint32_t* __restrict__ hz_;
  int nhz_;
  int lane_;
  int4 cur_, next_;
  int hz_i_ = 0, next_hz_;
  uint32_t state_ = 0;   // machine state entering the next chunk
  int pos_ = 0;          // output position entering the next chunk
};

}  // namespace cfs

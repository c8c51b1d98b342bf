// stream_decode.h -- wavefront-parallel RLE parser for concatenated segments.
//
// Restates RunLengthDecodeBinary (reference COMPRESSION.cpp:1198-1283) for a
// whole "stream": segments laid back to back, each decoding to exactly 512
// values.  One wave consumes the stream in chunks of 64*TPL tokens; lane l
// owns tokens [base + TPL*l, base + TPL*l + TPL).
//
// Per chunk:
//   1. segment ends inside the chunk become per-lane "last" bits;
//   2. every token gets a transition function of the VALUE/DUP/LEN machine
//      (the last token of a segment resets the machine to VALUE);
//   3. a wave scan of the composed functions gives every token's role;
//   4. VALUE tokens emit 1 value (singleton) or L values (run [v, v, L]);
//      a wave prefix sum of the emission counts gives the position P of each
//      VALUE token in the concatenated output, so segment = P >> 9 and
//      zigzag position = P & 511.
// Malformed streams raise ERR_* bits instead of reading or writing out of
// bounds: run lengths outside [2,512], segments that do not decode to exactly
// 512 values, runs cut off by the end of a segment, positions past the last
// segment.
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
  // tok must be 16-byte aligned; tokens outside [lo, hi) are ignored and the
  // ends (inclusive, in LDS) are offsets from tok.
  __device__ StreamParser(const int32_t* __restrict__ tok, int lo, int hi, const int* ends_lds, int nseg)
      : tok_(tok), lo_(lo), len_(hi), ends_(ends_lds), nseg_(nseg) {
    lane_ = lane_id();
    next_ = *reinterpret_cast<const int4*>(tok_ + TPL * lane_);
  }

  // Parses the chunk starting at token `base` (a multiple of CHUNK).
  __device__ __forceinline__ void parse(int base, Chunk& ch) {
    const int4 cur = next_;
    // Prefetch the next chunk.  The token array ends with GUARD_TOKENS zeros,
    // so this load stays in bounds even past the end of the last stream.
    next_ = *reinterpret_cast<const int4*>(tok_ + base + CHUNK + TPL * lane_);
    const uint32_t nx0 = readlane((uint32_t)next_.x, 0);
    const uint32_t nx1 = readlane((uint32_t)next_.y, 0);
    int t[TPL + 2];
    t[0] = cur.x;
    t[1] = cur.y;
    t[2] = cur.z;
    t[3] = cur.w;
    t[4] = (int)wave_shl1((uint32_t)cur.x, nx0);
    t[5] = (int)wave_shl1((uint32_t)cur.y, nx1);
    const int tid0 = base + TPL * lane_;

    // 1. "last token of a segment" bits for this lane's tokens.
    uint32_t lastm = 0;
    for (;;) {
      const int c = c_next_ + lane_;
      const int e = (c < nseg_) ? ends_[c] : 0x7fffffff;
      const bool in = e < base + CHUNK;
      uint64_t m = __ballot(in);
      const int cnt = __popcll(m);
      const int o = e - base;
      while (m) {
        const int l = __ffsll((long long)m) - 1;
        m &= m - 1;
        const int oo = readlane(o, l);
        if (lane_ == (oo >> 2)) lastm |= 1u << (oo & 3);
      }
      c_next_ += cnt;
      if (cnt < WAVE) break;
    }

    // 2. transition functions and the lane's composed function.
    uint32_t f[TPL];
    bool eq[TPL], valid[TPL];
#pragma unroll
    for (int i = 0; i < TPL; i++) {
      valid[i] = tid0 + i >= lo_ && tid0 + i < len_;
      const bool last = (lastm >> i) & 1u;
      eq[i] = valid[i] && !last && (tid0 + i + 1 < len_) && t[i] == t[i + 1];
      f[i] = !valid[i] ? FSM_ID : (last ? FSM_RESET : (eq[i] ? FSM_EQ : FSM_NE));
    }
    uint32_t F = f[0];
#pragma unroll
    for (int i = 1; i < TPL; i++) F = fsm_compose(f[i], F);

    // 3. roles.
    const uint32_t Fin = wave_scan_fsm(F);
    const uint32_t Fex = wave_shr1(Fin, FSM_ID);
    uint32_t s[TPL];
    s[0] = fsm_apply(Fex, state_);
#pragma unroll
    for (int i = 1; i < TPL; i++) s[i] = fsm_apply(f[i - 1], s[i - 1]);
    state_ = fsm_apply(readlane(Fin, WAVE - 1), state_);

    // 4. emissions and positions.
    int lane_sum = 0, pre[TPL];
#pragma unroll
    for (int i = 0; i < TPL; i++) {
      const bool isv = valid[i] && s[i] == 0;
      const int L = t[i + 2];
      const bool bad = isv && eq[i] && (L < 2 || L > BC);
      if (bad) err_ |= ERR_RUN_LENGTH;
      if (((lastm >> i) & 1u) && s[i] == 1) err_ |= ERR_TRUNCATED;
      ch.em[i] = isv ? ((eq[i] && !bad) ? L : 1) : 0;
      ch.v[i] = t[i];
      pre[i] = lane_sum;
      lane_sum += ch.em[i];
    }
    const int incl = wave_scan_add(lane_sum);
    const int p0 = pos_ + incl - lane_sum;
    pos_ += readlane(incl, WAVE - 1);
#pragma unroll
    for (int i = 0; i < TPL; i++) {
      ch.P[i] = p0 + pre[i];
      if (ch.em[i] > 0 && (ch.P[i] >> 9) >= nseg_) err_ |= ERR_SEG_INDEX;
      // A segment's last token must close a multiple of 512 values; with the
      // total checked in finish(), every segment then holds exactly 512.
      if ((lastm >> i) & 1u) {
        if (((ch.P[i] + ch.em[i]) & (BC - 1)) != 0) err_ |= ERR_SEG_SIZE;
      } else if (ch.em[i] > 0 && (ch.P[i] & (BC - 1)) + ch.em[i] > BC) {
        err_ |= ERR_SEG_SIZE;  // a run crossing into the next segment
      }
    }
  }

  // Called after the last chunk: every segment must have produced 512 values.
  // Returns the error bits of the whole wave (every lane gets the same value).
  __device__ uint32_t finish() {
    if (pos_ != nseg_ * BC || state_ != 0) err_ |= ERR_SEG_SIZE;
    uint32_t all = 0;
    for (uint32_t bit = 1; bit <= ERR_SEG_INDEX; bit <<= 1)
      if (__ballot((err_ & bit) != 0)) all |= bit;
    return all;
  }
  __device__ uint32_t errors() const { return err_; }

 private:
  const int32_t* __restrict__ tok_;
  int lo_, len_;
  const int* ends_;
  int nseg_;
  int lane_;
  int4 next_;
  int c_next_ = 0;       // first segment whose end has not been passed
  uint32_t state_ = 0;   // machine state entering the next chunk
  int pos_ = 0;          // output position entering the next chunk
  uint32_t err_ = 0;
};

}  // namespace cfs

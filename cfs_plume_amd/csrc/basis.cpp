// basis.cpp -- host side of the C ABI (include/cfs_plume.h): component-file
// parsing and validation, block-major repacking, upload, kernel dispatch.
//
// Parsing follows ReadBinaryFileToMemory (reference COMPRESSION.cpp:1456-1594):
// header int32[10] = version, nBits, X, Y, Z, Px, Py, Pz, r, nB; then sList
// (f64), gammaList (f64), blockLengths (i32), blockIndices (i32), each nB x r
// column-major (entry (b, c) at b + nB*c); then int32 tokens to the end of the
// file.  Where the reference asserts or silently reads out of range, the
// loader rejects the file with CFS_ERR_FORMAT.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <memory>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cfs_plume.h"
#include "basis_internal.h"
#include "cfs_device.h"

namespace cfs {
// kernels (step_kernels.hip, decode_kernels.hip)
hipError_t launch_step(const DevBasis& B, int mode, const float* v_in, const double* q_in, double* q_out,
                       float* field_out, const float* d_damp, const void* d_tab, int grid, hipStream_t s);
hipError_t step_occupancy(int mode, int r, int nk, int* wg_per_cu);
int step_max_cols();
int step_partial_rows(int grid);
size_t tables_bytes();
void fill_tables(void* host_tab, const float* dct, const uint16_t* zig, const uint16_t* rev, const uint8_t* dz);
void fill_mfma_frags(void* host_tab);
hipError_t launch_decode(const DevBasis& B, int mode, const DecodeArgs& a, const float* d_damp, const void* d_tab,
                         hipStream_t s);
hipError_t launch_untransform_svd(float* cols, long cells, int ncols, const double* V, hipStream_t s);
hipError_t launch_accumulate_columns(const float* cols, int ncols, long n3, const double* q, float* field, int X,
                                     int Y, int Z, int first, hipStream_t s);
}  // namespace cfs

using namespace cfs;

static thread_local std::string g_err;

namespace cfs {
int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}
int hip_fail(hipError_t e, const char* what) {
  return fail(CFS_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}
}  // namespace cfs

namespace {

int pad_of(int n) { return (BS - (n % BS)) % BS; }

struct FileView {
  int32_t nBits, X, Y, Z, Px, Py, Pz, r, nB;
  const double* s;
  const double* gamma;
  const int32_t* len;
  const int32_t* idx;
  const int32_t* tok;
  size_t ntok;
};

int parse_file(const uint8_t* d, size_t n, int comp, FileView* f) {
  const std::string who = "component " + std::to_string(comp) + ": ";
  if (n < 40) return fail(CFS_ERR_FORMAT, who + "truncated header");
  int32_t h[10];
  std::memcpy(h, d, 40);
  if (h[0] != 1)
    return fail(CFS_ERR_FORMAT, who + "unsupported version " + std::to_string(h[0]) +
                                    " (version 0 files store paddings in place of padded dims)");
  f->nBits = h[1];
  f->X = h[2]; f->Y = h[3]; f->Z = h[4];
  f->Px = h[5]; f->Py = h[6]; f->Pz = h[7];
  f->r = h[8]; f->nB = h[9];
  if (f->nBits < 1 || f->nBits > 32) return fail(CFS_ERR_FORMAT, who + "nBits out of range");
  if (f->X < 1 || f->Y < 1 || f->Z < 1) return fail(CFS_ERR_FORMAT, who + "bad dims");
  // dims are positive int32, so X + 7 and the padded dims fit in int64
  if ((int64_t)f->Px != (int64_t)f->X + pad_of(f->X) || (int64_t)f->Py != (int64_t)f->Y + pad_of(f->Y) ||
      (int64_t)f->Pz != (int64_t)f->Z + pad_of(f->Z))
    return fail(CFS_ERR_FORMAT, who + "padded dims do not match dims");
  // (P/8)^3 < 2^84 can overflow 64 bits: multiply step by step against the int32 nB
  {
    const uint64_t bx = (uint64_t)(f->Px / BS), by = (uint64_t)(f->Py / BS), bz = (uint64_t)(f->Pz / BS);
    const uint64_t lim = 0x7fffffffull;
    const bool fits = bx * by <= lim && bx * by * bz <= lim;
    if (f->nB < 1 || !fits || (uint64_t)f->nB != bx * by * bz)
      return fail(CFS_ERR_FORMAT, who + "numBlocks does not match padded dims");
  }
  if (f->r < 1) return fail(CFS_ERR_FORMAT, who + "numCols <= 0");
  if (f->r > step_max_cols())
    return fail(CFS_ERR_FORMAT, who + "numCols " + std::to_string(f->r) + " exceeds the supported maximum " +
                                    std::to_string(step_max_cols()));
  // r <= 256 and nB < 2^31, so m * 24 < 2^44 cannot wrap; the size check
  // divides instead of multiplying anyway
  const size_t m = (size_t)f->r * (size_t)f->nB;
  if ((n - 40) / 24 < m) return fail(CFS_ERR_FORMAT, who + "truncated metadata or token stream");
  const size_t meta = 40 + m * 24;
  if ((n - meta) % 4 != 0) return fail(CFS_ERR_FORMAT, who + "truncated metadata or token stream");
  // the arrays are 8- or 4-byte aligned inside the image because the header is 40 bytes
  f->s = reinterpret_cast<const double*>(d + 40);
  f->gamma = f->s + m;
  f->len = reinterpret_cast<const int32_t*>(f->gamma + m);
  f->idx = f->len + m;
  f->tok = f->idx + m;
  f->ntok = (n - meta) / 4;
  for (size_t i = 0; i < m; i++) {
    const int32_t L = f->len[i], o = f->idx[i];
    if (L < 1 || L > MAX_SEG_TOKENS) return fail(CFS_ERR_FORMAT, who + "segment length outside [1, 1024]");
    if (o < 0 || (size_t)o + (size_t)L > f->ntok) return fail(CFS_ERR_FORMAT, who + "segment past end of stream");
    const double g4 = 4.0 * f->gamma[i];
    if (!(g4 >= 0.0) || g4 > 4.0 * f->nBits || g4 != std::floor(g4))
      return fail(CFS_ERR_FORMAT, who + "gamma is not a quarter-integer in [0, nBits]");
    if (std::isnan(f->s[i]) || f->s[i] == 0.0) return fail(CFS_ERR_FORMAT, who + "scale s is NaN or zero");
  }
  return CFS_OK;
}

void parallel_for(size_t n, const std::function<void(size_t, size_t)>& fn) {
  unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (n < 4096) nt = 1;
  std::vector<std::thread> th;
  const size_t per = (n + nt - 1) / nt;
  for (unsigned t = 0; t < nt; t++) {
    const size_t a = t * per, b = std::min(n, a + per);
    if (a >= b) break;
    th.emplace_back([&fn, a, b] { fn(a, b); });
  }
  for (auto& x : th) x.join();
}

// Canonical RunLengthDecodeBinary parse (COMPRESSION.cpp:1198-1283) of one
// segment, checking what the reference only asserts (:1206-1207, :1267) or
// silently tolerates: run lengths in [2, 512], no run cut off by the end of
// the segment, exactly 512 values.  Returns nullptr if the segment is valid.
// Also returns vmax = max |int| * (1+u+v+w)^gamma over the segment, the bound
// of its dequantised values times s (damp_row: the 22 powers for its gamma).
const char* check_segment(const int32_t* t, int len, const double* damp_row, const uint8_t* dz,
                          bool* ends_in_singleton, double* vmax, std::vector<uint32_t>* long_runs,
                          uint32_t seg_off) {
  int i = 0, j = 0;
  bool single = false;
  double m = 0.0;
  while (i < len) {
    if (j >= BC) return "decodes to more than 512 values";
    const double a = std::fabs((double)t[i]);
    if (i + 1 < len && t[i] == t[i + 1]) {
      if (i + 2 >= len) return "ends inside a run";
      const int L = t[i + 2];
      if (L < 2 || L > BC) return "run length outside [2, 512]";
      if (j + L > BC) return "decodes to more than 512 values";
      if (a > 0) {
        for (int q = j; q < j + L; q++) m = std::max(m, a * damp_row[dz[q]]);
        if (L >= 4) long_runs->push_back(seg_off + (uint32_t)i + 2);  // the length token
      }
      j += L;
      i += 3;
      single = false;
    } else {
      m = std::max(m, a * damp_row[dz[j]]);
      j++;
      i++;
      single = true;
    }
  }
  if (j != BC) return "decodes to fewer than 512 values";
  *ends_in_singleton = single;
  *vmax = m;
  return nullptr;
}

// Marks the length token of every run of a segment that check_segment
// accepted: bit (o & 7) of bits[o >> 3] for stream offset o.  Segments of one
// block stream are marked by one thread, so plain byte writes are safe.
void mark_len_tokens(const int32_t* t, int len, uint32_t seg_off, uint8_t* bits) {
  for (int i = 0; i < len;) {
    if (i + 1 < len && t[i] == t[i + 1]) {
      const uint32_t o = seg_off + (uint32_t)i + 2;
      bits[o >> 3] |= (uint8_t)(1u << (o & 7));
      i += 3;
    } else {
      i++;
    }
  }
}

template <class T>
int dev_upload(T** dst, const T* src, size_t count, size_t* total) {
  const size_t bytes = std::max<size_t>(count * sizeof(T), 4);
  hipError_t e = hipMalloc(reinterpret_cast<void**>(dst), bytes);
  if (e != hipSuccess) return hip_fail(e, "hipMalloc");
  *total += bytes;
  if (src && count) {
    e = hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy");
  }
  return CFS_OK;
}

// Host copies of the tables every basis uses.
struct HostTables {
  uint16_t zig[BC], rev[BC];
  uint8_t dz[BC];
  float dct[64];
  HostTables() {
    int j = 0;
    for (int sum = 0; sum < 3 * BS; sum++)
      for (int z = 0; z < BS; z++)
        for (int y = 0; y < BS; y++)
          for (int x = 0; x < BS; x++)
            if (x + y + z == sum) zig[x + BS * y + BS * BS * z] = (uint16_t)j++;
    for (int i = 0; i < BC; i++) rev[zig[i]] = (uint16_t)i;
    for (int j2 = 0; j2 < BC; j2++) {
      const int i = rev[j2];
      dz[j2] = (uint8_t)((i & 7) + ((i >> 3) & 7) + (i >> 6));
    }
    for (int k = 0; k < BS; k++)
      for (int n = 0; n < BS; n++) {
        const double a = (k == 0) ? std::sqrt(0.125) : 0.5;
        dct[k * BS + n] = (float)(a * std::cos((M_PI * (double)((2 * n + 1) * k)) / 16.0));
      }
  }
};
const HostTables& host_tables() {
  static HostTables t;
  return t;
}

}  // namespace

struct cfs_basis {
  int device = 0;
  int nBits = 0, X = 0, Y = 0, Z = 0, Px = 0, Py = 0, Pz = 0, r = 0, nB = 0, nk = 0, kmin = 0;
  int64_t ntok[3] = {0, 0, 0};
  int64_t nhazards = 0;
  int64_t nlong = 0;  // nonzero runs with L >= 4
  int grid[4] = {0, 0, 0, 0};  // persistent grid per mode (index = mode)
  size_t device_bytes = 0;
  DevBasis dev{};
  float* d_damp = nullptr;
  void* d_tab = nullptr;
  std::vector<void*> allocs;
  ~cfs_basis() {
    for (void* p : allocs) (void)hipFree(p);
  }
};

namespace cfs {
int basis_device(const cfs_basis* b) { return b->device; }
const DevBasis& basis_dev(const cfs_basis* b) { return b->dev; }
const float* basis_damp(const cfs_basis* b) { return b->d_damp; }
const void* basis_tables(const cfs_basis* b) { return b->d_tab; }
int basis_num_blocks(const cfs_basis* b) { return b->nB; }
int64_t basis_total_tokens(const cfs_basis* b) { return b->ntok[0] + b->ntok[1] + b->ntok[2]; }
}  // namespace cfs

namespace {

int build_basis(const uint8_t* const bufs[3], const size_t sizes[3], cfs_basis** out) {
  if (!out) return fail(CFS_ERR_ARG, "null output handle");
  *out = nullptr;
  FileView f[3];
  for (int k = 0; k < 3; k++) {
    if (!bufs[k]) return fail(CFS_ERR_ARG, "null component buffer");
    int st = parse_file(bufs[k], sizes[k], k, &f[k]);
    if (st) return st;
  }
  for (int k = 1; k < 3; k++)
    if (f[k].nBits != f[0].nBits || f[k].X != f[0].X || f[k].Y != f[0].Y || f[k].Z != f[0].Z || f[k].r != f[0].r ||
        f[k].nB != f[0].nB)
      return fail(CFS_ERR_FORMAT, "component headers disagree");

  if (3.0 * (f[0].X + 2.0) * (f[0].Y + 2.0) * (f[0].Z + 2.0) >= 2147483647.0)
    return fail(CFS_ERR_FORMAT, "grid too large: a full-grid field must hold fewer than 2^31 floats");

  std::unique_ptr<cfs_basis> B(new cfs_basis);
  B->nBits = f[0].nBits;
  B->X = f[0].X; B->Y = f[0].Y; B->Z = f[0].Z;
  B->Px = f[0].Px; B->Py = f[0].Py; B->Pz = f[0].Pz;
  B->r = f[0].r; B->nB = f[0].nB;
  const int r = B->r, nB = B->nB;
  const size_t nseg = (size_t)3 * nB * r;

  // block-major layout: stream (k, b) holds segments (k, b, 0..r-1)
  std::vector<uint64_t> blk_base((size_t)3 * nB);
  std::vector<uint32_t> blk_len((size_t)3 * nB);
  std::vector<uint32_t> seg_lk(nseg), seg_off(nseg);
  std::vector<float> seg_sinv(nseg), seg_vmax(nseg), blk_bound((size_t)3 * nB);
  // damping-table rows actually used: k4 in [kmin, kmax]
  int kmin = 4 * B->nBits, kmax = 0;
  for (int k = 0; k < 3; k++)
    for (size_t i = 0; i < (size_t)nB * r; i++) {
      const int k4 = (int)(4.0 * f[k].gamma[i]);
      kmin = std::min(kmin, k4);
      kmax = std::max(kmax, k4);
    }
  B->kmin = kmin;
  B->nk = kmax - kmin + 1;
  uint64_t pos = 0;
  for (int k = 0; k < 3; k++) {
    B->ntok[k] = (int64_t)f[k].ntok;
    for (int b = 0; b < nB; b++) {
      const size_t kb = (size_t)k * nB + b;
      blk_base[kb] = pos;
      uint64_t off = 0;
      for (int c = 0; c < r; c++) {
        const size_t e2 = (size_t)b + (size_t)nB * c;  // file order
        const size_t si = kb * r + c;                   // device order
        const uint32_t L = (uint32_t)f[k].len[e2];
        const uint32_t k4 = (uint32_t)(4.0 * f[k].gamma[e2]) - (uint32_t)kmin;
        seg_lk[si] = L | (k4 << 16);
        seg_off[si] = (uint32_t)off;
        seg_sinv[si] = (float)(std::ldexp(1.0 / f[k].s[e2], FACTOR_SHIFT));
        off += L;
      }
      if (off > 0xffffffffull) return fail(CFS_ERR_FORMAT, "block stream too long");
      blk_len[kb] = (uint32_t)off;
      pos += (off + 3) & ~(uint64_t)3;
    }
  }
  const uint64_t total_tok = pos + GUARD_TOKENS;
  std::vector<int32_t> tokens;
  try {
    tokens.assign(total_tok, 0);
  } catch (...) {
    return fail(CFS_ERR_NOMEM, "host allocation for repacked tokens failed");
  }
  parallel_for((size_t)3 * nB, [&](size_t a, size_t bnd) {
    for (size_t kb = a; kb < bnd; kb++) {
      const int k = (int)(kb / nB), b = (int)(kb % nB);
      int32_t* dst = tokens.data() + blk_base[kb];
      for (int c = 0; c < r; c++) {
        const size_t e2 = (size_t)b + (size_t)nB * c;
        const int32_t L = f[k].len[e2];
        std::memcpy(dst, f[k].tok + f[k].idx[e2], (size_t)L * 4);
        dst += L;
      }
    }
  });
  // LEN-token bitmaps of the fused step: block stream kb owns 64-byte chunks
  // [mark_chunk[kb], mark_chunk[kb+1]) (one per 512 tokens), plus two guard
  // chunks for the parser's prefetch.
  constexpr int MARK_CHUNK = 64 * STEP_TPL;
  std::vector<uint64_t> mark_chunk((size_t)3 * nB + 1, 0);
  for (size_t kb = 0; kb < (size_t)3 * nB; kb++)
    mark_chunk[kb + 1] = mark_chunk[kb] + (blk_len[kb] + MARK_CHUNK - 1) / MARK_CHUNK;
  if (mark_chunk[(size_t)3 * nB] > 0xffffffffull) return fail(CFS_ERR_FORMAT, "too many token chunks");
  std::vector<uint8_t> len_bits;
  try {
    len_bits.assign((size_t)(mark_chunk[(size_t)3 * nB] + 2) * (MARK_CHUNK / 8), 0);
  } catch (...) {
    return fail(CFS_ERR_NOMEM, "host allocation for LEN-token bitmaps failed");
  }
  // Check every segment; record the "hazard" boundaries of each block stream:
  // a segment ending in a singleton equal to the next segment's first token.
  std::vector<std::vector<uint32_t>> hz_lists((size_t)3 * nB), lr_lists((size_t)3 * nB);
  std::vector<std::string> bad((size_t)3 * nB);
  std::vector<double> dpow((size_t)(4 * B->nBits + 1) * NDIAG);
  for (int k = 0; k <= 4 * B->nBits; k++)
    for (int d = 0; d < NDIAG; d++) dpow[(size_t)k * NDIAG + d] = std::pow((double)(1 + d), k / 4.0);
  const uint8_t* dzt = host_tables().dz;
  parallel_for((size_t)3 * nB, [&](size_t a, size_t bnd) {
    for (size_t kb = a; kb < bnd && bad[kb].empty(); kb++) {
      const int k = (int)(kb / nB), b = (int)(kb % nB);
      const int32_t* stream = tokens.data() + blk_base[kb];
      for (int c = 0; c < r; c++) {
        const size_t si = kb * r + c;
        const int L = (int)(seg_lk[si] & 0xffffu);
        bool single = false;
        double vmax = 0.0;
        const char* why = check_segment(stream + seg_off[si], L, dpow.data() + (size_t)((seg_lk[si] >> 16) + B->kmin) * NDIAG,
                                        dzt, &single, &vmax, &lr_lists[kb], seg_off[si]);
        if (why) {
          bad[kb] = "component " + std::to_string(k) + " block " + std::to_string(b) + " column " +
                    std::to_string(c) + ": segment " + why;
          break;
        }
        mark_len_tokens(stream + seg_off[si], L, seg_off[si], len_bits.data() + mark_chunk[kb] * (MARK_CHUNK / 8));
        if (single && c + 1 < r && stream[seg_off[si] + L - 1] == stream[seg_off[si] + L])
          hz_lists[kb].push_back(seg_off[si] + (uint32_t)L - 1);
        // in the scaled units of the device damping table, rounded up
        seg_vmax[si] = (float)(std::ldexp(vmax, -FACTOR_SHIFT) * (1.0 + 1e-6));
      }
    }
  });
  for (const auto& m : bad)
    if (!m.empty()) return fail(CFS_ERR_FORMAT, m);
  // per block stream: sum_c vmax_c / s_c, the bound of |sum_c q_c D_c[j]| / max|q|
  for (size_t kb = 0; kb < (size_t)3 * nB; kb++) {
    double acc = 0.0;
    for (int c = 0; c < r; c++) acc += (double)seg_vmax[kb * r + c] * (double)seg_sinv[kb * r + c];
    blk_bound[kb] = (float)(acc * (1.0 + 1e-5));
  }
  std::vector<uint32_t> hz_start((size_t)3 * nB + 1, 0), hz;
  for (size_t kb = 0; kb < (size_t)3 * nB; kb++) {
    hz_start[kb] = (uint32_t)hz.size();
    hz.insert(hz.end(), hz_lists[kb].begin(), hz_lists[kb].end());
  }
  hz_start[(size_t)3 * nB] = (uint32_t)hz.size();
  B->nhazards = (int64_t)hz.size();
  // length tokens of nonzero runs longer than 3 (their values past P+2 are
  // added outside the fused step's per-token pass)
  std::vector<uint32_t> lr_start((size_t)3 * nB + 1, 0), lr;
  for (size_t kb = 0; kb < (size_t)3 * nB; kb++) {
    lr_start[kb] = (uint32_t)lr.size();
    lr.insert(lr.end(), lr_lists[kb].begin(), lr_lists[kb].end());
  }
  lr_start[(size_t)3 * nB] = (uint32_t)lr.size();
  B->nlong = (int64_t)lr.size();
  // per segment: index of its first long run (segments are block-major, as lr)
  std::vector<uint32_t> seg_lr(nseg + 1);
  parallel_for((size_t)3 * nB, [&](size_t a, size_t bnd) {
    for (size_t kb = a; kb < bnd; kb++) {
      uint32_t li = lr_start[kb];
      for (int c = 0; c < r; c++) {
        const size_t si = kb * r + c;
        while (li < lr_start[kb + 1] && lr[li] < seg_off[si]) li++;
        seg_lr[si] = li;
      }
    }
  });
  seg_lr[nseg] = (uint32_t)lr.size();

  // work order: blocks with the longest streams first
  std::vector<int32_t> order(nB);
  std::iota(order.begin(), order.end(), 0);
  std::vector<uint64_t> work(nB);
  for (int b = 0; b < nB; b++) work[b] = (uint64_t)blk_len[b] + blk_len[(size_t)nB + b] + blk_len[(size_t)2 * nB + b];
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work[a] > work[b]; });

  // per-item descriptors of the fused step in schedule order (item i: block
  // order[i / 3], component i % 3), 12 words each
  std::vector<uint32_t> items((size_t)3 * nB * 12, 0);
  for (size_t it = 0; it < (size_t)3 * nB; it++) {
    const int b = order[it / 3], k = (int)(it % 3);
    const size_t kb = (size_t)k * nB + b;
    uint32_t* d = items.data() + it * 12;
    d[0] = (uint32_t)blk_base[kb];
    d[1] = (uint32_t)(blk_base[kb] >> 32);
    d[2] = blk_len[kb];
    d[3] = (uint32_t)kb;
    d[4] = hz_start[kb];
    d[5] = hz_start[kb + 1];
    d[6] = lr_start[kb];
    d[7] = lr_start[kb + 1];
    std::memcpy(&d[8], &blk_bound[kb], 4);
    d[9] = (uint32_t)b;
    d[10] = (uint32_t)k;
    d[11] = (uint32_t)mark_chunk[kb];
  }

  // damping table (1+d)^(k/4), FP64 -> FP32 (COMPRESSION_DATA.cpp:25-43),
  // scaled by 2^-FACTOR_SHIFT (and 1/s by 2^FACTOR_SHIFT) so that 22^32 fits FP32
  std::vector<float> damp((size_t)B->nk * NDIAG);
  for (int k = 0; k < B->nk; k++)
    for (int d = 0; d < NDIAG; d++)
      damp[(size_t)k * NDIAG + d] = (float)std::ldexp(std::pow((double)(1 + d), (k + kmin) / 4.0), -FACTOR_SHIFT);
  std::vector<uint8_t> tab(tables_bytes());
  const HostTables& T = host_tables();
  fill_tables(tab.data(), T.dct, T.zig, T.rev, T.dz);
  fill_mfma_frags(tab.data());

  // persistent grid sizes
  hipError_t e = hipGetDevice(&B->device);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  int cus = 0;
  e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, B->device);
  if (e != hipSuccess) return hip_fail(e, "hipDeviceGetAttribute");
  int maxgrid = 1;
  for (int mode = 1; mode <= 3; mode++) {
    int per = 0;
    e = step_occupancy(mode, r, B->nk, &per);
    if (e != hipSuccess) return hip_fail(e, "hipOccupancyMaxActiveBlocksPerMultiprocessor");
    if (per < 1) return fail(CFS_ERR_ARG, "numCols too large for the step kernel's LDS budget");
    B->grid[mode] = std::max(1, std::min(cus * per, 3 * nB));
    maxgrid = std::max(maxgrid, B->grid[mode]);
  }

  size_t tot = 0;
  DevBasis& D = B->dev;
  int32_t* d_tok = nullptr;
  uint64_t* d_bb = nullptr;
  uint32_t *d_bl = nullptr, *d_lk = nullptr, *d_off = nullptr, *d_err = nullptr, *d_cnt = nullptr;
  uint32_t *d_hzs = nullptr, *d_hz = nullptr, *d_lrs = nullptr, *d_lr = nullptr, *d_items = nullptr,
           *d_seglr = nullptr;
  float* d_vmax = nullptr;
  float* d_bound = nullptr;
  float *d_sinv = nullptr, *d_damp = nullptr;
  int32_t* d_order = nullptr;
  double* d_part = nullptr;
  uint8_t* d_tab = nullptr;
  uint8_t* d_marks = nullptr;
  int st = CFS_OK;
#define UP(ptr, src, cnt)                        \
  do {                                           \
    st = dev_upload(&ptr, src, cnt, &tot);       \
    if (ptr) B->allocs.push_back((void*)ptr);    \
    if (st) return st;                           \
  } while (0)
  UP(d_tok, tokens.data(), tokens.size());
  UP(d_bb, blk_base.data(), blk_base.size());
  UP(d_bl, blk_len.data(), blk_len.size());
  UP(d_lk, seg_lk.data(), seg_lk.size());
  UP(d_off, seg_off.data(), seg_off.size());
  UP(d_sinv, seg_sinv.data(), seg_sinv.size());
  UP(d_vmax, seg_vmax.data(), seg_vmax.size());
  UP(d_bound, blk_bound.data(), blk_bound.size());
  UP(d_damp, damp.data(), damp.size());
  UP(d_tab, tab.data(), tab.size());
  UP(d_order, order.data(), order.size());
  UP(d_hzs, hz_start.data(), hz_start.size());
  UP(d_hz, hz.data(), hz.size());
  UP(d_lrs, lr_start.data(), lr_start.size());
  UP(d_lr, lr.data(), lr.size());
  UP(d_items, items.data(), items.size());
  UP(d_seglr, seg_lr.data(), seg_lr.size());
  UP(d_marks, len_bits.data(), len_bits.size());
  const uint32_t zeros[4] = {0, 0, 0, 0};
  UP(d_err, zeros, 1);
  UP(d_cnt, zeros, 2);
  UP(d_part, (const double*)nullptr, (size_t)step_partial_rows(maxgrid) * r);
#undef UP
  D.tokens = d_tok;
  D.blk_base = d_bb;
  D.blk_len = d_bl;
  D.seg_lk = d_lk;
  D.seg_off = d_off;
  D.seg_sinv = d_sinv;
  D.seg_vmax = d_vmax;
  D.blk_bound = d_bound;
  D.damp = d_damp;
  D.order = d_order;
  D.hz_start = d_hzs;
  D.hz = d_hz;
  D.lr_start = d_lrs;
  D.lr = d_lr;
  D.items = reinterpret_cast<const uint4*>(d_items);
  D.seg_lr = d_seglr;
  D.len_bits = d_marks;
  D.err = d_err;
  D.counter = d_cnt;
  D.partials = d_part;
  D.X = B->X; D.Y = B->Y; D.Z = B->Z;
  D.nbx = B->Px / BS; D.nby = B->Py / BS; D.nbz = B->Pz / BS;
  D.nB = nB; D.r = r; D.nk = B->nk;
  B->d_damp = d_damp;
  B->d_tab = d_tab;
  B->device_bytes = tot;
  *out = B.release();
  return CFS_OK;
}

int read_file(const char* path, std::vector<uint8_t>* buf) {
  if (!path) return fail(CFS_ERR_ARG, "null path");
  FILE* fp = std::fopen(path, "rb");
  if (!fp) return fail(CFS_ERR_IO, std::string("cannot open ") + path);
  std::fseek(fp, 0, SEEK_END);
  const long n = std::ftell(fp);
  std::fseek(fp, 0, SEEK_SET);
  if (n < 0) {
    std::fclose(fp);
    return fail(CFS_ERR_IO, std::string("cannot size ") + path);
  }
  buf->resize((size_t)n);
  const size_t got = n ? std::fread(buf->data(), 1, (size_t)n, fp) : 0;
  std::fclose(fp);
  if (got != (size_t)n) return fail(CFS_ERR_IO, std::string("short read on ") + path);
  return CFS_OK;
}

}  // namespace

extern "C" {

const char* cfs_last_error(void) { return g_err.c_str(); }
int cfs_abi_version(void) { return CFS_ABI_VERSION; }

int cfs_basis_open_memory(const void* bx, size_t nx, const void* by, size_t ny, const void* bz, size_t nz,
                          cfs_basis** out) {
  g_err.clear();
  const uint8_t* bufs[3] = {(const uint8_t*)bx, (const uint8_t*)by, (const uint8_t*)bz};
  const size_t sizes[3] = {nx, ny, nz};
  return build_basis(bufs, sizes, out);
}

int cfs_basis_open_files(const char* px, const char* py, const char* pz, cfs_basis** out) {
  g_err.clear();
  std::vector<uint8_t> b[3];
  const char* p[3] = {px, py, pz};
  for (int k = 0; k < 3; k++) {
    int st = read_file(p[k], &b[k]);
    if (st) return st;
  }
  const uint8_t* bufs[3] = {b[0].data(), b[1].data(), b[2].data()};
  const size_t sizes[3] = {b[0].size(), b[1].size(), b[2].size()};
  return build_basis(bufs, sizes, out);
}

int cfs_basis_close(cfs_basis* basis) {
  if (!basis) return fail(CFS_ERR_ARG, "null basis");
  DeviceGuard g(basis->device);
  delete basis;
  return CFS_OK;
}

int cfs_basis_get_info(const cfs_basis* B, cfs_basis_info* info) {
  if (!B || !info) return fail(CFS_ERR_ARG, "null argument");
  info->n_bits = B->nBits;
  info->dims[0] = B->X; info->dims[1] = B->Y; info->dims[2] = B->Z;
  info->padded[0] = B->Px; info->padded[1] = B->Py; info->padded[2] = B->Pz;
  info->num_cols = B->r;
  info->num_blocks = B->nB;
  info->hazards = (int32_t)B->nhazards;
  for (int k = 0; k < 3; k++) info->tokens[k] = B->ntok[k];
  info->device_bytes = (int64_t)B->device_bytes;
  return CFS_OK;
}

int cfs_project(cfs_basis* B, const float* v, double* q, void* stream) {
  if (!B || !v || !q) return fail(CFS_ERR_ARG, "null argument");
  DeviceGuard g(B->device);
  hipError_t e = launch_step(B->dev, 1, v, nullptr, q, nullptr, B->d_damp, B->d_tab, B->grid[1], (hipStream_t)stream);
  return e == hipSuccess ? CFS_OK : hip_fail(e, "project launch");
}

int cfs_unproject(cfs_basis* B, const double* q, float* field, void* stream) {
  if (!B || !q || !field) return fail(CFS_ERR_ARG, "null argument");
  DeviceGuard g(B->device);
  hipError_t e =
      launch_step(B->dev, 2, nullptr, q, nullptr, field, B->d_damp, B->d_tab, B->grid[2], (hipStream_t)stream);
  return e == hipSuccess ? CFS_OK : hip_fail(e, "unproject launch");
}

int cfs_step(cfs_basis* pre, cfs_basis* fin, const float* v_in, double* q_out, const double* q_in, float* field,
             void* stream) {
  if (!pre || !fin || !v_in || !q_out || !q_in || !field) return fail(CFS_ERR_ARG, "null argument");
  if (pre == fin) {
    DeviceGuard g(pre->device);
    hipError_t e = launch_step(pre->dev, 3, v_in, q_in, q_out, field, pre->d_damp, pre->d_tab, pre->grid[3],
                               (hipStream_t)stream);
    return e == hipSuccess ? CFS_OK : hip_fail(e, "step launch");
  }
  int st = cfs_project(pre, v_in, q_out, stream);
  if (st) return st;
  return cfs_unproject(fin, q_in, field, stream);
}

static int decode_common(cfs_basis* B, int mode, int k0, int nk3, int col0, int ncol, int b0, int nblk, int point,
                         void* out, void* stream) {
  if (!B || !out) return fail(CFS_ERR_ARG, "null argument");
  if (k0 < 0 || k0 + nk3 > 3 || nk3 < 1) return fail(CFS_ERR_ARG, "component out of range");
  if (col0 < 0 || ncol < 0 || col0 + ncol > B->r) return fail(CFS_ERR_ARG, "column range out of bounds");
  if (b0 < 0 || nblk < 0 || b0 + nblk > B->nB) return fail(CFS_ERR_ARG, "block range out of bounds");
  DeviceGuard g(B->device);
  DecodeArgs a{k0, nk3, col0, ncol, b0, nblk, point, out};
  hipError_t e = launch_decode(B->dev, mode, a, B->d_damp, B->d_tab, (hipStream_t)stream);
  return e == hipSuccess ? CFS_OK : hip_fail(e, "decode launch");
}

int cfs_decode_block_ints(cfs_basis* B, int comp, int col, int block0, int nblocks, int32_t* out, void* stream) {
  return decode_common(B, 0, comp, 1, col, 1, block0, nblocks, 0, out, stream);
}
int cfs_decode_block_coeffs(cfs_basis* B, int comp, int col, int block0, int nblocks, float* out, void* stream) {
  return decode_common(B, 1, comp, 1, col, 1, block0, nblocks, 0, out, stream);
}
int cfs_decode_columns(cfs_basis* B, int col0, int ncols, float* out, void* stream) {
  if (!B) return fail(CFS_ERR_ARG, "null basis");
  return decode_common(B, 2, 0, 3, col0, ncols, 0, B->nB, 0, out, stream);
}
int cfs_get_submatrix(cfs_basis* B, int start_row, double* out, void* stream) {
  if (!B) return fail(CFS_ERR_ARG, "null basis");
  const long long rows = 3LL * B->X * B->Y * B->Z;
  if (start_row < 0 || start_row % 3 != 0 || start_row >= rows)
    return fail(CFS_ERR_ARG, "start_row must be a multiple of 3 inside the matrix");
  // ComputeBlockNumber (COMPRESSION.cpp:1712-1749)
  const int index = start_row / 3;
  const int z = index / (B->X * B->Y);
  const int rem = index - B->X * B->Y * z;
  const int y = rem / B->X;
  const int x = rem - B->X * y;
  const int point = (x % BS) + BS * (y % BS) + BS * BS * (z % BS);
  const int nbx = B->Px / BS, nby = B->Py / BS;
  const int blk = x / BS + (y / BS) * nbx + (z / BS) * nbx * nby;
  return decode_common(B, 3, 0, 3, 0, B->r, blk, 1, point, out, stream);
}

int cfs_unproject_spatial(cfs_basis* B, const double* q, float* field, float* scratch, int group, void* stream) {
  if (!B || !q || !field) return fail(CFS_ERR_ARG, "null argument");
  if (group < 1) return fail(CFS_ERR_ARG, "group must be at least one column");
  const DevBasis& D = B->dev;
  const int r = D.r;
  group = std::min(group, r);
  const long n3 = 3L * D.X * D.Y * D.Z;
  DeviceGuard g(B->device);
  const hipStream_t s = (hipStream_t)stream;
  float* buf = scratch;
  if (!buf) {
    const hipError_t e = hipMallocAsync((void**)&buf, (size_t)group * n3 * sizeof(float), s);
    if (e != hipSuccess) return hip_fail(e, "hipMallocAsync");
  }
  int st = CFS_OK;
  for (int c0 = 0; c0 < r && st == CFS_OK; c0 += group) {
    const int n = std::min(group, r - c0);
    st = cfs_decode_columns(B, c0, n, buf, stream);
    if (st == CFS_OK) {
      const hipError_t e = launch_accumulate_columns(buf, n, n3, q + c0, field, D.X, D.Y, D.Z, c0 == 0, s);
      if (e != hipSuccess) st = hip_fail(e, "accumulate launch");
    }
  }
  if (!scratch) (void)hipFreeAsync(buf, s);
  return st;
}

int cfs_untransform_svd(float* columns, int ncols, int64_t cells, const double* V, void* stream) {
  if (ncols < 0 || cells < 0) return fail(CFS_ERR_ARG, "negative size");
  if (ncols == 0 || cells == 0) return CFS_OK;
  if (!columns || !V) return fail(CFS_ERR_ARG, "null argument");
  if (ncols > 65535) return fail(CFS_ERR_ARG, "more than 65535 columns in one call");
  hipError_t e = launch_untransform_svd(columns, (long)cells, ncols, V, (hipStream_t)stream);
  return e == hipSuccess ? CFS_OK : hip_fail(e, "untransform launch");
}

int cfs_check_stream_errors(cfs_basis* B, uint32_t* bits) {
  if (!B || !bits) return fail(CFS_ERR_ARG, "null argument");
  DeviceGuard g(B->device);
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return hip_fail(e, "hipDeviceSynchronize");
  uint32_t v = 0;
  e = hipMemcpy(&v, B->dev.err, 4, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpy");
  e = hipMemset(B->dev.err, 0, 4);
  if (e != hipSuccess) return hip_fail(e, "hipMemset");
  *bits = v;
  return CFS_OK;
}

}  // extern "C"

// cfs_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A plain FP64 CPU restatement of the reference codec (aadjones/CFS_PLUME,
// src/compression/*) used as the parity checker for the HIP path.  Only the
// test suite, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may
// load this library; the product never links or calls it.
//
// Every function cites the reference file:line whose behaviour it restates.
// The reference itself cannot be compiled in this image (fftw3.h, png.h and
// GL/glut.h are absent), so the forward/inverse DCT is restated from the
// closed form of FFTW's REDFT10/REDFT01 plus the reference's unitary
// normalisation (COMPRESSION.cpp:173-272) instead of calling FFTW.
//
// Parity pins (see tests/test_oracle_*.py):
//   zigzag / damping / RLE / block geometry / cumsum : the reference's own
//       known-answer tests (tests/test_compression_data.cpp:1087-1460,1743-1814)
//   DCT : scipy.fft.dctn(type=2, norm='ortho') golden vectors + brute force
//   encoder : invariants (quarter-integer gamma, energy criterion, ties-even)
//   project / unproject : dense FP64 U^T v / U q on the spatially decoded basis
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off, single-threaded).
#include <algorithm>
#include <atomic>
#include <cmath>
#include <functional>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

namespace {

constexpr int BS = 8;            // SETTINGS.h:4 BLOCK_SIZE
constexpr int BC = BS * BS * BS; // cells per block

// status codes shared with the Python side (oracle/oracle.py)
enum {
  ORC_OK = 0,
  ORC_ERR_IO = -1,
  ORC_ERR_FORMAT = -2,
  ORC_ERR_RANGE = -3,
  ORC_ERR_STREAM = -4,
};

// ---------------------------------------------------------------------------
// Tables
// ---------------------------------------------------------------------------

// COMPRESSION_DATA.cpp:90-120 set_zigzagArray: index j enumerates (x,y,z) by
// increasing x+y+z, then z, then y, then x; reverse[j] = x + 8y + 64z.
void zigzag_tables(int32_t* zig, int32_t* rev) {
  int j = 0;
  for (int sum = 0; sum < 3 * BS; sum++)
    for (int z = 0; z < BS; z++)
      for (int y = 0; y < BS; y++)
        for (int x = 0; x < BS; x++)
          if (x + y + z == sum) zig[x + BS * y + BS * BS * z] = j++;
  for (int i = 0; i < BC; i++) rev[zig[i]] = i;
}

struct Tables {
  int32_t zig[BC], rev[BC];
  double damp[BC];
  Tables() {
    zigzag_tables(zig, rev);
    // COMPRESSION_DATA.cpp:71-88 set_dampingArray: damp(u,v,w) = 1+u+v+w
    for (int w = 0; w < BS; w++)
      for (int v = 0; v < BS; v++)
        for (int u = 0; u < BS; u++) damp[u + BS * v + BS * BS * w] = 1 + u + v + w;
  }
};
const Tables& tables() {
  static Tables t;
  return t;
}

// ---------------------------------------------------------------------------
// DCT (COMPRESSION.cpp:173-342): orthonormal DCT-II / DCT-III along x, y, z.
// C[k][n] = a_k cos(pi (2n+1) k / 16), a_0 = sqrt(1/8), a_k = 1/2.
// ---------------------------------------------------------------------------
struct DctMatrix {
  double c[BS][BS];
  DctMatrix() {
    for (int k = 0; k < BS; k++)
      for (int n = 0; n < BS; n++) {
        double a = (k == 0) ? std::sqrt(0.125) : 0.5;
        c[k][n] = a * std::cos((M_PI * (double)((2 * n + 1) * k)) / 16.0);
      }
  }
};
const DctMatrix& dctm() {
  static DctMatrix m;
  return m;
}

// One 8-point pass along the axis with element stride `stride`; forward uses
// C, inverse uses C^T.  Sums run in increasing index order.
void dct_pass(double* b, int stride, int outer_a, int sa, int outer_b, int sb, bool forward) {
  const DctMatrix& m = dctm();
  for (int i = 0; i < outer_a; i++)
    for (int j = 0; j < outer_b; j++) {
      double* p = b + i * sa + j * sb;
      double in[BS], out[BS];
      for (int n = 0; n < BS; n++) in[n] = p[n * stride];
      for (int k = 0; k < BS; k++) {
        double s = 0.0;
        for (int n = 0; n < BS; n++) s += (forward ? m.c[k][n] : m.c[n][k]) * in[n];
        out[k] = s;
      }
      for (int k = 0; k < BS; k++) p[k * stride] = out[k];
    }
}

// UnitaryBlockDCTEigen (COMPRESSION.cpp:536-555) on one block, x-fastest.
void dct3(double* b, bool forward) {
  dct_pass(b, 1, BS, BS * BS, BS, BS, forward);       // along x: lines over (z, y)
  dct_pass(b, BS, BS, BS * BS, BS, 1, forward);       // along y: lines over (z, x)
  dct_pass(b, BS * BS, BS, BS, BS, 1, forward);       // along z: lines over (y, x)
}

// ---------------------------------------------------------------------------
// Geometry (COMPRESSION.cpp:352-505, FIELD_3D.cpp:6255-6309)
// ---------------------------------------------------------------------------
int pad_of(int n) { return (BS - (n % BS)) % BS; }  // GetPaddings :352-364

// GetBlocks (:371-404, edge replication) / GetBlocksEigen (:410-451, zeros).
// F is X*Y*Z x-fastest; blocks are written block-major, each x-fastest.
void get_blocks(const double* F, int X, int Y, int Z, bool zero_pad, double* blocks) {
  int Px = X + pad_of(X), Py = Y + pad_of(Y), Pz = Z + pad_of(Z);
  int nbx = Px / BS, nby = Py / BS, nbz = Pz / BS;
  for (int bz = 0; bz < nbz; bz++)
    for (int by = 0; by < nby; by++)
      for (int bx = 0; bx < nbx; bx++) {
        double* out = blocks + (size_t)(bx + nbx * (by + nby * bz)) * BC;
        for (int z = 0; z < BS; z++)
          for (int y = 0; y < BS; y++)
            for (int x = 0; x < BS; x++) {
              int gx = bx * BS + x, gy = by * BS + y, gz = bz * BS + z;
              double v;
              if (zero_pad) {
                v = (gx < X && gy < Y && gz < Z) ? F[gx + (size_t)X * (gy + (size_t)Y * gz)] : 0.0;
              } else {
                int cx = gx < X ? gx : X - 1, cy = gy < Y ? gy : Y - 1, cz = gz < Z ? gz : Z - 1;
                v = F[cx + (size_t)X * (cy + (size_t)Y * cz)];
              }
              out[x + BS * y + BS * BS * z] = v;
            }
      }
}

// AssimilateBlocksEigen (:483-505) followed by subfield(0..X,0..Y,0..Z).
void assimilate_unpadded(const double* blocks, int X, int Y, int Z, double* F) {
  int Px = X + pad_of(X), Py = Y + pad_of(Y);
  int nbx = Px / BS, nby = Py / BS;
  for (int z = 0; z < Z; z++)
    for (int y = 0; y < Y; y++)
      for (int x = 0; x < X; x++) {
        int b = x / BS + nbx * (y / BS + nby * (z / BS));
        F[x + (size_t)X * (y + (size_t)Y * z)] =
            blocks[(size_t)b * BC + (x % BS) + BS * (y % BS) + BS * BS * (z % BS)];
      }
}

// ---------------------------------------------------------------------------
// RLE (COMPRESSION.cpp:1112-1283)
// ---------------------------------------------------------------------------

// RunLengthEncodeBinary :1112-1192 -- emit v, and v,L for runs of length L>1.
int rle_encode(const int32_t* zz, int32_t* out) {
  int n = 0;
  for (int i = 0; i < BC; i++) {
    int32_t v = zz[i];
    out[n++] = v;
    int run = 1;
    while (i + 1 < BC && zz[i] == zz[i + 1]) {
      run++;
      i++;
    }
    if (run > 1) {
      out[n++] = v;
      out[n++] = run;
    }
  }
  return n;
}

// RunLengthDecodeBinary :1198-1283 (canonical semantics) with the reference's
// asserts (:1206-1207, :1267) turned into status codes.  Positions past the
// stream fill with zeros as in the reference (parsedData->setZero()); a stream
// that decodes to fewer than 512 values sets *underfull (still ORC_OK).
int rle_decode(const int32_t* tok, int len, int32_t* out, int* underfull) {
  if (len < 0 || len > 2 * BC) return ORC_ERR_STREAM;
  for (int j = 0; j < BC; j++) out[j] = 0;
  int i = 0, j = 0;
  while (i < len) {
    if (j >= BC) return ORC_ERR_STREAM;
    out[j] = tok[i];
    if (i + 1 < len && tok[i] == tok[i + 1]) {
      if (i + 2 >= len) return ORC_ERR_STREAM;
      int run = tok[i + 2];
      if (run <= 1 || run > BC) return ORC_ERR_STREAM;
      if (j + run > BC) return ORC_ERR_STREAM;
      for (int m = 1; m < run; m++) out[j + m] = tok[i];
      i += 3;
      j += run;
    } else {
      i++;
      j++;
    }
  }
  if (underfull) *underfull = (j < BC);
  return ORC_OK;
}

// ---------------------------------------------------------------------------
// Encoder (COMPRESSION.cpp:733-949, FIELD_3D.cpp:1517-1637,2422)
// ---------------------------------------------------------------------------
struct BlockCode {
  double s, gamma;
  std::vector<int32_t> tokens;
};

// PreprocessBlock :733-759 + TuneGammaQuantized :841-916 + EncodeBlock :924-949
// + ZigzagFlatten :1402-1421 + RunLengthEncodeBinary.  F is the forward DCT of
// one block; it is scaled in place as the reference does.
void encode_block(double* F, int nBits, double percent, int maxIterations, BlockCode* bc) {
  const Tables& T = tables();
  // PreprocessBlock: absMax (FIELD_3D.cpp:1543-1552) then scale
  double fmax = std::fabs(F[0]);
  for (int i = 0; i < BC; i++)
    if (std::fabs(F[i]) > fmax) fmax = std::fabs(F[i]);
  double s = (std::pow(2, nBits - 1) - 1) / fmax;
  for (int i = 0; i < BC; i++) F[i] *= s;
  bc->s = s;

  // TuneGammaQuantized
  double damp[BC];
  auto set_power = [&](double g) {
    for (int i = 0; i < BC; i++) damp[i] = std::pow(T.damp[i], g);  // toPower :2422
  };
  auto energy = [&]() {
    // ((damp) * trunc(F / damp)).sumSq()  (FIELD_3D.cpp:700,743,1632,1517)
    double e = 0;
    for (int i = 0; i < BC; i++) {
      double d = std::trunc(F[i] / damp[i]);
      double p = damp[i] * d;
      e += p * p;
    }
    return e;
  };
  double lower = 0.0, upper = nBits, eps = 0.000001, gamma = 0.0;
  set_power(gamma);
  double total = 0;
  for (int i = 0; i < BC; i++) total += F[i] * F[i];
  double diff = std::fabs(total - energy());
  double pe = 1.0 - (diff / total);
  int it = 0;
  while (std::fabs(percent - pe) > eps && it < maxIterations) {
    if (pe < percent) upper = gamma; else lower = gamma;
    gamma = 0.5 * (upper + lower);
    set_power(gamma);
    diff = std::fabs(total - energy());
    pe = 1.0 - (diff / total);
    it++;
  }
  gamma = std::rint(4 * gamma) / 4.0;
  set_power(gamma);
  bc->gamma = gamma;

  // RoundFieldToInt(F / damp) :22-30 (rint, FE_TONEAREST), zigzag, RLE
  int32_t q[BC], zz[BC], tok[3 * BC];
  for (int i = 0; i < BC; i++) q[i] = (int32_t)std::rint(F[i] / damp[i]);
  for (int i = 0; i < BC; i++) zz[T.zig[i]] = q[i];
  int n = rle_encode(zz, tok);
  bc->tokens.assign(tok, tok + n);
}

// ---------------------------------------------------------------------------
// Component file (format v1, COMPRESSION.cpp:1456-1594, :2165-2253)
// ---------------------------------------------------------------------------
struct Component {
  int nBits = 0, X = 0, Y = 0, Z = 0, Px = 0, Py = 0, Pz = 0, r = 0, nB = 0;
  std::vector<double> s, gamma;          // nB x r column-major
  std::vector<int32_t> len, idx;         // nB x r column-major
  std::vector<int32_t> tokens;
  std::vector<double> dlist;             // (4 nBits + 1) x 512 damping list
};

int parse_component(const uint8_t* d, size_t size, Component* c) {
  if (size < 40) return ORC_ERR_FORMAT;
  int32_t h[10];
  std::memcpy(h, d, 40);
  if (h[0] != 1) return ORC_ERR_FORMAT;  // v0 computes paddings wrongly :1506-1514
  c->nBits = h[1]; c->X = h[2]; c->Y = h[3]; c->Z = h[4];
  c->Px = h[5]; c->Py = h[6]; c->Pz = h[7]; c->r = h[8]; c->nB = h[9];
  if (c->r <= 0 || c->nB <= 0 || c->nBits <= 0 || c->nBits > 32) return ORC_ERR_FORMAT;
  size_t n = (size_t)c->r * c->nB;
  size_t meta = 40 + n * 24;
  if (size < meta) return ORC_ERR_FORMAT;
  c->s.resize(n); c->gamma.resize(n); c->len.resize(n); c->idx.resize(n);
  std::memcpy(c->s.data(), d + 40, n * 8);
  std::memcpy(c->gamma.data(), d + 40 + n * 8, n * 8);
  std::memcpy(c->len.data(), d + 40 + n * 16, n * 4);
  std::memcpy(c->idx.data(), d + 40 + n * 20, n * 4);
  size_t nt = (size - meta) / 4;
  c->tokens.assign(nt + 4, 0);  // +4: the in-place decoders read 2 ints past a stream
  std::memcpy(c->tokens.data(), d + meta, nt * 4);
  for (size_t i = 0; i < n; i++)
    if (c->len[i] < 0 || c->len[i] > 2 * BC || c->idx[i] < 0 || (size_t)c->idx[i] + c->len[i] > nt)
      return ORC_ERR_FORMAT;
  // set_dampingArrayList COMPRESSION_DATA.cpp:25-43
  const Tables& T = tables();
  int nk = 4 * c->nBits + 1;
  c->dlist.resize((size_t)nk * BC);
  for (int k = 0; k < nk; k++)
    for (int i = 0; i < BC; i++) c->dlist[(size_t)k * BC + i] = std::pow(T.damp[i], k / 4.0);
  return ORC_OK;
}

// ReadBinaryFileToMemory COMPRESSION.cpp:1456-1594
int read_component(const char* path, Component* c) {
  FILE* f = std::fopen(path, "rb");
  if (!f) return ORC_ERR_IO;
  std::fseek(f, 0, SEEK_END);
  long n = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  if (n < 0) { std::fclose(f); return ORC_ERR_IO; }
  std::vector<uint8_t> buf((size_t)n);
  size_t got = n ? std::fread(buf.data(), 1, (size_t)n, f) : 0;
  std::fclose(f);
  if (got != (size_t)n) return ORC_ERR_IO;
  return parse_component(buf.data(), buf.size(), c);
}

int write_component(const char* path, const Component& c) {
  FILE* f = std::fopen(path, "wb");
  if (!f) return ORC_ERR_IO;
  int32_t h[10] = {1, c.nBits, c.X, c.Y, c.Z, c.Px, c.Py, c.Pz, c.r, c.nB};
  size_t n = (size_t)c.r * c.nB;
  std::fwrite(h, 4, 10, f);
  std::fwrite(c.s.data(), 8, n, f);
  std::fwrite(c.gamma.data(), 8, n, f);
  std::fwrite(c.len.data(), 4, n, f);
  std::fwrite(c.idx.data(), 4, n, f);
  std::fwrite(c.tokens.data(), 4, c.tokens.size(), f);
  std::fclose(f);
  return ORC_OK;
}

struct Basis {
  Component comp[3];
};

// RunLengthDecodeBinary + ZigzagUnflatten (:1427-1445): ints in u+8v+64w order
int decode_ints(const Component& c, int b, int col, int32_t* out, int* underfull) {
  if (b < 0 || b >= c.nB || col < 0 || col >= c.r) return ORC_ERR_RANGE;
  size_t e = (size_t)b + (size_t)c.nB * col;
  int32_t zz[BC];
  int st = rle_decode(c.tokens.data() + c.idx[e], c.len[e], zz, underfull);
  if (st) return st;
  const Tables& T = tables();
  for (int i = 0; i < BC; i++) out[i] = zz[T.zig[i]];
  return ORC_OK;
}

// DecodeBlockWithCompressionData :1003-1057: c = int * (damp^gamma * (1/s))
int decode_coeffs(const Component& c, int b, int col, double* out) {
  int32_t q[BC];
  int st = decode_ints(c, b, col, q, nullptr);
  if (st) return st;
  const Tables& T = tables();
  size_t e = (size_t)b + (size_t)c.nB * col;
  double sInv = 1.0 / c.s[e];
  double gamma = c.gamma[e];
  for (int i = 0; i < BC; i++) {
    double d = std::pow(T.damp[i], gamma);
    out[i] = (double)q[i] * (d * sInv);
  }
  return ORC_OK;
}

// DecodeBlockWithCompressionDataSparseQuantized :1066-1106 (table by 4*gamma)
int decode_coeffs_table(const Component& c, int b, int col, double* out) {
  int32_t q[BC];
  int st = decode_ints(c, b, col, q, nullptr);
  if (st) return st;
  size_t e = (size_t)b + (size_t)c.nB * col;
  double sInv = 1.0 / c.s[e];
  int k = (int)(4 * c.gamma[e]);
  if (k < 0 || k > 4 * c.nBits) return ORC_ERR_FORMAT;
  const double* row = c.dlist.data() + (size_t)k * BC;
  for (int i = 0; i < BC; i++) out[i] = q[i] ? (double)q[i] * (row[i] * sInv) : 0.0;
  return ORC_OK;
}

}  // namespace

// ===========================================================================
// extern "C" surface used by oracle/oracle.py
// ===========================================================================
extern "C" {

void orc_zigzag_tables(int32_t* zig, int32_t* rev) { zigzag_tables(zig, rev); }

// COMPRESSION_DATA.cpp:25-43
int orc_damping_list(int nBits, double* out) {
  const Tables& T = tables();
  int nk = 4 * nBits + 1;
  for (int k = 0; k < nk; k++)
    for (int i = 0; i < BC; i++) out[(size_t)k * BC + i] = std::pow(T.damp[i], k / 4.0);
  return nk;
}

// ModifiedCumSum COMPRESSION.cpp:109-125
void orc_modified_cumsum(const int32_t* in, int n, int32_t* out) {
  int acc = 0;
  if (n > 0) out[0] = 0;
  for (int i = 1; i < n; i++) {
    acc += in[i - 1];
    out[i] = acc;
  }
}

void orc_get_paddings(const int32_t* dims, int32_t* pad) {
  for (int i = 0; i < 3; i++) pad[i] = pad_of(dims[i]);
}

// ComputeBlockNumber COMPRESSION.cpp:1712-1749
int orc_compute_block_number(int row, const int32_t* dims, int32_t* blockNumber, int32_t* blockIndex) {
  int X = dims[0], Y = dims[1], Z = dims[2];
  if (row < 0 || row >= 3 * X * Y * Z) return ORC_ERR_RANGE;
  int index = row / 3;
  int z = index / (X * Y);
  int rem = index - X * Y * z;
  int y = rem / X;
  int x = rem - X * y;
  *blockIndex = (x % BS) + BS * (y % BS) + BS * BS * (z % BS);
  int Px = X + pad_of(X), Py = Y + pad_of(Y);
  *blockNumber = x / BS + (y / BS) * (Px / BS) + (z / BS) * (Px / BS) * (Py / BS);
  return ORC_OK;
}

int orc_rle_encode(const int32_t* zz, int32_t* out) { return rle_encode(zz, out); }
int orc_rle_decode(const int32_t* tok, int len, int32_t* out, int* underfull) {
  return rle_decode(tok, len, out, underfull);
}

void orc_zigzag_flatten(const int32_t* field, int32_t* zz) {
  const Tables& T = tables();
  for (int i = 0; i < BC; i++) zz[T.zig[i]] = field[i];
}
void orc_zigzag_unflatten(const int32_t* zz, int32_t* field) {
  const Tables& T = tables();
  for (int i = 0; i < BC; i++) field[i] = zz[T.zig[i]];
}

void orc_dct_matrix(double* c) {
  const DctMatrix& m = dctm();
  std::memcpy(c, m.c, sizeof(m.c));
}
// direction 1 = forward (DCT-II), -1 = inverse (DCT-III); nblocks blocks
void orc_dct3(double* blocks, int nblocks, int direction) {
  for (int b = 0; b < nblocks; b++) dct3(blocks + (size_t)b * BC, direction == 1);
}

void orc_get_blocks(const double* F, int X, int Y, int Z, int zero_pad, double* blocks) {
  get_blocks(F, X, Y, Z, zero_pad != 0, blocks);
}
void orc_assimilate_unpadded(const double* blocks, int X, int Y, int Z, double* F) {
  assimilate_unpadded(blocks, X, Y, Z, F);
}

// Encode one DCT-domain block (already forward-transformed).  Returns the
// token count; writes s, gamma and up to 1536 tokens.
int orc_encode_block(const double* Fin, int nBits, double percent, int maxIterations,
                     double* s, double* gamma, int32_t* tokens) {
  double F[BC];
  std::memcpy(F, Fin, sizeof(F));
  BlockCode bc;
  encode_block(F, nBits, percent, maxIterations, &bc);
  *s = bc.s;
  *gamma = bc.gamma;
  std::memcpy(tokens, bc.tokens.data(), bc.tokens.size() * 4);
  return (int)bc.tokens.size();
}

// CompressAndWriteMatrixComponents COMPRESSION.cpp:2271-2359 with the SVD
// transform disabled as on the live path (:2312).  U is column-major
// (rows = 3*X*Y*Z interleaved xyz, cols = r), as Eigen stores MatrixXd.
// Writes <prefix>0, <prefix>1, <prefix>2.
int orc_compress_matrix(const double* U, int X, int Y, int Z, int r, int nBits, double percent,
                        int maxIterations, const char* prefix) {
  int Px = X + pad_of(X), Py = Y + pad_of(Y), Pz = Z + pad_of(Z);
  int nB = (Px / BS) * (Py / BS) * (Pz / BS);
  size_t N = (size_t)X * Y * Z;
  Component comp[3];
  for (int k = 0; k < 3; k++) {
    Component& c = comp[k];
    c.nBits = nBits; c.X = X; c.Y = Y; c.Z = Z; c.Px = Px; c.Py = Py; c.Pz = Pz; c.r = r; c.nB = nB;
    c.s.assign((size_t)nB * r, 0); c.gamma.assign((size_t)nB * r, 0);
    c.len.assign((size_t)nB * r, 0); c.idx.assign((size_t)nB * r, 0);
  }
  std::vector<double> field(N), blocks((size_t)nB * BC);
  for (int col = 0; col < r; col++) {
    const double* u = U + (size_t)col * 3 * N;
    for (int k = 0; k < 3; k++) {
      Component& c = comp[k];
      for (size_t i = 0; i < N; i++) field[i] = u[3 * i + k];  // scalarField(k)
      get_blocks(field.data(), X, Y, Z, false, blocks.data());  // GetBlocks, edge pad
      for (int b = 0; b < nB; b++) {
        double* F = blocks.data() + (size_t)b * BC;
        dct3(F, true);
        BlockCode bc;
        encode_block(F, nBits, percent, maxIterations, &bc);
        size_t e = (size_t)b + (size_t)nB * col;
        c.s[e] = bc.s;
        c.gamma[e] = bc.gamma;
        c.len[e] = (int32_t)bc.tokens.size();
        c.tokens.insert(c.tokens.end(), bc.tokens.begin(), bc.tokens.end());
      }
    }
  }
  for (int k = 0; k < 3; k++) {
    Component& c = comp[k];
    orc_modified_cumsum(c.len.data(), (int)c.len.size(), c.idx.data());  // BuildBlockIndicesMatrix :1683
    std::string path = std::string(prefix) + (char)('0' + k);
    int st = write_component(path.c_str(), c);
    if (st) return st;
  }
  return ORC_OK;
}

// ----- basis handle -----------------------------------------------------------
void* orc_basis_open(const char* p0, const char* p1, const char* p2, int* status) {
  Basis* B = new Basis;
  const char* p[3] = {p0, p1, p2};
  for (int k = 0; k < 3; k++) {
    int st = read_component(p[k], &B->comp[k]);
    if (st) {
      delete B;
      *status = st;
      return nullptr;
    }
  }
  const Component& c0 = B->comp[0];
  for (int k = 1; k < 3; k++) {
    const Component& c = B->comp[k];
    if (c.X != c0.X || c.Y != c0.Y || c.Z != c0.Z || c.r != c0.r || c.nB != c0.nB) {
      delete B;
      *status = ORC_ERR_FORMAT;
      return nullptr;
    }
  }
  *status = ORC_OK;
  return B;
}
void orc_basis_close(void* h) { delete (Basis*)h; }

void* orc_basis_open_mem(const uint8_t* b0, size_t n0, const uint8_t* b1, size_t n1, const uint8_t* b2, size_t n2,
                         int* status) {
  Basis* B = new Basis;
  const uint8_t* b[3] = {b0, b1, b2};
  size_t n[3] = {n0, n1, n2};
  for (int k = 0; k < 3; k++) {
    int st = parse_component(b[k], n[k], &B->comp[k]);
    if (st) {
      delete B;
      *status = st;
      return nullptr;
    }
  }
  *status = ORC_OK;
  return B;
}

// info: nBits, X, Y, Z, Px, Py, Pz, r, nB, tokens0, tokens1, tokens2
void orc_basis_info(void* h, int64_t* info) {
  Basis* B = (Basis*)h;
  const Component& c = B->comp[0];
  int64_t v[9] = {c.nBits, c.X, c.Y, c.Z, c.Px, c.Py, c.Pz, c.r, c.nB};
  for (int i = 0; i < 9; i++) info[i] = v[i];
  for (int k = 0; k < 3; k++) info[9 + k] = (int64_t)B->comp[k].tokens.size() - 4;
}

int orc_decode_block_ints(void* h, int comp, int b, int col, int32_t* out, int* underfull) {
  Basis* B = (Basis*)h;
  if (comp < 0 || comp > 2) return ORC_ERR_RANGE;
  return decode_ints(B->comp[comp], b, col, out, underfull);
}
int orc_decode_block_coeffs(void* h, int comp, int b, int col, int use_table, double* out) {
  Basis* B = (Basis*)h;
  if (comp < 0 || comp > 2) return ORC_ERR_RANGE;
  return use_table ? decode_coeffs_table(B->comp[comp], b, col, out)
                   : decode_coeffs(B->comp[comp], b, col, out);
}

// DecodeVectorField COMPRESSION.cpp:2640-2670 (DecodeScalarField :2389-2445 x3):
// out is 3*X*Y*Z, interleaved xyz, peeled and unpadded.
int orc_decode_vector_field(void* h, int col, double* out) {
  Basis* B = (Basis*)h;
  const Component& c0 = B->comp[0];
  size_t N = (size_t)c0.X * c0.Y * c0.Z;
  std::vector<double> blocks((size_t)c0.nB * BC), F(N);
  for (int k = 0; k < 3; k++) {
    const Component& c = B->comp[k];
    for (int b = 0; b < c.nB; b++) {
      int st = decode_coeffs(c, b, col, blocks.data() + (size_t)b * BC);
      if (st) return st;
      dct3(blocks.data() + (size_t)b * BC, false);
    }
    assimilate_unpadded(blocks.data(), c.X, c.Y, c.Z, F.data());
    for (size_t i = 0; i < N; i++) out[3 * i + k] = F[i];
  }
  return ORC_OK;
}

// orc_decode_vector_field for columns [col0, col1) on `threads` host threads
// (one column per task; each column's arithmetic is unchanged), into
// out[(col - col0) * 3*X*Y*Z ...].
int orc_decode_vector_fields_mt(void* h, int col0, int col1, int threads, double* out) {
  Basis* B = (Basis*)h;
  const Component& c0 = B->comp[0];
  if (col0 < 0 || col1 > c0.r || col0 > col1) return ORC_ERR_RANGE;
  const size_t n = (size_t)3 * c0.X * c0.Y * c0.Z;
  std::atomic<int> next(col0), status(ORC_OK);
  std::vector<std::thread> th;
  for (int t = 0; t < std::max(1, threads); t++)
    th.emplace_back([&] {
      for (int c = next++; c < col1; c = next++) {
        const int st = orc_decode_vector_field(h, c, out + (size_t)(c - col0) * n);
        if (st) status = st;
      }
    });
  for (auto& x : th) x.join();
  return status;
}

// PeeledCompressedProjectTransformNoSVD COMPRESSION.cpp:2978-3016 for columns
// [col0, col1).  V is the full grid (X+2)(Y+2)(Z+2) interleaved xyz.
// q[col - col0] receives sum over X, Y, Z of per-block dots (block order).
int orc_project(void* h, const double* V, double* q, int col0, int col1) {
  Basis* B = (Basis*)h;
  const Component& c0 = B->comp[0];
  int X = c0.X, Y = c0.Y, Z = c0.Z, nB = c0.nB;
  if (col0 < 0 || col1 > c0.r || col0 > col1) return ORC_ERR_RANGE;
  size_t N = (size_t)X * Y * Z;
  int GX = X + 2, GY = Y + 2;
  // TransformDCT :2958-2974: peelBoundary, scalar fields, zero-pad blocks, DCT
  std::vector<double> vh[3];
  std::vector<double> F(N);
  for (int k = 0; k < 3; k++) {
    for (int z = 0; z < Z; z++)
      for (int y = 0; y < Y; y++)
        for (int x = 0; x < X; x++)
          F[x + (size_t)X * (y + (size_t)Y * z)] =
              V[3 * ((x + 1) + (size_t)GX * ((y + 1) + (size_t)GY * (z + 1))) + k];
    vh[k].resize((size_t)nB * BC);
    get_blocks(F.data(), X, Y, Z, true, vh[k].data());
    for (int b = 0; b < nB; b++) dct3(vh[k].data() + (size_t)b * BC, true);
  }
  double blk[BC];
  for (int col = col0; col < col1; col++) {
    double total = 0.0;
    for (int k = 0; k < 3; k++) {
      double sum = 0.0;  // GetDotProductSum :2937-2950
      for (int b = 0; b < nB; b++) {
        int st = decode_coeffs(B->comp[k], b, col, blk);
        if (st) return st;
        const double* w = vh[k].data() + (size_t)b * BC;
        double d = 0.0;
        for (int i = 0; i < BC; i++) d += blk[i] * w[i];
        sum += d;
      }
      total += sum;
    }
    q[col - col0] = total;
  }
  return ORC_OK;
}

// PeeledCompressedUnprojectTransform COMPRESSION.cpp:2739-2900 for columns
// [col0, col1) with q indexed by col - col0.  Frequency-domain accumulation
// over nonzeros, one inverse DCT per block, assimilate, strip padding, then
// setWithPeeled (VECTOR3_FIELD_3D.cpp:3415-3430): the boundary of V is kept.
int orc_unproject(void* h, const double* q, double* V, int col0, int col1) {
  Basis* B = (Basis*)h;
  const Component& c0 = B->comp[0];
  int X = c0.X, Y = c0.Y, Z = c0.Z, nB = c0.nB;
  if (col0 < 0 || col1 > c0.r || col0 > col1) return ORC_ERR_RANGE;
  size_t N = (size_t)X * Y * Z;
  int GX = X + 2, GY = Y + 2;
  std::vector<double> F(N);
  double blk[BC];
  for (int k = 0; k < 3; k++) {
    std::vector<double> acc((size_t)nB * BC, 0.0);
    for (int col = col0; col < col1; col++) {
      double qc = q[col - col0];
      for (int b = 0; b < nB; b++) {
        int st = decode_coeffs_table(B->comp[k], b, col, blk);
        if (st) return st;
        double* a = acc.data() + (size_t)b * BC;
        for (int i = 0; i < BC; i++)
          if (blk[i] != 0.0) a[i] += qc * blk[i];
      }
    }
    for (int b = 0; b < nB; b++) dct3(acc.data() + (size_t)b * BC, false);
    assimilate_unpadded(acc.data(), X, Y, Z, F.data());
    for (int z = 0; z < Z; z++)
      for (int y = 0; y < Y; y++)
        for (int x = 0; x < X; x++)
          V[3 * ((x + 1) + (size_t)GX * ((y + 1) + (size_t)GY * (z + 1))) + k] =
              F[x + (size_t)X * (y + (size_t)Y * z)];
  }
  return ORC_OK;
}

// Multi-threaded forms of orc_project / orc_unproject for the CPU baseline
// measured on all cores (the reference itself is single-threaded,
// projects/include.mk:15).  Work is split so that every sum keeps the
// single-threaded order: project over columns, unproject over blocks (each
// block's accumulator sums the columns in order), so the results are
// identical to orc_project / orc_unproject.
static void run_threads(int n, int threads, const std::function<void(int, int)>& fn) {
  threads = std::max(1, std::min(threads, n));
  std::vector<std::thread> th;
  for (int t = 0; t < threads; t++) {
    const int a = (int)((long)n * t / threads), b = (int)((long)n * (t + 1) / threads);
    th.emplace_back([&fn, a, b] { fn(a, b); });
  }
  for (auto& x : th) x.join();
}

int orc_project_mt(void* h, const double* V, double* q, int col0, int col1, int threads) {
  Basis* B = (Basis*)h;
  const Component& c0 = B->comp[0];
  int X = c0.X, Y = c0.Y, Z = c0.Z, nB = c0.nB;
  if (col0 < 0 || col1 > c0.r || col0 > col1) return ORC_ERR_RANGE;
  size_t N = (size_t)X * Y * Z;
  int GX = X + 2, GY = Y + 2;
  std::vector<double> vh[3];
  std::vector<double> F(N);
  for (int k = 0; k < 3; k++) {
    for (int z = 0; z < Z; z++)
      for (int y = 0; y < Y; y++)
        for (int x = 0; x < X; x++)
          F[x + (size_t)X * (y + (size_t)Y * z)] =
              V[3 * ((x + 1) + (size_t)GX * ((y + 1) + (size_t)GY * (z + 1))) + k];
    vh[k].resize((size_t)nB * BC);
    get_blocks(F.data(), X, Y, Z, true, vh[k].data());
    run_threads(nB, threads, [&](int a, int b) {
      for (int i = a; i < b; i++) dct3(vh[k].data() + (size_t)i * BC, true);
    });
  }
  int status = ORC_OK;
  run_threads(col1 - col0, threads, [&](int a, int b) {
    double blk[BC];
    for (int col = col0 + a; col < col0 + b; col++) {
      double total = 0.0;
      for (int k = 0; k < 3; k++) {
        double sum = 0.0;
        for (int bb = 0; bb < nB; bb++) {
          if (decode_coeffs(B->comp[k], bb, col, blk)) {
            status = ORC_ERR_FORMAT;
            return;
          }
          const double* w = vh[k].data() + (size_t)bb * BC;
          double d = 0.0;
          for (int i = 0; i < BC; i++) d += blk[i] * w[i];
          sum += d;
        }
        total += sum;
      }
      q[col - col0] = total;
    }
  });
  return status;
}

int orc_unproject_mt(void* h, const double* q, double* V, int col0, int col1, int threads) {
  Basis* B = (Basis*)h;
  const Component& c0 = B->comp[0];
  int X = c0.X, Y = c0.Y, Z = c0.Z, nB = c0.nB;
  if (col0 < 0 || col1 > c0.r || col0 > col1) return ORC_ERR_RANGE;
  size_t N = (size_t)X * Y * Z;
  int GX = X + 2, GY = Y + 2;
  std::vector<double> F(N);
  int status = ORC_OK;
  for (int k = 0; k < 3; k++) {
    std::vector<double> acc((size_t)nB * BC, 0.0);
    run_threads(nB, threads, [&](int a, int bend) {
      double blk[BC];
      for (int b = a; b < bend; b++) {
        double* ac = acc.data() + (size_t)b * BC;
        for (int col = col0; col < col1; col++) {
          const double qc = q[col - col0];
          if (decode_coeffs_table(B->comp[k], b, col, blk)) {
            status = ORC_ERR_FORMAT;
            return;
          }
          for (int i = 0; i < BC; i++)
            if (blk[i] != 0.0) ac[i] += qc * blk[i];
        }
        dct3(ac, false);
      }
    });
    if (status) return status;
    assimilate_unpadded(acc.data(), X, Y, Z, F.data());
    for (int z = 0; z < Z; z++)
      for (int y = 0; y < Y; y++)
        for (int x = 0; x < X; x++)
          V[3 * ((x + 1) + (size_t)GX * ((y + 1) + (size_t)GY * (z + 1))) + k] =
              F[x + (size_t)X * (y + (size_t)Y * z)];
  }
  return ORC_OK;
}

// GetSubmatrixNoSVDSparse COMPRESSION.cpp:1949-2158: the 3 x r block of U at
// startRow (a multiple of 3), column-major out[comp + 3*col].
int orc_get_submatrix(void* h, int startRow, double* out) {
  Basis* B = (Basis*)h;
  const Component& c0 = B->comp[0];
  if (startRow % 3 != 0) return ORC_ERR_RANGE;
  int32_t dims[3] = {c0.X, c0.Y, c0.Z};
  int32_t bn, bi;
  int st = orc_compute_block_number(startRow, dims, &bn, &bi);
  if (st) return st;
  double blk[BC];
  for (int col = 0; col < c0.r; col++)
    for (int k = 0; k < 3; k++) {
      st = decode_coeffs_table(B->comp[k], bn, col, blk);
      if (st) return st;
      dct3(blk, false);
      out[k + 3 * col] = blk[bi];
    }
  return ORC_OK;
}

// SVD coordinate transform, COMPRESSION.cpp:566-727.  The transform is off on
// the live path (:2312, :2667-2668); these restate its per-cell algebra.
// UntransformVectorFieldSVD :713-727: u <- V u' for every cell (V row-major 3x3).
void orc_untransform_svd(const double* Vm, double* field, int64_t cells) {
  for (int64_t i = 0; i < cells; i++) {
    double a = field[3 * i], b = field[3 * i + 1], c = field[3 * i + 2];
    for (int r = 0; r < 3; r++) field[3 * i + r] = Vm[3 * r] * a + Vm[3 * r + 1] * b + Vm[3 * r + 2] * c;
  }
}
// TransformVectorFieldSVDCached :623-643: u' <- V^T u.
void orc_transform_svd(const double* Vm, double* field, int64_t cells) {
  for (int64_t i = 0; i < cells; i++) {
    double a = field[3 * i], b = field[3 * i + 1], c = field[3 * i + 2];
    for (int r = 0; r < 3; r++) field[3 * i + r] = Vm[r] * a + Vm[3 + r] * b + Vm[6 + r] * c;
  }
}

}  // extern "C"

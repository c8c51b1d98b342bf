// encoder.cpp -- host encoder producing the reference component-file format.
//
// Restates CompressAndWriteMatrixComponents (reference COMPRESSION.cpp:2271-2359)
// with the SVD transform off as on the live path (:2312), per component and
// column:
//   GetBlocks (:371-404, edge-replicated padding FIELD_3D.cpp:6255-6281)
//   UnitaryBlockDCT(1) (:511-531), orthonormal DCT-II along x, y, z
//   PreprocessBlock (:733-759)      s = (2^(nBits-1)-1)/absMax, F *= s
//   TuneGammaQuantized (:841-916)   binary search on gamma, trunc energy test
//   EncodeBlock (:924-949)          rint(F / damp^gamma), ties to even
//   ZigzagFlatten (:1402-1421), RunLengthEncodeBinary (:1112-1192)
//   BuildBlockIndicesMatrix (:1683-1704), WriteMetaData (:2165-2228)
// Columns x components run in parallel on host threads; each block's
// arithmetic is sequential FP64 in the reference's order (built without FMA
// contraction), so the output does not depend on the thread count.  The
// damping powers use the fact that (1+u+v+w)^g takes only 22 distinct bases.
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/cfs_plume.h"

namespace {

constexpr int BS = 8;
constexpr int BC = 512;
constexpr int ND = 22;

int pad_of(int n) { return (BS - (n % BS)) % BS; }

struct Consts {
  int zig[BC];
  int diag[BC];   // u+v+w of cell x+8y+64z
  double c[BS][BS];
  Consts() {
    int j = 0;
    for (int sum = 0; sum < 3 * BS; sum++)
      for (int z = 0; z < BS; z++)
        for (int y = 0; y < BS; y++)
          for (int x = 0; x < BS; x++)
            if (x + y + z == sum) zig[x + BS * y + BS * BS * z] = j++;
    for (int i = 0; i < BC; i++) diag[i] = (i & 7) + ((i >> 3) & 7) + (i >> 6);
    for (int k = 0; k < BS; k++)
      for (int n = 0; n < BS; n++) {
        const double a = (k == 0) ? std::sqrt(0.125) : 0.5;
        c[k][n] = a * std::cos((M_PI * (double)((2 * n + 1) * k)) / 16.0);
      }
  }
};
const Consts& consts() {
  static Consts k;
  return k;
}

// one 8-point DCT-II pass; sums in increasing index order
inline void pass(double* b, int stride, int na, int sa, int nb, int sb, const Consts& K) {
  for (int i = 0; i < na; i++)
    for (int j = 0; j < nb; j++) {
      double* p = b + i * sa + j * sb;
      double in[BS], out[BS];
      for (int n = 0; n < BS; n++) in[n] = p[n * stride];
      for (int k = 0; k < BS; k++) {
        double s = 0.0;
        for (int n = 0; n < BS; n++) s += K.c[k][n] * in[n];
        out[k] = s;
      }
      for (int k = 0; k < BS; k++) p[k * stride] = out[k];
    }
}

inline void forward_dct(double* b, const Consts& K) {
  pass(b, 1, BS, BS * BS, BS, BS, K);
  pass(b, BS, BS, BS * BS, BS, 1, K);
  pass(b, BS * BS, BS, BS, BS, 1, K);
}

struct ColumnCode {
  std::vector<double> s, gamma;   // [nB]
  std::vector<int32_t> len;       // [nB]
  std::vector<int32_t> tokens;
};

// Encodes block F (DCT domain) in place; appends its tokens.
void encode_block(double* F, int nBits, double percent, int maxIterations, const Consts& K, double* s_out,
                  double* g_out, std::vector<int32_t>* tokens) {
  double fmax = std::fabs(F[0]);
  for (int i = 0; i < BC; i++)
    if (std::fabs(F[i]) > fmax) fmax = std::fabs(F[i]);
  const double s = (std::pow(2, nBits - 1) - 1) / fmax;
  for (int i = 0; i < BC; i++) F[i] *= s;
  *s_out = s;

  double pw[ND];
  auto set_power = [&](double g) {
    for (int d = 0; d < ND; d++) pw[d] = std::pow((double)(1 + d), g);
  };
  auto energy = [&]() {
    double e = 0;
    for (int i = 0; i < BC; i++) {
      const double dm = pw[K.diag[i]];
      const double t = std::trunc(F[i] / dm);
      const double p = dm * t;
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
  *g_out = gamma;

  int32_t zz[BC];
  for (int i = 0; i < BC; i++) zz[K.zig[i]] = (int32_t)std::rint(F[i] / pw[K.diag[i]]);
  for (int i = 0; i < BC; i++) {
    const int32_t v = zz[i];
    tokens->push_back(v);
    int run = 1;
    while (i + 1 < BC && zz[i] == zz[i + 1]) {
      run++;
      i++;
    }
    if (run > 1) {
      tokens->push_back(v);
      tokens->push_back(run);
    }
  }
}

// Encodes component k of column col of U.
void encode_column(const double* U, int X, int Y, int Z, int k, int nBits, double percent, int maxIterations,
                   ColumnCode* out) {
  const Consts& K = consts();
  const int Px = X + pad_of(X), Py = Y + pad_of(Y), Pz = Z + pad_of(Z);
  const int nbx = Px / BS, nby = Py / BS, nbz = Pz / BS;
  const int nB = nbx * nby * nbz;
  out->s.resize(nB);
  out->gamma.resize(nB);
  out->len.resize(nB);
  out->tokens.clear();
  out->tokens.reserve((size_t)nB * 96);
  double F[BC];
  for (int bz = 0; bz < nbz; bz++)
    for (int by = 0; by < nby; by++)
      for (int bx = 0; bx < nbx; bx++) {
        const int b = bx + nbx * (by + nby * bz);
        for (int z = 0; z < BS; z++)
          for (int y = 0; y < BS; y++)
            for (int x = 0; x < BS; x++) {
              int gx = bx * BS + x, gy = by * BS + y, gz = bz * BS + z;
              gx = gx < X ? gx : X - 1;
              gy = gy < Y ? gy : Y - 1;
              gz = gz < Z ? gz : Z - 1;
              F[x + BS * y + BS * BS * z] = U[3 * ((size_t)gx + (size_t)X * (gy + (size_t)Y * gz)) + k];
            }
        forward_dct(F, K);
        const size_t before = out->tokens.size();
        encode_block(F, nBits, percent, maxIterations, K, &out->s[b], &out->gamma[b], &out->tokens);
        out->len[b] = (int32_t)(out->tokens.size() - before);
      }
}

}  // namespace

struct cfs_encoded {
  std::vector<uint8_t> file[3];
};

extern "C" {

int cfs_encode_matrix(const double* U, int X, int Y, int Z, int r, int n_bits, double percent, int max_iterations,
                      int threads, cfs_encoded** out) {
  if (!U || !out || X < 1 || Y < 1 || Z < 1 || r < 1 || n_bits < 1 || n_bits > 32 || max_iterations < 0)
    return CFS_ERR_ARG;
  *out = nullptr;
  const int Px = X + pad_of(X), Py = Y + pad_of(Y), Pz = Z + pad_of(Z);
  const int nB = (Px / BS) * (Py / BS) * (Pz / BS);
  const size_t rows = (size_t)3 * X * Y * Z;
  std::vector<ColumnCode> codes((size_t)3 * r);
  unsigned nt = threads > 0 ? (unsigned)threads : std::thread::hardware_concurrency();
  if (nt < 1) nt = 1;
  std::atomic<int> next{0};
  auto worker = [&]() {
    for (;;) {
      const int t = next.fetch_add(1);
      if (t >= 3 * r) break;
      const int col = t / 3, k = t % 3;
      encode_column(U + (size_t)col * rows, X, Y, Z, k, n_bits, percent, max_iterations, &codes[(size_t)k * r + col]);
    }
  };
  std::vector<std::thread> pool;
  for (unsigned i = 1; i < nt; i++) pool.emplace_back(worker);
  worker();
  for (auto& th : pool) th.join();

  cfs_encoded* E = new cfs_encoded;
  const size_t m = (size_t)nB * r;
  for (int k = 0; k < 3; k++) {
    size_t ntok = 0;
    for (int c = 0; c < r; c++) ntok += codes[(size_t)k * r + c].tokens.size();
    std::vector<uint8_t>& f = E->file[k];
    f.resize(40 + m * 24 + ntok * 4);
    const int32_t h[10] = {1, n_bits, X, Y, Z, Px, Py, Pz, r, nB};
    std::memcpy(f.data(), h, 40);
    double* s = reinterpret_cast<double*>(f.data() + 40);
    double* g = s + m;
    int32_t* len = reinterpret_cast<int32_t*>(g + m);
    int32_t* idx = len + m;
    int32_t* tok = idx + m;
    int32_t acc = 0;  // ModifiedCumSum over the column-major lengths
    for (int c = 0; c < r; c++) {
      const ColumnCode& cc = codes[(size_t)k * r + c];
      for (int b = 0; b < nB; b++) {
        const size_t e = (size_t)b + (size_t)nB * c;
        s[e] = cc.s[b];
        g[e] = cc.gamma[b];
        len[e] = cc.len[b];
        idx[e] = acc;
        acc += cc.len[b];
      }
      std::memcpy(tok, cc.tokens.data(), cc.tokens.size() * 4);
      tok += cc.tokens.size();
    }
  }
  *out = E;
  return CFS_OK;
}

int cfs_encode_block(const double* F, int n_bits, double percent, int max_iterations, double* s, double* gamma,
                     int32_t* tokens, int* ntokens) {
  if (!F || !s || !gamma || !tokens || !ntokens || n_bits < 1 || n_bits > 32 || max_iterations < 0)
    return CFS_ERR_ARG;
  double B[BC];
  std::memcpy(B, F, sizeof(B));
  std::vector<int32_t> t;
  encode_block(B, n_bits, percent, max_iterations, consts(), s, gamma, &t);
  std::memcpy(tokens, t.data(), t.size() * 4);
  *ntokens = (int)t.size();
  return CFS_OK;
}

int cfs_encoded_buffer(const cfs_encoded* enc, int comp, const void** data, size_t* size) {
  if (!enc || comp < 0 || comp > 2 || !data || !size) return CFS_ERR_ARG;
  *data = enc->file[comp].data();
  *size = enc->file[comp].size();
  return CFS_OK;
}

int cfs_encoded_write(const cfs_encoded* enc, const char* prefix) {
  if (!enc || !prefix) return CFS_ERR_ARG;
  for (int k = 0; k < 3; k++) {
    const std::string path = std::string(prefix) + (char)('0' + k);
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return CFS_ERR_IO;
    const size_t n = enc->file[k].size();
    const size_t w = std::fwrite(enc->file[k].data(), 1, n, f);
    std::fclose(f);
    if (w != n) return CFS_ERR_IO;
  }
  return CFS_OK;
}

void cfs_encoded_free(cfs_encoded* enc) { delete enc; }

}  // extern "C"

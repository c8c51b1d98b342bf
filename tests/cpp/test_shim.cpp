// C++ tests of the reference-named host layer (include/cfs_plume.hpp) against
// the FP64 CPU oracle (oracle/cfs_oracle.cpp, test infrastructure), written in
// the style of the reference's tests/test_compression_data.cpp: a registry of
// TEST functions, ASSERT_MSG, exit status 1 on any failure.
//
//   test_shim            all tests (needs a HIP device)
//   test_shim --no-gpu   the host-only tests
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <string>
#include <unistd.h>

#include <vector>

#include "cfs_plume.hpp"

// oracle (libcfs_oracle.so)
extern "C" {
void* orc_basis_open(const char* p0, const char* p1, const char* p2, int* status);
void orc_basis_close(void* h);
int orc_project(void* h, const double* V, double* q, int col0, int col1);
int orc_unproject(void* h, const double* q, double* V, int col0, int col1);
int orc_decode_vector_field(void* h, int col, double* out);
int orc_get_submatrix(void* h, int startRow, double* out);
}

namespace {

struct Test {
  const char* name;
  bool gpu;
  std::function<bool()> fn;
};
std::vector<Test>& registry() {
  static std::vector<Test> t;
  return t;
}
struct Register {
  Register(const char* n, bool gpu, std::function<bool()> f) { registry().push_back({n, gpu, f}); }
};
#define TEST(name, gpu)                                 \
  static bool test_##name();                            \
  static Register reg_##name(#name, gpu, test_##name); \
  static bool test_##name()
#define ASSERT_MSG(cond, msg)                                            \
  do {                                                                   \
    if (!(cond)) {                                                       \
      std::printf("  FAILED %s:%d: %s\n", __FILE__, __LINE__, (msg).c_str()); \
      return false;                                                      \
    }                                                                    \
  } while (0)

double rel(const std::vector<double>& a, const std::vector<double>& b) {
  double num = 0, den = 0;
  for (size_t i = 0; i < a.size(); i++) {
    num += (a[i] - b[i]) * (a[i] - b[i]);
    den += b[i] * b[i];
  }
  return std::sqrt(num / (den > 0 ? den : 1e-300));
}

// A smooth synthetic basis: column c, component k = sum of 3 cosines.
const int X = 21, Y = 14, Z = 10, R = 5;
std::vector<double> make_U() {
  std::vector<double> U((size_t)3 * X * Y * Z * R);
  for (int c = 0; c < R; c++)
    for (int z = 0; z < Z; z++)
      for (int y = 0; y < Y; y++)
        for (int x = 0; x < X; x++)
          for (int k = 0; k < 3; k++) {
            double v = 0;
            for (int t = 1; t <= 3; t++)
              v += std::cos(2 * M_PI * ((c + t) * x / (double)X + (k + t) * y / (double)Y + t * z / (double)Z) +
                            0.3 * c + k) / t;
            U[(size_t)c * 3 * X * Y * Z + 3 * (x + X * (y + Y * z)) + k] = v;
          }
  return U;
}

std::string g_prefix;

const std::string& prefix() {
  if (g_prefix.empty()) {
    g_prefix = std::string(std::getenv("TMPDIR") ? std::getenv("TMPDIR") : "/tmp") + "/cfs_shim_test_" +
               std::to_string((long)getpid()) + ".U.component";
    cfs_plume::CompressAndWriteMatrixComponents(g_prefix, make_U(), X, Y, Z, R);
  }
  return g_prefix;
}

std::vector<double> full_field(int seed) {
  std::vector<double> V((size_t)3 * (X + 2) * (Y + 2) * (Z + 2), 0.0);
  unsigned s = 12345u + seed;
  for (int z = 1; z <= Z; z++)
    for (int y = 1; y <= Y; y++)
      for (int x = 1; x <= X; x++)
        for (int k = 0; k < 3; k++) {
          s = s * 1664525u + 1013904223u;
          V[3 * (x + (X + 2) * (y + (Y + 2) * z)) + k] = ((s >> 8) / 16777216.0) - 0.5;
        }
  return V;
}

struct Oracle {
  void* h = nullptr;
  Oracle() {
    int st = 0;
    const std::string& p = prefix();
    h = orc_basis_open((p + "0").c_str(), (p + "1").c_str(), (p + "2").c_str(), &st);
  }
  ~Oracle() {
    if (h) orc_basis_close(h);
  }
};

}  // namespace

TEST(missing_files_raise_io_error, false) {
  try {
    cfs_plume::MatrixCompressionData U("/nonexistent/a0", "/nonexistent/a1", "/nonexistent/a2");
  } catch (const cfs_plume::Error& e) {
    ASSERT_MSG(e.status() == CFS_ERR_IO, std::string("status ") + std::to_string(e.status()));
    return true;
  }
  ASSERT_MSG(false, std::string("no exception for missing files"));
}

TEST(loaded_dims_match, true) {
  cfs_plume::MatrixCompressionData U(prefix() + "0", prefix() + "1", prefix() + "2");
  ASSERT_MSG(U.xRes() == X && U.yRes() == Y && U.zRes() == Z, std::string("dims"));
  ASSERT_MSG(U.get_numCols() == R && U.get_nBits() == 32, std::string("numCols / nBits"));
  ASSERT_MSG(U.get_numBlocks() == 3 * 2 * 2, std::string("numBlocks ") + std::to_string(U.get_numBlocks()));
  return true;
}

TEST(project_matches_oracle, true) {
  cfs_plume::MatrixCompressionData U(prefix() + "0", prefix() + "1", prefix() + "2");
  Oracle o;
  ASSERT_MSG(o.h != nullptr, std::string("oracle open"));
  std::vector<double> V = full_field(1), q;
  cfs_plume::PeeledCompressedProjectTransformNoSVD(V, &U, &q);
  std::vector<double> V32(V.size()), want(R);
  for (size_t i = 0; i < V.size(); i++) V32[i] = (double)(float)V[i];
  orc_project(o.h, V32.data(), want.data(), 0, R);
  ASSERT_MSG(q.size() == (size_t)R, std::string("q resized"));
  ASSERT_MSG(rel(q, want) <= 1e-5, std::string("q rel-L2 ") + std::to_string(rel(q, want)));
  return true;
}

TEST(unproject_matches_oracle_and_keeps_boundary, true) {
  cfs_plume::MatrixCompressionData U(prefix() + "0", prefix() + "1", prefix() + "2");
  Oracle o;
  std::vector<double> q = {0.7, -1.1, 0.4, 2.0, -0.3};
  std::vector<double> V(3 * U.full_cells(), 2.5), want = V;
  cfs_plume::PeeledCompressedUnprojectTransform(&U, q, &V);
  orc_unproject(o.h, q.data(), want.data(), 0, R);
  ASSERT_MSG(rel(V, want) <= 1e-5, std::string("field rel-L2 ") + std::to_string(rel(V, want)));
  for (int z = 0; z < Z + 2; z++)
    for (int y = 0; y < Y + 2; y++)
      for (int x = 0; x < X + 2; x++) {
        const bool inner = x >= 1 && x <= X && y >= 1 && y <= Y && z >= 1 && z <= Z;
        if (!inner)
          for (int k = 0; k < 3; k++)
            ASSERT_MSG(V[3 * (x + (X + 2) * (y + (Y + 2) * z)) + k] == 2.5, std::string("boundary cell written"));
      }
  return true;
}

TEST(decode_vector_field_matches_oracle, true) {
  cfs_plume::MatrixCompressionData U(prefix() + "0", prefix() + "1", prefix() + "2");
  Oracle o;
  for (int c = 0; c < R; c++) {
    std::vector<double> got, want(3 * U.peeled_cells());
    cfs_plume::DecodeVectorField(&U, c, &got);
    orc_decode_vector_field(o.h, c, want.data());
    // MFMA inverse DCT with BF16 hi + lo split products
    ASSERT_MSG(rel(got, want) <= 3e-5, std::string("column ") + std::to_string(c) + " rel-L2 " +
                                           std::to_string(rel(got, want)));
  }
  return true;
}

TEST(submatrix_matches_oracle, true) {
  cfs_plume::MatrixCompressionData U(prefix() + "0", prefix() + "1", prefix() + "2");
  Oracle o;
  for (int cell : {0, 7, X * Y * Z - 1}) {
    std::vector<double> got, want(3 * R);
    cfs_plume::GetSubmatrixNoSVDSparse(3 * cell, &U, &got);
    orc_get_submatrix(o.h, 3 * cell, want.data());
    ASSERT_MSG(rel(got, want) <= 1e-5, std::string("cell ") + std::to_string(cell));
  }
  return true;
}

TEST(project_unproject_round_trip, true) {
  // the reference's bound for this round trip (tests/test_compression_data.cpp:1005-1081) is rel < 0.5
  cfs_plume::MatrixCompressionData U(prefix() + "0", prefix() + "1", prefix() + "2");
  std::vector<double> q = {1.0, 0.0, -0.5, 0.25, 0.0}, V(3 * U.full_cells(), 0.0), q2;
  cfs_plume::PeeledCompressedUnprojectTransform(&U, q, &V);
  cfs_plume::PeeledCompressedProjectTransformNoSVD(V, &U, &q2);
  // the basis is not orthonormal: compare U^T U q with itself through the oracle instead
  Oracle o;
  std::vector<double> Vw(V.size(), 0.0), qw(R);
  orc_unproject(o.h, q.data(), Vw.data(), 0, R);
  orc_project(o.h, Vw.data(), qw.data(), 0, R);
  ASSERT_MSG(rel(q2, qw) < 1e-4, std::string("U^T U q rel-L2 ") + std::to_string(rel(q2, qw)));
  return true;
}

int main(int argc, char** argv) {
  const bool gpu = !(argc > 1 && std::string(argv[1]) == "--no-gpu");
  int failed = 0, ran = 0;
  for (const Test& t : registry()) {
    if (t.gpu && !gpu) continue;
    bool ok = false;
    try {
      ok = t.fn();
    } catch (const std::exception& e) {
      std::printf("  EXCEPTION %s\n", e.what());
    }
    std::printf("[%s] %s\n", ok ? "PASS" : "FAIL", t.name);
    failed += !ok;
    ran++;
  }
  if (!g_prefix.empty())
    for (int k = 0; k < 3; k++) std::remove((g_prefix + std::to_string(k)).c_str());
  std::printf("%d of %d tests passed\n", ran - failed, ran);
  return failed ? 1 : 0;
}

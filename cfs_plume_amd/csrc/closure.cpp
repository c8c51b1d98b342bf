// closure.cpp -- C ABI of the reduced-advection step and its learned closure.
//
// cfs_reduced_advect: q' = alpha U^T A(U q) + f, the exact reduced
// semi-Lagrangian advection on the full grid.  It is the closure's fallback
// and the reference its accuracy is measured against.  The reference
// approximates the same projection with cubature key cells
// (reducedAdvectCompressionFriendly, SUBSPACE_FLUID_3D_COMPRESSED_EIGEN.cpp:
// 1781-1854) and advects with VECTOR3_FIELD_3D::advect (VECTOR3_FIELD_3D.cpp:
// 2001-2077).
//
// cfs_closure_*: the MLP r -> 256 -> 256 -> 256 -> r (GELU) predicting the
// advection increment, its MFMA forward, FP32 backward and Adam update, and
// the fused inference step q' = alpha (q + s_out * MLP(s_in * q)) + f.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "../../include/cfs_plume.h"
#include "basis_internal.h"

namespace cfs {
struct ClosureNet {
  const uint16_t* Wp[4];
  const float* b[4];
  const float* lin;
  const float* s_in;
  const float* s_out;
  int r, rp, H;
};
hipError_t launch_advect(const float* vel, float* out, int X, int Y, int Z, float dt0, hipStream_t s);
hipError_t launch_advect_scalar(const float* vel, const float* old, float* out, int X, int Y, int Z, float dt0,
                                hipStream_t s);
hipError_t launch_zero_words(void* p, long nwords, hipStream_t s);
hipError_t launch_axpb(double* q, int r, double alpha, const double* forcing, hipStream_t s);
hipError_t launch_mlp_fwd(const float* X, int ldx, int n, const uint16_t* Wp, const float* bias, int K, int M,
                          float* Z, float* Y, int act, hipStream_t s);
struct GemmProblem {
  const float* A;
  long sam, sak;
  const float* B;
  long sbk, sbn;
  float* C;
  const float* G;
  float* colsum;
  int ldc, M, N, K;
};
hipError_t launch_gemm_f32(const GemmProblem& p0, const GemmProblem* p1, hipStream_t s);
hipError_t launch_loss_grad(const float* Y, const float* T, int n, int M, int r, float* dY, double* loss,
                            hipStream_t s);
hipError_t launch_gather_rows(const float* src, const int32_t* idx, int n, int ld, float* dst, hipStream_t s);
hipError_t launch_adam(float* p, const float* g, float* m, float* v, long n, float lr, float b1, float b2, float eps,
                       float c1, float c2, hipStream_t s);
hipError_t launch_pack(const float* W, int K, int M, uint16_t* Wp, float* Wq, hipStream_t s);
hipError_t launch_closure_step(const ClosureNet& net, const double* q_in, double* q_out, int n, double alpha,
                               const double* forcing, hipStream_t s);
}  // namespace cfs

using namespace cfs;

namespace {
constexpr int HIDDEN = 256;
constexpr int LAYERS = 4;

#define HIP_TRY(expr, what)                     \
  do {                                          \
    hipError_t e_ = (expr);                     \
    if (e_ != hipSuccess) return hip_fail(e_, what); \
  } while (0)

struct DevBuf {
  void* p = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t bytes) {
    hipError_t e = hipMalloc(&p, std::max<size_t>(bytes, 16));
    if (e != hipSuccess) return hip_fail(e, "hipMalloc");
    e = hipMemset(p, 0, std::max<size_t>(bytes, 16));
    return e == hipSuccess ? CFS_OK : hip_fail(e, "hipMemset");
  }
  template <class T>
  T* as() const {
    return reinterpret_cast<T*>(p);
  }
};

uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}
}  // namespace

struct cfs_closure {
  int device = 0;
  int r = 0, rp = 0, H = HIDDEN, max_batch = 0;
  int K[LAYERS], M[LAYERS];
  size_t offW[LAYERS], offB[LAYERS], nparams = 0;
  long step = 0;  // Adam steps taken
  DevBuf params, grads, adam_m, adam_v, packed, wq, s_in, s_out, lin;
  bool has_lin = false;
  size_t offP[LAYERS], offQ[LAYERS];
  DevBuf xb, tb, idx, z[LAYERS], h[LAYERS], dy, dz0, dz1, loss;

  int init(int r_, int max_batch_) {
    r = r_;
    rp = (r + 63) / 64 * 64;
    max_batch = max_batch_;
    for (int l = 0; l < LAYERS; l++) {
      K[l] = l == 0 ? rp : H;
      M[l] = l == LAYERS - 1 ? rp : H;
    }
    size_t o = 0, op = 0;
    for (int l = 0; l < LAYERS; l++) {
      offW[l] = o;
      o += (size_t)K[l] * M[l];
      offB[l] = o;
      o += (size_t)M[l];
      offP[l] = op;
      offQ[l] = op;
      op += (size_t)K[l] * M[l];
    }
    nparams = o;
    int st;
    if ((st = params.alloc(o * 4)) || (st = grads.alloc(o * 4)) || (st = adam_m.alloc(o * 4)) ||
        (st = adam_v.alloc(o * 4)) || (st = packed.alloc(op * 2)) || (st = wq.alloc(op * 4)) ||
        (st = s_in.alloc((size_t)rp * 4)) || (st = s_out.alloc((size_t)rp * 4)) ||
        (st = lin.alloc((size_t)rp * rp * 4)))
      return st;
    const size_t nb = (size_t)max_batch;
    if ((st = xb.alloc(nb * rp * 4)) || (st = tb.alloc(nb * rp * 4)) || (st = idx.alloc(nb * 4)) ||
        (st = dy.alloc(nb * rp * 4)) || (st = dz0.alloc(nb * H * 4)) || (st = dz1.alloc(nb * H * 4)) ||
        (st = loss.alloc(8)))
      return st;
    for (int l = 0; l < LAYERS; l++)
      if ((st = z[l].alloc(nb * M[l] * 4)) || (st = h[l].alloc(nb * M[l] * 4))) return st;
    return CFS_OK;
  }

  ClosureNet net() const {
    ClosureNet n;
    for (int l = 0; l < LAYERS; l++) {
      n.Wp[l] = packed.as<uint16_t>() + offP[l];
      n.b[l] = params.as<float>() + offB[l];
    }
    n.lin = has_lin ? lin.as<float>() : nullptr;
    n.s_in = s_in.as<float>();
    n.s_out = s_out.as<float>();
    n.r = r;
    n.rp = rp;
    n.H = H;
    return n;
  }

  int repack(hipStream_t s) {
    for (int l = 0; l < LAYERS; l++)
      HIP_TRY(launch_pack(params.as<float>() + offW[l], K[l], M[l], packed.as<uint16_t>() + offP[l],
                          wq.as<float>() + offQ[l], s),
              "pack launch");
    return CFS_OK;
  }

  // forward of n rows of x (leading dimension rp) into h[LAYERS-1]; z saved
  int forward(const float* x, int n, hipStream_t s) {
    const float* in = x;
    for (int l = 0; l < LAYERS; l++) {
      HIP_TRY(launch_mlp_fwd(in, K[l], n, packed.as<uint16_t>() + offP[l], params.as<float>() + offB[l], K[l], M[l],
                             z[l].as<float>(), h[l].as<float>(), l < LAYERS - 1 ? 1 : 0, s),
              "mlp forward launch");
      in = h[l].as<float>();
    }
    return CFS_OK;
  }

  // gradients of the mean-squared loss of n rows (x, t) into grads; the sum
  // of squared errors is added to loss
  int backward(const float* x, const float* t, int n, hipStream_t s) {
    float* g = grads.as<float>();
    HIP_TRY(launch_loss_grad(h[LAYERS - 1].as<float>(), t, n, rp, r, dy.as<float>(), loss.as<double>(), s),
            "loss launch");
    const float* dZ = dy.as<float>();  // gradient at the pre-activation of layer l
    float* bufs[2] = {dz0.as<float>(), dz1.as<float>()};
    for (int l = LAYERS - 1; l >= 0; l--) {
      const float* X = l == 0 ? x : h[l - 1].as<float>();
      // dW = X^T dZ  (K x M), X(m = k_in, k = row), dZ(k = row, n = out);
      // db = column sums of dZ
      const GemmProblem dw{X, 1, K[l], dZ, M[l], 1, g + offW[l], nullptr, g + offB[l], M[l], K[l], M[l], n};
      if (l > 0) {
        // dZ_prev = (dZ W^T) * gelu'(Z_prev): dZ(m = row, k = out), Wq(k = out, n = in) at in * M + out
        float* out = bufs[l & 1];
        const GemmProblem dx{dZ, M[l], 1, wq.as<float>() + offQ[l], 1, M[l], out, z[l - 1].as<float>(), nullptr,
                             K[l], n, K[l], M[l]};
        HIP_TRY(launch_gemm_f32(dw, &dx, s), "gemm launch");
        dZ = out;
      } else {
        HIP_TRY(launch_gemm_f32(dw, nullptr, s), "gemm launch");
      }
    }
    return CFS_OK;
  }
};

extern "C" {

int cfs_closure_create(int r, int hidden, int max_batch, uint64_t seed, cfs_closure** out) {
  if (!out) return fail(CFS_ERR_ARG, "null argument");
  *out = nullptr;
  if (r < 1 || r > HIDDEN) return fail(CFS_ERR_ARG, "closure width r must be in [1, 256]");
  if (hidden != HIDDEN) return fail(CFS_ERR_ARG, "hidden width must be 256");
  if (max_batch < 1 || max_batch > (1 << 20)) return fail(CFS_ERR_ARG, "max_batch out of range");
  std::unique_ptr<cfs_closure> c(new cfs_closure);
  HIP_TRY(hipGetDevice(&c->device), "hipGetDevice");
  int st = c->init(r, max_batch);
  if (st) return st;
  // Glorot-uniform weights, zero biases; padding rows/columns stay zero
  std::vector<float> p(c->nparams, 0.0f);
  uint64_t s = seed;
  for (int l = 0; l < LAYERS; l++) {
    const int fin = l == 0 ? r : c->H, fout = l == LAYERS - 1 ? r : c->H;
    const double a = std::sqrt(6.0 / (double)(fin + fout));
    for (int k = 0; k < fin; k++)
      for (int m = 0; m < fout; m++) {
        const double u = (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
        p[c->offW[l] + (size_t)k * c->M[l] + m] = (float)((2.0 * u - 1.0) * a);
      }
  }
  HIP_TRY(hipMemcpy(c->params.p, p.data(), p.size() * 4, hipMemcpyHostToDevice), "hipMemcpy");
  std::vector<float> ones(c->rp, 0.0f);
  for (int j = 0; j < r; j++) ones[j] = 1.0f;
  HIP_TRY(hipMemcpy(c->s_in.p, ones.data(), ones.size() * 4, hipMemcpyHostToDevice), "hipMemcpy");
  HIP_TRY(hipMemcpy(c->s_out.p, ones.data(), ones.size() * 4, hipMemcpyHostToDevice), "hipMemcpy");
  if ((st = c->repack(nullptr))) return st;
  HIP_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
  *out = c.release();
  return CFS_OK;
}

int cfs_closure_destroy(cfs_closure* c) {
  delete c;
  return CFS_OK;
}

int cfs_closure_get_dims(const cfs_closure* c, int* r, int* r_padded, int* hidden, int64_t* n_params) {
  if (!c) return fail(CFS_ERR_ARG, "null argument");
  if (r) *r = c->r;
  if (r_padded) *r_padded = c->rp;
  if (hidden) *hidden = c->H;
  if (n_params) *n_params = (int64_t)c->nparams;
  return CFS_OK;
}

int cfs_closure_get_params(const cfs_closure* c, float* host) {
  if (!c || !host) return fail(CFS_ERR_ARG, "null argument");
  HIP_TRY(hipMemcpy(host, c->params.p, c->nparams * 4, hipMemcpyDeviceToHost), "hipMemcpy");
  return CFS_OK;
}

int cfs_closure_set_params(cfs_closure* c, const float* host) {
  if (!c || !host) return fail(CFS_ERR_ARG, "null argument");
  HIP_TRY(hipMemcpy(c->params.p, host, c->nparams * 4, hipMemcpyHostToDevice), "hipMemcpy");
  int st = c->repack(nullptr);
  if (st) return st;
  HIP_TRY(hipMemset(c->adam_m.p, 0, c->nparams * 4), "hipMemset");
  HIP_TRY(hipMemset(c->adam_v.p, 0, c->nparams * 4), "hipMemset");
  c->step = 0;
  HIP_TRY(hipDeviceSynchronize(), "hipDeviceSynchronize");
  return CFS_OK;
}

int cfs_closure_set_scales(cfs_closure* c, const float* s_in, const float* s_out) {
  if (!c || !s_in || !s_out) return fail(CFS_ERR_ARG, "null argument");
  std::vector<float> a(c->rp, 0.0f), b(c->rp, 0.0f);
  std::memcpy(a.data(), s_in, (size_t)c->r * 4);
  std::memcpy(b.data(), s_out, (size_t)c->r * 4);
  HIP_TRY(hipMemcpy(c->s_in.p, a.data(), a.size() * 4, hipMemcpyHostToDevice), "hipMemcpy");
  HIP_TRY(hipMemcpy(c->s_out.p, b.data(), b.size() * 4, hipMemcpyHostToDevice), "hipMemcpy");
  return CFS_OK;
}

int cfs_closure_set_linear(cfs_closure* c, const float* L) {
  if (!c) return fail(CFS_ERR_ARG, "null argument");
  if (!L) {
    c->has_lin = false;
    return CFS_OK;
  }
  std::vector<float> a((size_t)c->rp * c->rp, 0.0f);
  for (int k = 0; k < c->r; k++)
    for (int j = 0; j < c->r; j++) a[(size_t)k * c->rp + j] = L[(size_t)k * c->r + j];
  HIP_TRY(hipMemcpy(c->lin.p, a.data(), a.size() * 4, hipMemcpyHostToDevice), "hipMemcpy");
  c->has_lin = true;
  return CFS_OK;
}

int cfs_closure_get_scales(const cfs_closure* c, float* s_in, float* s_out) {
  if (!c || !s_in || !s_out) return fail(CFS_ERR_ARG, "null argument");
  HIP_TRY(hipMemcpy(s_in, c->s_in.p, (size_t)c->r * 4, hipMemcpyDeviceToHost), "hipMemcpy");
  HIP_TRY(hipMemcpy(s_out, c->s_out.p, (size_t)c->r * 4, hipMemcpyDeviceToHost), "hipMemcpy");
  return CFS_OK;
}

int cfs_closure_forward(cfs_closure* c, const float* x, float* y, int n, void* stream) {
  if (!c || !x || !y) return fail(CFS_ERR_ARG, "null argument");
  if (n < 1 || n > c->max_batch) return fail(CFS_ERR_ARG, "batch size outside [1, max_batch]");
  const hipStream_t s = (hipStream_t)stream;
  int st = c->forward(x, n, s);
  if (st) return st;
  HIP_TRY(hipMemcpyAsync(y, c->h[LAYERS - 1].p, (size_t)n * c->rp * 4, hipMemcpyDeviceToDevice, s), "hipMemcpyAsync");
  return CFS_OK;
}

int cfs_closure_gradients(cfs_closure* c, const float* x, const float* t, int n, float* grads, double* loss,
                          void* stream) {
  if (!c || !x || !t) return fail(CFS_ERR_ARG, "null argument");
  if (n < 1 || n > c->max_batch) return fail(CFS_ERR_ARG, "batch size outside [1, max_batch]");
  const hipStream_t s = (hipStream_t)stream;
  HIP_TRY(launch_zero_words(c->loss.p, 2, s), "zero launch");
  int st = c->forward(x, n, s);
  if (st) return st;
  if ((st = c->backward(x, t, n, s))) return st;
  if (grads)
    HIP_TRY(hipMemcpyAsync(grads, c->grads.p, c->nparams * 4, hipMemcpyDeviceToDevice, s), "hipMemcpyAsync");
  if (loss) HIP_TRY(hipMemcpyAsync(loss, c->loss.p, 8, hipMemcpyDeviceToDevice, s), "hipMemcpyAsync");
  return CFS_OK;
}

int cfs_closure_train_step(cfs_closure* c, const float* x, const float* t, const int32_t* rows, int n, float lr,
                           double* loss_sum, void* stream) {
  if (!c || !x || !t || !rows) return fail(CFS_ERR_ARG, "null argument");
  if (n < 1 || n > c->max_batch) return fail(CFS_ERR_ARG, "batch size outside [1, max_batch]");
  const hipStream_t s = (hipStream_t)stream;
  float* xb = c->xb.as<float>();
  float* tb = c->tb.as<float>();
  HIP_TRY(launch_gather_rows(x, rows, n, c->rp, xb, s), "gather launch");
  HIP_TRY(launch_gather_rows(t, rows, n, c->rp, tb, s), "gather launch");
  HIP_TRY(launch_zero_words(c->loss.p, 2, s), "zero launch");
  int st = c->forward(xb, n, s);
  if (st) return st;
  if ((st = c->backward(xb, tb, n, s))) return st;
  c->step++;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = (float)(1.0 / (1.0 - std::pow((double)b1, (double)c->step)));
  const float c2 = (float)(1.0 / (1.0 - std::pow((double)b2, (double)c->step)));
  HIP_TRY(launch_adam(c->params.as<float>(), c->grads.as<float>(), c->adam_m.as<float>(), c->adam_v.as<float>(),
                      (long)c->nparams, lr, b1, b2, 1e-8f, c1, c2, s),
          "adam launch");
  if ((st = c->repack(s))) return st;
  if (loss_sum) HIP_TRY(hipMemcpyAsync(loss_sum, c->loss.p, 8, hipMemcpyDeviceToDevice, s), "hipMemcpyAsync");
  return CFS_OK;
}

int cfs_closure_step(cfs_closure* c, const double* q_in, double* q_out, int n, double alpha, const double* forcing,
                     void* stream) {
  if (!c || !q_in || !q_out) return fail(CFS_ERR_ARG, "null argument");
  if (n < 1) return fail(CFS_ERR_ARG, "n must be positive");
  HIP_TRY(launch_closure_step(c->net(), q_in, q_out, n, alpha, forcing, (hipStream_t)stream), "closure step launch");
  return CFS_OK;
}

int cfs_advect_field(const float* vel, float* out, int X, int Y, int Z, float dt0, void* stream) {
  if (!vel || !out) return fail(CFS_ERR_ARG, "null argument");
  if (X < 4 || Y < 4 || Z < 4) return fail(CFS_ERR_ARG, "full-grid dims must be at least 4");
  if ((double)X * Y * Z * 3 >= 2147483647.0) return fail(CFS_ERR_ARG, "grid too large");
  HIP_TRY(launch_advect(vel, out, X, Y, Z, dt0, (hipStream_t)stream), "advect launch");
  return CFS_OK;
}

int cfs_advect_scalar(const float* vel, const float* old, float* out, int X, int Y, int Z, float dt0, void* stream) {
  if (!vel || !old || !out) return fail(CFS_ERR_ARG, "null argument");
  if (old == out) return fail(CFS_ERR_ARG, "the advected field must not be written in place");
  if (X < 4 || Y < 4 || Z < 4) return fail(CFS_ERR_ARG, "full-grid dims must be at least 4");
  if ((double)X * Y * Z * 3 >= 2147483647.0) return fail(CFS_ERR_ARG, "grid too large");
  HIP_TRY(launch_advect_scalar(vel, old, out, X, Y, Z, dt0, (hipStream_t)stream), "advect launch");
  return CFS_OK;
}

int cfs_reduced_advect(cfs_basis* U, const double* q_in, double* q_out, float* field_a, float* field_b, float dt0,
                       double alpha, const double* forcing, void* stream) {
  if (!U || !q_in || !q_out || !field_a || !field_b) return fail(CFS_ERR_ARG, "null argument");
  cfs_basis_info info;
  int st = cfs_basis_get_info(U, &info);
  if (st) return st;
  const int X = info.dims[0] + 2, Y = info.dims[1] + 2, Z = info.dims[2] + 2;
  const hipStream_t s = (hipStream_t)stream;
  HIP_TRY(launch_zero_words(field_a, (long)3 * X * Y * Z, s), "zero launch");
  if ((st = cfs_unproject(U, q_in, field_a, stream))) return st;
  if ((st = cfs_advect_field(field_a, field_b, X, Y, Z, dt0, stream))) return st;
  if ((st = cfs_project(U, field_b, q_out, stream))) return st;
  HIP_TRY(launch_axpb(q_out, info.num_cols, alpha, forcing, s), "axpb launch");
  return CFS_OK;
}

}  // extern "C"

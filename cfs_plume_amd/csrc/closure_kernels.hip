// closure_kernels.hip -- the reduced-advection step and its learned closure.
//
// Exact reduced advection (the quantity the reference's cubature step
// reducedAdvectCompressionFriendly, SUBSPACE_FLUID_3D_COMPRESSED_EIGEN.cpp:
// 1781-1854, approximates with key cells): q' = U^T A(U q), where A is the
// first-order semi-Lagrangian self-advection of the full-grid velocity field
// (VECTOR3_FIELD_3D::advect, VECTOR3_FIELD_3D.cpp:2001-2077: backtrace by
// dt * v, clamp to [1.5, res - 2.5], trilinear interpolation; the per-cell
// kernels are in grid_ops.h).
//
// Learned closure: an MLP r -> H -> H -> H -> r (GELU after each hidden
// layer) predicting the advection increment A(q) - q in scaled units.
// Weights are BF16 with FP32 master copies for training; the batched forward
// runs on MFMA (v_mfma_f32_16x16x16_bf16) with each activation split into a
// BF16 hi + lo pair and FP32 accumulation; the backward products are FP32
// MFMA (v_mfma_f32_16x16x4_f32); the optimiser is Adam.  The fused inference step evaluates the
// same network with the BF16 weights and FP32 arithmetic for a few samples
// per workgroup.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "grid_ops.h"
#include "mfma_idct.h"

namespace cfs {

// Clears n 32-bit words.  Used instead of hipMemsetAsync on the per-step
// paths, so that a caller who captures them in a HIP graph records kernel
// nodes only.
__global__ void __launch_bounds__(256) zero_words_kernel(uint32_t* __restrict__ p, long n) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = 0u;
}

// q_out = alpha * q_out + forcing (forcing may be null)
__global__ void axpb_kernel(double* __restrict__ q, int r, double alpha, const double* __restrict__ forcing) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < r) q[j] = alpha * q[j] + (forcing ? forcing[j] : 0.0);
}

// ---------------------------------------------------------------------------
// MLP pieces
// ---------------------------------------------------------------------------
__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// Y[n x M] = act(X[n x K] W + b) on MFMA.  X row-major FP32 (leading
// dimension ldx >= K, K a multiple of 16); W packed BF16: element (k, m) at
// Wp[((k / 4) * M + m) * 4 + k % 4], so a lane's four K values are one 8-byte
// load; M a multiple of 64.  Workgroup: 4 waves, wave w computes the 16 x 16
// tile of rows 16 blockIdx.x.., columns 64 blockIdx.y + 16 w...  A operand:
// lane l holds row l % 16, K = 4 (l / 16) .. +3, split into BF16 hi + lo (two
// products).  Z (pre-activation) is stored when non-null.
__global__ void __launch_bounds__(256) mlp_fwd_kernel(const float* __restrict__ X, int ldx, int n,
                                                      const uint16_t* __restrict__ Wp, const float* __restrict__ bias,
                                                      int K, int M, float* __restrict__ Z, float* __restrict__ Y,
                                                      int act) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int m0 = 16 * blockIdx.x, n0 = 64 * blockIdx.y + 16 * w;
  const int ar = m0 + (lane & 15), ak = 4 * (lane >> 4);
  const bool arow = ar < n;
  f32x4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
  for (int k0 = 0; k0 < K; k0 += 16) {
    float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (arow) a = *reinterpret_cast<const float4*>(X + (size_t)ar * ldx + k0 + ak);
    short h0, h1, h2, h3, l0, l1, l2, l3;
    split_bf16(a.x, h0, l0);
    split_bf16(a.y, h1, l1);
    split_bf16(a.z, h2, l2);
    split_bf16(a.w, h3, l3);
    const bf16x4_t ah = {h0, h1, h2, h3}, al = {l0, l1, l2, l3};
    const bf16x4_t b = *reinterpret_cast<const bf16x4_t*>(Wp + ((size_t)((k0 + ak) >> 2) * M + n0 + (lane & 15)) * 4);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah, b, acc, 0, 0, 0);
    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al, b, acc, 0, 0, 0);
  }
  // D layout: row 4 (l / 16) + i, column l % 16
  const int col = n0 + (lane & 15);
  const float bc = bias[col];
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int row = m0 + 4 * (lane >> 4) + i;
    if (row < n) {
      const float z = acc[i] + bc;
      if (Z) Z[(size_t)row * M + col] = z;
      Y[(size_t)row * M + col] = act ? gelu(z) : z;
    }
  }
}

// One product of a grouped GEMM launch: C[m, n] = sum_k A(m, k) B(k, n) with
// A(m, k) = A[m * sam + k * sak], B(k, n) = B[k * sbk + n * sbn], C row-major
// with leading dimension ldc; multiplied by gelu'(G[m * ldc + n]) when G is
// non-null; colsum[n] = sum_k B(k, n) as well when colsum is non-null.
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

// Two independent products in one launch (workgroups [0, tiles0) take p0,
// the rest p1).  FP32 on MFMA (v_mfma_f32_16x16x4_f32: FP32 products, FP32
// accumulation).  A workgroup computes one 16 x 16 tile of C; its 4 waves
// split K (wave w takes the 16-wide K steps w, w + 4, ...) and their
// accumulators meet in LDS in a fixed order, so results do not depend on
// scheduling.  In a K step a lane holds row (or column) l % 16 at
// k = k0 + 4 (l / 16) + i for the i-th of four MFMAs: every k meets its
// partner once, and a lane's four k are adjacent.
__global__ void __launch_bounds__(256) gemm_f32_kernel(GemmProblem p0, GemmProblem p1, int tiles0) {
  __shared__ f32x4_t part[4][64];
  __shared__ float cpart[4][64];
  const bool first = (int)blockIdx.x < tiles0;
  const GemmProblem& P = first ? p0 : p1;
  const int t = first ? blockIdx.x : blockIdx.x - tiles0;
  const int ntn = (P.N + 15) >> 4;
  const int m0 = 16 * (t / ntn), n0 = 16 * (t % ntn);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int am = m0 + (lane & 15), bn = n0 + (lane & 15), kq = 4 * (lane >> 4);
  const bool aok = am < P.M, bok = bn < P.N;
  const bool sums = P.colsum != nullptr && m0 == 0;
  f32x4_t acc = {0.0f, 0.0f, 0.0f, 0.0f};
  float cs = 0.0f;
  for (int k0 = 16 * w; k0 < P.K; k0 += 64) {
    float a[4], b[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int k = k0 + kq + i;
      a[i] = (aok && k < P.K) ? P.A[am * P.sam + k * P.sak] : 0.0f;
      b[i] = (bok && k < P.K) ? P.B[k * P.sbk + bn * P.sbn] : 0.0f;
    }
#pragma unroll
    for (int i = 0; i < 4; i++) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i], b[i], acc, 0, 0, 0);
    if (sums) cs += (b[0] + b[1]) + (b[2] + b[3]);
  }
  part[w][lane] = acc;
  cpart[w][lane] = cs;
  __syncthreads();
  if (w != 0) return;
  const f32x4_t s0 = part[0][lane], s1 = part[1][lane], s2 = part[2][lane], s3 = part[3][lane];
  if (!bok) return;
  // D layout: row 4 (l / 16) + i, column l % 16
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int row = m0 + kq + i;
    if (row < P.M) {
      float v = (s0[i] + s1[i]) + (s2[i] + s3[i]);
      if (P.G) v *= gelu_grad(P.G[(size_t)row * P.ldc + bn]);
      P.C[(size_t)row * P.ldc + bn] = v;
    }
  }
  if (sums && lane < 16) {
    float c = 0.0f;
#pragma unroll
    for (int ww = 0; ww < 4; ww++)
      c += (cpart[ww][lane] + cpart[ww][lane + 16]) + (cpart[ww][lane + 32] + cpart[ww][lane + 48]);
    P.colsum[bn] = c;
  }
}

// Mean-squared loss over the first r of the M output columns:
// dY = 2 (Y - T) / (n r) on those columns, 0 on the padding; the sum of
// squared errors is added to *loss (FP64).
__global__ void __launch_bounds__(256) loss_grad_kernel(const float* __restrict__ Y, const float* __restrict__ T,
                                                        int n, int M, int r, float* __restrict__ dY,
                                                        double* __restrict__ loss) {
  __shared__ double part[256];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double e2 = 0.0;
  if (i < n * M) {
    const int j = i % M;
    float g = 0.0f;
    if (j < r) {
      const float d = Y[i] - T[i];
      e2 = (double)d * (double)d;
      g = 2.0f * d / ((float)n * (float)r);
    }
    dY[i] = g;
  }
  part[threadIdx.x] = e2;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) part[threadIdx.x] += part[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0 && loss) atomicAdd(loss, part[0]);
}

// rows idx[0..n) of src (leading dimension ld) into dst
__global__ void gather_rows_kernel(const float* __restrict__ src, const int32_t* __restrict__ idx, int n, int ld,
                                   float* __restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * ld) return;
  const int row = i / ld, c = i % ld;
  dst[i] = src[(size_t)idx[row] * ld + c];
}

__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, long n, float lr, float b1, float b2, float eps, float c1,
                            float c2) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

// BF16 copies of layer weights W [K x M] (FP32 master): packed for the MFMA
// forward (Wp) and as FP32 values (Wq) for the backward pass.
__global__ void pack_kernel(const float* __restrict__ W, int K, int M, uint16_t* __restrict__ Wp,
                            float* __restrict__ Wq) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= K * M) return;
  const int k = i / M, m = i % M;
  const __bf16 h = (__bf16)W[i];
  Wq[i] = (float)h;
  Wp[((size_t)(k >> 2) * M + m) * 4 + (k & 3)] = (uint16_t)bf16_bits(h);
}

// Inference step for up to CL_NS samples per workgroup: q_out = alpha (q_in +
// s_out * (MLP(x) + x L)) + forcing with x = s_in * q_in (L the optional FP32
// linear skip; forcing per sample, [n][r], or null), FP64 q, FP32 network
// arithmetic with
// the BF16 weights.  Thread j computes output j of each layer for every
// sample of the workgroup (reading its packed weight column with 8-byte
// loads); activations live in LDS.
constexpr int CL_NS = 4;
constexpr int CL_H = 256;

struct ClosureNet {
  const uint16_t* Wp[4];
  const float* b[4];
  const float* lin;  // FP32 [rp][rp] linear skip added to the output (input-major), or null
  const float* s_in;
  const float* s_out;
  int r, rp, H;
};

// NS samples per workgroup; KS thread groups of CL_H split each dot product
// over K (partial sums combined through LDS), so that KS * CL_H threads keep
// weight loads in flight.
template <int NS, int KS>
__global__ void __launch_bounds__(CL_H * KS) closure_step_kernel(ClosureNet net, const double* __restrict__ q_in,
                                                                 double* __restrict__ q_out, int n, double alpha,
                                                                 const double* __restrict__ forcing) {
  __shared__ float act[2][NS][CL_H];
  __shared__ float part[KS > 1 ? KS - 1 : 1][NS][CL_H];
  const int j = threadIdx.x % CL_H, ks = threadIdx.x / CL_H;
  const int s0 = NS * blockIdx.x;
  const int ns = min(NS, n - s0);
  if (ks == 0)
    for (int s = 0; s < NS; s++)
      act[0][s][j] = (s < ns && j < net.r) ? (float)q_in[(size_t)(s0 + s) * net.r + j] * net.s_in[j] : 0.0f;
  __syncthreads();
  // linear skip x L (FP32), kept in registers until the output
  float lin[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) lin[s] = 0.0f;
  if (ks == 0 && net.lin && j < net.rp) {
    for (int k = 0; k < net.rp; k++) {
      const float a = net.lin[(size_t)k * net.rp + j];
#pragma unroll
      for (int s = 0; s < NS; s++) lin[s] += act[0][s][k] * a;
    }
  }
  int cur = 0;
  for (int l = 0; l < 4; l++) {
    const int K = l == 0 ? net.rp : net.H;
    const int M = l == 3 ? net.rp : net.H;
    float acc[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) acc[s] = 0.0f;
    if (j < M) {
      // this group's quarter (KS = 4) of the K range, in packed groups of 4
      const int q0 = ks * (K / 4) / KS, q1 = (ks + 1) * (K / 4) / KS;
      const uint16_t* wc = net.Wp[l] + (size_t)j * 4;
      for (int kq = q0; kq < q1; kq++) {
        const uint2 wbits = *reinterpret_cast<const uint2*>(wc + (size_t)kq * M * 4);
        const float w0 = __uint_as_float(wbits.x << 16), w1 = __uint_as_float(wbits.x & 0xffff0000u);
        const float w2 = __uint_as_float(wbits.y << 16), w3 = __uint_as_float(wbits.y & 0xffff0000u);
#pragma unroll
        for (int s = 0; s < NS; s++) {
          const float4 x = *reinterpret_cast<const float4*>(&act[cur][s][4 * kq]);
          acc[s] += x.x * w0 + x.y * w1 + x.z * w2 + x.w * w3;
        }
      }
    }
    if (KS > 1) {
      if (ks > 0)
#pragma unroll
        for (int s = 0; s < NS; s++) part[ks - 1][s][j] = acc[s];
      __syncthreads();
    }
    if (ks == 0 && j < M) {
      const float bj = net.b[l][j];
#pragma unroll
      for (int s = 0; s < NS; s++) {
        float z = acc[s];
#pragma unroll
        for (int g = 1; g < KS; g++) z += part[g - 1][s][j];
        z += bj;
        act[cur ^ 1][s][j] = l < 3 ? gelu(z) : z;
      }
    }
    __syncthreads();
    cur ^= 1;
  }
  if (ks == 0 && j < net.r) {
    for (int s = 0; s < ns; s++) {
      const size_t o = (size_t)(s0 + s) * net.r + j;
      const double qn = q_in[o] + (double)(act[cur][s][j] + lin[s]) * (double)net.s_out[j];
      q_out[o] = alpha * qn + (forcing ? forcing[o] : 0.0);
    }
  }
}

// ---------------------------------------------------------------------------
// launchers (called from closure.cpp)
// ---------------------------------------------------------------------------
hipError_t launch_advect(const float* vel, float* out, int X, int Y, int Z, float dt0, hipStream_t s) {
  const int n = X * Y * Z;
  hipLaunchKernelGGL(advect_vector_kernel<float>, dim3(grid_blocks(n)), dim3(GRID_TPB), 0, s, vel, out, X, Y, Z, dt0);
  return hipGetLastError();
}

hipError_t launch_advect_scalar(const float* vel, const float* old, float* out, int X, int Y, int Z, float dt0,
                                hipStream_t s) {
  const int n = X * Y * Z;
  hipLaunchKernelGGL(advect_scalar_field_kernel<float>, dim3(grid_blocks(n)), dim3(GRID_TPB), 0, s, vel, old, out, X,
                     Y, Z, dt0);
  return hipGetLastError();
}

hipError_t launch_zero_words(void* p, long nwords, hipStream_t s) {
  if (nwords <= 0) return hipSuccess;
  const long blocks = std::min<long>((nwords + 255) / 256, 4096);
  hipLaunchKernelGGL(zero_words_kernel, dim3((unsigned)blocks), dim3(256), 0, s, reinterpret_cast<uint32_t*>(p), nwords);
  return hipGetLastError();
}

hipError_t launch_axpb(double* q, int r, double alpha, const double* forcing, hipStream_t s) {
  hipLaunchKernelGGL(axpb_kernel, dim3((r + 255) / 256), dim3(256), 0, s, q, r, alpha, forcing);
  return hipGetLastError();
}

hipError_t launch_mlp_fwd(const float* X, int ldx, int n, const uint16_t* Wp, const float* bias, int K, int M,
                          float* Z, float* Y, int act, hipStream_t s) {
  hipLaunchKernelGGL(mlp_fwd_kernel, dim3((n + 15) / 16, M / 64), dim3(256), 0, s, X, ldx, n, Wp, bias, K, M, Z, Y,
                     act);
  return hipGetLastError();
}

// Grouped launch of the backward products of one layer: the weight gradient
// (with the bias gradient as its column sums) and, when p1 is non-null, the
// input gradient.
hipError_t launch_gemm_f32(const GemmProblem& p0, const GemmProblem* p1, hipStream_t s) {
  auto tiles = [](const GemmProblem& p) { return ((p.M + 15) / 16) * ((p.N + 15) / 16); };
  const int t0 = tiles(p0), t1 = p1 ? tiles(*p1) : 0;
  if (t0 + t1 == 0) return hipSuccess;
  hipLaunchKernelGGL(gemm_f32_kernel, dim3(t0 + t1), dim3(256), 0, s, p0, p1 ? *p1 : p0, t0);
  return hipGetLastError();
}


hipError_t launch_loss_grad(const float* Y, const float* T, int n, int M, int r, float* dY, double* loss,
                            hipStream_t s) {
  hipLaunchKernelGGL(loss_grad_kernel, dim3((n * M + 255) / 256), dim3(256), 0, s, Y, T, n, M, r, dY, loss);
  return hipGetLastError();
}

hipError_t launch_gather_rows(const float* src, const int32_t* idx, int n, int ld, float* dst, hipStream_t s) {
  hipLaunchKernelGGL(gather_rows_kernel, dim3((n * ld + 255) / 256), dim3(256), 0, s, src, idx, n, ld, dst);
  return hipGetLastError();
}

hipError_t launch_adam(float* p, const float* g, float* m, float* v, long n, float lr, float b1, float b2, float eps,
                       float c1, float c2, hipStream_t s) {
  hipLaunchKernelGGL(adam_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, g, m, v, n, lr, b1, b2, eps,
                     c1, c2);
  return hipGetLastError();
}

hipError_t launch_pack(const float* W, int K, int M, uint16_t* Wp, float* Wq, hipStream_t s) {
  hipLaunchKernelGGL(pack_kernel, dim3((K * M + 255) / 256), dim3(256), 0, s, W, K, M, Wp, Wq);
  return hipGetLastError();
}

hipError_t launch_closure_step(const ClosureNet& net, const double* q_in, double* q_out, int n, double alpha,
                               const double* forcing, hipStream_t s) {
  // one sample (the closure step of one run): one workgroup, the K range split
  // four ways; batches: four samples per workgroup
  if (n == 1)
    hipLaunchKernelGGL((closure_step_kernel<1, 4>), dim3(1), dim3(CL_H * 4), 0, s, net, q_in, q_out, n, alpha, forcing);
  else
    hipLaunchKernelGGL((closure_step_kernel<CL_NS, 1>), dim3((n + CL_NS - 1) / CL_NS), dim3(CL_H), 0, s, net, q_in,
                       q_out, n, alpha, forcing);
  return hipGetLastError();
}

}  // namespace cfs

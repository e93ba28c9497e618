// Learned warm start: the MLP 48 -> 256 -> 256 -> 256 -> 3 (GELU, fp32) that
// predicts a vertex's converged Newton displacement. Forward, backward, loss
// and Adam are hand-written kernels; one tiled GEMM kernel with fused
// epilogues covers every layer product.
//
// Parameters are one flat fp32 array: W1 [256][48], b1 [256], W2 [256][256],
// b2, W3 [256][256], b3, W4 [3][256], b4 [3] (W row-major, out x in).
// Activations of a forward pass with backward in mind ("work") hold
// Z1, H1, Z2, H2, Z3, H3, each [n][256] (pre-activation and GELU output).
#include <cmath>
#include <mutex>

#include "../../include/ipc_warp_amd.h"
#include "ipc_kernels.hpp"

// Activations of ipc_mlp_forward_noalloc: a file-scope device array (one
// copy per GPU, held until the process ends) instead of a caller or library
// buffer, so that inference allocates nothing. Host code must use the
// address hipGetSymbolAddress returns on each device (scratch_on_device),
// never this name: on the host it is a stand-in, not a device address.
static __device__ __align__(256) float scratch[3 * IPC_MLP_HIDDEN * IPC_MLP_NOALLOC_MAX_ROWS];

namespace ipcd {

constexpr int kIn = IPC_MLP_IN, kHid = IPC_MLP_HIDDEN, kOut = IPC_MLP_OUT;
constexpr long long kOffW1 = 0;
constexpr long long kOffB1 = kOffW1 + (long long)kHid * kIn;
constexpr long long kOffW2 = kOffB1 + kHid;
constexpr long long kOffB2 = kOffW2 + (long long)kHid * kHid;
constexpr long long kOffW3 = kOffB2 + kHid;
constexpr long long kOffB3 = kOffW3 + (long long)kHid * kHid;
constexpr long long kOffW4 = kOffB3 + kHid;
constexpr long long kOffB4 = kOffW4 + (long long)kOut * kHid;
constexpr long long kParams = kOffB4 + kOut;
static_assert(kParams == IPC_MLP_PARAMS, "parameter layout");

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
  // d/dz [z Phi(z)] = Phi(z) + z phi(z)
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

enum Epi { EPI_BIAS = 0, EPI_BIAS_GELU = 1, EPI_GELU_BACK = 2, EPI_ATOMIC = 3 };

constexpr int kTile = 64, kTk = 16, kGemmThreads = 256;

// C[r][c] = epilogue( sum_k A(r, k) B(k, c) ), 64 x 64 tiles, 16-deep K
// steps staged in LDS, 4 x 4 outputs per thread.
//   A(r, k) = TA ? A[k * lda + r] : A[r * lda + k]
//   B(k, c) = TB ? B[c * ldb + k] : B[k * ldb + c]
// EPI_BIAS:      C = acc + bias[c]
// EPI_BIAS_GELU: P[r][c] = acc + bias[c]; C = gelu(P)
// EPI_GELU_BACK: C = acc * gelu'(P[r][c])
// EPI_ATOMIC:    C += acc (atomically; blockIdx.z splits the K range)
template <bool TA, bool TB, int EPI>
__global__ void __launch_bounds__(kGemmThreads) gemm_kernel(int R, int Cn, int K, const float* __restrict__ A, int lda,
                                                            const float* __restrict__ B, int ldb, float* C, int ldc,
                                                            const float* __restrict__ bias, float* P, int k_chunk) {
  __shared__ float As[kTk][kTile + 4];
  __shared__ float Bs[kTk][kTile + 4];
  int tr = threadIdx.x / 16, tc = threadIdx.x % 16;
  int r0 = blockIdx.y * kTile, c0 = blockIdx.x * kTile;
  int k_begin = blockIdx.z * k_chunk;
  int k_end = min(K, k_begin + k_chunk);
  float acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = 0.f;
  for (int k0 = k_begin; k0 < k_end; k0 += kTk) {
    // stage A tile (64 rows x 16 k) and B tile (16 k x 64 cols): 1024 values each, 4 per thread
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      int idx = threadIdx.x + q * kGemmThreads;
      int rr, kk;
      if (TA) { kk = idx / kTile; rr = idx % kTile; }
      else { rr = idx / kTk; kk = idx % kTk; }
      int gr = r0 + rr, gk = k0 + kk;
      float v = 0.f;
      if (gr < R && gk < k_end) v = TA ? A[(long long)gk * lda + gr] : A[(long long)gr * lda + gk];
      As[kk][rr] = v;
      int cc, kb;
      if (TB) { cc = idx / kTk; kb = idx % kTk; }
      else { kb = idx / kTile; cc = idx % kTile; }
      int gc = c0 + cc, gkb = k0 + kb;
      float w = 0.f;
      if (gc < Cn && gkb < k_end) w = TB ? B[(long long)gc * ldb + gkb] : B[(long long)gkb * ldb + gc];
      Bs[kb][cc] = w;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kTk; ++kk) {
      float av[4], bv[4];
#pragma unroll
      for (int a = 0; a < 4; ++a) av[a] = As[kk][tr * 4 + a];
#pragma unroll
      for (int b = 0; b < 4; ++b) bv[b] = Bs[kk][tc * 4 + b];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = fmaf(av[a], bv[b], acc[a][b]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    int r = r0 + tr * 4 + a;
    if (r >= R) continue;
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      int c = c0 + tc * 4 + b;
      if (c >= Cn) continue;
      long long o = (long long)r * ldc + c;
      float v = acc[a][b];
      if (EPI == EPI_BIAS) {
        C[o] = v + bias[c];
      } else if (EPI == EPI_BIAS_GELU) {
        float z = v + bias[c];
        P[o] = z;
        C[o] = gelu(z);
      } else if (EPI == EPI_GELU_BACK) {
        C[o] = v * gelu_grad(P[o]);
      } else {
        unsafeAtomicAdd(C + o, v);
      }
    }
  }
}

// Column sums of D [n][m] added into out[m] (bias gradients).
__global__ void __launch_bounds__(256) colsum_kernel(int n, int m, const float* __restrict__ D, float* out) {
  int c = blockIdx.x * 64 + (threadIdx.x & 63);
  int part = threadIdx.x >> 6;  // 4 row partitions per block
  if (c >= m) return;
  float s = 0.f;
  for (int r = blockIdx.y * 4 + part; r < n; r += gridDim.y * 4) s += D[(long long)r * m + c];
  unsafeAtomicAdd(out + c, s);
}

// Weighted mean squared error over [n][3]: loss += sum w_i |y_i - t_i|^2 / W,
// dy = 2 w_i (y_i - t_i) / W with W = sum of weights (passed in).
__global__ void __launch_bounds__(256) mse_kernel(int n, const float* __restrict__ y, const float* __restrict__ t,
                                                  const float* __restrict__ w, float inv_wsum, float* __restrict__ dy,
                                                  double* loss) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  double l = 0.0;
  if (i < n) {
    float wi = w ? w[i] : 1.0f;
    for (int k = 0; k < kOut; ++k) {
      float d = y[3 * i + k] - t[3 * i + k];
      l += double(wi) * d * d;
      if (dy) dy[3 * i + k] = 2.0f * wi * d * inv_wsum;
    }
  }
  for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off, 64);
  __shared__ double part[4];
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = l;
  __syncthreads();
  if (threadIdx.x == 0 && loss) {
    double s = part[0] + part[1] + part[2] + part[3];
    unsafeAtomicAdd(loss, s * double(inv_wsum));
  }
}

// Adam (Kingma & Ba 2015) with bias correction, step t >= 1.
__global__ void __launch_bounds__(256) adam_kernel(long long n, float* __restrict__ p, const float* __restrict__ g,
                                                   float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                                                   float b2, float eps, float c1, float c2) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float gi = g[i];
  float mi = b1 * m[i] + (1.0f - b1) * gi;
  float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void __launch_bounds__(256) zero_f32_kernel(long long n, float* p) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

// --------------------------------------------------------------- launchers ----
static inline unsigned gblocks(long long n, int bs) { return unsigned((n + bs - 1) / bs); }

template <bool TA, bool TB, int EPI>
static hipError_t gemm(int R, int Cn, int K, const float* A, int lda, const float* B, int ldb, float* C, int ldc,
                       const float* bias, float* P, hipStream_t s, int splits = 1) {
  if (R <= 0 || Cn <= 0) return hipSuccess;
  int chunk = (K + splits - 1) / splits;
  chunk = (chunk + kTk - 1) / kTk * kTk;
  int zs = (K + chunk - 1) / chunk;
  dim3 grid(gblocks(Cn, kTile), gblocks(R, kTile), zs < 1 ? 1 : zs);
  hipLaunchKernelGGL((gemm_kernel<TA, TB, EPI>), grid, dim3(kGemmThreads), 0, s, R, Cn, K, A, lda, B, ldb, C, ldc,
                     bias, P, chunk);
  return hipGetLastError();
}

#define IPC_TRY(expr)                \
  do {                               \
    hipError_t _e = (expr);          \
    if (_e != hipSuccess) return _e; \
  } while (0)

hipError_t mlp_forward(const float* prm, int n, const float* x, float* y, float* work, float* tmp, hipStream_t s) {
  // Without a work buffer (inference) the activations ping-pong in tmp.
  long long L = (long long)n * kHid;
  float *Z1, *H1, *Z2, *H2, *Z3, *H3;
  if (work) {
    Z1 = work; H1 = work + L; Z2 = work + 2 * L; H2 = work + 3 * L; Z3 = work + 4 * L; H3 = work + 5 * L;
  } else {
    Z1 = Z2 = Z3 = tmp;
    H1 = H3 = tmp + L;
    H2 = tmp + 2 * L;
  }
  IPC_TRY((gemm<false, true, EPI_BIAS_GELU>(n, kHid, kIn, x, kIn, prm + kOffW1, kIn, H1, kHid, prm + kOffB1, Z1, s)));
  IPC_TRY((gemm<false, true, EPI_BIAS_GELU>(n, kHid, kHid, H1, kHid, prm + kOffW2, kHid, H2, kHid, prm + kOffB2, Z2, s)));
  IPC_TRY((gemm<false, true, EPI_BIAS_GELU>(n, kHid, kHid, H2, kHid, prm + kOffW3, kHid, H3, kHid, prm + kOffB3, Z3, s)));
  return gemm<false, true, EPI_BIAS>(n, kOut, kHid, H3, kHid, prm + kOffW4, kHid, y, kOut, prm + kOffB4, nullptr, s);
}

hipError_t mlp_backward(const float* prm, int n, const float* x, const float* work, const float* dy, float* grad,
                        float* tmp, hipStream_t s) {
  long long L = (long long)n * kHid;
  const float *Z1 = work, *H1 = work + L, *Z2 = work + 2 * L, *H2 = work + 3 * L, *Z3 = work + 4 * L,
              *H3 = work + 5 * L;
  float *dA = tmp, *dB = tmp + L;
  int splits = n / 2048 + 1;
  if (splits > 64) splits = 64;
  // layer 4: dW4 += dy^T H3, db4 += sum dy, dZ3 = (dy W4) * gelu'(Z3)
  IPC_TRY((gemm<true, false, EPI_ATOMIC>(kOut, kHid, n, dy, kOut, H3, kHid, grad + kOffW4, kHid, nullptr, nullptr, s,
                                         splits)));
  hipLaunchKernelGGL(colsum_kernel, dim3(gblocks(kOut, 64), 32), dim3(256), 0, s, n, kOut, dy, grad + kOffB4);
  IPC_TRY((gemm<false, false, EPI_GELU_BACK>(n, kHid, kOut, dy, kOut, prm + kOffW4, kHid, dA, kHid, nullptr,
                                             const_cast<float*>(Z3), s)));
  // layer 3
  IPC_TRY((gemm<true, false, EPI_ATOMIC>(kHid, kHid, n, dA, kHid, H2, kHid, grad + kOffW3, kHid, nullptr, nullptr, s,
                                         splits)));
  hipLaunchKernelGGL(colsum_kernel, dim3(gblocks(kHid, 64), 32), dim3(256), 0, s, n, kHid, dA, grad + kOffB3);
  IPC_TRY((gemm<false, false, EPI_GELU_BACK>(n, kHid, kHid, dA, kHid, prm + kOffW3, kHid, dB, kHid, nullptr,
                                             const_cast<float*>(Z2), s)));
  // layer 2
  IPC_TRY((gemm<true, false, EPI_ATOMIC>(kHid, kHid, n, dB, kHid, H1, kHid, grad + kOffW2, kHid, nullptr, nullptr, s,
                                         splits)));
  hipLaunchKernelGGL(colsum_kernel, dim3(gblocks(kHid, 64), 32), dim3(256), 0, s, n, kHid, dB, grad + kOffB2);
  IPC_TRY((gemm<false, false, EPI_GELU_BACK>(n, kHid, kHid, dB, kHid, prm + kOffW2, kHid, dA, kHid, nullptr,
                                             const_cast<float*>(Z1), s)));
  // layer 1
  IPC_TRY((gemm<true, false, EPI_ATOMIC>(kHid, kIn, n, dA, kHid, x, kIn, grad + kOffW1, kIn, nullptr, nullptr, s,
                                         splits)));
  hipLaunchKernelGGL(colsum_kernel, dim3(gblocks(kHid, 64), 32), dim3(256), 0, s, n, kHid, dA, grad + kOffB1);
  return hipGetLastError();
}

hipError_t mlp_mse(const float* y, const float* t, const float* w, int n, float wsum, float* dy, double* loss,
                   hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(mse_kernel, dim3(gblocks(n, 256)), dim3(256), 0, s, n, y, t, w, 1.0f / wsum, dy, loss);
  return hipGetLastError();
}

hipError_t mlp_adam(float* p, const float* g, float* m, float* v, long long n, float lr, float b1, float b2, float eps,
                    int step, hipStream_t s) {
  float c1 = 1.0f - powf(b1, float(step)), c2 = 1.0f - powf(b2, float(step));
  hipLaunchKernelGGL(adam_kernel, dim3(gblocks(n, 256)), dim3(256), 0, s, n, p, g, m, v, lr, b1, b2, eps, c1, c2);
  return hipGetLastError();
}

// Device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept; nullptr with *err set when the
// lookup fails.
static float* scratch_on_device(hipError_t* err) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static float* addr[kMaxDevices] = {};
  int dev = 0;
  *err = hipGetDevice(&dev);
  if (*err != hipSuccess) return nullptr;
  if (dev < 0 || dev >= kMaxDevices) {
    *err = hipErrorInvalidDevice;
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void* p = nullptr;
    *err = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (*err != hipSuccess) return nullptr;
    addr[dev] = static_cast<float*>(p);
  }
  return addr[dev];
}

hipError_t mlp_forward_noalloc(const float* prm, int n, const float* x, float* y, hipStream_t s) {
  hipError_t err = hipSuccess;
  float* tmp = scratch_on_device(&err);
  if (!tmp) return err;
  return mlp_forward(prm, n, x, y, nullptr, tmp, s);
}

hipError_t zero_f32(float* p, long long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(zero_f32_kernel, dim3(gblocks(n, 256)), dim3(256), 0, s, n, p);
  return hipGetLastError();
}

}  // namespace ipcd

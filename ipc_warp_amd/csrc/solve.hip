// Block-sparse products and conjugate gradients over the BSR matrices of
// bsr.hip (3x3 fp32 blocks): the linear solve of the Newton system
// (warp.optim.linear.cg at time_integrator_base.py:500-507 and
// affine_body_dynamics.py:580-592).
//
// y = alpha A x + beta y: eight lanes per block row, each lane every eighth
// block of the row, the three sums reduced across the eight lanes.
//
// CG keeps every scalar on the device: an iteration is five launches
// (q = A p with per-block partial sums of p.q; one block adds them in a
// fixed order and forms alpha; x, r, z updated with partial sums of r.z and
// r.r; one block forms beta; p = z + beta p), so no host round trip happens
// between convergence checks and the result does not depend on scheduling.
// The iterations between two checks are captured once as a HIP graph and
// replayed (a launch-bound loop: five small launches per iteration).
#include <cmath>
#include <cstdio>

#include "../../include/ipc_warp_amd.h"
#include "cull.hpp"
#include "ipc_kernels.hpp"

using namespace ipcd;

struct ipc_cg {
  DevBuf r, z, p, q, dinv;  // float [3 n]
  DevBuf part;              // double [3][kGrid]: partial sums (p.q | r.z, r.r, b.b)
  DevBuf scal;              // double [8]: r.z (two slots), r.r (two slots), b.b, -, alpha, beta
  hipGraphExec_t graph = nullptr;  // check_every iterations, captured for the buffers below
  hipStream_t capture = nullptr;   // private stream the graph is captured on (the caller's may be the null stream)
  const void* graph_key[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  long long graph_n = -1;
  int graph_iters = 0, graph_jacobi = -1;
  double* host = nullptr;   // pinned readback
};

namespace {

constexpr int kB = 256;
constexpr int kGroup = 8;           // lanes per block row in the products
constexpr int kRowsPerBlock = kB / kGroup;
constexpr int kGrid = 8192;         // most blocks of the launches that leave partial sums

int cg_fail(int code, const char* what) {
  set_error(what);
  return code;
}

int cg_hip(hipError_t e, const char* where) {
  if (e == hipSuccess) return IPC_OK;
  char buf[512];
  std::snprintf(buf, sizeof(buf), "%s: %s", where, hipGetErrorString(e));
  set_error(buf);
  return -int(e);
}

inline int nblk(long long n, int per) { return int((n + per - 1) / per); }

__device__ inline double block_sum(double v, double* sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int d = kB / 2; d > 0; d >>= 1) {
    if (threadIdx.x < d) sh[threadIdx.x] += sh[threadIdx.x + d];
    __syncthreads();
  }
  const double s = sh[0];
  __syncthreads();
  return s;
}

// Row r of A x for the calling 8-lane group (all its lanes get the sums).
__device__ inline void row_product(int r, int j, const int* off, const int* col, const float* val, const float* x,
                                   float& a0, float& a1, float& a2) {
  a0 = a1 = a2 = 0.0f;
  const int k1 = off[r + 1];
  for (int k = off[r] + j; k < k1; k += kGroup) {
    const float* v = val + 9ll * k;
    const float* xc = x + 3ll * col[k];
    const float x0 = xc[0], x1 = xc[1], x2 = xc[2];
    a0 += v[0] * x0 + v[1] * x1 + v[2] * x2;
    a1 += v[3] * x0 + v[4] * x1 + v[5] * x2;
    a2 += v[6] * x0 + v[7] * x1 + v[8] * x2;
  }
#pragma unroll
  for (int m = kGroup / 2; m > 0; m >>= 1) {
    a0 += __shfl_xor(a0, m, kGroup);
    a1 += __shfl_xor(a1, m, kGroup);
    a2 += __shfl_xor(a2, m, kGroup);
  }
}

__global__ __launch_bounds__(kB) void bsr_mv_kernel(int nr, const int* off, const int* col, const float* val,
                                                    const float* x, float* y, float alpha, float beta) {
  const int j = threadIdx.x & (kGroup - 1);
  for (int r = (blockIdx.x * kB + threadIdx.x) / kGroup; r < nr; r += gridDim.x * kRowsPerBlock) {
    float a0, a1, a2;
    row_product(r, j, off, col, val, x, a0, a1, a2);
    if (j == 0) {
      float* yr = y + 3ll * r;
      if (beta == 0.0f) {
        yr[0] = alpha * a0;
        yr[1] = alpha * a1;
        yr[2] = alpha * a2;
      } else {
        yr[0] = alpha * a0 + beta * yr[0];
        yr[1] = alpha * a1 + beta * yr[1];
        yr[2] = alpha * a2 + beta * yr[2];
      }
    }
  }
}

// q = A p and the per-block partial sums of p.q.
__global__ __launch_bounds__(kB) void cg_apply_kernel(int nr, const int* off, const int* col, const float* val,
                                                      const float* p, float* q, double* part_pq) {
  __shared__ double sh[kB];
  const int j = threadIdx.x & (kGroup - 1);
  double pq = 0.0;
  for (int r = (blockIdx.x * kB + threadIdx.x) / kGroup; r < nr; r += gridDim.x * kRowsPerBlock) {
    float a0, a1, a2;
    row_product(r, j, off, col, val, p, a0, a1, a2);
    if (j == 0) {
      float* qr = q + 3ll * r;
      const float* pr = p + 3ll * r;
      qr[0] = a0;
      qr[1] = a1;
      qr[2] = a2;
      pq += double(pr[0]) * a0 + double(pr[1]) * a1 + double(pr[2]) * a2;
    }
  }
  const double s = block_sum(pq, sh);
  if (threadIdx.x == 0) part_pq[blockIdx.x] = s;
}

// One block: the partial sums of the previous launch added in a fixed order
// (each thread a strided subset in order, then a fixed tree), and the CG
// coefficient derived from them. mode 0: alpha = rz / pq into scal[6];
// mode 1: rz_new and rr_new into their slots and beta = rz_new / rz_old
// into scal[7].
__global__ __launch_bounds__(kB) void cg_scalars_kernel(int mode, const double* part_a, const double* part_b, int np,
                                                        double* scal, int par) {
  __shared__ double sh[kB];
  double a = 0.0, b = 0.0;
  for (int i = threadIdx.x; i < np; i += kB) {
    a += part_a[i];
    if (part_b) b += part_b[i];
  }
  a = block_sum(a, sh);
  if (part_b) b = block_sum(b, sh);
  if (threadIdx.x != 0) return;
  const double rz_old = scal[par];
  if (mode == 0) {
    scal[6] = (a > 0.0 && std::isfinite(a)) ? rz_old / a : 0.0;
  } else {
    scal[1 - par] = a;
    scal[2 + (1 - par)] = b;
    const double beta = rz_old != 0.0 ? a / rz_old : 0.0;
    scal[7] = std::isfinite(beta) ? beta : 0.0;
  }
}

// x += alpha p, r -= alpha q, z = D^-1 r; partial sums of r.z and r.r.
__global__ __launch_bounds__(kB) void cg_update_kernel(long long n, float* x, float* r, const float* p, const float* q,
                                                       const float* dinv, float* z, const double* scal,
                                                       double* part_rz, double* part_rr) {
  __shared__ double sh[kB];
  const double alpha = scal[6];
  double rz = 0.0, rr = 0.0;
  for (long long i = blockIdx.x * (long long)kB + threadIdx.x; i < n; i += (long long)gridDim.x * kB) {
    x[i] = float(double(x[i]) + alpha * double(p[i]));
    const float ri = float(double(r[i]) - alpha * double(q[i]));
    r[i] = ri;
    const float zi = dinv ? dinv[i] * ri : ri;
    if (dinv) z[i] = zi;
    rz += double(ri) * zi;
    rr += double(ri) * ri;
  }
  const double a = block_sum(rz, sh), b = block_sum(rr, sh);
  if (threadIdx.x == 0) {
    part_rz[blockIdx.x] = a;
    part_rr[blockIdx.x] = b;
  }
}

// p = z + beta p.
__global__ __launch_bounds__(kB) void cg_direction_kernel(long long n, float* p, const float* z, const double* scal) {
  const double beta = scal[7];
  for (long long i = blockIdx.x * (long long)kB + threadIdx.x; i < n; i += (long long)gridDim.x * kB)
    p[i] = float(double(z[i]) + beta * double(p[i]));
}

// r = b - q (q = A x), z = D^-1 r, p = z; partial sums of r.z, r.r, b.b.
__global__ __launch_bounds__(kB) void cg_init_kernel(long long n, const float* b, const float* q, float* r, float* z,
                                                     float* p, const float* dinv, double* part_rz, double* part_rr,
                                                     double* part_bb) {
  __shared__ double sh[kB];
  double rz = 0.0, rr = 0.0, bb = 0.0;
  for (long long i = blockIdx.x * (long long)kB + threadIdx.x; i < n; i += (long long)gridDim.x * kB) {
    const float ri = b[i] - q[i];
    r[i] = ri;
    const float zi = dinv ? dinv[i] * ri : ri;
    if (dinv) z[i] = zi;
    p[i] = zi;
    rz += double(ri) * zi;
    rr += double(ri) * ri;
    bb += double(b[i]) * b[i];
  }
  const double a = block_sum(rz, sh), c = block_sum(rr, sh), d = block_sum(bb, sh);
  if (threadIdx.x == 0) {
    part_rz[blockIdx.x] = a;
    part_rr[blockIdx.x] = c;
    part_bb[blockIdx.x] = d;
  }
}

__global__ __launch_bounds__(kB) void cg_init_sums_kernel(const double* part_rz, const double* part_rr,
                                                          const double* part_bb, int np, double* scal) {
  __shared__ double sh[kB];
  double a = 0.0, b = 0.0, c = 0.0;
  for (int i = threadIdx.x; i < np; i += kB) {
    a += part_rz[i];
    b += part_rr[i];
    c += part_bb[i];
  }
  a = block_sum(a, sh);
  b = block_sum(b, sh);
  c = block_sum(c, sh);
  if (threadIdx.x == 0) {
    scal[0] = a;
    scal[2] = b;
    scal[4] = c;
  }
}

// Jacobi: the inverse diagonal of each row's diagonal block (1 where it is
// missing, zero or not finite).
__global__ void cg_diag_kernel(int nr, const int* off, const int* col, const float* val, float* dinv) {
  const int r = blockIdx.x * kB + threadIdx.x;
  if (r >= nr) return;
  float d[3] = {1.0f, 1.0f, 1.0f};
  for (int k = off[r]; k < off[r + 1]; ++k)
    if (col[k] == r) {
      const float* v = val + 9ll * k;
      d[0] = v[0];
      d[1] = v[4];
      d[2] = v[8];
      break;
    }
#pragma unroll
  for (int c = 0; c < 3; ++c) dinv[3ll * r + c] = (d[c] != 0.0f && std::isfinite(d[c])) ? 1.0f / d[c] : 1.0f;
}

int grid_for(long long items, int per) {
  const int g = nblk(items, per);
  return g < 1 ? 1 : (g < kGrid ? g : kGrid);
}

}  // namespace

extern "C" {

int ipc_bsr_mv(int n_rows, const int* offsets, const int* columns, const float* values, const float* x, float* y,
               float alpha, float beta, void* stream) {
  if (n_rows < 0) return cg_fail(IPC_EINVAL, "bsr_mv: negative n_rows");
  if (n_rows == 0) return IPC_OK;
  if (!offsets || !x || !y) return cg_fail(IPC_EINVAL, "bsr_mv: null pointer");
  const int g = nblk(n_rows, kRowsPerBlock) < 8192 ? nblk(n_rows, kRowsPerBlock) : 8192;
  hipLaunchKernelGGL(bsr_mv_kernel, dim3(g), dim3(kB), 0, static_cast<hipStream_t>(stream), n_rows, offsets, columns,
                     values, x, y, alpha, beta);
  return cg_hip(hipGetLastError(), "bsr_mv");
}

int ipc_cg_create(ipc_cg** out) {
  if (!out) return cg_fail(IPC_EINVAL, "cg_create: null output");
  ipc_cg* w = new (std::nothrow) ipc_cg();
  if (!w) return cg_fail(IPC_EINVAL, "cg_create: out of host memory");
  hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&w->host), 8 * sizeof(double));
  if (e != hipSuccess) {
    delete w;
    return cg_hip(e, "cg_create");
  }
  *out = w;
  return IPC_OK;
}

int ipc_cg_destroy(ipc_cg* w) {
  if (!w) return IPC_OK;
  if (w->graph) hipGraphExecDestroy(w->graph);
  if (w->capture) hipStreamDestroy(w->capture);
  for (DevBuf* d : {&w->r, &w->z, &w->p, &w->q, &w->dinv, &w->part, &w->scal}) d->release();
  if (w->host) hipHostFree(w->host);
  delete w;
  return IPC_OK;
}

int ipc_bsr_cg(ipc_cg* w, int n_rows, const int* offsets, const int* columns, const float* values, const float* b,
               float* x, float tol, int maxiter, int check_every, int jacobi, int* iterations, double* residual,
               void* stream) {
  if (!w) return cg_fail(IPC_EINVAL, "bsr_cg: null workspace");
  if (n_rows < 0) return cg_fail(IPC_EINVAL, "bsr_cg: negative n_rows");
  if (!(tol >= 0.0f)) return cg_fail(IPC_EINVAL, "bsr_cg: tol must be >= 0");
  if (iterations) *iterations = 0;
  if (residual) *residual = 0.0;
  if (n_rows == 0) return IPC_OK;
  if (!offsets || !b || !x) return cg_fail(IPC_EINVAL, "bsr_cg: null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const long long n = 3ll * n_rows;
  hipError_t e = hipSuccess;
  for (DevBuf* d : {&w->r, &w->z, &w->p, &w->q}) {
    if (e == hipSuccess) e = d->ensure(sizeof(float) * size_t(n));
  }
  if (e == hipSuccess && jacobi) e = w->dinv.ensure(sizeof(float) * size_t(n));
  if (e == hipSuccess) e = w->part.ensure(sizeof(double) * 3 * kGrid);
  if (e == hipSuccess) e = w->scal.ensure(sizeof(double) * 8);
  if (e != hipSuccess) return cg_hip(e, "bsr_cg: workspace");
  float* r = w->r.as<float>();
  float* p = w->p.as<float>();
  float* q = w->q.as<float>();
  float* z = jacobi ? w->z.as<float>() : r;
  const float* dinv = jacobi ? w->dinv.as<float>() : nullptr;
  double* part = w->part.as<double>();
  double* scal = w->scal.as<double>();
  const int gr = grid_for(n_rows, kRowsPerBlock), gv = grid_for(n, kB);
  if (jacobi)
    hipLaunchKernelGGL(cg_diag_kernel, dim3(nblk(n_rows, kB)), dim3(kB), 0, s, n_rows, offsets, columns, values,
                       w->dinv.as<float>());
  hipLaunchKernelGGL(cg_apply_kernel, dim3(gr), dim3(kB), 0, s, n_rows, offsets, columns, values, x, q, part);
  hipLaunchKernelGGL(cg_init_kernel, dim3(gv), dim3(kB), 0, s, n, b, q, r, z, p, dinv, part, part + kGrid,
                     part + 2 * kGrid);
  hipLaunchKernelGGL(cg_init_sums_kernel, dim3(1), dim3(kB), 0, s, part, part + kGrid, part + 2 * kGrid, gv, scal);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(w->host, scal, 8 * sizeof(double), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return cg_hip(e, "bsr_cg: start");
  const double bb = w->host[4];
  const double goal = double(tol) * double(tol) * bb;
  double rr = w->host[2];
  if (residual) *residual = std::sqrt(rr);
  if (rr <= goal) return IPC_OK;
  if (maxiter <= 0) maxiter = 3 * n_rows;
  if (check_every <= 0) check_every = 10;
  // One iteration: five launches with every scalar on the device.
  auto iteration = [&](hipStream_t cs, int par) {
    hipLaunchKernelGGL(cg_apply_kernel, dim3(gr), dim3(kB), 0, cs, n_rows, offsets, columns, values, p, q, part);
    hipLaunchKernelGGL(cg_scalars_kernel, dim3(1), dim3(kB), 0, cs, 0, part, (const double*)nullptr, gr, scal, par);
    hipLaunchKernelGGL(cg_update_kernel, dim3(gv), dim3(kB), 0, cs, n, x, r, p, q, dinv, z, scal, part + kGrid,
                       part + 2 * kGrid);
    hipLaunchKernelGGL(cg_scalars_kernel, dim3(1), dim3(kB), 0, cs, 1, part + kGrid, part + 2 * kGrid, gv, scal, par);
    hipLaunchKernelGGL(cg_direction_kernel, dim3(gv), dim3(kB), 0, cs, n, p, z, scal);
  };
  // check_every iterations (an even count keeps the two r.z slots in step)
  // as one captured graph, replayed between convergence checks
  const int chunk = (check_every + 1) & ~1;
  const void* key[6] = {offsets, columns, values, b, x, w};
  bool same = w->graph && w->graph_n == n && w->graph_iters == chunk && w->graph_jacobi == (jacobi ? 1 : 0);
  for (int k = 0; k < 6 && same; ++k) same = w->graph_key[k] == key[k];
  if (!same) {
    if (w->graph) {
      hipGraphExecDestroy(w->graph);
      w->graph = nullptr;
    }
    hipGraph_t g = nullptr;
    if (!w->capture) e = hipStreamCreate(&w->capture);
    if (e == hipSuccess) e = hipStreamBeginCapture(w->capture, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      for (int i = 0; i < chunk; ++i) iteration(w->capture, i & 1);
      e = hipStreamEndCapture(w->capture, &g);
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&w->graph, g, nullptr, nullptr, 0);
    if (g) hipGraphDestroy(g);
    if (e != hipSuccess) {
      w->graph = nullptr;
      return cg_hip(e, "bsr_cg: graph capture");
    }
    w->graph_n = n;
    w->graph_iters = chunk;
    w->graph_jacobi = jacobi ? 1 : 0;
    for (int k = 0; k < 6; ++k) w->graph_key[k] = key[k];
  }
  int done = 0;
  while (done < maxiter) {
    e = hipGraphLaunch(w->graph, s);
    if (e == hipSuccess) e = hipMemcpyAsync(w->host, scal, 8 * sizeof(double), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return cg_hip(e, "bsr_cg: iterations");
    done += chunk;
    rr = w->host[2];  // after an even number of iterations the current sums are in slot 0
    if (iterations) *iterations = done;
    if (residual) *residual = std::sqrt(rr);
    if (rr <= goal || !std::isfinite(rr)) break;
  }
  return IPC_OK;
}

}  // extern "C"

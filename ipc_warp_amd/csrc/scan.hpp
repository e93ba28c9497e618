// Device-wide int helpers shared by the assembly and active-set code: a fill
// kernel and an exclusive prefix sum in tiles of 4096 (256 lanes x 16),
// then the tile sums, then the tile offsets added. Internal linkage: every
// translation unit that includes this keeps its own copy of the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "cull.hpp"

namespace ipcd {
namespace {

constexpr int kScanB = 256;
constexpr int kScanPer = 16;
constexpr int kScanTile = kScanB * kScanPer;

__global__ void fill_i32_kernel(int* p, int n, int v) {
  for (int i = blockIdx.x * kScanB + threadIdx.x; i < n; i += gridDim.x * kScanB) p[i] = v;
}

hipError_t fill_i32(int* p, int n, int v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int blocks = (n + kScanB - 1) / kScanB;
  const int g = blocks < 1024 ? blocks : 1024;
  hipLaunchKernelGGL(fill_i32_kernel, dim3(g), dim3(kScanB), 0, s, p, n, v);
  return hipGetLastError();
}

__device__ int block_exclusive(int v, int* sh, int& total) {
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int d = 1; d < kScanB; d <<= 1) {
    const int u = tid >= d ? sh[tid - d] : 0;
    __syncthreads();
    sh[tid] += u;
    __syncthreads();
  }
  total = sh[kScanB - 1];
  const int r = sh[tid] - v;
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(kScanB) void scan_tiles_kernel(const int* in, int n, int* out, int* tile_sum) {
  __shared__ int sh[kScanB];
  const long long base = (long long)blockIdx.x * kScanTile + threadIdx.x * kScanPer;
  int v[kScanPer];
  int acc = 0;
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    v[k] = base + k < n ? in[base + k] : 0;
    acc += v[k];
  }
  int total;
  int run = block_exclusive(acc, sh, total);
#pragma unroll
  for (int k = 0; k < kScanPer; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 0) tile_sum[blockIdx.x] = total;
}

__global__ __launch_bounds__(kScanB) void scan_sums_kernel(int* tile_sum, int nt) {
  __shared__ int sh[kScanB];
  int carry = 0;
  for (int b = 0; b < nt; b += kScanB) {
    const int i = b + threadIdx.x;
    const int v = i < nt ? tile_sum[i] : 0;
    int total;
    const int e = block_exclusive(v, sh, total);
    if (i < nt) tile_sum[i] = carry + e;
    carry += total;
  }
}

__global__ void scan_add_kernel(int* out, int n, const int* tile_sum) {
  const int tile = blockIdx.x + 1;
  const int add = tile_sum[tile];
  const long long base = (long long)tile * kScanTile;
  for (int k = threadIdx.x; k < kScanTile; k += kScanB)
    if (base + k < n) out[base + k] += add;
}

// out[0..n) = exclusive prefix sums of in[0..n); in and out distinct. `part`
// holds one int per tile.
hipError_t exclusive_scan(const int* in, int n, int* out, DevBuf& part, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int nt = (n + kScanTile - 1) / kScanTile;
  hipError_t e = part.ensure(sizeof(int) * size_t(nt));
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(scan_tiles_kernel, dim3(nt), dim3(kScanB), 0, s, in, n, out, part.as<int>());
  if (nt > 1) {
    hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(kScanB), 0, s, part.as<int>(), nt);
    hipLaunchKernelGGL(scan_add_kernel, dim3(nt - 1), dim3(kScanB), 0, s, out, n, part.as<int>());
  }
  return hipGetLastError();
}

}  // namespace
}  // namespace ipcd

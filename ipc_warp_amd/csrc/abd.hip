// Affine-body contact reduction (affine_body_dynamics.py:144-190): vertex
// contact Hessian blocks and forces carried onto the affine degrees of
// freedom q_b = [p, a0, a1, a2] of each body, with x_i = p + r_i.x a0 +
// r_i.y a1 + r_i.z a2 for the rest-frame coordinates r_i. A vertex triplet
// (v_i, v_j, H) adds c_a(v_i) c_c(v_j) H to the 3x3 block (4 b_i + a,
// 4 b_j + c) for a, c in 0..3, with c_0 = 1 and c_k = r_k-1; a block touching
// a fixed body is zero.
//
// The reference writes those 16 blocks per vertex triplet as new triplets
// and sorts them into BSR; with ~1e7 vertex triplets onto a handful of
// bodies that is ~1.6e8 triplets falling into a few rows. Here the blocks
// are summed where they are produced:
//   mark     one bit per (body, body) pair that receives a nonzero block,
//            each workgroup setting the distinct pairs of its triplets once;
//   pattern  the pairs as a BSR over bodies (rows sorted) and its DOF-level
//            BSR: body pair (A, B) owns the 16 blocks (4A + a, 4B + c);
//   reduce   each wave walks a contiguous run of vertex triplets, each lane
//            a fixed stride of it, summing its blocks in fp32 registers
//            while its body pair stays the same; on a change the wave adds
//            the lanes' sums per body pair (lane order, fp32) into one of
//            up to 64 fp64 accumulator copies (wave index modulo the copy
//            count) with one atomic per value;
//   finish   values = scale * (sum of the copies) rounded to fp32.
// The forces reduce the same way per vertex (one body per vertex).
#include <cstdio>

#include "../../include/ipc_warp_amd.h"
#include "cull.hpp"
#include "ipc_kernels.hpp"

using namespace ipcd;

struct ipc_abd {
  DevBuf bits;     // uint32 [n_b][words]: body pairs with a nonzero block
  DevBuf pcnt;     // int [n_b + 1]
  DevBuf poff;     // int [n_b + 1]: body-level offsets
  DevBuf pcol;     // int [pairs]: body-level columns
  DevBuf part;     // scan partials
  DevBuf acc;      // double [16 * pairs * 9]
  DevBuf facc;     // double [4 * n_b * 3]
  int* host = nullptr;
};

namespace {

constexpr int kB = 256;
constexpr int kWaves = kB / 64;
constexpr int kScanTile = kB * 16;
constexpr int kMaxBodies = 8192;

int abd_fail(int code, const char* what) {
  set_error(what);
  return code;
}

int abd_hip(hipError_t e, const char* where) {
  if (e == hipSuccess) return IPC_OK;
  char buf[512];
  std::snprintf(buf, sizeof(buf), "%s: %s", where, hipGetErrorString(e));
  set_error(buf);
  return -int(e);
}

inline int nblk(long long n, int per) { return int((n + per - 1) / per); }
inline int words_of(int nb) { return (nb + 31) >> 5; }

__global__ void fill_u32_kernel(unsigned* p, long long n) {
  for (long long i = blockIdx.x * (long long)kB + threadIdx.x; i < n; i += (long long)gridDim.x * kB) p[i] = 0u;
}

__global__ void fill_f64_kernel(double* p, long long n) {
  for (long long i = blockIdx.x * (long long)kB + threadIdx.x; i < n; i += (long long)gridDim.x * kB) p[i] = 0.0;
}

hipError_t zero_u32(unsigned* p, long long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int g = nblk(n, kB) < 2048 ? nblk(n, kB) : 2048;
  hipLaunchKernelGGL(fill_u32_kernel, dim3(g), dim3(kB), 0, s, p, n);
  return hipGetLastError();
}

hipError_t zero_f64(double* p, long long n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int g = nblk(n, kB) < 2048 ? nblk(n, kB) : 2048;
  hipLaunchKernelGGL(fill_f64_kernel, dim3(g), dim3(kB), 0, s, p, n);
  return hipGetLastError();
}

// exclusive scan of n ints (one tile per block, tile sums, tile offsets)
__device__ int block_exclusive(int v, int* sh, int& total) {
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int d = 1; d < kB; d <<= 1) {
    const int u = tid >= d ? sh[tid - d] : 0;
    __syncthreads();
    sh[tid] += u;
    __syncthreads();
  }
  total = sh[kB - 1];
  const int r = sh[tid] - v;
  __syncthreads();
  return r;
}

__global__ __launch_bounds__(kB) void scan_tiles_kernel(const int* in, int n, int* out, int* tile_sum) {
  __shared__ int sh[kB];
  const long long base = (long long)blockIdx.x * kScanTile + threadIdx.x * 16;
  int v[16];
  int acc = 0;
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    v[k] = base + k < n ? in[base + k] : 0;
    acc += v[k];
  }
  int total;
  int run = block_exclusive(acc, sh, total);
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 0) tile_sum[blockIdx.x] = total;
}

__global__ __launch_bounds__(kB) void scan_sums_kernel(int* tile_sum, int nt) {
  __shared__ int sh[kB];
  int carry = 0;
  for (int b = 0; b < nt; b += kB) {
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
  for (int k = threadIdx.x; k < kScanTile; k += kB)
    if (base + k < n) out[base + k] += add;
}

hipError_t exclusive_scan(const int* in, int n, int* out, DevBuf& part, hipStream_t s) {
  const int nt = nblk(n, kScanTile);
  hipError_t e = part.ensure(sizeof(int) * size_t(nt));
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(scan_tiles_kernel, dim3(nt), dim3(kB), 0, s, in, n, out, part.as<int>());
  if (nt > 1) {
    hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(kB), 0, s, part.as<int>(), nt);
    hipLaunchKernelGGL(scan_add_kernel, dim3(nt - 1), dim3(kB), 0, s, out, n, part.as<int>());
  }
  return hipGetLastError();
}

// The body pair of a kept vertex triplet, or -1: in range, both bodies free,
// some value nonzero.
__device__ inline bool triplet_pair(const int* rows, const int* cols, const float* vals, int t, int nv, int nb,
                                    const int* body, const int* fixed, int& bi, int& bj) {
  const int r = rows[t], c = cols[t];
  if (unsigned(r) >= unsigned(nv) || unsigned(c) >= unsigned(nv)) return false;
  bi = body[r];
  bj = body[c];
  if (unsigned(bi) >= unsigned(nb) || unsigned(bj) >= unsigned(nb) || fixed[bi] || fixed[bj]) return false;
  bool nz = false;
#pragma unroll
  for (int k = 0; k < 9; ++k) nz |= !(vals[9ll * t + k] == 0.0f);
  return nz;
}

constexpr int kMarkPer = 64;     // triplets per thread of the mark kernel
constexpr int kMarkSet = 256;    // LDS set of body pairs per workgroup

// Bits of the body pairs receiving a nonzero block. Each workgroup collects
// the distinct pairs of its kB * kMarkPer triplets in an LDS set and sets
// each bit once; a pair that finds the set full sets its bit itself.
__global__ __launch_bounds__(kB) void mark_kernel(const int* rows, const int* cols, const float* vals, int m, int nv,
                                                  int nb, const int* body, const int* fixed, int words,
                                                  unsigned* bits) {
  __shared__ unsigned set[kMarkSet];
  for (int i = threadIdx.x; i < kMarkSet; i += kB) set[i] = 0xffffffffu;
  __syncthreads();
  const long long base = (long long)blockIdx.x * kB * kMarkPer;
  for (int k = 0; k < kMarkPer; ++k) {
    const long long t = base + (long long)k * kB + threadIdx.x;
    int bi, bj;
    if (t >= m || !triplet_pair(rows, cols, vals, int(t), nv, nb, body, fixed, bi, bj)) continue;
    const unsigned key = unsigned(bi) * unsigned(nb) + unsigned(bj);
    unsigned h = (key * 2654435761u) >> 24;  // 8 bits: kMarkSet slots
    bool placed = false;
    for (int probe = 0; probe < 16 && !placed; ++probe) {
      const unsigned slot = (h + probe) & (kMarkSet - 1);
      const unsigned old = atomicCAS(&set[slot], 0xffffffffu, key);
      placed = old == 0xffffffffu || old == key;
    }
    if (!placed) atomicOr(bits + (long long)bi * words + (bj >> 5), 1u << (bj & 31));
  }
  __syncthreads();
  for (int i = threadIdx.x; i < kMarkSet; i += kB) {
    const unsigned key = set[i];
    if (key == 0xffffffffu) continue;
    const int bi = int(key / unsigned(nb)), bj = int(key % unsigned(nb));
    atomicOr(bits + (long long)bi * words + (bj >> 5), 1u << (bj & 31));
  }
}

__global__ void row_count_kernel(const unsigned* bits, int nb, int words, int* cnt) {
  const int r = blockIdx.x * kB + threadIdx.x;
  if (r > nb) return;
  int n = 0;
  if (r < nb)
    for (int k = 0; k < words; ++k) n += __popc(bits[(long long)r * words + k]);
  cnt[r] = n;
}

__global__ void row_fill_kernel(const unsigned* bits, int nb, int words, const int* off, int* col) {
  const int r = blockIdx.x * kB + threadIdx.x;
  if (r >= nb) return;
  int o = off[r];
  for (int k = 0; k < words; ++k) {
    unsigned w = bits[(long long)r * words + k];
    while (w) {
      const int j = __ffs(int(w)) - 1;
      col[o++] = 32 * k + j;
      w &= w - 1;
    }
  }
}

// DOF-level pattern: row 4A + a holds, for each body B of body row A in
// order, the blocks (4A + a, 4B + c), c = 0..3.
__global__ void dof_pattern_kernel(int nb, const int* poff, const int* pcol, int* offsets, int* columns) {
  const int R = blockIdx.x * kB + threadIdx.x;
  if (R > 4 * nb) return;
  if (R == 4 * nb) {
    offsets[R] = 16 * poff[nb];
    return;
  }
  const int A = R >> 2, a = R & 3;
  const int p0 = poff[A], len = poff[A + 1] - p0;
  const int o = 16 * p0 + 4 * a * len;
  offsets[R] = o;
  for (int j = 0; j < len; ++j) {
    const int B = pcol[p0 + j];
    if (columns) {
#pragma unroll
      for (int c = 0; c < 4; ++c) columns[o + 4 * j + c] = 4 * B + c;
    }
  }
}

// index of block (4A + a, 4B + c) in the DOF-level arrays
__device__ inline long long dof_block(const int* poff, const int* pcol, int A, int B, int a, int c) {
  const int p0 = poff[A], len = poff[A + 1] - p0;
  int lo = 0, hi = len - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (pcol[p0 + mid] < B) lo = mid + 1;
    else hi = mid;
  }
  return 16ll * p0 + 4ll * a * len + 4 * lo + c;
}

__device__ inline float wave_sum(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

constexpr int kPad = 65;    // LDS row pitch of the flush table (conflict-free both ways)
constexpr int kChunk = 48;  // values per LDS pass of a flush (3 passes of the 144)

// Adds every pending lane sum into the accumulator, one body pair at a time:
// per pass of 48 values the lanes of that pair write their sums to LDS
// (others write zeros), then lane k < 48 adds value k over the 64 lanes in
// lane order and issues the fp64 atomic. `out` is this wave's accumulator
// copy.
__device__ void flush_pairs(int& cur_a, int cur_b, float* acc, int lane, const int* poff, const int* pcol,
                            float* sh, double* out) {
  while (true) {
    const unsigned long long pend = __ballot(cur_a >= 0);
    if (!pend) break;
    const int leader = __ffsll(static_cast<long long>(pend)) - 1;
    const int A = __shfl(cur_a, leader), B = __shfl(cur_b, leader);
    const bool mine = cur_a == A && cur_b == B;
#pragma unroll
    for (int c = 0; c < 144 / kChunk; ++c) {
#pragma unroll
      for (int k = 0; k < kChunk; ++k) sh[k * kPad + lane] = mine ? acc[c * kChunk + k] : 0.0f;
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      if (lane < kChunk) {
        float sum = 0.0f;
        for (int j = 0; j < 64; ++j) sum += sh[lane * kPad + j];
        const int k = c * kChunk + lane;
        const int blk = k / 9, e = k - 9 * blk;
        const long long o = dof_block(poff, pcol, A, B, blk >> 2, blk & 3);
        atomicAdd(out + 9 * o + e, static_cast<double>(sum));
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    }
    if (mine) {
      cur_a = -1;
#pragma unroll
      for (int k = 0; k < 144; ++k) acc[k] = 0.0f;
    }
  }
}

// One wave per block, each a contiguous run of per_wave triplets.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void reduce_hessian_kernel(const int* rows, const int* cols, const float* vals, int m,
                                                            int per_wave, int nv, int nb, const int* body,
                                                            const int* fixed, const float* rest, const int* poff,
                                                            const int* pcol, double* acc_reps, int reps,
                                                            long long rep_stride) {
  __shared__ float sh[kChunk * kPad];
  const int lane = threadIdx.x;
  // waves spread their atomics over `reps` copies of the accumulator, so a
  // body pair shared by every wave is not one chain of atomics per value
  double* out = acc_reps + (blockIdx.x % reps) * rep_stride;
  const long long w0 = (long long)blockIdx.x * per_wave;
  if (w0 >= m) return;
  const long long w1 = w0 + per_wave < m ? w0 + per_wave : m;
  float acc[144];
#pragma unroll
  for (int k = 0; k < 144; ++k) acc[k] = 0.0f;
  int cur_a = -1, cur_b = -1;
  for (long long base = w0; base < w1; base += 64) {
    const long long t = base + lane;
    int bi = -1, bj = -1;
    const bool kept = t < w1 && triplet_pair(rows, cols, vals, int(t), nv, nb, body, fixed, bi, bj);
    // a lane moving to another body pair first hands its sums to the wave
    if (__ballot(kept && cur_a >= 0 && (bi != cur_a || bj != cur_b)))
      flush_pairs(cur_a, cur_b, acc, lane, poff, pcol, sh, out);
    if (kept) {
      cur_a = bi;
      cur_b = bj;
      const int r = rows[t], c = cols[t];
      const float ca[4] = {1.0f, rest[3ll * r], rest[3ll * r + 1], rest[3ll * r + 2]};
      const float cc[4] = {1.0f, rest[3ll * c], rest[3ll * c + 1], rest[3ll * c + 2]};
      float h[9];
#pragma unroll
      for (int k = 0; k < 9; ++k) h[k] = vals[9 * t + k];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const float w = ca[a] * cc[b];
#pragma unroll
          for (int k = 0; k < 9; ++k) acc[(a * 4 + b) * 9 + k] += w * h[k];
        }
    }
  }
  flush_pairs(cur_a, cur_b, acc, lane, poff, pcol, sh, out);
}

__global__ void finish_values_kernel(const double* acc, long long n, int reps, double scale, float* values) {
  for (long long i = blockIdx.x * (long long)kB + threadIdx.x; i < n; i += (long long)gridDim.x * kB) {
    double v = 0.0;
    for (int r = 0; r < reps; ++r) v += acc[r * n + i];
    values[i] = static_cast<float>(scale * v);
  }
}

// Forces: vertex i of a free body adds (f, r.x f, r.y f, r.z f) to its
// body's four rows; the wave adds its lanes' terms per body.
__global__ __launch_bounds__(kB) void reduce_force_kernel(const float* force, const float* rest, const int* body,
                                                          const int* fixed, int nv, int nb, double* out) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * kB + threadIdx.x;
  int b = -1;
  float v[12];
  if (i < nv) {
    b = body[i];
    if (unsigned(b) >= unsigned(nb) || fixed[b]) b = -1;
  }
  if (b >= 0) {
    const float f[3] = {force[3ll * i], force[3ll * i + 1], force[3ll * i + 2]};
    const float c[4] = {1.0f, rest[3ll * i], rest[3ll * i + 1], rest[3ll * i + 2]};
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int k = 0; k < 3; ++k) v[3 * a + k] = c[a] * f[k];
  } else {
#pragma unroll
    for (int k = 0; k < 12; ++k) v[k] = 0.0f;
  }
  while (true) {
    const unsigned long long pend = __ballot(b >= 0);
    if (!pend) break;
    const int B = __shfl(b, __ffsll(static_cast<long long>(pend)) - 1);
    const bool mine = b == B;
    float keep = 0.0f;
#pragma unroll
    for (int k = 0; k < 12; ++k) {
      const float s = wave_sum(mine ? v[k] : 0.0f);
      if (lane == k) keep = s;
    }
    if (lane < 12) atomicAdd(out + 12ll * B + lane, static_cast<double>(keep));
    if (mine) b = -1;
  }
}

__global__ void finish_force_kernel(const double* acc, int n, float* out) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i < n) out[i] += static_cast<float>(acc[i]);
}

}  // namespace

extern "C" {

int ipc_abd_create(ipc_abd** out) {
  if (!out) return abd_fail(IPC_EINVAL, "abd_create: null output");
  ipc_abd* w = new (std::nothrow) ipc_abd();
  if (!w) return abd_fail(IPC_EINVAL, "abd_create: out of host memory");
  hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&w->host), sizeof(int));
  if (e != hipSuccess) {
    delete w;
    return abd_hip(e, "abd_create");
  }
  *out = w;
  return IPC_OK;
}

int ipc_abd_destroy(ipc_abd* w) {
  if (!w) return IPC_OK;
  for (DevBuf* d : {&w->bits, &w->pcnt, &w->poff, &w->pcol, &w->part, &w->acc, &w->facc}) d->release();
  if (w->host) hipHostFree(w->host);
  delete w;
  return IPC_OK;
}

int ipc_abd_reduce_contacts(ipc_abd* w, int n_bodies, const int* body, const float* rest, const int* fixed,
                            int n_vertices, const int* rows, const int* cols, const float* vals, int n_triplets,
                            const float* force, double scale, int* offsets, int* columns, float* values, int capacity,
                            float* reduced_force, void* stream) {
  if (!w) return abd_fail(IPC_EINVAL, "abd_reduce_contacts: null workspace");
  if (n_bodies <= 0 || n_bodies > kMaxBodies)
    return abd_fail(IPC_EINVAL, "abd_reduce_contacts: n_bodies must be in [1, 8192]");
  if (n_vertices < 0 || n_triplets < 0 || capacity < 0) return abd_fail(IPC_EINVAL, "abd_reduce_contacts: negative size");
  if (!body || !rest || !fixed || !offsets) return abd_fail(IPC_EINVAL, "abd_reduce_contacts: null pointer");
  if (n_triplets > 0 && (!rows || !cols || !vals)) return abd_fail(IPC_EINVAL, "abd_reduce_contacts: null triplets");
  if (capacity > 0 && (!columns || !values)) return abd_fail(IPC_EINVAL, "abd_reduce_contacts: null columns/values");
  if (force && !reduced_force) return abd_fail(IPC_EINVAL, "abd_reduce_contacts: force given without reduced_force");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nb = n_bodies, words = words_of(nb);
  hipError_t e = w->bits.ensure(sizeof(unsigned) * size_t(nb) * words);
  if (e == hipSuccess) e = w->pcnt.ensure(sizeof(int) * (size_t(nb) + 1));
  if (e == hipSuccess) e = w->poff.ensure(sizeof(int) * (size_t(nb) + 1));
  if (e != hipSuccess) return abd_hip(e, "abd_reduce_contacts: workspace");
  e = zero_u32(w->bits.as<unsigned>(), (long long)nb * words, s);
  if (e == hipSuccess && n_triplets > 0)
    hipLaunchKernelGGL(mark_kernel, dim3(nblk(n_triplets, kB * kMarkPer)), dim3(kB), 0, s, rows, cols, vals, n_triplets,
                       n_vertices, nb, body, fixed, words, w->bits.as<unsigned>());
  if (e == hipSuccess) e = hipGetLastError();
  if (e == hipSuccess)
    hipLaunchKernelGGL(row_count_kernel, dim3(nblk(nb + 1, kB)), dim3(kB), 0, s, w->bits.as<unsigned>(), nb, words,
                       w->pcnt.as<int>());
  if (e == hipSuccess) e = hipGetLastError();
  if (e == hipSuccess) e = exclusive_scan(w->pcnt.as<int>(), nb + 1, w->poff.as<int>(), w->part, s);
  if (e == hipSuccess) e = hipMemcpyAsync(w->host, w->poff.as<int>() + nb, sizeof(int), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return abd_hip(e, "abd_reduce_contacts: body pattern");
  const int pairs = *w->host;
  const long long blocks = 16ll * pairs;
  if (blocks > 0x7fffffffll) return abd_fail(IPC_EINVAL, "abd_reduce_contacts: pattern too large");
  // accumulator replicas: up to 64, within 64 MiB
  long long reps = (64ll << 20) / (72 * (blocks > 0 ? blocks : 1));
  reps = reps < 1 ? 1 : (reps > 64 ? 64 : reps);
  e = w->pcol.ensure(sizeof(int) * size_t(pairs > 0 ? pairs : 1));
  if (e == hipSuccess) e = w->acc.ensure(sizeof(double) * size_t(blocks > 0 ? blocks : 1) * 9 * size_t(reps));
  if (e == hipSuccess) e = w->facc.ensure(sizeof(double) * size_t(nb) * 12);
  if (e != hipSuccess) return abd_hip(e, "abd_reduce_contacts: accumulators");
  hipLaunchKernelGGL(row_fill_kernel, dim3(nblk(nb, kB)), dim3(kB), 0, s, w->bits.as<unsigned>(), nb, words,
                     w->poff.as<int>(), w->pcol.as<int>());
  e = hipGetLastError();
  if (e == hipSuccess)
    hipLaunchKernelGGL(dof_pattern_kernel, dim3(nblk(4 * nb + 1, kB)), dim3(kB), 0, s, nb, w->poff.as<int>(),
                       w->pcol.as<int>(), offsets, blocks <= capacity ? columns : nullptr);
  if (e == hipSuccess) e = hipGetLastError();
  if (e != hipSuccess) return abd_hip(e, "abd_reduce_contacts: pattern");
  if (blocks > capacity) return int(blocks);
  if (force) {
    e = zero_f64(w->facc.as<double>(), 12ll * nb, s);
    if (e == hipSuccess && n_vertices > 0)
      hipLaunchKernelGGL(reduce_force_kernel, dim3(nblk(n_vertices, kB)), dim3(kB), 0, s, force, rest, body, fixed,
                         n_vertices, nb, w->facc.as<double>());
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess)
      hipLaunchKernelGGL(finish_force_kernel, dim3(nblk(12 * nb, kB)), dim3(kB), 0, s, w->facc.as<double>(), 12 * nb,
                         reduced_force);
    if (e == hipSuccess) e = hipGetLastError();
    if (e != hipSuccess) return abd_hip(e, "abd_reduce_contacts: force");
  }
  if (blocks == 0) return IPC_OK;
  e = zero_f64(w->acc.as<double>(), 9 * blocks * reps, s);
  if (e == hipSuccess && n_triplets > 0) {
    // each wave a contiguous run of at least 64 triplets, about 4096 waves
    long long per = (n_triplets + 4095) / 4096;
    per = (per + 63) / 64 * 64;
    const long long waves = (n_triplets + per - 1) / per;
    hipLaunchKernelGGL(reduce_hessian_kernel, dim3(int(waves)), dim3(64), 0, s, rows, cols, vals, n_triplets,
                       int(per), n_vertices, nb, body, fixed, rest, w->poff.as<int>(), w->pcol.as<int>(),
                       w->acc.as<double>(), int(reps), 9 * blocks);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    const long long n = 9 * blocks;
    const int g = nblk(n, kB) < 4096 ? nblk(n, kB) : 4096;
    hipLaunchKernelGGL(finish_values_kernel, dim3(g), dim3(kB), 0, s, w->acc.as<double>(), n, int(reps), scale,
                       values);
    e = hipGetLastError();
  }
  return abd_hip(e, "abd_reduce_contacts: reduce");
}

}  // extern "C"

// Block-sparse (BSR, 3x3 fp32 blocks) assembly of the Newton system: the
// triplet -> BSR conversion the contact Hessian goes through, a value-only
// refill over a kept pattern, and the sum of two BSR matrices
// (warp.sparse.bsr_from_triplets / bsr_axpy at dynamic_contacts.py:290-297,
// affine_body_dynamics.py:500-503, augmented_lagrangian.py:631-632).
//
// Triplet -> BSR in row buckets instead of a global key sort:
//   count    one lane per triplet: drop it (row or column out of range, or
//            all nine values zero), else take its slot in the row's bucket
//            with an atomic on the row's counter;
//   scan     bucket starts = exclusive scan of the row counts;
//   scatter  (column, triplet) into the row's bucket;
//   sort     one wave per row: sort the bucket by (column, triplet) — by
//            rank in registers up to 64 entries, by rank against the row
//            staged in LDS up to 1024 — and count the distinct columns; the
//            (rare) longer rows go to a segmented radix sort (hipCUB);
//   scan     row offsets = exclusive scan of the distinct counts;
//   emit     one wave per row: the first entry of each column run writes the
//            column and the fp32 sum of its run, added in ascending triplet
//            order from +0, and remembers where its run starts (the pattern
//            a refill reuses).
// Every step reads or writes each triplet a bounded number of times; the
// result does not depend on the order the atomics hand out slots.
#include <cstdio>

#include <hipcub/hipcub.hpp>

#include "../../include/ipc_warp_amd.h"
#include "cull.hpp"
#include "ipc_kernels.hpp"
#include "scan.hpp"

using namespace ipcd;

struct ipc_bsr {
  DevBuf cnt;      // int [n_rows + 1]: per-row kept triplets, then distinct columns
  DevBuf start;    // int [n_rows + 1]: bucket start of each row
  DevBuf slot;     // int [n_triplets]: slot in the row's bucket, -1 if dropped
  DevBuf bucket;   // u64 [n_triplets]: column << 32 | triplet, by row, unsorted
  DevBuf sorted;   // u64 [n_triplets]: by row, then column, then triplet
  DevBuf big;      // int [3 * big_cap + 1]: rows longer than kSortCap (begin, end, row), count
  DevBuf temp;     // segmented radix sort workspace
  DevBuf seg;      // int [capacity + 1]: first sorted entry of each block
  DevBuf part;     // scan partial sums
  DevBuf nnz;      // int [1]: blocks of the kept pattern
  int* host = nullptr;  // pinned readback of the block count
  int n_rows = 0, n_triplets = 0, max_blocks = 0;
  bool planned = false;
};

namespace {

constexpr int kB = 256;
constexpr int kWaves = kB / 64;
constexpr int kSortCap = 1024;  // entries of one row staged in LDS by its wave
constexpr int kWin = 256;       // triplet values staged per window and wave

int bsr_fail(int code, const char* what) {
  set_error(what);
  return code;
}

int bsr_hip(hipError_t e, const char* where) {
  if (e == hipSuccess) return IPC_OK;
  char buf[512];
  std::snprintf(buf, sizeof(buf), "%s: %s", where, hipGetErrorString(e));
  set_error(buf);
  return -int(e);
}

inline int nblk(long long n, int per) { return int((n + per - 1) / per); }

// ---------------------------------------------------- triplets -> BSR ---
__device__ inline bool block_nonzero(const float* v) {
  bool nz = false;
#pragma unroll
  for (int k = 0; k < 9; ++k) nz |= !(v[k] == 0.0f);  // NaN counts as nonzero
  return nz;
}

// Lanes of a wave whose triplets are kept in one row take consecutive slots
// with one atomic per run of neighbouring lanes (the contact Hessian writes
// the four blocks of a pair's block row next to each other).
__global__ __launch_bounds__(kB) void bsr_count_kernel(const int* rows, const int* cols, const float* vals, int m,
                                                       int nr, int nc, int* cnt, int* slot) {
  const int t = blockIdx.x * kB + threadIdx.x;
  const int lane = threadIdx.x & 63;
  int key = -1;
  if (t < m) {
    const int r = rows[t], c = cols[t];
    if (unsigned(r) < unsigned(nr) && unsigned(c) < unsigned(nc) && block_nonzero(vals + 9ll * t)) key = r;
  }
  const bool keep = key >= 0;
  const int prev = __shfl_up(key, 1);
  const bool head = keep && (lane == 0 || prev != key);
  const unsigned long long hm = __ballot(head), km = __ballot(keep);
  // this lane's run starts at the last head at or below it and ends before
  // the next head or dropped lane
  const unsigned long long upto = lane == 63 ? ~0ull : (2ull << lane) - 1;
  const int h = keep ? 63 - __clzll(hm & upto) : lane;
  const unsigned long long brk = (hm | ~km) & ~((h == 63 ? ~0ull : (2ull << h) - 1));
  const int end = brk ? __ffsll(static_cast<long long>(brk)) - 1 : 64;
  int base = 0;
  if (head) base = atomicAdd(cnt + key, end - h);
  base = __shfl(base, h);
  if (t < m) slot[t] = keep ? base + (lane - h) : -1;
}

__global__ void bsr_scatter_kernel(const int* rows, const int* cols, const int* slot, const int* start, int m,
                                   unsigned long long* bucket) {
  const int t = blockIdx.x * kB + threadIdx.x;
  if (t >= m) return;
  const int s = slot[t];
  if (s >= 0) bucket[start[rows[t]] + s] = (static_cast<unsigned long long>(unsigned(cols[t])) << 32) | unsigned(t);
}

__device__ inline int key_col(unsigned long long k) { return int(k >> 32); }
__device__ inline int key_triplet(unsigned long long k) { return int(unsigned(k)); }

// Rank of K keys per lane (entries g + lane + 64 k of the row) among the
// row's L keys, read one at a time at a wave-uniform index from `at`. A key
// starts its column's run (a head) when its rank equals the number of keys
// with a smaller column.
template <int K, typename At>
__device__ inline int rank_pass(At at, int L, int g, int lane, int b, const unsigned long long* bucket,
                                unsigned long long* sorted) {
  unsigned long long key[K];
  unsigned col[K];
  int rank[K], below[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = g + lane + 64 * k;
    key[k] = i < L ? bucket[b + i] : ~0ull;
    col[k] = unsigned(key[k] >> 32);
    rank[k] = 0;
    below[k] = 0;
  }
#pragma unroll 4
  for (int q = 0; q < L; ++q) {
    const unsigned long long kq = at(q);
    const unsigned cq = unsigned(kq >> 32);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      rank[k] += kq < key[k];
      below[k] += cq < col[k];
    }
  }
  int heads = 0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const bool in = g + lane + 64 * k < L;
    if (in) sorted[b + rank[k]] = key[k];
    heads += __popcll(__ballot(in && rank[k] == below[k]));
  }
  return heads;
}

template <typename At>
__device__ inline int rank_row(At at, int L, int lane, int b, const unsigned long long* bucket,
                               unsigned long long* sorted) {
  int heads = 0;
  int g = 0;
  for (; L - g > 192; g += 256) heads += rank_pass<4>(at, L, g, lane, b, bucket, sorted);
  if (L - g > 128) heads += rank_pass<3>(at, L, g, lane, b, bucket, sorted);
  else if (L - g > 64) heads += rank_pass<2>(at, L, g, lane, b, bucket, sorted);
  else if (L - g > 0) heads += rank_pass<1>(at, L, g, lane, b, bucket, sorted);
  return heads;
}

// One wave per row: the row's bucket sorted by (column, triplet) into
// `sorted`, its distinct column count into rnnz. Keys are distinct (the
// triplet index is), so each key's rank is its sorted position. Rows longer
// than kSortCap are listed in `big` (begin, end, row) for the segmented radix
// sort instead.
__global__ __launch_bounds__(kB) void bsr_sort_rows_kernel(const int* start, int nr, const unsigned long long* bucket,
                                                           unsigned long long* sorted, int* rnnz, int* big,
                                                           int big_cap) {
  __shared__ unsigned long long sk[kWaves][kSortCap];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  unsigned long long* s = sk[w];
  for (int r = blockIdx.x * kWaves + w; r < nr; r += gridDim.x * kWaves) {
    const int b = start[r], L = start[r + 1] - b;
    int heads = 0;
    if (L == 0) {
    } else if (L <= 64) {
      const unsigned long long k = lane < L ? bucket[b + lane] : ~0ull;
      const int khi = int(k >> 32), klo = int(unsigned(k));
      int rank = 0, below = 0;
      // q is wave-uniform: each key is read from its lane into a scalar
      // register (v_readlane), not through the LDS crossbar
#pragma unroll 4
      for (int q = 0; q < L; ++q) {
        const int hi = __builtin_amdgcn_readlane(khi, q), lo = __builtin_amdgcn_readlane(klo, q);
        const unsigned long long kq = (static_cast<unsigned long long>(unsigned(hi)) << 32) | unsigned(lo);
        rank += kq < k;
        below += unsigned(hi) < unsigned(khi);
      }
      if (lane < L) sorted[b + rank] = k;
      heads = __popcll(__ballot(lane < L && rank == below));
    } else if (L <= kSortCap) {
      for (int i = lane; i < L; i += 64) s[i] = bucket[b + i];
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      heads = rank_row([&](int q) { return s[q]; }, L, lane, b, bucket, sorted);
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    } else {
      // at most n_triplets / (kSortCap + 1) rows are this long: a slot exists
      if (lane == 0) {
        const int i = atomicAdd(big + 3 * big_cap, 1);
        big[i] = b;
        big[big_cap + i] = b + L;
        big[2 * big_cap + i] = r;
      }
      continue;
    }
    if (lane == 0) rnnz[r] = heads;
  }
}

// One workgroup per long row after its segmented sort: its distinct columns.
__global__ __launch_bounds__(kB) void bsr_big_heads_kernel(const int* big, int big_cap,
                                                           const unsigned long long* sorted, int* rnnz) {
  __shared__ int sh[kB];
  const int i = blockIdx.x;
  if (i >= big[3 * big_cap]) return;
  const int b = big[i], e = big[big_cap + i], r = big[2 * big_cap + i];
  int heads = 0;
  for (int q = b + threadIdx.x; q < e; q += kB) heads += q == b || key_col(sorted[q]) != key_col(sorted[q - 1]);
  sh[threadIdx.x] = heads;
  __syncthreads();
  for (int d = kB / 2; d > 0; d >>= 1) {
    if (threadIdx.x < d) sh[threadIdx.x] += sh[threadIdx.x + d];
    __syncthreads();
  }
  if (threadIdx.x == 0) rnnz[r] = sh[0];
}

// One wave per row: the first entry of each column run writes its block's
// column and where its run starts; the last row also closes the run list.
__global__ __launch_bounds__(kB) void bsr_pattern_kernel(const int* start, const int* offsets, int nr,
                                                         const unsigned long long* sorted, int cap, int* columns,
                                                         int* seg,
                                                         int* nnz_out) {
  const int total = offsets[nr];
  if (blockIdx.x == 0 && threadIdx.x == 0) *nnz_out = total <= cap ? total : -1;
  if (total > cap) return;
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const unsigned long long below = (1ull << lane) - 1;
  for (int r = blockIdx.x * kWaves + w; r < nr; r += gridDim.x * kWaves) {
    const int b = start[r], L = start[r + 1] - b;
    int base = offsets[r];
    for (int i0 = 0; i0 < L; i0 += 64) {
      const int i = i0 + lane;
      int c = -1;
      bool head = false;
      if (i < L) {
        c = key_col(sorted[b + i]);
        head = i == 0 || key_col(sorted[b + i - 1]) != c;
      }
      const unsigned long long hb = __ballot(head);
      if (head) {
        const int o = base + __popcll(hb & below);
        columns[o] = c;
        seg[o] = b + i;
      }
      base += __popcll(hb);
    }
    if (r == nr - 1 && lane == 0) seg[total] = start[nr];
  }
}

// Block values: one wave per 64 consecutive blocks. The triplet values of
// their runs (one contiguous range of the sorted entries) are gathered into
// LDS a window at a time by the whole wave, then each lane adds its block's
// entries of the window in order, so every block is the fp32 sum of its
// triplets in ascending triplet order from +0 however the windows fall.
__global__ __launch_bounds__(kB) void bsr_sum_kernel(const int* nnz_p, const int* seg, const unsigned long long* sorted,
                                                     const float* vals, float* values) {
  __shared__ float sv[kWaves][kWin * 9];
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  float* s = sv[w];
  const int n = *nnz_p;
  for (int o0 = (blockIdx.x * kWaves + w) * 64; o0 < n; o0 += gridDim.x * kWaves * 64) {
    const int o = o0 + lane;
    const bool active = o < n;
    const int mb = active ? seg[o] : 0, me = active ? seg[o + 1] : 0;
    const int e0 = seg[o0], e1 = seg[o0 + 64 < n ? o0 + 64 : n];
    float acc[9] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int w0 = e0; w0 < e1; w0 += kWin) {
      const int cnt = e1 - w0 < kWin ? e1 - w0 : kWin;
      for (int k = lane; k < cnt; k += 64) {
        const float* v = vals + 9ll * key_triplet(sorted[w0 + k]);
#pragma unroll
        for (int j = 0; j < 9; ++j) s[k * 9 + j] = v[j];
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
      const int lo = mb > w0 ? mb : w0, hi = me < w0 + cnt ? me : w0 + cnt;
      for (int q = lo; q < hi; ++q) {
#pragma unroll
        for (int j = 0; j < 9; ++j) acc[j] += s[(q - w0) * 9 + j];
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    }
    if (active) {
#pragma unroll
      for (int j = 0; j < 9; ++j) values[9ll * o + j] = acc[j];
    }
  }
}

hipError_t launch_sum(const ipc_bsr* w, int max_blocks, const float* vals, float* values, hipStream_t s) {
  const int need = nblk(max_blocks, kB);
  const int g = need < 1 ? 1 : (need < 4096 ? need : 4096);
  hipLaunchKernelGGL(bsr_sum_kernel, dim3(g), dim3(kB), 0, s, w->nnz.as<int>(), w->seg.as<int>(),
                     w->sorted.as<unsigned long long>(), vals, values);
  return hipGetLastError();
}

// ------------------------------------------------------------ axpy ---
// Row r of z = alpha x + beta y over the union of the two rows' columns
// (both sorted and distinct per row); a block missing from one side counts
// as zero there.
__global__ void bsr_union_count_kernel(int nr, const int* xo, const int* xc, const int* yo, const int* yc, int* cnt) {
  const int r = blockIdx.x * kB + threadIdx.x;
  if (r > nr) return;
  if (r == nr) {
    cnt[r] = 0;
    return;
  }
  int i = xo[r], j = yo[r];
  const int ie = xo[r + 1], je = yo[r + 1];
  int n = 0;
  while (i < ie && j < je) {
    const int a = xc[i], c = yc[j];
    ++n;
    i += a <= c;
    j += c <= a;
  }
  cnt[r] = n + (ie - i) + (je - j);
}

// Eight lanes per row: the group's first lane merges the two column lists,
// up to kUnionChunk union entries at a time, into LDS (source block of each
// side or -1, column); then the group writes those entries' columns and
// values with consecutive lanes on consecutive floats, so that a row's
// output is written as one contiguous stream. (One lane per row writing its
// own blocks 36 bytes at a time left partial cache lines across the wave:
// the fill moved ten times its algorithmic bytes.)
constexpr int kRowGroup = 8;
constexpr int kUnionChunk = 64;
constexpr int kGroups = kB / kRowGroup;

__global__ __launch_bounds__(kB) void bsr_union_fill_kernel(int nr, const int* xo, const int* xc, const float* xv,
                                                            float alpha, const int* yo, const int* yc, const float* yv,
                                                            float beta, const int* zo, int* zc, float* zv, int cap,
                                                            int* nnz_out) {
  __shared__ int sx[kGroups][kUnionChunk], sy[kGroups][kUnionChunk], sc[kGroups][kUnionChunk];
  const int total = zo[nr];
  if (blockIdx.x == 0 && threadIdx.x == 0 && nnz_out) *nnz_out = total;
  if (total > cap) return;
  const int g = threadIdx.x / kRowGroup, l = threadIdx.x % kRowGroup;
  const int r = blockIdx.x * kGroups + g;
  if (r >= nr) return;  // the whole group
  const int leader = (threadIdx.x & 63) & ~(kRowGroup - 1);
  int i = xo[r], j = yo[r], o = zo[r];
  const int ie = xo[r + 1], je = yo[r + 1];
  for (;;) {
    int m = 0;
    if (l == 0) {
      while (m < kUnionChunk && (i < ie || j < je)) {
        const int a = i < ie ? xc[i] : 0x7fffffff, c = j < je ? yc[j] : 0x7fffffff;
        const bool tx = a <= c, ty = c <= a;
        sx[g][m] = tx ? i : -1;
        sy[g][m] = ty ? j : -1;
        sc[g][m] = tx ? a : c;
        ++m;
        i += tx;
        j += ty;
      }
    }
    m = __shfl(m, leader, 64);
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    for (int k = l; k < m; k += kRowGroup) zc[o + k] = sc[g][k];
    for (int f = l; f < 9 * m; f += kRowGroup) {
      const int k = f / 9, e = f - 9 * k;
      const int ix = sx[g][k], iy = sy[g][k];
      const float xs = ix >= 0 ? xv[9ll * ix + e] : 0.0f;
      const float ys = iy >= 0 ? yv[9ll * iy + e] : 0.0f;
      zv[9ll * o + f] = alpha * xs + beta * ys;
    }
    o += m;
    const int more = __shfl(int(i < ie || j < je), leader, 64);
    if (!more) break;
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");  // the entries are read before the next chunk
  }
}

}  // namespace

// ============================================================== C ABI ===
extern "C" {

int ipc_bsr_create(ipc_bsr** out) {
  if (!out) return bsr_fail(IPC_EINVAL, "bsr_create: null output");
  ipc_bsr* w = new (std::nothrow) ipc_bsr();
  if (!w) return bsr_fail(IPC_EINVAL, "bsr_create: out of host memory");
  hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&w->host), sizeof(int));
  if (e != hipSuccess) {
    delete w;
    return bsr_hip(e, "bsr_create");
  }
  *out = w;
  return IPC_OK;
}

int ipc_bsr_destroy(ipc_bsr* w) {
  if (!w) return IPC_OK;
  for (DevBuf* d : {&w->cnt, &w->start, &w->slot, &w->bucket, &w->sorted, &w->seg, &w->part, &w->nnz, &w->big, &w->temp})
    d->release();
  if (w->host) hipHostFree(w->host);
  delete w;
  return IPC_OK;
}

int ipc_bsr_from_triplets(ipc_bsr* w, int n_rows, int n_cols, const int* rows, const int* cols, const float* vals,
                          int n_triplets, int* offsets, int* columns, float* values, int capacity, void* stream) {
  if (!w) return bsr_fail(IPC_EINVAL, "bsr_from_triplets: null workspace");
  if (n_rows < 0 || n_cols < 0 || n_triplets < 0 || capacity < 0)
    return bsr_fail(IPC_EINVAL, "bsr_from_triplets: negative size");
  if (!offsets) return bsr_fail(IPC_EINVAL, "bsr_from_triplets: null offsets");
  if (n_triplets > 0 && (!rows || !cols || !vals)) return bsr_fail(IPC_EINVAL, "bsr_from_triplets: null triplets");
  if (capacity > 0 && (!columns || !values)) return bsr_fail(IPC_EINVAL, "bsr_from_triplets: null columns/values");
  hipStream_t s = static_cast<hipStream_t>(stream);
  w->planned = false;
  const int m = n_triplets, nr = n_rows;
  hipError_t e = hipSuccess;
  for (DevBuf* d : {&w->cnt, &w->start}) {
    if (e == hipSuccess) e = d->ensure(sizeof(int) * (size_t(nr) + 1));
  }
  if (e == hipSuccess) e = w->slot.ensure(sizeof(int) * size_t(m > 0 ? m : 1));
  if (e == hipSuccess) e = w->bucket.ensure(sizeof(unsigned long long) * size_t(m > 0 ? m : 1));
  if (e == hipSuccess) e = w->sorted.ensure(sizeof(unsigned long long) * size_t(m > 0 ? m : 1));
  // rows longer than kSortCap go through a segmented radix sort; there are
  // at most m / (kSortCap + 1) of them
  const int big_cap = m / (kSortCap + 1);
  size_t temp_bytes = 0;
  if (e == hipSuccess) e = w->big.ensure(sizeof(int) * (3 * size_t(big_cap) + 1));
  if (e == hipSuccess && big_cap > 0)
    e = hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, temp_bytes, (const unsigned long long*)nullptr,
                                                   (unsigned long long*)nullptr, m, big_cap, (const int*)nullptr,
                                                   (const int*)nullptr, 0, 64, s);
  if (e == hipSuccess && big_cap > 0) e = w->temp.ensure(temp_bytes);
  if (e == hipSuccess) e = w->seg.ensure(sizeof(int) * (size_t(capacity) + 1));
  if (e == hipSuccess) e = w->nnz.ensure(sizeof(int));
  if (e != hipSuccess) return bsr_hip(e, "bsr_from_triplets: workspace");
  int* cnt = w->cnt.as<int>();
  int* start = w->start.as<int>();
  e = fill_i32(cnt, nr + 1, 0, s);
  if (e == hipSuccess && m > 0)
    hipLaunchKernelGGL(bsr_count_kernel, dim3(nblk(m, kB)), dim3(kB), 0, s, rows, cols, vals, m, nr, n_cols, cnt,
                       w->slot.as<int>());
  if (e == hipSuccess) e = hipGetLastError();
  if (e == hipSuccess) e = exclusive_scan(cnt, nr + 1, start, w->part, s);
  if (e == hipSuccess && m > 0)
    hipLaunchKernelGGL(bsr_scatter_kernel, dim3(nblk(m, kB)), dim3(kB), 0, s, rows, cols, w->slot.as<int>(), start, m,
                       w->bucket.as<unsigned long long>());
  const int rg = nr > 0 ? (nblk(nr, kWaves) < 65536 ? nblk(nr, kWaves) : 65536) : 0;
  if (e == hipSuccess) e = hipGetLastError();
  if (e == hipSuccess && nr > 0) {
    // cnt[nr] stays 0 from the fill: the row's distinct counts go to cnt[0..nr)
    int* big = w->big.as<int>();
    e = fill_i32(big, 3 * big_cap + 1, 0, s);
    if (e == hipSuccess)
      hipLaunchKernelGGL(bsr_sort_rows_kernel, dim3(rg), dim3(kB), 0, s, start, nr,
                         w->bucket.as<unsigned long long>(), w->sorted.as<unsigned long long>(), cnt, big, big_cap);
    if (e == hipSuccess) e = hipGetLastError();
    if (e == hipSuccess && big_cap > 0) {
      // unused slots are empty segments (begin = end = 0)
      e = hipcub::DeviceSegmentedRadixSort::SortKeys(w->temp.p, temp_bytes, w->bucket.as<unsigned long long>(),
                                                     w->sorted.as<unsigned long long>(), m, big_cap, big,
                                                     big + big_cap, 0, 64, s);
      if (e == hipSuccess)
        hipLaunchKernelGGL(bsr_big_heads_kernel, dim3(big_cap), dim3(kB), 0, s, big, big_cap,
                           w->sorted.as<unsigned long long>(), cnt);
      if (e == hipSuccess) e = hipGetLastError();
    }
  }
  if (e == hipSuccess) e = exclusive_scan(cnt, nr + 1, offsets, w->part, s);
  if (e == hipSuccess) {
    if (nr > 0) {
      hipLaunchKernelGGL(bsr_pattern_kernel, dim3(rg), dim3(kB), 0, s, start, offsets, nr,
                         w->sorted.as<unsigned long long>(),
                         capacity, columns, w->seg.as<int>(), w->nnz.as<int>());
      e = hipGetLastError();
      if (e == hipSuccess && m > 0) e = launch_sum(w, capacity < m ? capacity : m, vals, values, s);
    } else {
      e = fill_i32(w->nnz.as<int>(), 1, 0, s);
      if (e == hipSuccess) e = fill_i32(w->seg.as<int>(), 1, 0, s);
    }
  }
  if (e == hipSuccess) e = hipGetLastError();
  if (e != hipSuccess) return bsr_hip(e, "bsr_from_triplets");
  w->n_rows = nr;
  w->n_triplets = m;
  w->max_blocks = capacity < m ? capacity : m;
  if (capacity >= m) {
    w->planned = true;
    return IPC_OK;
  }
  // the pattern may not fit: read the block count back
  e = hipMemcpyAsync(w->host, offsets + nr, sizeof(int), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return bsr_hip(e, "bsr_from_triplets: block count readback");
  if (*w->host > capacity) return *w->host;
  w->planned = true;
  return IPC_OK;
}

int ipc_bsr_refill(ipc_bsr* w, const float* vals, int n_triplets, float* values, void* stream) {
  if (!w) return bsr_fail(IPC_EINVAL, "bsr_refill: null workspace");
  if (!w->planned) return bsr_fail(IPC_EINVAL, "bsr_refill: no pattern (ipc_bsr_from_triplets has not succeeded)");
  if (n_triplets != w->n_triplets) return bsr_fail(IPC_EINVAL, "bsr_refill: triplet count differs from the pattern's");
  if (n_triplets == 0) return IPC_OK;
  if (!vals || !values) return bsr_fail(IPC_EINVAL, "bsr_refill: null pointer");
  hipError_t e = launch_sum(w, w->max_blocks, vals, values, static_cast<hipStream_t>(stream));
  return bsr_hip(e, "bsr_refill");
}

int ipc_bsr_axpy(ipc_bsr* w, int n_rows, const int* x_offsets, const int* x_columns, const float* x_values, int x_nnz,
                 float alpha, const int* y_offsets, const int* y_columns, const float* y_values, int y_nnz, float beta,
                 int* z_offsets, int* z_columns, float* z_values, int capacity, void* stream) {
  if (!w) return bsr_fail(IPC_EINVAL, "bsr_axpy: null workspace");
  if (n_rows < 0 || x_nnz < 0 || y_nnz < 0 || capacity < 0) return bsr_fail(IPC_EINVAL, "bsr_axpy: negative size");
  if (!x_offsets || !y_offsets || !z_offsets) return bsr_fail(IPC_EINVAL, "bsr_axpy: null offsets");
  if ((x_nnz > 0 && (!x_columns || !x_values)) || (y_nnz > 0 && (!y_columns || !y_values)) ||
      (capacity > 0 && (!z_columns || !z_values)))
    return bsr_fail(IPC_EINVAL, "bsr_axpy: null columns/values");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipError_t e = w->cnt.ensure(sizeof(int) * (size_t(n_rows) + 1));
  if (e != hipSuccess) return bsr_hip(e, "bsr_axpy: workspace");
  // the row counts share the triplet path's count buffer; a kept pattern
  // (sorted entries, runs, block count) stays valid
  int* cnt = w->cnt.as<int>();
  hipLaunchKernelGGL(bsr_union_count_kernel, dim3(nblk(n_rows + 1, kB)), dim3(kB), 0, s, n_rows, x_offsets, x_columns,
                     y_offsets, y_columns, cnt);
  e = hipGetLastError();
  if (e == hipSuccess) e = exclusive_scan(cnt, n_rows + 1, z_offsets, w->part, s);
  if (e == hipSuccess && n_rows > 0)
    hipLaunchKernelGGL(bsr_union_fill_kernel, dim3(nblk(n_rows, kGroups)), dim3(kB), 0, s, n_rows, x_offsets, x_columns,
                       x_values, alpha, y_offsets, y_columns, y_values, beta, z_offsets, z_columns, z_values, capacity,
                       nullptr);
  if (e == hipSuccess) e = hipGetLastError();
  if (e != hipSuccess) return bsr_hip(e, "bsr_axpy");
  if (static_cast<long long>(x_nnz) + y_nnz <= capacity) return IPC_OK;
  e = hipMemcpyAsync(w->host, z_offsets + n_rows, sizeof(int), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return bsr_hip(e, "bsr_axpy: block count readback");
  return *w->host > capacity ? *w->host : IPC_OK;
}

}  // extern "C"

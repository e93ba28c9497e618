// Candidate culling of libipc_warp_amd: linear BVH build and AABB queries for
// DCD (edge_edge_collision / point_triangle_collision, contact.py:199-336)
// and for the CCD broad phase (edge_edge_toi / point_triangle_toi,
// ccd/ccd.py:166-210,360-427).
//
// The BVH is rebuilt on every call from the current positions (the reference
// refits a BVH built once; a rebuild keeps query cost independent of how far
// the mesh has moved and costs a handful of small kernels):
//   1. one AABB per primitive (fp32, the reference's inflation),
//   2. 30-bit Morton codes of box centres in the scene bounds, made unique by
//      the primitive id in the low 32 bits of a 64-bit key,
//   3. hipCUB radix sort of the keys,
//   4. Karras (2012) hierarchy: internal node i spans a key range and splits
//      at the highest differing bit,
//   5. bottom-up boxes: each internal node stores the boxes of its two
//      children (64 B per node), so a traversal step tests both children
//      with one 64 B load.
// Queries keep a 64-entry stack in LDS (a Karras tree over 64-bit unique keys
// is at most 64 levels deep), run a wave-uniform loop (one node per lane per
// iteration) and append accepted pairs with one atomicAdd per wave per
// iteration (ballot + prefix count).
#include <hipcub/hipcub.hpp>

#include "cull.hpp"
#include "ipc_device.hpp"

namespace ipcd {

// ------------------------------------------------------------ utilities ----
__device__ __forceinline__ int ord_of(float f) {
  int i = __float_as_int(f);
  return i >= 0 ? i : i ^ 0x7fffffff;
}
__device__ __forceinline__ float float_of(int k) { return __int_as_float(k >= 0 ? k : k ^ 0x7fffffff); }

__global__ void init_bounds_kernel(int* bounds) {
  int t = threadIdx.x;
  if (t < 3) bounds[t] = 0x7fffffff;           // min
  else if (t < 6) bounds[t] = (int)0x80000000;  // max
}

__device__ __forceinline__ unsigned expand_bits10(unsigned v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

struct Box { float lo[3], hi[3]; };

__device__ __forceinline__ bool overlaps(const float qlo[3], const float qhi[3], float4 lo, float4 hi) {
  // Inclusive test (a lower bound above the other's upper bound separates).
  return !(qlo[0] > hi.x || qlo[1] > hi.y || qlo[2] > hi.z || lo.x > qhi[0] || lo.y > qhi[1] || lo.z > qhi[2]);
}

// --------------------------------------------------------- primitive boxes ----
// mode 0: edge box +- pad (edge_aabb, contact.py:30-37)
// mode 1: triangle bounds +- pad (wp.Mesh bounds when pad = 0)
// mode 2: swept edge box over x0, x1, +- pad (swept_edge_aabbs, ccd/ccd.py:166-186)
// mode 3: swept triangle box over x0, x1, +- pad (swept_triangle_aabbs, ccd/ccd.py:188-210)
__global__ void __launch_bounds__(256) prim_boxes_kernel(int mode, int n, const int* __restrict__ prims,
                                                         const float* __restrict__ x0, const float* __restrict__ x1,
                                                         float pad, float4* __restrict__ lo4, float4* __restrict__ hi4,
                                                         int* __restrict__ bounds) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  float c[3] = {0.f, 0.f, 0.f};
  bool live = i < n;
  if (live) {
    int k = (mode == 0 || mode == 2) ? 2 : 3;
    V3<float> lo, hi;
    int v0 = prims[k * i];
    lo = hi = ldx(x0, v0);
    for (int j = 1; j < k; ++j) {
      V3<float> p = ldx(x0, prims[k * i + j]);
      lo = vmin(lo, p);
      hi = vmax(hi, p);
    }
    if (mode >= 2) {
      for (int j = 0; j < k; ++j) {
        V3<float> p = ldx(x1, prims[k * i + j]);
        lo = vmin(lo, p);
        hi = vmax(hi, p);
      }
    }
    V3<float> pv = mk(pad, pad, pad);
    lo = lo - pv;
    hi = hi + pv;
    lo4[i] = make_float4(lo.x, lo.y, lo.z, 0.f);
    hi4[i] = make_float4(hi.x, hi.y, hi.z, 0.f);
    c[0] = 0.5f * (lo.x + hi.x);
    c[1] = 0.5f * (lo.y + hi.y);
    c[2] = 0.5f * (lo.z + hi.z);
  }
  // block-wide min / max of centres, then one atomic per block and axis
  __shared__ int smin_[3][4], smax_[3][4];
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int a = 0; a < 3; ++a) {
    int mn = live ? ord_of(c[a]) : 0x7fffffff;
    int mx = live ? ord_of(c[a]) : (int)0x80000000;
    for (int off = 32; off > 0; off >>= 1) {
      mn = min(mn, __shfl_xor(mn, off, 64));
      mx = max(mx, __shfl_xor(mx, off, 64));
    }
    if (lane == 0) { smin_[a][wave] = mn; smax_[a][wave] = mx; }
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    int a = threadIdx.x;
    int mn = smin_[a][0], mx = smax_[a][0];
    for (int w = 1; w < int(blockDim.x >> 6); ++w) { mn = min(mn, smin_[a][w]); mx = max(mx, smax_[a][w]); }
    atomicMin(bounds + a, mn);
    atomicMax(bounds + 3 + a, mx);
  }
}

__global__ void __launch_bounds__(256) morton_kernel(int n, const float4* __restrict__ lo4,
                                                     const float4* __restrict__ hi4, const int* __restrict__ bounds,
                                                     unsigned long long* __restrict__ keys) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 lo = lo4[i], hi = hi4[i];
  float c[3] = {0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z)};
  unsigned q[3];
  for (int a = 0; a < 3; ++a) {
    float b0 = float_of(bounds[a]), b1 = float_of(bounds[3 + a]);
    float ext = b1 - b0;
    float u = ext > 0.f ? (c[a] - b0) / ext : 0.5f;
    float s = u * 1024.f;
    s = s < 0.f ? 0.f : (s > 1023.f ? 1023.f : s);
    q[a] = unsigned(s);
  }
  unsigned m = (expand_bits10(q[0]) << 2) | (expand_bits10(q[1]) << 1) | expand_bits10(q[2]);
  keys[i] = (static_cast<unsigned long long>(m) << 32) | static_cast<unsigned>(i);
}

// ------------------------------------------------------------ hierarchy ----
__device__ __forceinline__ int delta(const unsigned long long* keys, int n, int i, int j) {
  if (j < 0 || j >= n) return -1;
  return __clzll(keys[i] ^ keys[j]);
}

// Internal node i of a Karras tree over n sorted unique keys. Children are
// internal node ids (>= 0) or ~primitive id (< 0) for leaves.
__global__ void __launch_bounds__(256) karras_kernel(int n, const unsigned long long* __restrict__ keys,
                                                     int2* __restrict__ child, int* __restrict__ parent,
                                                     int* __restrict__ leaf_parent) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n - 1) return;
  int d = (delta(keys, n, i, i + 1) - delta(keys, n, i, i - 1)) >= 0 ? 1 : -1;
  int dmin = delta(keys, n, i, i - d);
  int lmax = 2;
  while (delta(keys, n, i, i + lmax * d) > dmin) lmax <<= 1;
  int l = 0;
  for (int t = lmax >> 1; t >= 1; t >>= 1)
    if (delta(keys, n, i, i + (l + t) * d) > dmin) l += t;
  int j = i + l * d;
  int dnode = delta(keys, n, i, j);
  int s = 0, t = l;
  do {
    t = (t + 1) >> 1;
    if (delta(keys, n, i, i + (s + t) * d) > dnode) s += t;
  } while (t > 1);
  int gamma = i + s * d + (d < 0 ? -1 : 0);
  int lo = i < j ? i : j, hi = i < j ? j : i;
  int left, right;
  if (lo == gamma) {
    left = ~static_cast<int>(keys[gamma] & 0xffffffffu);
    leaf_parent[gamma] = i;
  } else {
    left = gamma;
    parent[gamma] = i;
  }
  if (hi == gamma + 1) {
    right = ~static_cast<int>(keys[gamma + 1] & 0xffffffffu);
    leaf_parent[gamma + 1] = i;
  } else {
    right = gamma + 1;
    parent[gamma + 1] = i;
  }
  child[i] = make_int2(left, right);
  if (i == 0) parent[0] = -1;
}

__global__ void __launch_bounds__(256) zero_u32_kernel(unsigned* p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// Bottom-up boxes. Node q's 64 B record is (left lo, left hi, right lo,
// right hi). The second child to arrive at a node (atomic counter) merges
// the two slots and continues to the parent.
__global__ void __launch_bounds__(256) refit_kernel(int n, const unsigned long long* __restrict__ keys,
                                                    const float4* __restrict__ lo4, const float4* __restrict__ hi4,
                                                    const int2* __restrict__ child, const int* __restrict__ parent,
                                                    const int* __restrict__ leaf_parent, float4* box,
                                                    unsigned* flags) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  int prim = static_cast<int>(keys[k] & 0xffffffffu);
  float4 lo = lo4[prim], hi = hi4[prim];
  int self = ~prim;
  int q = leaf_parent[k];
  while (q >= 0) {
    int slot = (child[q].x == self) ? 0 : 1;
    box[4 * q + 2 * slot] = lo;
    box[4 * q + 2 * slot + 1] = hi;
    __threadfence();
    unsigned old = atomicAdd(flags + q, 1u);
    if (old == 0u) return;
    __threadfence();
    const volatile float4* vb = box + 4 * q;
    float4 l0 = make_float4(vb[0].x, vb[0].y, vb[0].z, 0.f), h0 = make_float4(vb[1].x, vb[1].y, vb[1].z, 0.f);
    float4 l1 = make_float4(vb[2].x, vb[2].y, vb[2].z, 0.f), h1 = make_float4(vb[3].x, vb[3].y, vb[3].z, 0.f);
    lo = make_float4(fminf(l0.x, l1.x), fminf(l0.y, l1.y), fminf(l0.z, l1.z), 0.f);
    hi = make_float4(fmaxf(h0.x, h1.x), fmaxf(h0.y, h1.y), fmaxf(h0.z, h1.z), 0.f);
    self = q;
    q = parent[q];
  }
}

// ------------------------------------------------------------------ queries ----
// Wave-aggregated append: every lane with `want` gets a distinct slot.
__device__ __forceinline__ int wave_append(bool want, int* count) {
  unsigned long long m = __ballot(want);
  if (m == 0ull) return -1;
  int lane = __lane_id();
  int leader = __ffsll(static_cast<long long>(m)) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(count, __popcll(m));
  base = __shfl(base, leader, 64);
  if (!want) return -1;
  return base + __popcll(m & ((1ull << lane) - 1ull));
}

constexpr int kQueryBlock = 128;
constexpr int kStackDepth = 64;

// Query mode:
//  0 DCD EE: query edge i's box against the edge BVH; accept i < j, not
//    connected, not same body (if excluded), fp32 segment distance < 2 th.
//    Record (a1, a2, b1, b2), (i, j)            contact.py:199-229
//  1 DCD PT: query x_i +- 2(th + buf) against triangle bounds; accept not
//    incident, not same body, fp32 point-triangle distance < 2 th.
//    Record (i, t[1], t[0], t[2]), (i, j)       contact.py:304-336
//  2 CCD EE: swept edge box against swept edge BVH; i < j, not connected,
//    not same body. Record (a0, a1, b0, b1)     ccd/ccd.py:397-427
//  3 CCD PT: swept point box +- pad against swept triangle BVH; not
//    incident, not same body. Record (i, t0, t1, t2)  ccd/ccd.py:360-394
struct QueryArgs {
  int mode;
  int n_query;
  const int* edges;
  const int* tris;
  const int* body;
  const float* x0;
  const float* x1;
  const float4* qlo;  // query boxes (modes 0, 2: the leaves' own boxes)
  const float4* qhi;
  float pad;          // mode 1: 2 (th + buf); mode 3: CCD padding
  float thickness;
  int exclude_same_body;
  // BVH
  int n_leaves;
  const int2* child;
  const float4* box;
  const float4* leaf_lo;  // single-leaf trees
  const float4* leaf_hi;
  // output
  int4* rec;
  int2* ids;
  int capacity;
  int* count;
};

__device__ __forceinline__ bool accept_pair(const QueryArgs& A, int i, int j, int4& rec) {
  if (A.mode == 0 || A.mode == 2) {
    if (!(i < j)) return false;
    int a0 = A.edges[2 * i], a1 = A.edges[2 * i + 1], b0 = A.edges[2 * j], b1 = A.edges[2 * j + 1];
    if (a0 == b0 || a0 == b1 || a1 == b0 || a1 == b1) return false;
    if (A.exclude_same_body && A.body[a0] == A.body[b0]) return false;
    rec = make_int4(a0, a1, b0, b1);
    if (A.mode == 2) return true;
    EdgeClosest ec = closest_point_edge_edge(ldx(A.x0, a0), ldx(A.x0, a1), ldx(A.x0, b0), ldx(A.x0, b1), 1.0e-6f);
    return ec.dist < A.thickness * 2.0f;
  }
  int t0 = A.tris[3 * j], t1 = A.tris[3 * j + 1], t2 = A.tris[3 * j + 2];
  if (t0 == i || t1 == i || t2 == i) return false;
  if (A.exclude_same_body && A.body[i] == A.body[t0]) return false;
  if (A.mode == 3) {
    rec = make_int4(i, t0, t1, t2);
    return true;
  }
  rec = make_int4(i, t1, t0, t2);
  TriClosest tc = closest_point_triangle(ldx(A.x0, i), ldx(A.x0, t0), ldx(A.x0, t1), ldx(A.x0, t2));
  return tc.dist < A.thickness * 2.0f;
}

__global__ void __launch_bounds__(kQueryBlock) query_kernel(QueryArgs A) {
  __shared__ int stack[kStackDepth * kQueryBlock];
  int tid = threadIdx.x;
  int i = blockIdx.x * blockDim.x + tid;
  bool live = i < A.n_query;
  float qlo[3] = {0.f, 0.f, 0.f}, qhi[3] = {-1.f, -1.f, -1.f};
  if (live) {
    if (A.mode == 0 || A.mode == 2) {
      float4 l = A.qlo[i], h = A.qhi[i];
      qlo[0] = l.x; qlo[1] = l.y; qlo[2] = l.z;
      qhi[0] = h.x; qhi[1] = h.y; qhi[2] = h.z;
    } else {
      V3<float> p = ldx(A.x0, i), lo = p, hi = p;
      if (A.mode == 3) {
        V3<float> q = ldx(A.x1, i);
        lo = vmin(p, q);
        hi = vmax(p, q);
      }
      V3<float> pv = mk(A.pad, A.pad, A.pad);
      lo = lo - pv;
      hi = hi + pv;
      qlo[0] = lo.x; qlo[1] = lo.y; qlo[2] = lo.z;
      qhi[0] = hi.x; qhi[1] = hi.y; qhi[2] = hi.z;
    }
  }
  int sp = 0;
  // hits of this iteration (primitive ids, -1 = none)
  int h0 = -1, h1 = -1;
  if (live) {
    if (A.n_leaves == 1) {
      if (overlaps(qlo, qhi, A.leaf_lo[0], A.leaf_hi[0])) h0 = 0;
    } else if (A.n_leaves > 1) {
      stack[tid] = 0;
      sp = 1;
    }
  }
  while (true) {
    if (sp > 0) {
      int node = stack[(--sp) * kQueryBlock + tid];
      int2 ch = A.child[node];
      const float4* bx = A.box + 4 * node;
      float4 l0 = bx[0], u0 = bx[1], l1 = bx[2], u1 = bx[3];
      // A Karras tree over 62-bit unique keys is at most 61 internal levels
      // deep, so sp never reaches kStackDepth; the guard only keeps LDS
      // writes in bounds.
      if (overlaps(qlo, qhi, l0, u0)) {
        if (ch.x < 0) h0 = ~ch.x;
        else if (sp < kStackDepth) stack[(sp++) * kQueryBlock + tid] = ch.x;
      }
      if (overlaps(qlo, qhi, l1, u1)) {
        if (ch.y < 0) h1 = ~ch.y;
        else if (sp < kStackDepth) stack[(sp++) * kQueryBlock + tid] = ch.y;
      }
    }
    int4 r0, r1;
    bool w0 = h0 >= 0 && accept_pair(A, i, h0, r0);
    bool w1 = h1 >= 0 && accept_pair(A, i, h1, r1);
    int s0 = wave_append(w0, A.count);
    if (s0 >= 0 && s0 < A.capacity) {
      A.rec[s0] = r0;
      if (A.ids) A.ids[s0] = make_int2(i, h0);
    }
    int s1 = wave_append(w1, A.count);
    if (s1 >= 0 && s1 < A.capacity) {
      A.rec[s1] = r1;
      if (A.ids) A.ids[s1] = make_int2(i, h1);
    }
    h0 = h1 = -1;
    if (!__any(sp > 0)) break;
  }
}

// x1 = x - dx (_add_dx, contact.py:49-52)
__global__ void __launch_bounds__(256) sub_kernel(int n3, const float* __restrict__ x, const float* __restrict__ dx,
                                                  float* __restrict__ out) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n3) out[i] = x[i] - dx[i];
}

__global__ void fill_u64_kernel(unsigned long long* p, unsigned long long v) { p[threadIdx.x] = v; }

// out[0] = min, out[1] = max over a[0..n) (out initialised to INT_MAX, INT_MIN).
__global__ void __launch_bounds__(256) index_range_kernel(const int* __restrict__ a, long long n, int* out) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int mn = 0x7fffffff, mx = (int)0x80000000;
  for (; i < n; i += (long long)gridDim.x * blockDim.x) {
    int v = a[i];
    mn = min(mn, v);
    mx = max(mx, v);
  }
  for (int off = 32; off > 0; off >>= 1) {
    mn = min(mn, __shfl_xor(mn, off, 64));
    mx = max(mx, __shfl_xor(mx, off, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMin(out, mn);
    atomicMax(out + 1, mx);
  }
}

__global__ void init_range_kernel(int* out) {
  out[0] = 0x7fffffff;
  out[1] = (int)0x80000000;
}
__global__ void zero_i32_kernel(int* p, int n) {
  int i = threadIdx.x;
  if (i < n) p[i] = 0;
}

// ------------------------------------------------------------- host side ----
static inline int nblocks(long long n, int bs) { return int((n + bs - 1) / bs); }

hipError_t DevBuf::ensure(size_t b) {
  if (b <= bytes) return hipSuccess;
  if (p) {
    hipError_t e = hipFree(p);
    if (e != hipSuccess) return e;
    p = nullptr;
    bytes = 0;
  }
  size_t want = b < 256 ? 256 : b;
  hipError_t e = hipMalloc(&p, want);
  if (e == hipSuccess) bytes = want;
  return e;
}

void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

void Bvh::release() {
  lo.release(); hi.release(); keys.release(); keys_sorted.release(); child.release(); parent.release();
  leaf_parent.release(); box.release(); flags.release(); temp.release(); bounds.release();
  n = 0;
}

#define IPC_TRY(expr)                    \
  do {                                   \
    hipError_t _e = (expr);              \
    if (_e != hipSuccess) return _e;     \
  } while (0)

hipError_t build_bvh(Bvh& B, int mode, int n, const int* prims, const float* x0, const float* x1, float pad,
                     hipStream_t s) {
  B.n = n;
  if (n <= 0) return hipSuccess;
  IPC_TRY(B.lo.ensure(sizeof(float4) * n));
  IPC_TRY(B.hi.ensure(sizeof(float4) * n));
  IPC_TRY(B.bounds.ensure(sizeof(int) * 8));
  hipLaunchKernelGGL(init_bounds_kernel, dim3(1), dim3(64), 0, s, B.bounds.as<int>());
  hipLaunchKernelGGL(prim_boxes_kernel, dim3(nblocks(n, 256)), dim3(256), 0, s, mode, n, prims, x0, x1, pad,
                     B.lo.as<float4>(), B.hi.as<float4>(), B.bounds.as<int>());
  IPC_TRY(hipGetLastError());
  if (n == 1) return hipSuccess;
  IPC_TRY(B.keys.ensure(sizeof(unsigned long long) * n));
  IPC_TRY(B.keys_sorted.ensure(sizeof(unsigned long long) * n));
  hipLaunchKernelGGL(morton_kernel, dim3(nblocks(n, 256)), dim3(256), 0, s, n, B.lo.as<float4>(), B.hi.as<float4>(),
                     B.bounds.as<int>(), B.keys.as<unsigned long long>());
  IPC_TRY(hipGetLastError());
  size_t temp_bytes = 0;
  IPC_TRY(hipcub::DeviceRadixSort::SortKeys(nullptr, temp_bytes, B.keys.as<unsigned long long>(),
                                            B.keys_sorted.as<unsigned long long>(), n, 0, 62, s));
  IPC_TRY(B.temp.ensure(temp_bytes));
  temp_bytes = B.temp.bytes;
  IPC_TRY(hipcub::DeviceRadixSort::SortKeys(B.temp.p, temp_bytes, B.keys.as<unsigned long long>(),
                                            B.keys_sorted.as<unsigned long long>(), n, 0, 62, s));
  IPC_TRY(B.child.ensure(sizeof(int2) * (n - 1)));
  IPC_TRY(B.parent.ensure(sizeof(int) * (n - 1)));
  IPC_TRY(B.leaf_parent.ensure(sizeof(int) * n));
  IPC_TRY(B.box.ensure(sizeof(float4) * 4 * (n - 1)));
  IPC_TRY(B.flags.ensure(sizeof(unsigned) * (n - 1)));
  hipLaunchKernelGGL(karras_kernel, dim3(nblocks(n - 1, 256)), dim3(256), 0, s, n,
                     B.keys_sorted.as<unsigned long long>(), B.child.as<int2>(), B.parent.as<int>(),
                     B.leaf_parent.as<int>());
  hipLaunchKernelGGL(zero_u32_kernel, dim3(nblocks(n - 1, 256)), dim3(256), 0, s, B.flags.as<unsigned>(), n - 1);
  hipLaunchKernelGGL(refit_kernel, dim3(nblocks(n, 256)), dim3(256), 0, s, n, B.keys_sorted.as<unsigned long long>(),
                     B.lo.as<float4>(), B.hi.as<float4>(), B.child.as<int2>(), B.parent.as<int>(),
                     B.leaf_parent.as<int>(), B.box.as<float4>(), B.flags.as<unsigned>());
  return hipGetLastError();
}

hipError_t run_query(const QueryLaunch& q, hipStream_t s) {
  if (q.n_query <= 0) return hipSuccess;
  QueryArgs A;
  A.mode = q.mode;
  A.n_query = q.n_query;
  A.edges = q.edges;
  A.tris = q.tris;
  A.body = q.body;
  A.x0 = q.x0;
  A.x1 = q.x1;
  A.qlo = q.qlo;
  A.qhi = q.qhi;
  A.pad = q.pad;
  A.thickness = q.thickness;
  A.exclude_same_body = q.exclude_same_body;
  A.n_leaves = q.bvh->n;
  A.child = q.bvh->child.as<int2>();
  A.box = q.bvh->box.as<float4>();
  A.leaf_lo = q.bvh->lo.as<float4>();
  A.leaf_hi = q.bvh->hi.as<float4>();
  A.rec = q.rec;
  A.ids = q.ids;
  A.capacity = q.capacity;
  A.count = q.count;
  if (A.n_leaves <= 0) return hipSuccess;
  hipLaunchKernelGGL(query_kernel, dim3(nblocks(q.n_query, kQueryBlock)), dim3(kQueryBlock), 0, s, A);
  return hipGetLastError();
}

hipError_t launch_sub(int n3, const float* x, const float* dx, float* out, hipStream_t s) {
  if (n3 <= 0) return hipSuccess;
  hipLaunchKernelGGL(sub_kernel, dim3(nblocks(n3, 256)), dim3(256), 0, s, n3, x, dx, out);
  return hipGetLastError();
}

hipError_t launch_fill_u64(unsigned long long* p, int n, unsigned long long v, hipStream_t s) {
  hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(n), 0, s, p, v);
  return hipGetLastError();
}

hipError_t index_range(const int* a, long long n, int* out_dev, hipStream_t s) {
  hipLaunchKernelGGL(init_range_kernel, dim3(1), dim3(1), 0, s, out_dev);
  if (n > 0) {
    long long blocks = (n + 255) / 256;
    if (blocks > 1024) blocks = 1024;
    hipLaunchKernelGGL(index_range_kernel, dim3(unsigned(blocks)), dim3(256), 0, s, a, n, out_dev);
  }
  return hipGetLastError();
}

hipError_t launch_zero_i32(int* p, int n, hipStream_t s) {
  hipLaunchKernelGGL(zero_i32_kernel, dim3(1), dim3(64), 0, s, p, n);
  return hipGetLastError();
}

}  // namespace ipcd

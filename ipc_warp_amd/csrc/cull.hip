// Candidate culling of libipc_warp_amd: BVH build and AABB queries for DCD
// (edge_edge_collision / point_triangle_collision, contact.py:199-336) and
// for the CCD broad phase (edge_edge_toi / point_triangle_toi,
// ccd/ccd.py:166-210,360-427).
//
// The BVH is rebuilt on every call from the current positions (the reference
// refits a BVH built once; a rebuild keeps query cost independent of how far
// the mesh has moved):
//   1. one AABB per primitive (fp32, the reference's inflation),
//   2. 30-bit Morton codes of box centres in the scene bounds, made unique by
//      the primitive id in the low 32 bits of a 64-bit key, sorted with the
//      hipCUB radix sort on the first build and every kResortEvery builds
//      (in between, the leaf order is kept and only the boxes are refreshed),
//   3. an implicit 8-ary hierarchy over the sorted leaves: node k of level L
//      bounds nodes 8k .. 8k+7 of level L-1 (level 0 = leaves). Levels 1-3
//      are reduced by one kernel in LDS (a 64-lane workgroup owns 512 leaves),
//      the levels above by a single workgroup. There is no bottom-up
//      propagation across workgroups, hence no device-scope fence: on a
//      multi-XCD part such a fence writes back and invalidates the L2.
// Nodes also carry the (min, max) body id below them, so a query that
// excludes its own body skips single-body subtrees. Queries run eight lanes
// per query: visiting a node, lane r tests the node's child r, and the
// group's ballot of overlapping children becomes one stack entry (level,
// first child, 8-bit mask) in LDS. Accepted pairs are appended with one
// atomicAdd per wave per emission round (ballot + prefix count).
#include <hipcub/hipcub.hpp>

#include "cull.hpp"
#include "ipc_device.hpp"
#include "traverse.hpp"

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


// --------------------------------------------------------- primitive boxes ----
// mode 0: edge box +- pad (edge_aabb, contact.py:30-37)
// mode 1: triangle bounds +- pad (wp.Mesh bounds when pad = 0)
// mode 2: swept edge box over x0, x1, +- pad (swept_edge_aabbs, ccd/ccd.py:166-186)
// mode 3: swept triangle box over x0, x1, +- pad (swept_triangle_aabbs, ccd/ccd.py:188-210)
__device__ __forceinline__ void prim_box(int mode, int p, const int* __restrict__ prims, const float* __restrict__ x0,
                                         const float* __restrict__ x1, float pad, float4& lo4, float4& hi4) {
  int k = (mode == 0 || mode == 2) ? 2 : 3;
  V3<float> lo, hi;
  lo = hi = ldx(x0, prims[k * p]);
  for (int j = 1; j < k; ++j) {
    V3<float> q = ldx(x0, prims[k * p + j]);
    lo = vmin(lo, q);
    hi = vmax(hi, q);
  }
  if (mode >= 2) {
    for (int j = 0; j < k; ++j) {
      V3<float> q = ldx(x1, prims[k * p + j]);
      lo = vmin(lo, q);
      hi = vmax(hi, q);
    }
  }
  V3<float> pv = mk(pad, pad, pad);
  lo = lo - pv;
  hi = hi + pv;
  lo4 = make_float4(lo.x, lo.y, lo.z, 0.f);
  hi4 = make_float4(hi.x, hi.y, hi.z, 0.f);
}

// Scene bounds of the primitive box centres (ordered-int atomics), for the
// Morton quantisation of a re-sort.
__global__ void __launch_bounds__(256) centre_bounds_kernel(int mode, int n, const int* __restrict__ prims,
                                                            const float* __restrict__ x0, const float* __restrict__ x1,
                                                            float pad, int* __restrict__ bounds) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  float c[3] = {0.f, 0.f, 0.f};
  bool live = i < n;
  if (live) {
    float4 lo, hi;
    prim_box(mode, i, prims, x0, x1, pad, lo, hi);
    c[0] = 0.5f * (lo.x + hi.x);
    c[1] = 0.5f * (lo.y + hi.y);
    c[2] = 0.5f * (lo.z + hi.z);
  }
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

__global__ void __launch_bounds__(256) morton_kernel(int mode, int n, const int* __restrict__ prims,
                                                     const float* __restrict__ x0, const float* __restrict__ x1,
                                                     float pad, const int* __restrict__ bounds,
                                                     unsigned long long* __restrict__ keys) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 lo, hi;
  prim_box(mode, i, prims, x0, x1, pad, lo, hi);
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
constexpr int kReduceBlock = 64;                      // threads of the leaf/level kernel (one wave)
constexpr int kLeavesPerBlock = kFan * kReduceBlock;  // 512 = 8^3 leaves per workgroup
constexpr int kInBlockLevels = 3;                     // levels 1..3 reduced inside the workgroup

__device__ __forceinline__ void box_merge(float4& lo, float4& hi, float4 l, float4 h) {
  lo = make_float4(fminf(lo.x, l.x), fminf(lo.y, l.y), fminf(lo.z, l.z), 0.f);
  hi = make_float4(fmaxf(hi.x, h.x), fmaxf(hi.y, h.y), fmaxf(hi.z, h.z), 0.f);
}

__device__ __forceinline__ void box_empty(float4& lo, float4& hi) {
  lo = make_float4(INFINITY, INFINITY, INFINITY, 0.f);
  hi = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.f);
}

// Leaves in Morton order and levels 1..3 of the hierarchy. Workgroup b owns
// leaves [512 b, 512 (b + 1)); each leaf's box is computed here from the
// topology and positions of primitive keys[k] (low 32 bits).
__global__ void __launch_bounds__(kReduceBlock) leaf_levels_kernel(
    int mode, int n, const int* __restrict__ prims, const int* __restrict__ body, const float* __restrict__ x0,
    const float* __restrict__ x1, float pad, const unsigned long long* __restrict__ keys,
    float4* __restrict__ leaf_lo, float4* __restrict__ leaf_hi, int* __restrict__ leaf_id, int* __restrict__ leaf_body,
    float4* __restrict__ node_lo, float4* __restrict__ node_hi, int2* __restrict__ node_body, LevelInfo info) {
  __shared__ float4 slo[kReduceBlock], shi[kReduceBlock];
  __shared__ int2 sbody[kReduceBlock];
  int t = threadIdx.x;
  int g = blockIdx.x * kReduceBlock + t;  // level-1 node
  int stride = (mode == 0 || mode == 2) ? 2 : 3;
  float4 blo, bhi;
  box_empty(blo, bhi);
  int2 br = make_int2(0x7fffffff, (int)0x80000000);
#pragma unroll
  for (int c = 0; c < kFan; ++c) {
    int k = g * kFan + c;
    if (k < n) {
      int prim = static_cast<int>(keys[k] & 0xffffffffu);
      float4 l, h;
      prim_box(mode, prim, prims, x0, x1, pad, l, h);
      int bd = body[prims[stride * prim]];
      leaf_lo[k] = l;
      leaf_hi[k] = h;
      leaf_id[k] = prim;
      leaf_body[k] = bd;
      box_merge(blo, bhi, l, h);
      br.x = min(br.x, bd);
      br.y = max(br.y, bd);
    }
  }
  if (g < info.count[1]) {
    node_lo[info.offset[1] + g] = blo;
    node_hi[info.offset[1] + g] = bhi;
    node_body[info.offset[1] + g] = br;
  }
  slo[t] = blo;
  shi[t] = bhi;
  sbody[t] = br;
  int width = kReduceBlock;
  for (int L = 2; L <= kInBlockLevels; ++L) {
    __syncthreads();
    width /= kFan;
    float4 l, h;
    box_empty(l, h);
    int2 r2 = make_int2(0x7fffffff, (int)0x80000000);
    if (t < width) {
#pragma unroll
      for (int c = 0; c < kFan; ++c) {
        box_merge(l, h, slo[t * kFan + c], shi[t * kFan + c]);
        r2.x = min(r2.x, sbody[t * kFan + c].x);
        r2.y = max(r2.y, sbody[t * kFan + c].y);
      }
    }
    __syncthreads();
    if (t < width) {
      slo[t] = l;
      shi[t] = h;
      sbody[t] = r2;
      int gl = blockIdx.x * width + t;
      if (L <= info.top && gl < info.count[L]) {
        node_lo[info.offset[L] + gl] = l;
        node_hi[info.offset[L] + gl] = h;
        node_body[info.offset[L] + gl] = r2;
      }
    }
  }
}

// Levels kInBlockLevels + 1 .. top in one workgroup.
__global__ void __launch_bounds__(1024) reduce_top_kernel(float4* __restrict__ node_lo, float4* __restrict__ node_hi,
                                                         int2* __restrict__ node_body, LevelInfo info) {
  for (int L = kInBlockLevels + 1; L <= info.top; ++L) {
    for (int g = threadIdx.x; g < info.count[L]; g += blockDim.x) {
      float4 l, h;
      box_empty(l, h);
      int2 r2 = make_int2(0x7fffffff, (int)0x80000000);
      for (int c = 0; c < kFan; ++c) {
        int k = g * kFan + c;
        if (k < info.count[L - 1]) {
          box_merge(l, h, node_lo[info.offset[L - 1] + k], node_hi[info.offset[L - 1] + k]);
          int2 cb = node_body[info.offset[L - 1] + k];
          r2.x = min(r2.x, cb.x);
          r2.y = max(r2.y, cb.y);
        }
      }
      node_lo[info.offset[L] + g] = l;
      node_hi[info.offset[L] + g] = h;
      node_body[info.offset[L] + g] = r2;
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ queries ----


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
  int n_query;
  const int* edges;
  const int* tris;
  const int* body;
  const float* x0;
  const float* x1;
  const float4* qlo;  // modes 0, 2: query box of slot k (the tree's leaf k)
  const float4* qhi;
  const int* order;   // optional query order (slot -> query primitive)
  float pad;          // mode 1: 2 (th + buf); mode 3: CCD padding
  float thickness;
  int exclude_same_body;
  // hierarchy
  int n_leaves;
  LevelInfo info;
  const float4* node_lo;
  const float4* node_hi;
  const int2* node_body;
  const float4* leaf_lo;
  const float4* leaf_hi;
  const int* leaf_id;
  const int* leaf_body;
  // output
  int4* rec;
  int2* ids;
  int capacity;
  int* count;
};

template <int MODE>
__device__ __forceinline__ bool accept_pair(const QueryArgs& A, int i, int j, int4& rec) {
  if (MODE == Q_ISECT) {
    // edge i against triangle j: a crossing, not sharing a vertex
    int a0 = A.edges[2 * i], a1 = A.edges[2 * i + 1];
    int t0 = A.tris[3 * j], t1 = A.tris[3 * j + 1], t2 = A.tris[3 * j + 2];
    if (a0 == t0 || a0 == t1 || a0 == t2 || a1 == t0 || a1 == t1 || a1 == t2) return false;
    rec = make_int4(a0, a1, t0, t1);
    return segment_hits_triangle(to_d(ldx(A.x0, a0)), to_d(ldx(A.x0, a1)), to_d(ldx(A.x0, t0)), to_d(ldx(A.x0, t1)),
                                 to_d(ldx(A.x0, t2)));
  }
  if (MODE == 0 || MODE == 2) {
    if (!(i < j)) return false;
    int a0 = A.edges[2 * i], a1 = A.edges[2 * i + 1], b0 = A.edges[2 * j], b1 = A.edges[2 * j + 1];
    if (a0 == b0 || a0 == b1 || a1 == b0 || a1 == b1) return false;
    if (A.exclude_same_body && A.body[a0] == A.body[b0]) return false;
    rec = make_int4(a0, a1, b0, b1);
    if (MODE == 2) return true;
    EdgeClosest ec = closest_point_edge_edge(ldx(A.x0, a0), ldx(A.x0, a1), ldx(A.x0, b0), ldx(A.x0, b1), 1.0e-6f);
    return ec.dist < A.thickness * 2.0f;
  }
  int t0 = A.tris[3 * j], t1 = A.tris[3 * j + 1], t2 = A.tris[3 * j + 2];
  if (t0 == i || t1 == i || t2 == i) return false;
  if (A.exclude_same_body && A.body[i] == A.body[t0]) return false;
  if (MODE == 3) {
    rec = make_int4(i, t0, t1, t2);
    return true;
  }
  rec = make_int4(i, t1, t0, t2);
  TriClosest tc = closest_point_triangle(ldx(A.x0, i), ldx(A.x0, t0), ldx(A.x0, t1), ldx(A.x0, t2));
  return tc.dist < A.thickness * 2.0f;
}


// Eight lanes share one query. Visiting a node, lane r tests the node's
// child r; the group ballot of overlapping children becomes one stack entry
// (level, first child, mask). Every lane of a group computes and writes the
// same stack values, so each lane only ever reads back its own LDS writes.
template <int MODE>
__global__ void __launch_bounds__(kQueryBlock) query_kernel(QueryArgs A) {
  __shared__ unsigned long long stack[kGroupsPerBlock][kMaskStack];
  // Accepted pairs are staged per block and written out with one global
  // atomicAdd per block (one counter shared by every wave serialises).
  __shared__ int4 out_rec[kBlockOut];
  __shared__ int2 out_ids[kBlockOut];
  __shared__ int out_n, out_base;
  int tid = threadIdx.x;
  if (tid == 0) out_n = 0;
  __syncthreads();
  int grp = tid / kFan, r = tid % kFan;
  int lane = __lane_id();
  int group_shift = lane & ~(kFan - 1);  // first lane of this group in the wave
  int slot = blockIdx.x * kGroupsPerBlock + grp;
  bool live = slot < A.n_query && A.n_leaves > 0;
  int i = 0;
  float qlo[3] = {0.f, 0.f, 0.f}, qhi[3] = {-1.f, -1.f, -1.f};
  int qbody = -1;
  if (live) {
    i = A.order ? A.order[slot] : slot;
    if (MODE == Q_ISECT) {
      V3<float> p = ldx(A.x0, A.edges[2 * i]), q = ldx(A.x0, A.edges[2 * i + 1]);
      V3<float> lo = vmin(p, q), hi = vmax(p, q);
      qlo[0] = lo.x; qlo[1] = lo.y; qlo[2] = lo.z;
      qhi[0] = hi.x; qhi[1] = hi.y; qhi[2] = hi.z;
    } else if (MODE == 0 || MODE == 2) {
      float4 l = A.qlo[slot], h = A.qhi[slot];
      qlo[0] = l.x; qlo[1] = l.y; qlo[2] = l.z;
      qhi[0] = h.x; qhi[1] = h.y; qhi[2] = h.z;
      if (A.exclude_same_body) qbody = A.body[A.edges[2 * i]];
    } else {
      V3<float> p = ldx(A.x0, i), lo = p, hi = p;
      if (MODE == 3) {
        V3<float> q = ldx(A.x1, i);
        lo = vmin(p, q);
        hi = vmax(p, q);
      }
      V3<float> pv = mk(A.pad, A.pad, A.pad);
      lo = lo - pv;
      hi = hi + pv;
      qlo[0] = lo.x; qlo[1] = lo.y; qlo[2] = lo.z;
      qhi[0] = hi.x; qhi[1] = hi.y; qhi[2] = hi.z;
      if (A.exclude_same_body) qbody = A.body[i];
    }
  }
  int sp = 0;
  if (live) {
    stack[grp][0] = mask_entry(A.info.top, 1u, 0);  // the root as the only candidate of its level
    sp = 1;
  }
  while (true) {
    bool want = false;
    int j = -1;
    int4 rec;
    bool active = sp > 0;
    int L = 0, node = 0;
    if (active) {
      unsigned long long e = stack[grp][sp - 1];
      L = int(e >> 40);
      unsigned mask = unsigned(e >> 32) & 0xffu;
      int first = int(e & 0xffffffffu);
      int c = __ffs(mask) - 1;
      node = first + c;
      mask &= mask - 1u;
      if (mask == 0u) --sp;
      else stack[grp][sp - 1] = mask_entry(L, mask, first);
    }
    // the node's children: internal nodes of level L-1, or leaves when L == 1
    bool overlap_child = false;
    int child = node * kFan + r;
    if (active && L >= 2) {
      int lv = L - 1;
      if (child < A.info.count[lv]) {
        int at = A.info.offset[lv] + child;
        overlap_child = overlaps(qlo, qhi, A.node_lo[at], A.node_hi[at]);
        if (overlap_child && qbody >= 0) {
          int2 br = A.node_body[at];
          overlap_child = !(br.x == qbody && br.y == qbody);
        }
      }
    } else if (active && L == 1) {
      if (child < A.n_leaves && overlaps(qlo, qhi, A.leaf_lo[child], A.leaf_hi[child]) &&
          !(qbody >= 0 && A.leaf_body[child] == qbody)) {
        j = A.leaf_id[child];
        want = accept_pair<MODE>(A, i, j, rec);
      }
    }
    unsigned long long ball = __ballot(overlap_child);
    unsigned gmask = unsigned(ball >> group_shift) & 0xffu;
    if (active && L >= 2 && gmask != 0u && sp < kMaskStack) {
      stack[grp][sp] = mask_entry(L - 1, gmask, node * kFan);
      ++sp;
    }
    int slot_out = wave_append(want, &out_n);  // LDS counter
    if (want) {
      if (slot_out < kBlockOut) {
        out_rec[slot_out] = rec;
        out_ids[slot_out] = make_int2(i, j);
      } else {
        int out = atomicAdd(A.count, 1);  // staging full: straight to the output
        if (out < A.capacity) {
          if (A.rec) A.rec[out] = rec;
          if (A.ids) A.ids[out] = make_int2(i, j);
        }
      }
    }
    if (!__any(sp > 0)) break;
  }
  __syncthreads();
  int staged = out_n < kBlockOut ? out_n : kBlockOut;
  if (tid == 0) out_base = staged ? atomicAdd(A.count, staged) : 0;
  __syncthreads();
  for (int k = tid; k < staged; k += kQueryBlock) {
    int out = out_base + k;
    if (out < A.capacity) {
      if (A.rec) A.rec[out] = out_rec[k];
      if (A.ids) A.ids[out] = out_ids[k];
    }
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
  keys.release(); keys_sorted.release(); leaf_lo.release(); leaf_hi.release();
  leaf_id.release(); node_lo.release(); node_hi.release(); leaf_body.release(); node_body.release();
  temp.release(); bounds.release();
  n = 0;
  builds_since_sort = -1;
}

#define IPC_TRY(expr)                    \
  do {                                   \
    hipError_t _e = (expr);              \
    if (_e != hipSuccess) return _e;     \
  } while (0)

hipError_t build_bvh(Bvh& B, int mode, int n, const int* prims, const int* body, const float* x0, const float* x1,
                     float pad, hipStream_t s) {
  B.n = n;
  if (n <= 0) return hipSuccess;
  if (n > (1 << 27)) return hipErrorInvalidValue;
  // level sizes: count[0] = n leaves, count[L] = ceil(count[L-1] / 8) up to a single root
  LevelInfo& I = B.info;
  for (int L = 0; L < kMaxLevels; ++L) { I.count[L] = 0; I.offset[L] = 0; }
  I.count[0] = n;
  int L = 0, total = 0;
  do {
    ++L;
    I.count[L] = (I.count[L - 1] + kFan - 1) / kFan;
    I.offset[L] = total;
    total += I.count[L];
  } while (I.count[L] > 1);
  I.top = L;
  IPC_TRY(B.bounds.ensure(sizeof(int) * 8));
  IPC_TRY(B.keys.ensure(sizeof(unsigned long long) * n));
  IPC_TRY(B.keys_sorted.ensure(sizeof(unsigned long long) * n));
  IPC_TRY(B.leaf_lo.ensure(sizeof(float4) * n));
  IPC_TRY(B.leaf_hi.ensure(sizeof(float4) * n));
  IPC_TRY(B.leaf_id.ensure(sizeof(int) * n));
  IPC_TRY(B.node_lo.ensure(sizeof(float4) * total));
  IPC_TRY(B.node_hi.ensure(sizeof(float4) * total));
  IPC_TRY(B.leaf_body.ensure(sizeof(int) * n));
  IPC_TRY(B.node_body.ensure(sizeof(int2) * total));
  bool resort = B.builds_since_sort < 0 || B.builds_since_sort >= kResortEvery || B.n_sorted != n;
  if (resort) {
    hipLaunchKernelGGL(init_bounds_kernel, dim3(1), dim3(64), 0, s, B.bounds.as<int>());
    hipLaunchKernelGGL(centre_bounds_kernel, dim3(nblocks(n, 256)), dim3(256), 0, s, mode, n, prims, x0, x1, pad,
                       B.bounds.as<int>());
    hipLaunchKernelGGL(morton_kernel, dim3(nblocks(n, 256)), dim3(256), 0, s, mode, n, prims, x0, x1, pad,
                       B.bounds.as<int>(), B.keys.as<unsigned long long>());
    IPC_TRY(hipGetLastError());
    size_t temp_bytes = 0;
    IPC_TRY(hipcub::DeviceRadixSort::SortKeys(nullptr, temp_bytes, B.keys.as<unsigned long long>(),
                                              B.keys_sorted.as<unsigned long long>(), n, 0, 62, s));
    IPC_TRY(B.temp.ensure(temp_bytes));
    temp_bytes = B.temp.bytes;
    IPC_TRY(hipcub::DeviceRadixSort::SortKeys(B.temp.p, temp_bytes, B.keys.as<unsigned long long>(),
                                              B.keys_sorted.as<unsigned long long>(), n, 0, 62, s));
    B.builds_since_sort = 0;
    B.n_sorted = n;
  } else {
    ++B.builds_since_sort;
  }
  hipLaunchKernelGGL(leaf_levels_kernel, dim3(nblocks(n, kLeavesPerBlock)), dim3(kReduceBlock), 0, s, mode, n, prims,
                     body, x0, x1, pad, B.keys_sorted.as<unsigned long long>(), B.leaf_lo.as<float4>(),
                     B.leaf_hi.as<float4>(), B.leaf_id.as<int>(), B.leaf_body.as<int>(), B.node_lo.as<float4>(),
                     B.node_hi.as<float4>(), B.node_body.as<int2>(), I);
  if (I.top > kInBlockLevels)
    hipLaunchKernelGGL(reduce_top_kernel, dim3(1), dim3(1024), 0, s, B.node_lo.as<float4>(), B.node_hi.as<float4>(),
                       B.node_body.as<int2>(), I);
  return hipGetLastError();
}

hipError_t run_query(const QueryLaunch& q, hipStream_t s) {
  if (q.n_query <= 0 || q.bvh->n <= 0) return hipSuccess;
  QueryArgs A;
  A.n_query = q.n_query;
  A.edges = q.edges;
  A.tris = q.tris;
  A.body = q.body;
  A.x0 = q.x0;
  A.x1 = q.x1;
  A.qlo = q.qlo;
  A.qhi = q.qhi;
  A.order = q.order;
  A.pad = q.pad;
  A.thickness = q.thickness;
  A.exclude_same_body = q.exclude_same_body;
  A.n_leaves = q.bvh->n;
  A.info = q.bvh->info;
  A.node_lo = q.bvh->node_lo.as<float4>();
  A.node_hi = q.bvh->node_hi.as<float4>();
  A.node_body = q.bvh->node_body.as<int2>();
  A.leaf_lo = q.bvh->leaf_lo.as<float4>();
  A.leaf_hi = q.bvh->leaf_hi.as<float4>();
  A.leaf_id = q.bvh->leaf_id.as<int>();
  A.leaf_body = q.bvh->leaf_body.as<int>();
  A.rec = q.rec;
  A.ids = q.ids;
  A.capacity = q.capacity;
  A.count = q.count;
  dim3 grid(nblocks(q.n_query, kGroupsPerBlock)), block(kQueryBlock);
  switch (q.mode) {
    case Q_DCD_EE: hipLaunchKernelGGL(query_kernel<Q_DCD_EE>, grid, block, 0, s, A); break;
    case Q_DCD_PT: hipLaunchKernelGGL(query_kernel<Q_DCD_PT>, grid, block, 0, s, A); break;
    case Q_CCD_EE: hipLaunchKernelGGL(query_kernel<Q_CCD_EE>, grid, block, 0, s, A); break;
    case Q_CCD_PT: hipLaunchKernelGGL(query_kernel<Q_CCD_PT>, grid, block, 0, s, A); break;
    case Q_ISECT: hipLaunchKernelGGL(query_kernel<Q_ISECT>, grid, block, 0, s, A); break;
    default: return hipErrorInvalidValue;
  }
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

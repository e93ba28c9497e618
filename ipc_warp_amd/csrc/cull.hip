// Candidate culling of libipc_warp_amd: BVH build and AABB queries for DCD
// (edge_edge_collision / point_triangle_collision, contact.py:199-336) and
// for the CCD broad phase (edge_edge_toi / point_triangle_toi,
// ccd/ccd.py:166-210,360-427).
//
// The BVH is rebuilt on every call from the current positions (the reference
// refits a BVH built once; a rebuild keeps query cost independent of how far
// the mesh has moved):
//   1. one AABB per primitive (fp32, the reference's inflation),
//   2. 30-bit Morton codes of box centres in the scene bounds, the leaves
//      ordered by Morton cell (a counting sort on the top 18 code bits) on
//      the first build and every kResortEvery builds (in between, the leaf
//      order is kept and only the boxes are refreshed),
//   3. an implicit 8-ary hierarchy over the sorted leaves: node k of level L
//      bounds nodes 8k .. 8k+7 of level L-1 (level 0 = leaves). Levels 1-3
//      are reduced by one kernel, one thread per leaf (a 512-thread
//      workgroup owns 512 leaves: lane shuffles for levels 1 and 2, LDS for
//      level 3), the levels above by a single workgroup. There is no bottom-up
//      propagation across workgroups, hence no device-scope fence: on a
//      multi-XCD part such a fence writes back and invalidates the L2.
// Nodes also carry the (min, max) body id below them, so a query that
// excludes its own body skips single-body subtrees. Queries run eight lanes
// per query: visiting a node, lane r tests the node's child r, and the
// group's ballot of overlapping children becomes one stack entry (level,
// first child, 8-bit mask) in LDS. Accepted pairs are staged in LDS and
// each workgroup reserves its output range with one global atomicAdd. The
// paired launches of a DCD or CCD call dispatch their workgroups heaviest
// first, by the step counts of the previous call (QueryOrder).
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

// The scene bounds are kept in kBoundCopies copies of (min xyz, max xyz),
// workgroup b of the bounds launch folding into copy b % 8 (its XCD under
// round-robin dispatch): atomics of every workgroup on the same six words
// serialised and crossed XCDs. The Morton launch folds the copies.
constexpr int kBoundCopies = 8;
constexpr int kBoundInts = 6 * kBoundCopies;

// Leaf order cells: the top 18 of the 30 Morton code bits (64 per axis).
constexpr int kCellShift = 12;
constexpr int kCells = 1 << (30 - kCellShift);
// masked: a code built from a NaN centre may carry bits above 30
__device__ __forceinline__ int cell_of(unsigned code) { return static_cast<int>((code >> kCellShift) & (kCells - 1)); }

// Resets the bound copies and clears the kCells cell counts of a re-sort.
constexpr int kInitBlocks = 256;
__global__ void __launch_bounds__(256) init_bounds_kernel(int* __restrict__ bounds, int* __restrict__ count) {
  const int t = threadIdx.x;
  if (blockIdx.x == 0 && t < kBoundInts) bounds[t] = t % 6 < 3 ? 0x7fffffff : (int)0x80000000;  // min xyz, max xyz
  for (int i = blockIdx.x * blockDim.x + t; i < kCells; i += kInitBlocks * 256) count[i] = 0;
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
// mode 4: the point x0[p] (query ordering only)
// Vertex i at the end of a CCD step: x1[i], or, with dx, x0[i] - dx[i] (the
// fp32 subtraction of _add_dx, contact.py:49-52, so the same value).
__device__ __forceinline__ V3<float> end_pos(const float* __restrict__ x0, const float* __restrict__ x1,
                                             const float* __restrict__ dx, int i) {
  if (dx) {
    const V3<float> a = ldx(x0, i), d = ldx(dx, i);
    return mk(a.x - d.x, a.y - d.y, a.z - d.z);
  }
  return ldx(x1, i);
}

// The box of a primitive with vertex ids v[0 .. k) (k = 2 for an edge, 3 for
// a triangle) in box mode 0..3.
// DCD box slack. The reference inflates each DCD edge box by thickness +
// buffer (and each query point by twice that) but keeps only pairs whose fp32
// distance is below 2 thickness; a box needs only thickness plus the
// rounding of that distance test. With M the largest coordinate magnitude of
// the primitive, the fp32 distance of two primitives within 2 thickness of
// each other is at most about 12 ulp(M) below their exact distance, and
// sigma = thickness 2^-10 + 2^-16 M (256 ulp(M) and more) covers it. A DCD
// box is inflated by thickness + min(buffer, sigma): every pair the distance
// test accepts still overlaps, and where sigma >= buffer the box is the
// reference's, bit for bit, so the set found is the reference's either way.
// (a NaN buffer, or a NaN or infinite coordinate, selects the buffer: the
// reference's box exactly)
__device__ __forceinline__ float dcd_pad(float thickness, float buffer, float m) {
  const float sigma = thickness * 0x1p-10f + m * 0x1p-16f;
  return thickness + (buffer >= sigma ? sigma : buffer);
}
__device__ __forceinline__ float abs_max(V3<float> a, V3<float> b) {
  return fmaxf(fmaxf(fmaxf(fabsf(a.x), fabsf(a.y)), fmaxf(fabsf(a.z), fabsf(b.x))), fmaxf(fabsf(b.y), fabsf(b.z)));
}

// With mode BOX_EDGE (the DCD edge boxes), pad is the thickness and cap the
// buffer; the other modes are inflated by pad.
__device__ __forceinline__ void vertex_box(int mode, int k, const int v[3], const float* __restrict__ x0,
                                           const float* __restrict__ x1, const float* __restrict__ dx, float pad,
                                           float cap, float4& lo4, float4& hi4) {
  V3<float> lo, hi;
  lo = hi = ldx(x0, v[0]);
  for (int j = 1; j < k; ++j) {
    V3<float> q = ldx(x0, v[j]);
    lo = vmin(lo, q);
    hi = vmax(hi, q);
  }
  if (mode >= 2) {
    for (int j = 0; j < k; ++j) {
      V3<float> q = end_pos(x0, x1, dx, v[j]);
      lo = vmin(lo, q);
      hi = vmax(hi, q);
    }
  }
  if (mode == BOX_EDGE) pad = dcd_pad(pad, cap, abs_max(lo, hi));
  V3<float> pv = mk(pad, pad, pad);
  lo = lo - pv;
  hi = hi + pv;
  lo4 = make_float4(lo.x, lo.y, lo.z, 0.f);
  hi4 = make_float4(hi.x, hi.y, hi.z, 0.f);
}

__device__ __forceinline__ void prim_box(int mode, int p, const int* __restrict__ prims, const float* __restrict__ x0,
                                         const float* __restrict__ x1, const float* __restrict__ dx, float pad,
                                         float4& lo4, float4& hi4) {
  if (mode == BOX_POINT) {
    V3<float> q = ldx(x0, p);
    lo4 = hi4 = make_float4(q.x, q.y, q.z, 0.f);
    return;
  }
  const int k = (mode == 0 || mode == 2) ? 2 : 3;
  const int v[3] = {prims[k * p], prims[k * p + 1], k == 3 ? prims[k * p + 2] : 0};
  vertex_box(mode, k, v, x0, x1, dx, pad, 0.f, lo4, hi4);
}

// Scene bounds of the primitive box centres (ordered-int atomics), for the
// Morton quantisation of a re-sort.
__global__ void __launch_bounds__(256) centre_bounds_kernel(int mode, int n, const int* __restrict__ prims,
                                                            const float* __restrict__ x0, const float* __restrict__ x1,
                                                            const float* __restrict__ dx, float pad,
                                                            int* __restrict__ bounds) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  float c[3] = {0.f, 0.f, 0.f};
  bool live = i < n;
  if (live) {
    float4 lo, hi;
    prim_box(mode, i, prims, x0, x1, dx, pad, lo, hi);
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
    int* copy = bounds + 6 * (blockIdx.x % kBoundCopies);
    atomicMin(copy + a, mn);
    atomicMax(copy + 3 + a, mx);
  }
}

// Adds one for every lane with `want` to counters[cell], with one atomic per
// distinct cell of the wave (the lanes of a wave hold neighbouring
// primitives, and with cubic cells a few dozen of them share a cell: one
// atomic per lane serialised on those addresses). Returns the lane's slot,
// the counter's old value plus its rank among the wave's lanes of the same
// cell, or -1 without `want`. Called by every lane of the wave.
__device__ __forceinline__ int wave_cell_add(int* __restrict__ counters, int cell, bool want) {
  const int lane = __lane_id();
  unsigned long long left = __ballot(want);
  int slot = -1;
  while (left != 0ull) {
    const int leader = __ffsll(static_cast<long long>(left)) - 1;
    const int lc = __builtin_amdgcn_readlane(cell, leader);
    const bool mine = want && cell == lc;
    const unsigned long long same = __ballot(mine);
    int base = 0;
    if (lane == leader) base = atomicAdd(&counters[lc], __popcll(same));
    base = __builtin_amdgcn_readlane(base, leader);
    if (mine) slot = base + __popcll(same & ((1ull << lane) - 1ull));
    left &= ~same;
  }
  return slot;
}

__global__ void __launch_bounds__(256) morton_kernel(int mode, int n, const int* __restrict__ prims,
                                                     const float* __restrict__ x0, const float* __restrict__ x1,
                                                     const float* __restrict__ dx, float pad,
                                                     const int* __restrict__ bounds, unsigned* __restrict__ codes,
                                                     int* __restrict__ count) {
  __shared__ int sb[6];
  if (threadIdx.x < 6) {  // fold the copies
    const int t = threadIdx.x;
    int v = bounds[t];
    for (int k = 1; k < kBoundCopies; ++k) v = t < 3 ? min(v, bounds[6 * k + t]) : max(v, bounds[6 * k + t]);
    sb[t] = v;
  }
  __syncthreads();
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i < n;
  if (!live) i = n - 1;  // a valid primitive; its code is neither stored nor counted
  float4 lo, hi;
  prim_box(mode, i, prims, x0, x1, dx, pad, lo, hi);
  float c[3] = {0.5f * (lo.x + hi.x), 0.5f * (lo.y + hi.y), 0.5f * (lo.z + hi.z)};
  // Cubic cells: every axis is quantised by the scene's largest extent, so
  // that 8 consecutive codes group primitives that are close in every
  // direction (per-axis scaling stretches the cells of an elongated scene:
  // 8:1 on the rod bundle, where it made EE queries pop 43 nodes each
  // instead of 30).
  float ext = 0.f;
  for (int a = 0; a < 3; ++a) ext = fmaxf(ext, float_of(sb[3 + a]) - float_of(sb[a]));
  unsigned q[3];
  for (int a = 0; a < 3; ++a) {
    float b0 = float_of(sb[a]);
    float u = ext > 0.f ? (c[a] - b0) / ext : 0.5f;
    float s = u * 1024.f;
    s = s < 0.f ? 0.f : (s > 1023.f ? 1023.f : s);
    q[a] = unsigned(s);
  }
  const unsigned code = (expand_bits10(q[0]) << 2) | (expand_bits10(q[1]) << 1) | expand_bits10(q[2]);
  if (live) codes[i] = code;
  wave_cell_add(count, cell_of(code), live);
}

// Leaf order by Morton code: primitives are bucketed by the top 18 of the 30
// code bits (64 cells per axis) with a counting sort, cells in Morton order,
// then each cell's primitives are put in (code, id) order by ranking each
// within its cell. Where two surfaces lie close together (the screw and its
// nut, 1e-3 apart) a cell holds a few dozen primitives of both bodies, and
// the scatter's atomic order would give 8-leaf nodes cell-sized boxes and
// mixed body ranges. A cell of more than kRankCap primitives keeps the
// atomic order (the ranking is quadratic in the cell's size); the order only
// affects speed, never the pairs found. The counts are cleared in the
// bounds launch and counted in the Morton launch, so the sort adds three
// launches (scan, scatter, rank), against a block sort and about seven merge
// launches for a full radix sort of the codes at these sizes.
__global__ void __launch_bounds__(256) cell_scatter_kernel(int n, const unsigned* __restrict__ codes,
                                                           int* __restrict__ cursor,
                                                           unsigned long long* __restrict__ keys_out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = i < n;
  const unsigned code = live ? codes[i] : 0u;
  const int at = wave_cell_add(cursor, cell_of(code), live);
  if (live) keys_out[at] = (static_cast<unsigned long long>(code) << 32) | unsigned(i);
}

constexpr int kRankCap = 512;

// Slot s of the scattered order: its primitive goes to its cell's start plus
// the number of the cell's (code, id) keys below its own; the cell's keys lie
// contiguously in the scattered order, so a workgroup's threads of one cell
// read the same run. After the scatter, cursor[c] is cell c's end and
// count[c] its size.
__global__ void __launch_bounds__(256) cell_rank_kernel(int n, const int* __restrict__ count,
                                                        const int* __restrict__ cursor,
                                                        const unsigned long long* __restrict__ scattered,
                                                        int* __restrict__ ids_out) {
  const int slot = blockIdx.x * blockDim.x + threadIdx.x;
  if (slot >= n) return;
  const unsigned long long key = scattered[slot];
  const int i = static_cast<int>(key & 0xffffffffull);
  const int c = cell_of(static_cast<unsigned>(key >> 32));
  const int hi = cursor[c], len = count[c], lo = hi - len;
  if (len > kRankCap) {
    ids_out[slot] = i;
    return;
  }
  int rank = 0;
  for (int k = lo; k < hi; ++k) rank += scattered[k] < key;
  ids_out[lo + rank] = i;
}

// temp of a re-sort: cell counts (cleared by init_bounds_kernel, counted by
// morton_kernel), cell cursors, the scan's workspace, the scattered keys (n
// of 8 bytes). Each part starts on a 256-byte boundary: the scan's workspace
// holds its look-back state, which hipCUB expects aligned.
struct CellBufs {
  int* count;
  int* cursor;
  unsigned long long* scattered;
  void* scan_temp;
  size_t scan_bytes;
};

static hipError_t cell_buffers(DevBuf& temp, int n, CellBufs& c) {
  c.scan_bytes = 0;
  hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, c.scan_bytes, static_cast<const int*>(nullptr),
                                                  static_cast<int*>(nullptr), kCells, hipStream_t(nullptr));
  auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t cell_bytes = sizeof(int) * kCells;  // a multiple of 256
  const size_t scan_at = 2 * cell_bytes, order_at = scan_at + up(c.scan_bytes);
  if (e == hipSuccess) e = temp.ensure(order_at + sizeof(unsigned long long) * size_t(n > 0 ? n : 1));
  if (e != hipSuccess) return e;
  c.count = temp.as<int>();
  c.cursor = c.count + kCells;
  c.scan_temp = static_cast<char*>(temp.p) + scan_at;
  c.scattered = reinterpret_cast<unsigned long long*>(static_cast<char*>(temp.p) + order_at);
  return hipSuccess;
}

// ids_out receives the n primitive ids in Morton order (cells by counting
// sort, each cell of up to kRankCap ranked by (code, id)).
static hipError_t sort_codes(const CellBufs& c, const unsigned* codes, int n, int* ids_out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  size_t scan_bytes = c.scan_bytes;
  hipError_t e = hipcub::DeviceScan::ExclusiveSum(c.scan_temp, scan_bytes, c.count, c.cursor, kCells, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(cell_scatter_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, codes, c.cursor, c.scattered);
  hipLaunchKernelGGL(cell_rank_kernel, dim3((n + 255) / 256), dim3(256), 0, s, n, c.count, c.cursor, c.scattered,
                     ids_out);
  return hipGetLastError();
}

// ------------------------------------------------------------ hierarchy ----
constexpr int kLeafBlock = kFan * kFan * kFan;  // 512 leaves (and threads) per workgroup of the leaf kernel
constexpr int kInBlockLevels = 3;               // levels 1..3 reduced inside that workgroup

// A node is a pair of float4: xyz the box, and the w fields the bit patterns
// of the (min, max) body id below it, so that one step of a query reads the
// box and the body range in two 16-byte loads per child.
__device__ __forceinline__ int body_min(float4 lo) { return __float_as_int(lo.w); }
__device__ __forceinline__ int body_max(float4 hi) { return __float_as_int(hi.w); }

__device__ __forceinline__ void box_merge(float4& lo, float4& hi, float4 l, float4 h) {
  lo = make_float4(fminf(lo.x, l.x), fminf(lo.y, l.y), fminf(lo.z, l.z), __int_as_float(min(body_min(lo), body_min(l))));
  hi = make_float4(fmaxf(hi.x, h.x), fmaxf(hi.y, h.y), fmaxf(hi.z, h.z), __int_as_float(max(body_max(hi), body_max(h))));
}

// An empty box and an empty body range.
__device__ __forceinline__ void box_empty(float4& lo, float4& hi) {
  lo = make_float4(INFINITY, INFINITY, INFINITY, __int_as_float(0x7fffffff));
  hi = make_float4(-INFINITY, -INFINITY, -INFINITY, __int_as_float((int)0x80000000));
}

// Merge a node with the lane `off` away (xor), so that after offsets 1, 2, 4
// every lane of an aligned group of 8 holds the group's merge.
__device__ __forceinline__ void lane_merge(float4& lo, float4& hi, int off) {
  float4 l = make_float4(__shfl_xor(lo.x, off, 64), __shfl_xor(lo.y, off, 64), __shfl_xor(lo.z, off, 64),
                         __shfl_xor(lo.w, off, 64));
  float4 h = make_float4(__shfl_xor(hi.x, off, 64), __shfl_xor(hi.y, off, 64), __shfl_xor(hi.z, off, 64),
                         __shfl_xor(hi.w, off, 64));
  box_merge(lo, hi, l, h);
}

// One tree of a build launch (device side).
struct LeafJob {
  int mode, n;
  const int* prims;
  const float* x1;
  const float* dx;  // with dx, end positions are x0 - dx (x1 unused)
  float pad, cap;  // cap: the buffer of a BOX_EDGE job
  const int* ids;  // sorted primitive ids (leaf order)
  int* leaf_id;
  int4* leaf_vid;
  float4* node_lo;
  float4* node_hi;
  LevelInfo info;
  int reuse_vid;  // leaf_vid and the leaf bodies hold this leaf order's records
};

struct LeafSmem {
  float4 wlo[kLeafBlock / 64], whi[kLeafBlock / 64];
};

// Leaves in Morton order and levels 1..3 of the hierarchy, one thread per
// leaf slot. Workgroup b of the job owns leaf slots [512 b, 512 (b + 1)):
// level-1 nodes are merged across 8-lane groups, level-2 nodes across a
// wave, the level-3 node across the workgroup's 8 waves in LDS. Leaf k's box
// is computed from the topology and positions of primitive ids[k]; slots
// past n hold empty boxes. leaf_vid[k] holds the leaf's vertex
// ids (v2 = -1 for an edge) and primitive id, so that a query reads one
// record per leaf instead of chasing the topology.
__device__ __forceinline__ void leaf_block(const LeafJob& J, int bid, const int* __restrict__ body,
                                           const float* __restrict__ x0, LeafSmem& sm) {
  const LevelInfo& info = J.info;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int k = bid * kLeafBlock + t;
  const int stride = (J.mode == 0 || J.mode == 2) ? 2 : 3;
  float4 lo, hi;
  box_empty(lo, hi);
  int4 vid = make_int4(-1, -1, -1, -1);
  if (k < J.n) {
    int bd;
    if (J.reuse_vid) {
      // the leaf order is the previous build's: its record and body stand
      vid = J.leaf_vid[k];
      bd = __float_as_int(J.node_lo[k].w);
    } else {
      const int prim = J.ids[k];
      const int v0 = J.prims[stride * prim], v1 = J.prims[stride * prim + 1];
      const int v2 = stride == 3 ? J.prims[stride * prim + 2] : -1;
      bd = body[v0];
      J.leaf_id[k] = prim;
      vid = make_int4(v0, v1, v2, prim);
    }
    const int v[3] = {vid.x, vid.y, stride == 3 ? vid.z : 0};
    vertex_box(J.mode, stride, v, x0, J.x1, J.dx, J.pad, J.cap, lo, hi);
    lo.w = __int_as_float(bd);
    hi.w = __int_as_float(bd);
  }
  if (k < info.slots[0]) {  // offset[0] = 0
    J.node_lo[k] = lo;
    J.node_hi[k] = hi;
    if (!J.reuse_vid) J.leaf_vid[k] = vid;
  }
  // level 1: groups of 8 lanes
  lane_merge(lo, hi, 1);
  lane_merge(lo, hi, 2);
  lane_merge(lo, hi, 4);
  int g1 = bid * (kLeafBlock / kFan) + t / kFan;
  if ((lane & (kFan - 1)) == 0 && g1 < info.slots[1]) {
    J.node_lo[info.offset[1] + g1] = lo;
    J.node_hi[info.offset[1] + g1] = hi;
  }
  // level 2: the wave's 8 level-1 nodes
  lane_merge(lo, hi, 8);
  lane_merge(lo, hi, 16);
  lane_merge(lo, hi, 32);
  int g2 = bid * (kLeafBlock / 64) + wave;
  if (lane == 0) {
    if (g2 < info.slots[2]) {
      J.node_lo[info.offset[2] + g2] = lo;
      J.node_hi[info.offset[2] + g2] = hi;
    }
    sm.wlo[wave] = lo;
    sm.whi[wave] = hi;
  }
  __syncthreads();
  // level 3: the workgroup's 8 level-2 nodes
  if (t == 0 && bid < info.slots[3]) {
    float4 l = sm.wlo[0], h = sm.whi[0];
    for (int w = 1; w < kLeafBlock / 64; ++w) box_merge(l, h, sm.wlo[w], sm.whi[w]);
    J.node_lo[info.offset[3] + bid] = l;
    J.node_hi[info.offset[3] + bid] = h;
  }
}

// Workgroups [0, blocks0) build job 0, the rest job 1. The first workgroup
// also runs the launch's prologue (counters of the queries that follow).
// With end.out set, job 0's workgroups also write end.out[k] = x0[k] - dx[k]
// for the 3 nv floats of the positions (the CCD end positions that the
// queries and the TOI launch read next).
struct EndWrite {
  float* out;
  const float* dx;
  int n3;
};

__global__ void __launch_bounds__(kLeafBlock) leaf_levels_kernel(LeafJob J0, LeafJob J1, int blocks0,
                                                                 const int* __restrict__ body,
                                                                 const float* __restrict__ x0, BuildPrologue pro,
                                                                 EndWrite end) {
  __shared__ LeafSmem sm;
  run_prologue(pro);
  if (end.out && int(blockIdx.x) < blocks0)
    for (int f = blockIdx.x * kLeafBlock + threadIdx.x; f < end.n3; f += blocks0 * kLeafBlock)
      end.out[f] = x0[f] - end.dx[f];
  if (int(blockIdx.x) < blocks0) leaf_block(J0, blockIdx.x, body, x0, sm);
  else leaf_block(J1, blockIdx.x - blocks0, body, x0, sm);
}

// The prologue alone, for a launch whose trees are all empty.
__global__ void prologue_kernel(BuildPrologue pro) { run_prologue(pro); }

// Levels kInBlockLevels + 1 .. top of one tree in one workgroup, and the
// padding slots of level kInBlockLevels past the leaf kernel's one node per
// workgroup.
struct TopJob {
  float4* node_lo;
  float4* node_hi;
  LevelInfo info;
};

__device__ __forceinline__ void top_block(const TopJob& T) {
  const LevelInfo& info = T.info;
  if (info.top <= kInBlockLevels) return;
  float4 el, eh;
  box_empty(el, eh);
  for (int g = info.count[kInBlockLevels] + threadIdx.x; g < info.slots[kInBlockLevels]; g += blockDim.x) {
    T.node_lo[info.offset[kInBlockLevels] + g] = el;
    T.node_hi[info.offset[kInBlockLevels] + g] = eh;
  }
  __syncthreads();
  for (int L = kInBlockLevels + 1; L <= info.top; ++L) {
    for (int g = threadIdx.x; g < info.slots[L]; g += blockDim.x) {
      float4 l = el, h = eh;
      if (g < info.count[L]) {
        for (int c = 0; c < kFan; ++c) {
          int k = info.offset[L - 1] + g * kFan + c;
          box_merge(l, h, T.node_lo[k], T.node_hi[k]);
        }
      }
      T.node_lo[info.offset[L] + g] = l;
      T.node_hi[info.offset[L] + g] = h;
    }
    __syncthreads();
  }
}

// ------------------------------------------------------- dispatch order ----
// A launch's workgroups go to the XCDs round-robin: raw workgroup b runs on
// XCD b % 8, as that XCD's (b / 8)-th. In a paired query launch XCD x serves
// a[x] EE workgroups (one contiguous range of the Morton-ordered queries,
// from sa[x]; as many as there are raw b < n_a with b % 8 == x) and b[x] PT
// workgroups from sb[x] likewise, so each XCD's L2 holds its region's part
// of the trees.
struct XcdShare {
  int a[kXcds], b[kXcds], sa[kXcds], sb[kXcds];
};

__device__ __forceinline__ void xcd_share(int n_a, int n_b, XcdShare& X) {
  const int n = n_a + n_b;
  int sa = 0, sb = 0;
  for (int x = 0; x < kXcds; ++x) {
    const int cnt = (n + kXcds - 1 - x) / kXcds, a = (n_a + kXcds - 1 - x) / kXcds;
    X.a[x] = a;
    X.b[x] = cnt - a;
    X.sa[x] = sa;
    X.sb[x] = sb;
    sa += a;
    sb += cnt - a;
  }
}

constexpr int kOrderBuckets = 16;  // 4 steps per bucket; 60 steps and more share the first

struct OrderSmem {
  int hist[kOrderBuckets];
  XcdShare X;
};

// perm[raw] for XCD x's share of a paired launch, from the previous launch's
// per-wave step counts: a counting sort of the XCD's workgroups by
// decreasing cost (the largest of a workgroup's four waves'), written to
// the XCD's raw slots x, x + 8, ... in order. One workgroup per XCD.
__device__ __forceinline__ void order_block(const QueryOrder& O, int x, OrderSmem& sm) {
  const int t = threadIdx.x, nt = blockDim.x;
  if (t == 0) xcd_share(O.n_a, O.n_b, sm.X);
  if (t < kOrderBuckets) sm.hist[t] = 0;
  __syncthreads();
  const int a = sm.X.a[x], cnt = a + sm.X.b[x], sa = sm.X.sa[x], sb = O.n_a + sm.X.sb[x];
  auto item = [&](int i) { return i < a ? sa + i : sb + (i - a); };  // the XCD's i-th logical workgroup
  auto bucket = [&](int c) {
    const int4 w = reinterpret_cast<const int4*>(O.cost)[c];
    const int m = min(max(max(max(w.x, w.y), max(w.z, w.w)), 0) / 4, kOrderBuckets - 1);
    return kOrderBuckets - 1 - m;  // the most steps first
  };
  for (int i = t; i < cnt; i += nt) atomicAdd(&sm.hist[bucket(item(i))], 1);
  __syncthreads();
  if (t == 0) {  // exclusive scan
    int run = 0;
    for (int k = 0; k < kOrderBuckets; ++k) {
      const int v = sm.hist[k];
      sm.hist[k] = run;
      run += v;
    }
  }
  __syncthreads();
  for (int i = t; i < cnt; i += nt) {
    const int c = item(i);
    const int j = atomicAdd(&sm.hist[bucket(c)], 1);  // < cnt, so j * 8 + x < n_a + n_b
    O.perm[j * kXcds + x] = c;
  }
}

// Workgroup j < n_top reduces the top levels of job j; workgroups n_top +
// x, when launched, compute the dispatch order O for XCD x.
__global__ void __launch_bounds__(1024) reduce_top_kernel(TopJob T0, TopJob T1, int n_top, QueryOrder O) {
  if (int(blockIdx.x) >= n_top) {
    __shared__ OrderSmem sm;
    order_block(O, int(blockIdx.x) - n_top, sm);
    return;
  }
  if (blockIdx.x == 0) top_block(T0);
  else top_block(T1);
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
  float pad;          // mode 1: the thickness; mode 3: CCD padding
  float cap;          // mode 1: the buffer
  float thickness;
  int exclude_same_body;
  // modes 0, 2: query slot k is leaf k of the same tree; each unordered pair
  // is then taken by the query at the lower leaf position only
  int self_query;
  // hierarchy (every level in the node arrays, leaves from slot 0)
  int n_leaves;
  LevelInfo info;
  const float4* node_lo;  // xyz box, w body id bits (see body_min / body_max)
  const float4* node_hi;
  const int4* leaf_vid;
  // output
  int4* rec;
  int2* ids;
  int capacity;
  int* count;
};

// The query's own vertices (q0, q1; q1 unused for a point) and positions
// at x0 (p0, p1; DCD and crossing modes) are loaded once per query. The
// same-body test was applied to the leaf before (leaf_body is the body of
// the leaf's first vertex, qbody that of the query's first vertex).
struct QueryPrim {
  int i, q0, q1;
  V3<float> p0, p1;
};

// rec and ids receive the pair's record and primitive ids. An EE pair is
// written as the reference orders it, the lower edge id first (its distance
// is evaluated in that argument order too), whichever of the two is the
// query.
template <int MODE>
__device__ __forceinline__ bool accept_pair(const QueryArgs& A, const QueryPrim& Q, int4 lv, int4& rec, int2& ids) {
  ids = make_int2(Q.i, lv.w);
  if (MODE == Q_ISECT) {
    // edge i against triangle j: a crossing, not sharing a vertex
    int a0 = Q.q0, a1 = Q.q1, t0 = lv.x, t1 = lv.y, t2 = lv.z;
    if (a0 == t0 || a0 == t1 || a0 == t2 || a1 == t0 || a1 == t1 || a1 == t2) return false;
    rec = make_int4(a0, a1, t0, t1);
    return segment_hits_triangle(to_d(Q.p0), to_d(Q.p1), to_d(ldx(A.x0, t0)), to_d(ldx(A.x0, t1)),
                                 to_d(ldx(A.x0, t2)));
  }
  if (MODE == 0 || MODE == 2) {
    const int j = lv.w;
    const bool qlow = Q.i < j;  // the query edge has the lower id
    if (!A.self_query && !qlow) return false;
    int a0 = qlow ? Q.q0 : lv.x, a1 = qlow ? Q.q1 : lv.y, b0 = qlow ? lv.x : Q.q0, b1 = qlow ? lv.y : Q.q1;
    if (a0 == b0 || a0 == b1 || a1 == b0 || a1 == b1) return false;
    rec = make_int4(a0, a1, b0, b1);
    ids = qlow ? make_int2(Q.i, j) : make_int2(j, Q.i);
    if (MODE == 2) return true;
    V3<float> l0 = ldx(A.x0, lv.x), l1 = ldx(A.x0, lv.y);
    EdgeClosest ec = closest_point_edge_edge(qlow ? Q.p0 : l0, qlow ? Q.p1 : l1, qlow ? l0 : Q.p0, qlow ? l1 : Q.p1,
                                             1.0e-6f);
    return ec.dist < A.thickness * 2.0f;
  }
  int i = Q.i, t0 = lv.x, t1 = lv.y, t2 = lv.z;
  if (t0 == i || t1 == i || t2 == i) return false;
  if (MODE == 3) {
    rec = make_int4(i, t0, t1, t2);
    return true;
  }
  rec = make_int4(i, t1, t0, t2);
  TriClosest tc = closest_point_triangle(Q.p0, ldx(A.x0, t0), ldx(A.x0, t1), ldx(A.x0, t2));
  return tc.dist < A.thickness * 2.0f;
}

// Modes 0 and 1 test the distance of a candidate away from the traversal:
// pair_candidate applies the id filters and orients the record exactly as
// accept_pair does, and pair_distance_test evaluates accept_pair's distance
// test from the record's vertex ids (the same arguments in the same order,
// so the same result).
template <int MODE>
__device__ __forceinline__ bool pair_candidate(const QueryArgs& A, const QueryPrim& Q, int4 lv, int4& rec, int2& ids) {
  ids = make_int2(Q.i, lv.w);
  if (MODE == 0) {
    const int j = lv.w;
    const bool qlow = Q.i < j;
    if (!A.self_query && !qlow) return false;
    int a0 = qlow ? Q.q0 : lv.x, a1 = qlow ? Q.q1 : lv.y, b0 = qlow ? lv.x : Q.q0, b1 = qlow ? lv.y : Q.q1;
    if (a0 == b0 || a0 == b1 || a1 == b0 || a1 == b1) return false;
    rec = make_int4(a0, a1, b0, b1);
    ids = qlow ? make_int2(Q.i, j) : make_int2(j, Q.i);
    return true;
  }
  int i = Q.i, t0 = lv.x, t1 = lv.y, t2 = lv.z;
  if (t0 == i || t1 == i || t2 == i) return false;
  rec = make_int4(i, t1, t0, t2);
  return true;
}

template <int MODE>
__device__ __forceinline__ bool pair_distance_test(const QueryArgs& A, int4 rec) {
  if (MODE == 0) {
    EdgeClosest ec = closest_point_edge_edge(ldx(A.x0, rec.x), ldx(A.x0, rec.y), ldx(A.x0, rec.z), ldx(A.x0, rec.w),
                                             1.0e-6f);
    return ec.dist < A.thickness * 2.0f;
  }
  TriClosest tc = closest_point_triangle(ldx(A.x0, rec.x), ldx(A.x0, rec.z), ldx(A.x0, rec.y), ldx(A.x0, rec.w));
  return tc.dist < A.thickness * 2.0f;
}

constexpr int kCandQueue = 128;  // candidate pairs a wave holds for its dense distance tests
// A DCD launch stages fewer accepted pairs than the other queries, so that
// its LDS (stack, stage and candidate queues) stays within 20 KiB: eight
// workgroups of four waves per CU, the most a CU runs of them.
constexpr int kDcdBlockOut = 192;

// Eight lanes share one query. A stack entry (level L, mask, cbase) stands
// for the level-L nodes picked by the mask bits c of one parent; node c's
// children sit at slots cbase + 8 c + 0..7 of level L - 1 (cbase is an
// absolute slot index into the node arrays). A step takes up to kPopNodes
// nodes of the top entry; lane r tests child r of each, with every load of
// the step issued up front (padded levels keep them in bounds), and the
// group's ballot of overlapping children of a node becomes one new entry.
// Every lane of a group computes and writes the same stack values, so each
// lane only ever reads back its own LDS writes.
// Q = kCandQueue for launches with a DCD mode (each wave's candidate pairs
// awaiting the distance test), 1 otherwise (no LDS spent on it).
template <int Q>
struct QuerySmemT {
  unsigned long long stack[kGroupsPerBlock][kMaskStack];
  int level_off[kMaxLevels], level_cnt[kMaxLevels];
  BlockStageN<Q == 1 ? kBlockOut : kDcdBlockOut> stage;
  int4 cand_rec[kQueryBlock / 64][Q];
  int2 cand_ids[kQueryBlock / 64][Q];
};
using QuerySmem = QuerySmemT<1>;
template <int MA, int MB>
using QuerySmemFor = QuerySmemT<(MA == 0 || MA == 1 || MB == 0 || MB == 1) ? kCandQueue : 1>;

// Workgroup `bid` of a query launch over A. With wcost, each wave stores
// its number of steps in wcost[wave].
template <int MODE, typename Smem>
__device__ __forceinline__ void query_block(const QueryArgs& A, int bid, Smem& sm, int* wcost) {
  auto& stack = sm.stack;
  int* level_off = sm.level_off;
  int* level_cnt = sm.level_cnt;
  auto& stage = sm.stage;
  const StageOut out{A.rec, A.ids, A.count, A.capacity};
  int tid = threadIdx.x;
  if (tid < kMaxLevels) {
    level_off[tid] = A.info.offset[tid];
    level_cnt[tid] = A.info.count[tid];
  }
  stage_begin(stage);
  int grp = tid / kFan, r = tid % kFan;
  int lane = __lane_id();
  int group_shift = lane & ~(kFan - 1);  // first lane of this group in the wave
  int slot = bid * kGroupsPerBlock + grp;
  bool live = slot < A.n_query && A.n_leaves > 0;
  const bool self = (MODE == 0 || MODE == 2) && A.self_query;
  QueryPrim Q;
  Q.i = Q.q0 = Q.q1 = 0;
  Q.p0 = Q.p1 = mk(0.f, 0.f, 0.f);
  float qlo[3] = {0.f, 0.f, 0.f}, qhi[3] = {-1.f, -1.f, -1.f};
  int qbody = -1;
  if (live) {
    int i = A.order ? A.order[slot] : slot;
    Q.i = i;
    if (MODE == Q_ISECT) {
      Q.q0 = A.edges[2 * i];
      Q.q1 = A.edges[2 * i + 1];
      Q.p0 = ldx(A.x0, Q.q0);
      Q.p1 = ldx(A.x0, Q.q1);
      V3<float> lo = vmin(Q.p0, Q.p1), hi = vmax(Q.p0, Q.p1);
      qlo[0] = lo.x; qlo[1] = lo.y; qlo[2] = lo.z;
      qhi[0] = hi.x; qhi[1] = hi.y; qhi[2] = hi.z;
    } else if (MODE == 0 || MODE == 2) {
      float4 l = A.qlo[slot], h = A.qhi[slot];
      qlo[0] = l.x; qlo[1] = l.y; qlo[2] = l.z;
      qhi[0] = h.x; qhi[1] = h.y; qhi[2] = h.z;
      Q.q0 = A.edges[2 * i];
      Q.q1 = A.edges[2 * i + 1];
      if (MODE == 0) {
        Q.p0 = ldx(A.x0, Q.q0);
        Q.p1 = ldx(A.x0, Q.q1);
      }
      if (A.exclude_same_body) qbody = A.body[Q.q0];
    } else {
      Q.q0 = i;
      V3<float> p = ldx(A.x0, i), lo = p, hi = p;
      Q.p0 = p;
      if (MODE == 3) {
        V3<float> q = ldx(A.x1, i);
        lo = vmin(p, q);
        hi = vmax(p, q);
      }
      // DCD: 2 (thickness + min(buffer, sigma)); the reference's 2 (thickness
      // + buffer) where sigma >= buffer
      const float pe = MODE == 1 ? dcd_pad(A.pad, A.cap, abs_max(p, p)) * 2.0f : A.pad;
      V3<float> pv = mk(pe, pe, pe);
      lo = lo - pv;
      hi = hi + pv;
      qlo[0] = lo.x; qlo[1] = lo.y; qlo[2] = lo.z;
      qhi[0] = hi.x; qhi[1] = hi.y; qhi[2] = hi.z;
      if (A.exclude_same_body) qbody = A.body[i];
    }
  }
  int sp = 0;
  if (live) {
    stack[grp][0] = mask_entry(A.info.top, 1u, level_off[A.info.top - 1]);  // the root
    sp = 1;
  }
  int steps = 0;
  constexpr bool kDense = (MODE == 0 || MODE == 1) && sizeof(sm.cand_rec[0]) >= sizeof(int4) * 64;
  int4* cq_rec = sm.cand_rec[tid >> 6];
  int2* cq_ids = sm.cand_ids[tid >> 6];
  int qn = 0;  // queued candidates (wave-uniform)
  while (true) {
    ++steps;
    bool active = sp > 0;
    int L = 0, cbase = 0;
    unsigned take = 0u, rest = 0u;
    if (active) {
      unsigned long long e = stack[grp][sp - 1];
      L = int(e >> 40) & 0xf;
      unsigned mask = unsigned(e >> 32) & 0xffu;
      cbase = int(e & 0xffffffffu);
      rest = mask;
#pragma unroll
      for (int p = 0; p < kPopNodes; ++p) rest &= rest - 1u;
      take = mask ^ rest;  // the kPopNodes lowest nodes
      --sp;                // the rest goes back below, or to an idle group
    }
    // Work sharing: a group whose popped entry has nodes left hands them, with
    // its query, to a group of the same wave that has run out of work (the
    // k-th idle group takes the k-th such entry) and starts on them in the
    // next step. A wave otherwise runs as long as its slowest query while
    // the groups that finished wait. Every (query, node) is still visited
    // exactly once.
    bool gave = false;
    {
      constexpr unsigned long long kLead = 0x0101010101010101ull;  // the first lane of each group
      const unsigned long long idle_b = __ballot(!active) & kLead;
      const unsigned long long give_b = __ballot(rest != 0u) & kLead;
      // (at least two idle groups: on a launch that keeps the VALUs busy a
      // hand-off to a single idle group costs about the step it saves)
      if (__popcll(idle_b) >= 2 && give_b != 0ull) {
        const unsigned long long below = (1ull << group_shift) - 1ull;
        const int pairs = min(__popcll(idle_b), __popcll(give_b));
        const int idle_rank = __popcll(idle_b & below), give_rank = __popcll(give_b & below);
        const bool adopt = !active && idle_rank < pairs;
        gave = rest != 0u && give_rank < pairs;
        unsigned long long m = give_b;
        for (int j = 0; j < idle_rank && j < kFan; ++j) m &= m - 1ull;
        const int src = (adopt ? __ffsll(static_cast<long long>(m)) - 1 : group_shift) + r;
        const unsigned long long ent = mask_entry(L, rest, cbase);
        const unsigned ent_lo = __shfl(unsigned(ent & 0xffffffffu), src, 64);
        const unsigned ent_hi = __shfl(unsigned(ent >> 32), src, 64);
        float s_lo[3], s_hi[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          s_lo[a] = __shfl(qlo[a], src, 64);
          s_hi[a] = __shfl(qhi[a], src, 64);
        }
        const int s_body = __shfl(qbody, src, 64), s_slot = __shfl(slot, src, 64);
        const int s_i = __shfl(Q.i, src, 64), s_q0 = __shfl(Q.q0, src, 64), s_q1 = __shfl(Q.q1, src, 64);
        const V3<float> s_p0 = mk(__shfl(Q.p0.x, src, 64), __shfl(Q.p0.y, src, 64), __shfl(Q.p0.z, src, 64));
        const V3<float> s_p1 = mk(__shfl(Q.p1.x, src, 64), __shfl(Q.p1.y, src, 64), __shfl(Q.p1.z, src, 64));
        if (adopt) {
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            qlo[a] = s_lo[a];
            qhi[a] = s_hi[a];
          }
          qbody = s_body;
          slot = s_slot;
          Q.i = s_i;
          Q.q0 = s_q0;
          Q.q1 = s_q1;
          Q.p0 = s_p0;
          Q.p1 = s_p1;
          stack[grp][0] = (static_cast<unsigned long long>(ent_hi) << 32) | ent_lo;
          sp = 1;
        }
      }
    }
    if (rest != 0u && !gave) {
      stack[grp][sp] = mask_entry(L, rest, cbase);
      ++sp;
    }
    int cpos[kPopNodes], at[kPopNodes];
    {
      unsigned tk = take;
#pragma unroll
      for (int p = 0; p < kPopNodes; ++p) {
        cpos[p] = tk ? __ffs(tk) - 1 : -1;
        at[p] = tk ? cbase + cpos[p] * kFan + r : 0;
        tk &= tk - 1u;
      }
    }
    // The leaf record is fetched by lanes at the leaf level only: an
    // internal step moves the two box halves, 32 B per lane.
    float4 clo[kPopNodes], chi[kPopNodes];
    int4 cvid[kPopNodes];
#pragma unroll
    for (int p = 0; p < kPopNodes; ++p) {
      clo[p] = A.node_lo[at[p]];
      chi[p] = A.node_hi[at[p]];
      cvid[p] = make_int4(0, 0, 0, 0);
      if (L == 1) cvid[p] = A.leaf_vid[at[p]];
    }
    // Only real children count: a query box with NaN or infinite bounds
    // also "overlaps" the empty boxes of padding slots.
    const bool internal = L >= 2;
    const int child_off = level_off[L >= 1 ? L - 1 : 0], child_cnt = level_cnt[L >= 1 ? L - 1 : 0];
    const int grand_off = level_off[internal ? L - 2 : 0];
    // Self-queries skip children whose leaves all sit at positions <= the
    // query's own (one past the child's last leaf is (rel + 1) 8^(L-1)).
    // With at most 2^27 leaves (build_bvhs refuses more) a tree has at most
    // 10 levels above them and (rel + 1) 8^(L-1) stays below 2^31, so the
    // test runs in 32-bit unsigned arithmetic (a lane without a node has a
    // meaningless rel; its hit is false through cpos anyway).
    const unsigned shift = L >= 1 ? 3u * unsigned(L - 1) : 0u;
    auto beyond = [&](int rel) { return !self || (unsigned(rel + 1) << shift) > unsigned(slot) + 1u; };
    // (bitwise, so that no load is sunk behind a branch on another term)
    bool hit[kPopNodes];
#pragma unroll
    for (int p = 0; p < kPopNodes; ++p)
      hit[p] = (cpos[p] >= 0) & (at[p] - child_off < child_cnt) & overlaps(qlo, qhi, clo[p], chi[p]) &
               !((qbody >= 0) & (body_min(clo[p]) == qbody) & (body_max(chi[p]) == qbody)) & beyond(at[p] - child_off);
    // (A branch on exclude_same_body around the body test made the screw and
    // nut's launches, which exclude same-body pairs, 15 % slower.)
    // internal level: one entry per node with overlapping children, the
    // first node's pushed last so that it is visited first
#pragma unroll
    for (int p = kPopNodes - 1; p >= 0; --p) {
      unsigned gmask = unsigned(__ballot(internal && hit[p]) >> group_shift) & 0xffu;
      if (internal && gmask != 0u && sp < kMaskStack) {
        int rel = cbase - child_off + cpos[p] * kFan;  // first child's index within level L - 1
        stack[grp][sp] = mask_entry(L - 1, gmask, grand_off + rel * kFan);
        ++sp;
      }
    }
    // leaf level: test and stage the pairs. The DCD modes queue the
    // candidates that pass the id filters and test their distances 64 at a
    // time with every lane of the wave busy, instead of inside the step
    // where only the lanes with a leaf hit work.
#pragma unroll
    for (int p = 0; p < kPopNodes; ++p) {
      int4 rec = make_int4(0, 0, 0, 0);
      int2 ids = make_int2(Q.i, cvid[p].w);
      if (kDense) {
        const bool cand = L == 1 && hit[p] && pair_candidate<MODE>(A, Q, cvid[p], rec, ids);
        const unsigned long long cm = __ballot(cand);
        if (cand) {
          const int at_q = qn + __popcll(cm & ((1ull << lane) - 1ull));
          cq_rec[at_q] = rec;
          cq_ids[at_q] = ids;
        }
        qn += __popcll(cm);
        if (qn >= 64) {
          __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
          const int k = qn - 64 + lane;
          const int4 crec = cq_rec[k];
          const int2 cids = cq_ids[k];
          __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
          qn -= 64;
          stage_append(stage, out, pair_distance_test<MODE>(A, crec), crec, cids);
        }
      } else {
        bool want = false;
        if (L == 1 && hit[p]) want = accept_pair<MODE>(A, Q, cvid[p], rec, ids);
        stage_append(stage, out, want, rec, ids);
      }
    }
    if (!__any(sp > 0)) break;
  }
  if (kDense) {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    const bool mine = lane < qn;
    int4 crec = make_int4(0, 0, 0, 0);
    int2 cids = make_int2(0, 0);
    if (mine) {
      crec = cq_rec[lane];
      cids = cq_ids[lane];
    }
    stage_append(stage, out, mine && pair_distance_test<MODE>(A, crec), crec, cids);
  }
  if (wcost && lane == 0) wcost[tid >> 6] = steps;
  stage_flush(stage, out);
}

template <int MODE>
__global__ void __launch_bounds__(kQueryBlock) query_kernel(QueryArgs A) {
  __shared__ QuerySmemFor<MODE, MODE> sm;
  query_block<MODE>(A, xcd_local(blockIdx.x, gridDim.x), sm, nullptr);
}

// Two queries in one launch: logical workgroups [0, blocks_a) serve A, the
// rest B. With perm (a permutation of the launch's workgroups, see
// QueryOrder), raw workgroup b serves logical workgroup perm[b] and records
// its waves' step counts in cost; without, each half is spread over the
// XCDs in contiguous ranges.
template <int MA, int MB>
__global__ void __launch_bounds__(kQueryBlock) query_pair_kernel(QueryArgs A, QueryArgs B, int blocks_a,
                                                                 const int* __restrict__ perm, int* cost) {
  __shared__ QuerySmemFor<MA, MB> sm;
  const int n = gridDim.x, raw = blockIdx.x;
  int c = raw < blocks_a ? xcd_local(raw, blocks_a) : blocks_a + xcd_local(raw - blocks_a, n - blocks_a);
  if (perm) {
    const int pc = perm[raw];
    if (pc >= 0 && pc < n) c = pc;
  }
  int* wcost = cost ? cost + (kQueryBlock / 64) * c : nullptr;
  if (c < blocks_a) query_block<MA>(A, c, sm, wcost);
  else query_block<MB>(B, c - blocks_a, sm, wcost);
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
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
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
  keys.release(); leaf_order.release();
  leaf_id.release(); node_lo.release(); node_hi.release(); leaf_vid.release();
  temp.release(); bounds.release();
  n = 0;
  builds_since_sort = -1;
  vid_prims = nullptr;
  vid_body = nullptr;
  reuse_vid = false;
}

#define IPC_TRY(expr)                    \
  do {                                   \
    hipError_t _e = (expr);              \
    if (_e != hipSuccess) return _e;     \
  } while (0)

namespace {

// Level sizes of a tree of n leaves: count[0] = n, count[L] = ceil(count[L-1]
// / 8) up to a single root; level L occupies slots[L] = 8 count[L+1] slots
// from offset[L] (the root level one slot). Returns the total slot count.
long long plan_levels(LevelInfo& I, int n) {
  for (int L = 0; L < kMaxLevels; ++L) { I.count[L] = 0; I.slots[L] = 0; I.offset[L] = 0; }
  I.count[0] = n;
  int L = 0;
  do {
    ++L;
    I.count[L] = (I.count[L - 1] + kFan - 1) / kFan;
  } while (I.count[L] > 1);
  I.top = L;
  long long total = 0;
  for (int l = 0; l <= I.top; ++l) {
    I.slots[l] = l < I.top ? kFan * I.count[l + 1] : 1;
    I.offset[l] = static_cast<int>(total);
    total += I.slots[l];
  }
  return total;
}

// Buffers of one job and, when due, the Morton re-sort of its leaf order.
hipError_t prepare_job(const BuildJob& j, const int* body, const float* x0, hipStream_t s) {
  Bvh& B = *j.bvh;
  const int n = j.n;
  long long total = plan_levels(B.info, n);
  IPC_TRY(B.bounds.ensure(sizeof(int) * kBoundInts));
  IPC_TRY(B.keys.ensure(sizeof(unsigned) * n));
  IPC_TRY(B.leaf_order.ensure(sizeof(int) * n));
  IPC_TRY(B.leaf_id.ensure(sizeof(int) * n));
  IPC_TRY(B.leaf_vid.ensure(sizeof(int4) * B.info.slots[0]));
  IPC_TRY(B.node_lo.ensure(sizeof(float4) * total));
  IPC_TRY(B.node_hi.ensure(sizeof(float4) * total));
  bool resort = B.builds_since_sort < 0 || B.builds_since_sort >= kResortEvery || B.n_sorted != n;
  B.reuse_vid = !resort && B.vid_prims == j.prims && B.vid_body == body;
  B.vid_prims = j.prims;
  B.vid_body = body;
  if (resort) {
    CellBufs cb;
    IPC_TRY(cell_buffers(B.temp, n, cb));
    hipLaunchKernelGGL(init_bounds_kernel, dim3(kInitBlocks), dim3(256), 0, s, B.bounds.as<int>(), cb.count);
    hipLaunchKernelGGL(centre_bounds_kernel, dim3(nblocks(n, 256)), dim3(256), 0, s, j.mode, n, j.prims, x0, j.x1, j.dx,
                       j.pad, B.bounds.as<int>());
    hipLaunchKernelGGL(morton_kernel, dim3(nblocks(n, 256)), dim3(256), 0, s, j.mode, n, j.prims, x0, j.x1, j.dx, j.pad,
                       B.bounds.as<int>(), B.keys.as<unsigned>(), cb.count);
    IPC_TRY(hipGetLastError());
    IPC_TRY(sort_codes(cb, B.keys.as<unsigned>(), n, B.leaf_order.as<int>(), s));
    B.builds_since_sort = 0;
    B.n_sorted = n;
  } else {
    ++B.builds_since_sort;
  }
  return hipSuccess;
}

LeafJob leaf_job(const BuildJob& j) {
  Bvh& B = *j.bvh;
  LeafJob L{};
  L.mode = j.mode;
  L.n = j.n;
  L.prims = j.prims;
  L.x1 = j.x1;
  L.dx = j.dx;
  L.pad = j.pad;
  L.cap = j.cap;
  L.ids = B.leaf_order.as<int>();
  L.leaf_id = B.leaf_id.as<int>();
  L.leaf_vid = B.leaf_vid.as<int4>();
  L.node_lo = B.node_lo.as<float4>();
  L.node_hi = B.node_hi.as<float4>();
  L.info = B.info;
  L.reuse_vid = B.reuse_vid ? 1 : 0;
  return L;
}

TopJob top_job(const Bvh& B) {
  return TopJob{B.node_lo.as<float4>(), B.node_hi.as<float4>(), B.info};
}

}  // namespace

hipError_t build_bvhs(const BuildJob* jobs, int n_jobs, const int* body, const float* x0, const BuildPrologue& pro,
                      hipStream_t s, const QueryOrder* order) {
  if (n_jobs < 0 || n_jobs > 2 || pro.n_zero < 0 || pro.n_zero > 4 || pro.n_range < 0 || pro.n_range > 64)
    return hipErrorInvalidValue;
  const BuildJob* live[2] = {nullptr, nullptr};
  int n_live = 0;
  for (int k = 0; k < n_jobs; ++k) {
    jobs[k].bvh->n = jobs[k].n;
    if (jobs[k].n <= 0) continue;
    if (jobs[k].n > (1 << 27)) return hipErrorInvalidValue;
    IPC_TRY(prepare_job(jobs[k], body, x0, s));
    live[n_live++] = &jobs[k];
  }
  // end positions x0 - dx, written by the leaf launch (or on their own when
  // no tree is built)
  const BuildJob* ej = nullptr;
  for (int k = 0; k < n_jobs; ++k)
    if (jobs[k].dx && jobs[k].x1 && jobs[k].n_end > 0) ej = &jobs[k];
  if (n_live == 0) {
    if (pro.n_zero > 0 || pro.n_range > 0 || pro.toi) hipLaunchKernelGGL(prologue_kernel, dim3(1), dim3(64), 0, s, pro);
    if (ej) IPC_TRY(launch_sub(3 * ej->n_end, x0, ej->dx, const_cast<float*>(ej->x1), s));
    return hipGetLastError();
  }
  const EndWrite end = ej ? EndWrite{const_cast<float*>(ej->x1), ej->dx, 3 * ej->n_end} : EndWrite{nullptr, nullptr, 0};
  LeafJob j0 = leaf_job(*live[0]);
  LeafJob j1 = n_live > 1 ? leaf_job(*live[1]) : j0;
  int blocks0 = nblocks(live[0]->n, kLeafBlock);
  int blocks1 = n_live > 1 ? nblocks(live[1]->n, kLeafBlock) : 0;
  hipLaunchKernelGGL(leaf_levels_kernel, dim3(blocks0 + blocks1), dim3(kLeafBlock), 0, s, j0, j1, blocks0, body, x0,
                     pro, end);
  bool top0 = live[0]->bvh->info.top > kInBlockLevels;
  bool top1 = n_live > 1 && live[1]->bvh->info.top > kInBlockLevels;
  const bool ord = order && order->cost && order->perm && order->n_a > 0 && order->n_b > 0;
  if (top0 || top1 || ord) {
    TopJob t0 = top_job(*live[0]->bvh);
    TopJob t1 = n_live > 1 ? top_job(*live[1]->bvh) : t0;
    hipLaunchKernelGGL(reduce_top_kernel, dim3(n_live + (ord ? kXcds : 0)), dim3(1024), 0, s, t0, t1, n_live,
                       ord ? *order : QueryOrder{});
  }
  return hipGetLastError();
}

void PointOrder::release() {
  keys.release(); order.release(); temp.release(); bounds.release();
  n = 0;
  age = -1;
}

hipError_t point_order(PointOrder& P, int n, const float* x, hipStream_t s, const int** order) {
  *order = nullptr;
  if (n <= 0) return hipSuccess;
  if (P.age >= 0 && P.age < kResortEvery && P.n == n) {
    ++P.age;
    *order = P.order.as<int>();
    return hipSuccess;
  }
  IPC_TRY(P.bounds.ensure(sizeof(int) * kBoundInts));
  IPC_TRY(P.keys.ensure(sizeof(unsigned) * n));
  IPC_TRY(P.order.ensure(sizeof(int) * n));
  CellBufs cb;
  IPC_TRY(cell_buffers(P.temp, n, cb));
  hipLaunchKernelGGL(init_bounds_kernel, dim3(kInitBlocks), dim3(256), 0, s, P.bounds.as<int>(), cb.count);
  hipLaunchKernelGGL(centre_bounds_kernel, dim3(nblocks(n, 256)), dim3(256), 0, s, int(BOX_POINT), n,
                     static_cast<const int*>(nullptr), x, static_cast<const float*>(nullptr),
                     static_cast<const float*>(nullptr), 0.0f, P.bounds.as<int>());
  hipLaunchKernelGGL(morton_kernel, dim3(nblocks(n, 256)), dim3(256), 0, s, int(BOX_POINT), n,
                     static_cast<const int*>(nullptr), x, static_cast<const float*>(nullptr),
                     static_cast<const float*>(nullptr), 0.0f, P.bounds.as<int>(), P.keys.as<unsigned>(), cb.count);
  IPC_TRY(hipGetLastError());
  IPC_TRY(sort_codes(cb, P.keys.as<unsigned>(), n, P.order.as<int>(), s));
  P.n = n;
  P.age = 0;
  *order = P.order.as<int>();
  return hipSuccess;
}

hipError_t build_bvh(Bvh& B, int mode, int n, const int* prims, const int* body, const float* x0, const float* x1,
                     float pad, hipStream_t s) {
  BuildJob j{&B, mode, n, prims, x1, pad};
  return build_bvhs(&j, 1, body, x0, BuildPrologue{}, s);
}

namespace {

QueryArgs query_args(const QueryLaunch& q) {
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
  A.cap = q.cap;
  A.thickness = q.thickness;
  A.exclude_same_body = q.exclude_same_body;
  A.self_query = q.self_query;
  A.n_leaves = q.bvh->n;
  A.info = q.bvh->info;
  A.node_lo = q.bvh->node_lo.as<float4>();
  A.node_hi = q.bvh->node_hi.as<float4>();
  A.leaf_vid = q.bvh->leaf_vid.as<int4>();
  A.rec = q.rec;
  A.ids = q.ids;
  A.capacity = q.capacity;
  A.count = q.count;
  return A;
}

}  // namespace

int query_blocks(int n_query) { return n_query > 0 ? nblocks(n_query, kGroupsPerBlock) : 0; }

hipError_t run_queries(const QueryLaunch* q, int n_q, hipStream_t s, const QueryOrder* order) {
  const QueryLaunch* live[2] = {nullptr, nullptr};
  int n_live = 0;
  for (int k = 0; k < n_q && k < 2; ++k)
    if (q[k].n_query > 0 && q[k].bvh->n > 0) live[n_live++] = &q[k];
  if (n_live == 0) return hipSuccess;
  dim3 block(kQueryBlock);
  if (n_live == 1) {
    QueryArgs A = query_args(*live[0]);
    dim3 grid(nblocks(live[0]->n_query, kGroupsPerBlock));
    switch (live[0]->mode) {
      case Q_DCD_EE: hipLaunchKernelGGL(query_kernel<Q_DCD_EE>, grid, block, 0, s, A); break;
      case Q_DCD_PT: hipLaunchKernelGGL(query_kernel<Q_DCD_PT>, grid, block, 0, s, A); break;
      case Q_CCD_EE: hipLaunchKernelGGL(query_kernel<Q_CCD_EE>, grid, block, 0, s, A); break;
      case Q_CCD_PT: hipLaunchKernelGGL(query_kernel<Q_CCD_PT>, grid, block, 0, s, A); break;
      case Q_ISECT: hipLaunchKernelGGL(query_kernel<Q_ISECT>, grid, block, 0, s, A); break;
      default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
  // the two halves of a detection or CCD call: EE then PT
  QueryArgs A = query_args(*live[0]), B = query_args(*live[1]);
  int blocks_a = nblocks(live[0]->n_query, kGroupsPerBlock), blocks_b = nblocks(live[1]->n_query, kGroupsPerBlock);
  dim3 grid(blocks_a + blocks_b);
  const bool use = order && order->cost && order->perm && order->n_a == blocks_a && order->n_b == blocks_b;
  const int* perm = use ? order->perm : nullptr;
  int* cost = use ? order->cost : nullptr;
  if (live[0]->mode == Q_DCD_EE && live[1]->mode == Q_DCD_PT)
    hipLaunchKernelGGL((query_pair_kernel<Q_DCD_EE, Q_DCD_PT>), grid, block, 0, s, A, B, blocks_a, perm, cost);
  else if (live[0]->mode == Q_CCD_EE && live[1]->mode == Q_CCD_PT)
    hipLaunchKernelGGL((query_pair_kernel<Q_CCD_EE, Q_CCD_PT>), grid, block, 0, s, A, B, blocks_a, perm, cost);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t run_query(const QueryLaunch& q, hipStream_t s) { return run_queries(&q, 1, s); }

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
  if (n <= 0) return hipSuccess;
  const int blocks = n <= 64 ? 1 : (nblocks(n, 256) < 1024 ? nblocks(n, 256) : 1024);
  hipLaunchKernelGGL(zero_i32_kernel, dim3(blocks), dim3(n <= 64 ? 64 : 256), 0, s, p, n);
  return hipGetLastError();
}

}  // namespace ipcd

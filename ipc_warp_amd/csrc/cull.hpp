// Internal declarations of the culling layer (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

namespace ipcd {

// Grow-only device allocation.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  hipError_t ensure(size_t b);
  void release();
  template <typename T> T* as() const { return static_cast<T*>(p); }
};

constexpr int kFan = 8;        // children per hierarchy node
constexpr int kMaxLevels = 12;  // 8^10 > 2^27 primitives

// Sizes of the implicit 8-ary hierarchy: level 0 holds the n leaves in
// Morton order, level L holds count[L] = ceil(count[L-1] / 8) nodes, the
// root is node 0 of level `top`. Every level, the leaves included, is stored
// in the same node arrays from offset[L] (offset[0] = 0) and padded to
// slots[L] = 8 count[L+1] entries (slots[top] = 1) with empty boxes, so the
// eight children of any node are in bounds.
struct LevelInfo {
  int top;
  int count[kMaxLevels];
  int slots[kMaxLevels];
  int offset[kMaxLevels];
};

// BVH over n primitives (see cull.hip).
struct Bvh {
  int n = 0;
  LevelInfo info{};
  DevBuf keys;                // n uint32 Morton codes of the last re-sort
  DevBuf leaf_order;          // primitive id per leaf slot from the last re-sort (Morton cell order)
  DevBuf node_lo, node_hi;    // float4 per slot of every level (leaves first): xyz the box,
                              // w the bits of the (min, max) body id below; (b, b) for a leaf
  DevBuf leaf_id;             // primitive id per leaf
  DevBuf leaf_vid;            // int4 per leaf slot: vertex ids (v2 = -1 for an edge), primitive id
  DevBuf temp;                // re-sort workspace: cell counts, cursors, scan
  DevBuf bounds;              // ordered ints: copies of (min xyz, max xyz)
  int builds_since_sort = -1; // -1: no leaf order yet
  int n_sorted = 0;
  // leaf_vid and the leaves' body ids were written for this leaf order from
  // these topology and body arrays (a build that keeps the order reads them
  // instead of following keys -> topology -> body)
  const int* vid_prims = nullptr;
  const int* vid_body = nullptr;
  bool reuse_vid = false;     // this build reads them
  void release();
};

// Leaf order is recomputed (by Morton cell) on the first build and then every
// kResortEvery builds; in between, builds reuse it and only recompute boxes.
constexpr int kResortEvery = 16;

enum BoxMode { BOX_EDGE = 0, BOX_TRI = 1, BOX_SWEPT_EDGE = 2, BOX_SWEPT_TRI = 3, BOX_POINT = 4 };

// Morton order of the points of a soup, the query order of the point
// queries (a wave's eight queries then sit close together and walk the same
// subtrees). Recomputed on the first call and every kResortEvery calls; the
// order only affects speed, never the pairs found.
struct PointOrder {
  DevBuf keys, order, temp, bounds;
  int n = 0, age = -1;
  void release();
};

// Refreshes P from positions x when due; *order receives slot -> point.
hipError_t point_order(PointOrder& P, int n, const float* x, hipStream_t s, const int** order);
enum QueryMode { Q_DCD_EE = 0, Q_DCD_PT = 1, Q_CCD_EE = 2, Q_CCD_PT = 3, Q_ISECT = 4 };

// One tree of a build launch: n primitives `prims` (2 or 3 vertex ids each
// by mode) boxed at x0 (and x1 for the swept modes), +- pad.
struct BuildJob {
  Bvh* bvh;
  int mode;  // BoxMode
  int n;
  const int* prims;
  const float* x1;
  float pad;
  // With dx, the swept modes take the end positions as x0 - dx, and job 0
  // of the launch also writes them to x1 (3 n_end floats) for the launches
  // that follow.
  const float* dx = nullptr;
  int n_end = 0;
  float cap = 0.f;  // BOX_EDGE: the buffer (pad is the thickness)
};

// Device writes folded into the first workgroup of a call's first launch,
// for the queries and TOI kernels that follow on the same stream:
// *zero[k] = 0 for k < n_zero, range[0 .. n_range) = 0, and *toi = toi_init
// when toi is set.
struct BuildPrologue {
  int* zero[4] = {nullptr, nullptr, nullptr, nullptr};
  int n_zero = 0;
  int* range = nullptr;
  int n_range = 0;  // <= 64
  unsigned long long* toi = nullptr;
  unsigned long long toi_init = 0;
};

// Applied by the first 64 threads of the launch's workgroup 0.
__device__ __forceinline__ void run_prologue(const BuildPrologue& p) {
  if (blockIdx.x != 0) return;
  int t = threadIdx.x;
  if (t < p.n_zero) *p.zero[t] = 0;
  if (t < p.n_range) p.range[t] = 0;
  if (t == 0 && p.toi) *p.toi = p.toi_init;
}

// Dispatch order of a paired query launch (n_a workgroups of EE queries,
// then n_b of PT queries). The launch records each wave's step count in
// cost[4 c + w] (c the logical workgroup); the next build's top-level launch
// turns those counts into perm[raw workgroup] = logical workgroup, ordering
// each XCD's share of the workgroups by decreasing cost. A query launch's
// length is set by its slowest waves: started first, they overlap the rest.
// Any permutation returns the same pairs.
struct QueryOrder {
  int* cost = nullptr;  // 4 (n_a + n_b) ints, zero before the first launch
  int* perm = nullptr;  // n_a + n_b ints
  int n_a = 0, n_b = 0;
};

// Builds up to two trees with one leaf launch and one top-level launch
// (plus a Morton re-sort for a tree whose leaf order is due). With `order`,
// the top-level launch also computes order->perm from order->cost.
hipError_t build_bvhs(const BuildJob* jobs, int n_jobs, const int* body, const float* x0, const BuildPrologue& pro,
                      hipStream_t s, const QueryOrder* order = nullptr);
hipError_t build_bvh(Bvh& B, int mode, int n, const int* prims, const int* body, const float* x0, const float* x1,
                     float pad, hipStream_t s);

struct QueryLaunch {
  int mode;
  int n_query;
  const int* edges;
  const int* tris;
  const int* body;
  const float* x0;
  const float* x1;
  const float4* qlo;  // modes 0, 2: query box per slot (leaf boxes of the same tree)
  const float4* qhi;
  const int* order;   // optional slot -> query primitive (Morton order of the queries)
  float pad;          // DCD PT: the thickness; CCD PT: the padding
  float cap = 0.f;    // DCD PT: the buffer
  float thickness;
  int exclude_same_body;
  int self_query;  // EE: slot k is leaf k of bvh (order = its leaf_id, qlo/qhi = its leaf boxes)
  const Bvh* bvh;
  int4* rec;
  int2* ids;
  int capacity;
  int* count;
};

hipError_t run_query(const QueryLaunch& q, hipStream_t s);
// Up to two queries in one launch (the EE and PT halves of a DCD or CCD
// call: modes Q_DCD_EE + Q_DCD_PT or Q_CCD_EE + Q_CCD_PT, in that order).
// `order` (from the same call's build_bvhs) applies when both halves run
// and its block counts match the launch (query_blocks of each n_query).
hipError_t run_queries(const QueryLaunch* q, int n_q, hipStream_t s, const QueryOrder* order = nullptr);
// Workgroups of a world query launch over n queries.
int query_blocks(int n_query);
hipError_t launch_sub(int n3, const float* x, const float* dx, float* out, hipStream_t s);
hipError_t launch_fill_u64(unsigned long long* p, int n, unsigned long long v, hipStream_t s);
hipError_t launch_zero_i32(int* p, int n, hipStream_t s);
// Min and max of an int array into out_dev[0], out_dev[1] (device).
hipError_t index_range(const int* a, long long n, int* out_dev, hipStream_t s);

}  // namespace ipcd

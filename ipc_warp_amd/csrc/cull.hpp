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
  DevBuf keys, keys_sorted;   // uint64 Morton | id
  DevBuf node_lo, node_hi;    // float4 per slot of every level (leaves first)
  DevBuf node_body;           // int2 (min, max) body id per slot; (b, b) for a leaf
  DevBuf leaf_id;             // primitive id per leaf
  DevBuf leaf_vid;            // int4 per leaf slot: vertex ids (v2 = -1 for an edge), primitive id
  DevBuf temp;                // radix sort workspace
  DevBuf bounds;              // 6 ordered ints
  int builds_since_sort = -1; // -1: no leaf order yet
  int n_sorted = 0;
  void release();
};

// Leaf order is recomputed (Morton sort) on the first build and then every
// kResortEvery builds; in between, builds reuse it and only recompute boxes.
constexpr int kResortEvery = 16;

enum BoxMode { BOX_EDGE = 0, BOX_TRI = 1, BOX_SWEPT_EDGE = 2, BOX_SWEPT_TRI = 3 };
enum QueryMode { Q_DCD_EE = 0, Q_DCD_PT = 1, Q_CCD_EE = 2, Q_CCD_PT = 3, Q_ISECT = 4 };

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
  float pad;
  float thickness;
  int exclude_same_body;
  const Bvh* bvh;
  int4* rec;
  int2* ids;
  int capacity;
  int* count;
};

hipError_t run_query(const QueryLaunch& q, hipStream_t s);
hipError_t launch_sub(int n3, const float* x, const float* dx, float* out, hipStream_t s);
hipError_t launch_fill_u64(unsigned long long* p, int n, unsigned long long v, hipStream_t s);
hipError_t launch_zero_i32(int* p, int n, hipStream_t s);
// Min and max of an int array into out_dev[0], out_dev[1] (device).
hipError_t index_range(const int* a, long long n, int* out_dev, hipStream_t s);

}  // namespace ipcd

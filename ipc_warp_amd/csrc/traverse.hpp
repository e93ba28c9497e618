// Device helpers shared by the BVH queries (cull.hip) and the per-body
// forest queries (forest.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "cull.hpp"

namespace ipcd {

constexpr int kQueryBlock = 256;
constexpr int kGroupsPerBlock = kQueryBlock / kFan;  // one query per 8 lanes
constexpr int kPopNodes = 1;  // sibling nodes a query group takes from its stack per step
// A step pops from the top entry and pushes at most kPopNodes entries one
// level down, so a level never holds more than kPopNodes entries; n <= 2^27
// primitives gives at most 9 levels above the leaves.
constexpr int kMaskStack = kPopNodes * 9 + 1;
constexpr int kBlockOut = 512;  // accepted pairs a query block stages in LDS

// Workgroups are dispatched to the 8 XCDs round-robin (workgroup w runs on
// XCD w mod 8), each XCD with its own L2. xcd_local renumbers the `nblk`
// workgroups of a launch so that each XCD serves one contiguous range: with
// Morton-ordered queries, each L2 then holds its region's share of a tree's
// lower levels instead of the whole tree. A bijection of [0, nblk).
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_local(int bid, int nblk) {
  const int q = nblk / kXcds, r = nblk % kXcds;
  const int x = bid % kXcds, k = bid / kXcds;
  return x * q + (x < r ? x : r) + k;
}

// Inclusive box test (a lower bound above the other's upper bound
// separates). The terms are combined bitwise: with short-circuit operators
// the compiler splits the test into branches and defers each component's
// load behind the previous compare, one memory round trip per component.
__device__ __forceinline__ bool overlaps(const float qlo[3], const float qhi[3], float4 lo, float4 hi) {
  return !((qlo[0] > hi.x) | (qlo[1] > hi.y) | (qlo[2] > hi.z) | (lo.x > qhi[0]) | (lo.y > qhi[1]) |
           (lo.z > qhi[2]));
}

// Stack entry: level (bits 40..43), node mask (bits 32..39), and an index
// (bits 0..31) whose meaning each query kernel states.
__device__ __forceinline__ unsigned long long mask_entry(int level, unsigned mask, int first) {
  return (static_cast<unsigned long long>(level) << 40) | (static_cast<unsigned long long>(mask & 0xffu) << 32) |
         static_cast<unsigned>(first);
}

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

// Accepted pairs staged in LDS per workgroup, N of them. Appends from every
// wave go to an LDS counter; the workgroup reserves its output range with
// one global atomicAdd at the end (a single global counter hit by every wave
// serialises). Pairs past N go straight to the output, one global atomicAdd
// per wave.
template <int N>
struct BlockStageN {
  static constexpr int kCap = N;
  int4 rec[N];
  int2 ids[N];
  int n, base;
};
using BlockStage = BlockStageN<kBlockOut>;

struct StageOut {
  int4* rec;  // may be null (count only)
  int2* ids;  // may be null
  int* count;
  int capacity;
};

// Called by the whole workgroup before the first stage_append.
template <typename Stage>
__device__ __forceinline__ void stage_begin(Stage& S) {
  if (threadIdx.x == 0) S.n = 0;
  __syncthreads();
}

// Called by every lane of a wave together (ballots inside).
template <typename Stage>
__device__ __forceinline__ void stage_append(Stage& S, const StageOut& O, bool want, int4 rec, int2 id) {
  int s = wave_append(want, &S.n);
  bool spill = want && s >= Stage::kCap;
  int g = wave_append(spill, O.count);
  if (want && !spill) {
    S.rec[s] = rec;
    S.ids[s] = id;
  } else if (spill && g < O.capacity) {
    if (O.rec) O.rec[g] = rec;
    if (O.ids) O.ids[g] = id;
  }
}

// Called by the whole workgroup once every wave has finished appending.
template <typename Stage>
__device__ __forceinline__ void stage_flush(Stage& S, const StageOut& O) {
  __syncthreads();
  int staged = S.n < Stage::kCap ? S.n : Stage::kCap;
  if (threadIdx.x == 0) S.base = staged ? atomicAdd(O.count, staged) : 0;
  __syncthreads();
  for (int k = threadIdx.x; k < staged; k += blockDim.x) {
    int out = S.base + k;
    if (out < O.capacity) {
      if (O.rec) O.rec[out] = S.rec[k];
      if (O.ids) O.ids[out] = S.ids[k];
    }
  }
}

}  // namespace ipcd

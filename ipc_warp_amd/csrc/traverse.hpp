// Device helpers shared by the BVH queries (cull.hip) and the per-body
// forest queries (forest.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "cull.hpp"

namespace ipcd {

constexpr int kQueryBlock = 256;
constexpr int kGroupsPerBlock = kQueryBlock / kFan;  // one query per 8 lanes
constexpr int kMaskStack = 12;                       // > kMaxLevels - 1 entries per query
constexpr int kBlockOut = 512;                       // accepted pairs a query block stages in LDS

// Inclusive box test (a lower bound above the other's upper bound separates).
__device__ __forceinline__ bool overlaps(const float qlo[3], const float qhi[3], float4 lo, float4 hi) {
  return !(qlo[0] > hi.x || qlo[1] > hi.y || qlo[2] > hi.z || lo.x > qhi[0] || lo.y > qhi[1] || lo.z > qhi[2]);
}

// Stack entry: level (bits 40..43), child mask (bits 32..39), first child
// index at that level (bits 0..31).
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

}  // namespace ipcd

// Per-vertex input features of the learned Newton warm start (gfx950).
//
// features[v] (48 floats) = base[v] (12 floats, supplied by the caller)
// followed by the vertex's 6 nearest contact records, nearest first, each as
//   (partner - self) / dhat   (3)  closest point on the other primitive minus
//                                  the closest point on the vertex's own one
//   dist / dhat               (1)
//   w                         (1)  the vertex's weight in its own closest
//                                  point (1 for the point of a PT record)
//   is_pt                     (1)
// Unused slots are zero. There is no counterpart in the reference.
//
// Selection is a lock-free sorted insert per vertex: every (record, vertex)
// pair carries a 64-bit key (fp32 distance bits << 32 | record id) down the
// vertex's 6 slots with atomicMin, keeping the smaller key in each slot and
// carrying the larger one on. Each slot ends with the minimum of the keys that
// reached it, so the slots hold the 6 smallest keys in order. Record ids are
// 4 * index + vertex position (EE records first, then PT), so the result does
// not depend on the order in which lanes run.
#include "ipc_device.hpp"
#include "ipc_kernels.hpp"

namespace ipcd {

constexpr int kWsBase = 12, kWsSlots = 6, kWsSlotDim = 6, kWsDim = kWsBase + kWsSlots * kWsSlotDim;
static_assert(kWsDim == 48, "warm-start feature width");
constexpr unsigned long long kEmpty = ~0ull;

__device__ __forceinline__ bool ws_valid4(int4 c, int nv) {
  return (unsigned)c.x < (unsigned)nv && (unsigned)c.y < (unsigned)nv && (unsigned)c.z < (unsigned)nv &&
         (unsigned)c.w < (unsigned)nv;
}

__global__ void __launch_bounds__(256) ws_clear_kernel(unsigned long long* __restrict__ slots, long long n) {
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
    slots[i] = kEmpty;
}

__device__ __forceinline__ void ws_insert(unsigned long long* __restrict__ s, unsigned long long key) {
#pragma unroll
  for (int k = 0; k < kWsSlots; ++k) {
    unsigned long long old = atomicMin(s + k, key);
    if (old == kEmpty) return;  // the slot was free: placed
    if (old > key) key = old;   // displaced a farther record: carry it on
  }
}

template <bool EE>
__global__ void __launch_bounds__(256) ws_rank_kernel(const float* __restrict__ x, int nv, const int4* __restrict__ c4,
                                                      int n, unsigned rec_base, unsigned long long* __restrict__ slots) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  int4 c = c4[i];
  if (!ws_valid4(c, nv)) return;
  V3<float> x0 = ldx(x, c.x), x1 = ldx(x, c.y), x2 = ldx(x, c.z), x3 = ldx(x, c.w);
  float dist = EE ? closest_point_edge_edge(x0, x1, x2, x3, 1.0e-6f).dist : closest_point_triangle(x0, x1, x2, x3).dist;
  if (!(dist >= 0.0f)) return;  // NaN coordinates: no record
  unsigned long long hi = (unsigned long long)__float_as_uint(dist) << 32;
  unsigned r = rec_base + 4u * unsigned(i);
  ws_insert(slots + 6ll * c.x, hi | r);
  ws_insert(slots + 6ll * c.y, hi | (r + 1u));
  ws_insert(slots + 6ll * c.z, hi | (r + 2u));
  ws_insert(slots + 6ll * c.w, hi | (r + 3u));
}

__global__ void __launch_bounds__(256) ws_write_kernel(const float* __restrict__ x, int nv,
                                                       const float* __restrict__ base, const int4* __restrict__ ee,
                                                       int n_ee, const int4* __restrict__ pt, float inv_dhat,
                                                       const unsigned long long* __restrict__ slots,
                                                       float* __restrict__ out) {
  int v = blockIdx.x * 256 + threadIdx.x;
  if (v >= nv) return;
  float* o = out + (long long)kWsDim * v;
#pragma unroll
  for (int q = 0; q < kWsBase; ++q) o[q] = base ? base[(long long)kWsBase * v + q] : 0.0f;
  unsigned ee_end = 4u * unsigned(n_ee);
  for (int s = 0; s < kWsSlots; ++s) {
    unsigned long long key = slots[6ll * v + s];
    float f[kWsSlotDim] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    if (key != kEmpty) {
      unsigned r = unsigned(key);
      bool is_ee = r < ee_end;
      unsigned rr = is_ee ? r : r - ee_end;
      int i = int(rr >> 2), k = int(rr & 3u);
      int4 c = is_ee ? ee[i] : pt[i];
      V3<float> x0 = ldx(x, c.x), x1 = ldx(x, c.y), x2 = ldx(x, c.z), x3 = ldx(x, c.w);
      V3<float> self, partner;
      float w, dist;
      if (is_ee) {
        EdgeClosest ec = closest_point_edge_edge(x0, x1, x2, x3, 1.0e-6f);
        V3<float> pa = x0 * (1.0f - ec.s) + x1 * ec.s;
        V3<float> pb = x2 * (1.0f - ec.t) + x3 * ec.t;
        self = k < 2 ? pa : pb;
        partner = k < 2 ? pb : pa;
        w = k == 0 ? 1.0f - ec.s : (k == 1 ? ec.s : (k == 2 ? 1.0f - ec.t : ec.t));
        dist = ec.dist;
      } else {
        TriClosest tc = closest_point_triangle(x0, x1, x2, x3);
        float wc = 1.0f - tc.wa - tc.wb;
        V3<float> q = tc.wa * x1 + tc.wb * x2 + wc * x3;
        self = k == 0 ? x0 : q;
        partner = k == 0 ? q : x0;
        w = k == 0 ? 1.0f : (k == 1 ? tc.wa : (k == 2 ? tc.wb : wc));
        dist = tc.dist;
      }
      V3<float> off = partner - self;
      f[0] = off.x * inv_dhat;
      f[1] = off.y * inv_dhat;
      f[2] = off.z * inv_dhat;
      f[3] = dist * inv_dhat;
      f[4] = w;
      f[5] = is_ee ? 0.0f : 1.0f;
    }
#pragma unroll
    for (int q = 0; q < kWsSlotDim; ++q) o[kWsBase + kWsSlotDim * s + q] = f[q];
  }
}

hipError_t warmstart_features(const float* x, int nv, const float* base, const int* ee, int n_ee, const int* pt,
                              int n_pt, float dhat, unsigned long long* slots, float* out, hipStream_t s) {
  if (nv <= 0) return hipSuccess;
  long long ns = 6ll * nv;
  int gc = int((ns + 255) / 256);
  if (gc > 4096) gc = 4096;
  hipLaunchKernelGGL(ws_clear_kernel, dim3(gc), dim3(256), 0, s, slots, ns);
  const int4* e4 = reinterpret_cast<const int4*>(ee);
  const int4* p4 = reinterpret_cast<const int4*>(pt);
  if (n_ee > 0)
    hipLaunchKernelGGL(ws_rank_kernel<true>, dim3((n_ee + 255) / 256), dim3(256), 0, s, x, nv, e4, n_ee, 0u, slots);
  if (n_pt > 0)
    hipLaunchKernelGGL(ws_rank_kernel<false>, dim3((n_pt + 255) / 256), dim3(256), 0, s, x, nv, p4, n_pt,
                       4u * unsigned(n_ee), slots);
  hipLaunchKernelGGL(ws_write_kernel, dim3((nv + 255) / 256), dim3(256), 0, s, x, nv, base, e4, n_ee, p4, 1.0f / dhat,
                     slots, out);
  return hipGetLastError();
}

}  // namespace ipcd

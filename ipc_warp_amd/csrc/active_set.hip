// The augmented-Lagrangian active set (augmented_lagrangian.py:42-197,
// 506-547): constraints kept in a hash table of `table_size` buckets keyed
// by e0e1, stored contiguously by bucket (cell_start[b] .. cell_start[b+1])
// so that every later pass over the set is a plain sweep over its entries.
//
//   vertex TOI  per vertex, the earliest alpha of the detected pairs on it
//               (an atomic minimum on the IEEE bits);
//   create      the detected pairs that are some vertex's earliest (up to
//               1e-12) become fresh entries (lam 0, gamma 1, mollifier 1);
//   merge       old entries whose gamma is still >= 0.01, plus the new ones
//               whose key the old table lacks, into a new table.
//
// create and merge share one pipeline: count (one lane per input, an
// atomic per accepted input on its bucket), exclusive scan of the counts
// (cell_start), scatter (each accepted input takes a slot in its bucket),
// gather (each slot finds its input's final place by ranking it within the
// bucket, then the workgroup copies its 256 records word by word, so the
// stores of a wave are contiguous). Within a bucket the entries are ordered
// by decreasing input index — what the reference's slot-handing atomics give
// when its threads run in index order — so the result does not depend on
// the order the atomics hand out slots.
#include <cstdio>

#include "../../include/ipc_warp_amd.h"
#include "cull.hpp"
#include "ipc_device.hpp"
#include "ipc_kernels.hpp"
#include "scan.hpp"

using namespace ipcd;

struct ipc_active_set {
  DevBuf cnt;     // int [table_size + 1]: accepted inputs per bucket
  DevBuf bucket;  // int [inputs]: bucket of each input, -1 when not accepted
  DevBuf src;     // int [inputs]: input index of each slot, unordered within a bucket
  DevBuf part;    // scan partial sums
};

namespace {

constexpr int kB = 256;

int as_fail(int code, const char* what) {
  set_error(what);
  return code;
}

int as_hip(hipError_t e, const char* where) {
  if (e == hipSuccess) return IPC_OK;
  char buf[512];
  std::snprintf(buf, sizeof(buf), "%s: %s", where, hipGetErrorString(e));
  set_error(buf);
  return -int(e);
}

inline int nblk(long long n, int per) { return int((n + per - 1) / per); }

// Record layouts in 4-byte words, `scalar` S = float or double:
//   XConstraint  a1a2b1b2 0-3, e0e1 4-5, then l0, alpha, lam (9 or 12 words);
//   ALConstraint a1a2b1b2 0-3, e0e1 4-5, then l0, alpha, lam, gamma, slack,
//                mollifier, the int distance_type (and 4 bytes of padding
//                with doubles), d0, grad[12] (26 or 46 words).
template <typename S>
struct Layout {
  static constexpr int kSw = sizeof(S) / 4;                 // words per scalar
  static constexpr int kXWords = 6 + 3 * kSw;               // XConstraint
  static constexpr int kAlphaWord = 6 + kSw;                // alpha in both records
  static constexpr int kGammaWord = 6 + 3 * kSw;            // gamma in ALConstraint
  static constexpr int kMollifierWord = 6 + 5 * kSw;
  static constexpr int kTypeWord = 6 + 6 * kSw;
  static constexpr int kD0Word = kSw == 1 ? 13 : 20;
  static constexpr int kAWords = kD0Word + 13 * kSw;        // ALConstraint
  static constexpr int kCopied = 6 + 2 * kSw;               // a1a2b1b2, e0e1, l0, alpha
  static constexpr unsigned kOneHi = kSw == 1 ? 0x3F800000u : 0x3FF00000u;  // high word of 1.0
};
static_assert(Layout<float>::kAWords * 4 == IPC_ALCONSTRAINT_BYTES_F32, "fp32 ALConstraint");
static_assert(Layout<double>::kAWords * 4 == IPC_ALCONSTRAINT_BYTES_F64, "fp64 ALConstraint");
static_assert(Layout<float>::kXWords * 4 == IPC_XCONSTRAINT_BYTES_F32, "fp32 XConstraint");
static_assert(Layout<double>::kXWords * 4 == IPC_XCONSTRAINT_BYTES_F64, "fp64 XConstraint");

template <typename S>
__device__ inline S load_scalar(const unsigned* rec, int word) {
  return *reinterpret_cast<const S*>(rec + word);
}

// hash_pos (augmented_lagrangian.py:68-73) in int32 arithmetic: the two
// products wrap, the XOR's magnitude is taken, then the remainder. The one
// value whose magnitude int32 cannot hold (the XOR equal to INT_MIN) is
// taken as 2^31 here.
__device__ inline int hash_pos(int e0, int e1, int table_size) {
  const unsigned v = (unsigned(e0) * 73856093u) ^ (unsigned(e1) * 19349663u);
  const unsigned mag = int(v) < 0 ? 0u - v : v;
  return int(mag % unsigned(table_size));
}

// atomic minimum of a float / double on its IEEE bits: non-negative values
// order as signed integers, negative ones reversed as unsigned integers.
__device__ inline void atomic_min_scalar(float* p, float v) {
  const int bits = __float_as_int(v);
  if (bits >= 0)
    atomicMin(reinterpret_cast<int*>(p), bits);
  else
    atomicMax(reinterpret_cast<unsigned*>(p), unsigned(bits));
}

__device__ inline void atomic_min_scalar(double* p, double v) {
  const long long bits = __double_as_longlong(v);
  if (bits >= 0)
    atomicMin(reinterpret_cast<long long*>(p), bits);
  else
    atomicMax(reinterpret_cast<unsigned long long*>(p), static_cast<unsigned long long>(bits));
}

// accumulate_vertex_toi (augmented_lagrangian.py:105-113); vertices out of
// range are skipped.
template <typename S>
__global__ __launch_bounds__(kB) void vertex_toi_kernel(const unsigned* rec, int n, S* vertex_toi, int nv) {
  using L = Layout<S>;
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  const unsigned* r = rec + static_cast<long long>(i) * L::kXWords;
  const S alpha = load_scalar<S>(r, L::kAlphaWord);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int v = int(r[k]);
    if (unsigned(v) < unsigned(nv)) atomic_min_scalar(vertex_toi + v, alpha);
  }
}

// count_constraints (augmented_lagrangian.py:116-134): filter_accept keeps
// a pair when its alpha is within 1e-12 of the earliest TOI of one of its
// vertices (compared in `scalar`).
template <typename S>
__global__ __launch_bounds__(kB) void count_create_kernel(const unsigned* rec, int n, const S* vertex_toi, int nv,
                                                          int table_size, int* cnt, int* bucket) {
  using L = Layout<S>;
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  const unsigned* r = rec + static_cast<long long>(i) * L::kXWords;
  const S alpha = load_scalar<S>(r, L::kAlphaWord);
  bool accept = false;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int v = int(r[k]);
    if (unsigned(v) < unsigned(nv)) accept |= alpha <= vertex_toi[v] + S(1.0e-12);
  }
  int b = -1;
  if (accept) {
    b = hash_pos(int(r[4]), int(r[5]), table_size);
    atomicAdd(cnt + b, 1);
  }
  bucket[i] = b;
}

// count_merge (augmented_lagrangian.py:152-179): inputs [0, n_old) are the
// old entries, kept while gamma >= 0.01 (keep_old); inputs [n_old, n_old +
// n_new) are the new entries, kept when the old table has no entry with
// their key (is_new_constraint, probing the old table's bucket; its range
// is clamped to the old entries).
template <typename S>
__global__ __launch_bounds__(kB) void count_merge_kernel(const unsigned* old_e, int n_old, const int* old_start,
                                                         int old_table_size, const unsigned* new_e, int n_new,
                                                         int table_size, int* cnt, int* bucket) {
  using L = Layout<S>;
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= n_old + n_new) return;
  bool accept;
  const unsigned* r;
  if (i < n_old) {
    r = old_e + static_cast<long long>(i) * L::kAWords;
    accept = load_scalar<S>(r, L::kGammaWord) >= S(0.01);
  } else {
    r = new_e + static_cast<long long>(i - n_old) * L::kAWords;
    accept = true;
    if (n_old > 0) {
      const int ob = hash_pos(int(r[4]), int(r[5]), old_table_size);
      const int lo = max(old_start[ob], 0), hi = min(old_start[ob + 1], n_old);
      for (int k = lo; k < hi; ++k) {
        const unsigned* o = old_e + static_cast<long long>(k) * L::kAWords;
        if (o[4] == r[4] && o[5] == r[5]) {
          accept = false;
          break;
        }
      }
    }
  }
  int b = -1;
  if (accept) {
    b = hash_pos(int(r[4]), int(r[5]), table_size);
    atomicAdd(cnt + b, 1);
  }
  bucket[i] = b;
}

// Each accepted input takes a slot in its bucket; the counts return to 0.
__global__ __launch_bounds__(kB) void scatter_kernel(const int* bucket, int m, const int* start, int* cnt, int* src) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= m) return;
  const int b = bucket[i];
  if (b < 0) return;
  const int k = atomicSub(cnt + b, 1) - 1;
  src[start[b] + k] = i;
}

// Slot s of [0, capacity): below the entry count, the input in slot s goes
// to start[b] + (inputs of its bucket with a larger index); the slots past
// the count are zeroed. The 256 records of a workgroup are then written one
// word per lane. CREATE builds fresh entries from XConstraint records
// (to_al_constraint, augmented_lagrangian.py:88-102); otherwise entries
// are copied whole from the old ([0, n_old)) or the new table.
template <typename S, bool CREATE>
__global__ __launch_bounds__(kB) void gather_kernel(const unsigned* in_a, int n_a, const unsigned* in_b,
                                                    const int* src, const int* bucket, const int* start,
                                                    int table_size, int capacity, unsigned* out) {
  using L = Layout<S>;
  __shared__ const unsigned* rec_src[kB];
  __shared__ int rec_dst[kB];
  const int total = start[table_size];
  const int base = blockIdx.x * kB;
  const int s = base + threadIdx.x;
  if (s < capacity) {
    const unsigned* p = nullptr;
    int dst = s;
    if (s < total) {
      const int i = src[s];
      const int b = bucket[i];
      const int lo = start[b], hi = start[b + 1];
      int rank = 0;
      for (int k = lo; k < hi; ++k) rank += src[k] > i;
      dst = lo + rank;
      if (CREATE)
        p = in_a + static_cast<long long>(i) * L::kXWords;
      else
        p = i < n_a ? in_a + static_cast<long long>(i) * L::kAWords
                    : in_b + static_cast<long long>(i - n_a) * L::kAWords;
    }
    rec_src[threadIdx.x] = p;
    rec_dst[threadIdx.x] = dst;
  }
  __syncthreads();
  const int nrec = min(kB, capacity - base);
  for (int f = threadIdx.x; f < nrec * L::kAWords; f += kB) {
    const int r = f / L::kAWords;
    const int w = f - r * L::kAWords;
    const unsigned* p = rec_src[r];
    unsigned v = 0u;
    if (p) {
      if (!CREATE || w < L::kCopied)
        v = p[w];
      else if (w == L::kGammaWord + L::kSw - 1 || w == L::kMollifierWord + L::kSw - 1)
        v = L::kOneHi;  // gamma = mollifier = 1; lam, slack, type, d0, grad = 0
    }
    out[static_cast<long long>(rec_dst[r]) * L::kAWords + w] = v;
  }
}

// constraint_value (augmented_lagrangian.py:76-85): c = d0 + sum over the
// four vertices of dot(grad_k, x_k), each dot x*x + y*y + z*z in that order.
// False when a vertex is out of range (the entry is then left unchanged).
template <typename S>
__device__ inline bool constraint_value(const unsigned* e, const S* x, int nv, S& c) {
  using L = Layout<S>;
  c = load_scalar<S>(e, L::kD0Word);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int v = int(e[k]);
    if (unsigned(v) >= unsigned(nv)) return false;
    const int g = L::kD0Word + L::kSw * (1 + 3 * k);
    const S d = load_scalar<S>(e, g) * x[3ll * v] + load_scalar<S>(e, g + L::kSw) * x[3ll * v + 1] +
                load_scalar<S>(e, g + 2 * L::kSw) * x[3ll * v + 2];
    c = c + d;
  }
  return true;
}

template <typename S>
__device__ inline void store_scalar(unsigned* rec, int word, S v) {
  *reinterpret_cast<S*>(rec + word) = v;
}

// update_slack_variables (augmented_lagrangian.py:326-331): slack =
// max(0, c(x_hat) - lam / mu).
template <typename S>
__global__ __launch_bounds__(kB) void update_slack_kernel(unsigned* entries, int n, const S* x, int nv, S mu) {
  using L = Layout<S>;
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  unsigned* e = entries + static_cast<long long>(i) * L::kAWords;
  S c;
  if (!constraint_value<S>(e, x, nv, c)) return;
  const S lam = load_scalar<S>(e, 6 + 2 * L::kSw);
  const S v = c - lam / mu;
  store_scalar<S>(e, 6 + 4 * L::kSw, v > S(0) ? v : S(0));  // NaN and -0 give +0
}

// update_active_set_forces (augmented_lagrangian.py:334-345): a constraint
// with zero slack updates its multiplier, lam -= mu c(x_hat), and resets
// gamma to 1; otherwise lam = 0 and gamma decays by 0.9.
template <typename S>
__global__ __launch_bounds__(kB) void update_forces_kernel(unsigned* entries, int n, const S* x, int nv, S mu) {
  using L = Layout<S>;
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  unsigned* e = entries + static_cast<long long>(i) * L::kAWords;
  S c;
  if (!constraint_value<S>(e, x, nv, c)) return;
  const S lam = load_scalar<S>(e, 6 + 2 * L::kSw), gamma = load_scalar<S>(e, L::kGammaWord);
  if (load_scalar<S>(e, 6 + 4 * L::kSw) == S(0)) {
    store_scalar<S>(e, 6 + 2 * L::kSw, lam - mu * c);
    store_scalar<S>(e, L::kGammaWord, S(1));
  } else {
    store_scalar<S>(e, 6 + 2 * L::kSw, S(0));
    store_scalar<S>(e, L::kGammaWord, gamma * S(0.9));
  }
}

// ---------------------------------------------------------- linearise ---
// point_line_distance_squared (augmented_lagrangian.py:199-204), in S.
template <typename S>
__device__ inline S point_line(V3<S> p, V3<S> e0, V3<S> e1, S& alpha) {
  const V3<S> edge = e1 - e0;
  const S ee = dot(edge, edge);
  alpha = dot(p - e0, edge) / (ee > S(1.0e-30) ? ee : S(1.0e-30));
  const V3<S> off = p - (e0 + alpha * edge);
  return dot(off, off);
}

// The dispatch of edge_edge_distance_gradient_hessian
// (dynamic_contacts.py:48-108) on fp32 closest-point parameters s, t, the
// comparisons in S; `parallel` is its near-parallel test.
template <typename S>
__device__ inline Feature ee_dispatch(S s, S t, bool parallel) {
  const S fe = lit<S>(1.0e-6f, 1.0e-6), one = S(1);
  const bool sb = s <= fe || s >= one - fe, tb = t <= fe || t >= one - fe;
  if (sb && tb) return {F_PP, s >= one - fe ? 1 : 0, t >= one - fe ? 3 : 2, -1, -1};
  if (sb) return {F_PE, s >= one - fe ? 1 : 0, 2, 3, -1};
  if (tb) return {F_PE, t >= one - fe ? 3 : 2, 0, 1, -1};
  if (parallel) {
    const S ds0 = s, ds1 = one - s, dt0 = t, dt1 = one - t;
    const S m = fmin(fmin(ds0, ds1), fmin(dt0, dt1));
    if (m == ds0) return {F_PE, 0, 2, 3, -1};
    if (m == ds1) return {F_PE, 1, 2, 3, -1};
    if (m == dt0) return {F_PE, 2, 0, 1, -1};
    return {F_PE, 3, 0, 1, -1};
  }
  return {F_EE, 0, 1, 2, 3};
}

// linearize_constraints (augmented_lagrangian.py:270-323): at x_safe, the
// constraint's feature (PT: closest_point_triangle's region; EE: the
// parallel point-edge choice of :208-267 where the mollifier is below the
// threshold, else edge_edge_distance_gradient_hessian's dispatch) gives the
// gradient of the squared distance; grad = that / (2 d) and d0 = d - l0 -
// grad . x_safe, so that c(x) = d0 + grad . x is the distance's first-order
// model minus l0. Classification in fp32 (the reference's vec3 casts),
// mollifier and point-line distances in S, the feature gradient in FP64.
template <typename S>
__global__ __launch_bounds__(kB) void linearize_kernel(unsigned* entries, int n, const S* __restrict__ xs,
                                                       const S* __restrict__ xrest, int nv, int is_pt, S pe_threshold) {
  using L = Layout<S>;
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  unsigned* e = entries + static_cast<long long>(i) * L::kAWords;
  int id[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    id[k] = int(e[k]);
    if (unsigned(id[k]) >= unsigned(nv)) return;
  }
  V3<S> x[4];
  V3<float> xf[4];
  V3<double> xr[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    x[k] = ldx(xs, id[k]);
    xf[k] = mk(float(x[k].x), float(x[k].y), float(x[k].z));
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) xr[k] = to_d(x[k]) - to_d(x[0]);
  Feature f;
  S d2;
  S mollifier = S(1);
  int type = 0;
  if (is_pt) {
    const TriClosestT<float> tc = closest_point_triangle(xf[0], xf[1], xf[2], xf[3]);
    f = pt_feature(tc.type);
    d2 = S(tc.dist) * S(tc.dist);
  } else {
    const V3<S> ra = ldx(xrest, id[1]) - ldx(xrest, id[0]), rb = ldx(xrest, id[3]) - ldx(xrest, id[2]);
    const S eps_x = lit<S>(1.0e-3f, 1.0e-3) * dot(ra, ra) * dot(rb, rb);
    const V3<S> cr = cross(x[1] - x[0], x[3] - x[2]);
    const S c = dot(cr, cr);
    if (c < eps_x) {
      const S q = c / eps_x;
      mollifier = q * (S(2) - q);
    }
    if (mollifier < pe_threshold) {
      S a[4], dist[4];
      dist[0] = point_line(x[0], x[2], x[3], a[0]);
      dist[1] = point_line(x[1], x[2], x[3], a[1]);
      dist[2] = point_line(x[2], x[0], x[1], a[2]);
      dist[3] = point_line(x[3], x[0], x[1], a[3]);
      const S large = S(1.0e30);
      S score[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) score[k] = (a[k] < S(0) || a[k] > S(1)) ? large : dist[k];
      if (fmin(fmin(score[0], score[1]), fmin(score[2], score[3])) == large) {
#pragma unroll
        for (int k = 0; k < 4; ++k) score[k] = dist[k];
      }
      int best = 0;
      S b = score[0];
#pragma unroll
      for (int k = 1; k < 4; ++k)
        if (score[k] < b) {
          b = score[k];
          best = k;
        }
      type = best + 1;
      d2 = best == 0 ? dist[0] : (best == 1 ? dist[1] : (best == 2 ? dist[2] : dist[3]));
      f = best == 0 ? Feature{F_PE, 0, 2, 3, -1}
                    : (best == 1 ? Feature{F_PE, 1, 2, 3, -1}
                                 : (best == 2 ? Feature{F_PE, 2, 0, 1, -1} : Feature{F_PE, 3, 0, 1, -1}));
    } else {
      const EdgeClosestT<float> ec = closest_point_edge_edge(xf[0], xf[1], xf[2], xf[3], 1.0e-6f);
      const bool parallel = eps_x > S(0) && c < eps_x;
      f = ee_dispatch<S>(S(ec.s), S(ec.t), parallel);
      d2 = S(ec.dist) * S(ec.dist);
    }
  }
  const FeatureGeom g = feature_geom(f, xr);
  const S dist = sqrt(d2 > S(1.0e-30) ? d2 : S(1.0e-30));
  const double inv = 1.0 / (2.0 * double(dist));
  S grad[12];
  S off = dist - load_scalar<S>(e, 6);
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const double s2 = 2.0 * g.w[k];
    grad[3 * k] = S(s2 * g.r.x * inv);
    grad[3 * k + 1] = S(s2 * g.r.y * inv);
    grad[3 * k + 2] = S(s2 * g.r.z * inv);
    off = off - (grad[3 * k] * x[k].x + grad[3 * k + 1] * x[k].y + grad[3 * k + 2] * x[k].z);
  }
  store_scalar<S>(e, L::kD0Word, off);
#pragma unroll
  for (int j = 0; j < 12; ++j) store_scalar<S>(e, L::kD0Word + L::kSw * (1 + j), grad[j]);
  store_scalar<S>(e, L::kMollifierWord, mollifier);
  e[L::kTypeWord] = unsigned(type);
}

int scalar_of(int bytes, int f32, int f64) { return bytes == f32 ? 4 : (bytes == f64 ? 8 : 0); }

inline bool aligned(const void* p, unsigned long long a) { return (reinterpret_cast<unsigned long long>(p) & (a - 1)) == 0; }

// Shared tail of create and merge: scan the counts into cell_start, hand out
// the slots, write the entries.
template <typename S, bool CREATE>
int finish(ipc_active_set* ws, int m, const unsigned* in_a, int n_a, const unsigned* in_b, int table_size,
           int* cell_start, void* entries, int capacity, hipStream_t s, const char* who) {
  int* cnt = ws->cnt.as<int>();
  int* bucket = ws->bucket.as<int>();
  int* src = ws->src.as<int>();
  hipError_t e = exclusive_scan(cnt, table_size + 1, cell_start, ws->part, s);
  if (e == hipSuccess && m > 0) {
    hipLaunchKernelGGL(scatter_kernel, dim3(nblk(m, kB)), dim3(kB), 0, s, bucket, m, cell_start, cnt, src);
    e = hipGetLastError();
  }
  if (e == hipSuccess && capacity > 0) {
    hipLaunchKernelGGL((gather_kernel<S, CREATE>), dim3(nblk(capacity, kB)), dim3(kB), 0, s, in_a, n_a, in_b, src,
                       bucket, cell_start, table_size, capacity, static_cast<unsigned*>(entries));
    e = hipGetLastError();
  }
  return as_hip(e, who);
}

hipError_t reserve(ipc_active_set* ws, int table_size, int m) {
  hipError_t e = ws->cnt.ensure(sizeof(int) * (size_t(table_size) + 1));
  if (e == hipSuccess) e = ws->bucket.ensure(sizeof(int) * size_t(m > 0 ? m : 1));
  if (e == hipSuccess) e = ws->src.ensure(sizeof(int) * size_t(m > 0 ? m : 1));
  return e;
}

}  // namespace

extern "C" {

int ipc_active_set_create(ipc_active_set** ws) {
  if (!ws) return as_fail(IPC_EINVAL, "active_set_create: null handle pointer");
  *ws = new ipc_active_set();
  return IPC_OK;
}

int ipc_active_set_destroy(ipc_active_set* ws) {
  if (!ws) return IPC_OK;
  ws->cnt.release();
  ws->bucket.release();
  ws->src.release();
  ws->part.release();
  delete ws;
  return IPC_OK;
}

int ipc_al_vertex_toi(const void* constraints, int record_bytes, int n, void* vertex_toi, int n_vertices,
                      void* stream) {
  const int sb = scalar_of(record_bytes, IPC_XCONSTRAINT_BYTES_F32, IPC_XCONSTRAINT_BYTES_F64);
  if (!sb) return as_fail(IPC_EINVAL, "al_vertex_toi: record_bytes must be 36 (fp32 scalars) or 48 (fp64 scalars)");
  if (n < 0 || n_vertices < 0) return as_fail(IPC_EINVAL, "al_vertex_toi: negative n or n_vertices");
  if (n == 0) return IPC_OK;
  if (!constraints || !vertex_toi) return as_fail(IPC_EINVAL, "al_vertex_toi: null pointer");
  if (!aligned(constraints, sb) || !aligned(vertex_toi, sb)) return as_fail(IPC_EINVAL, "al_vertex_toi: misaligned pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const unsigned* rec = static_cast<const unsigned*>(constraints);
  if (sb == 4)
    hipLaunchKernelGGL(vertex_toi_kernel<float>, dim3(nblk(n, kB)), dim3(kB), 0, s, rec, n,
                       static_cast<float*>(vertex_toi), n_vertices);
  else
    hipLaunchKernelGGL(vertex_toi_kernel<double>, dim3(nblk(n, kB)), dim3(kB), 0, s, rec, n,
                       static_cast<double*>(vertex_toi), n_vertices);
  return as_hip(hipGetLastError(), "al_vertex_toi");
}

int ipc_al_create_hash(ipc_active_set* ws, const void* constraints, int record_bytes, int n, const void* vertex_toi,
                       int n_vertices, int table_size, int* cell_start, void* entries, int capacity, void* stream) {
  const int sb = scalar_of(record_bytes, IPC_XCONSTRAINT_BYTES_F32, IPC_XCONSTRAINT_BYTES_F64);
  if (!ws) return as_fail(IPC_EINVAL, "al_create_hash: null workspace");
  if (!sb) return as_fail(IPC_EINVAL, "al_create_hash: record_bytes must be 36 (fp32 scalars) or 48 (fp64 scalars)");
  if (n < 0 || n_vertices < 0 || table_size < 1 || capacity < 0)
    return as_fail(IPC_EINVAL, "al_create_hash: negative n, n_vertices or capacity, or table_size < 1");
  if (capacity < n) return as_fail(IPC_EINVAL, "al_create_hash: capacity below the number of constraints");
  if (!cell_start || (capacity > 0 && !entries) || (n > 0 && (!constraints || !vertex_toi)))
    return as_fail(IPC_EINVAL, "al_create_hash: null pointer");
  if (!aligned(constraints, sb) || !aligned(vertex_toi, sb) || !aligned(entries, sb) || !aligned(cell_start, 4))
    return as_fail(IPC_EINVAL, "al_create_hash: misaligned pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipError_t e = reserve(ws, table_size, n);
  if (e == hipSuccess) e = fill_i32(ws->cnt.as<int>(), table_size + 1, 0, s);
  const unsigned* rec = static_cast<const unsigned*>(constraints);
  if (e == hipSuccess && n > 0) {
    if (sb == 4)
      hipLaunchKernelGGL(count_create_kernel<float>, dim3(nblk(n, kB)), dim3(kB), 0, s, rec, n,
                         static_cast<const float*>(vertex_toi), n_vertices, table_size, ws->cnt.as<int>(),
                         ws->bucket.as<int>());
    else
      hipLaunchKernelGGL(count_create_kernel<double>, dim3(nblk(n, kB)), dim3(kB), 0, s, rec, n,
                         static_cast<const double*>(vertex_toi), n_vertices, table_size, ws->cnt.as<int>(),
                         ws->bucket.as<int>());
    e = hipGetLastError();
  }
  if (e != hipSuccess) return as_hip(e, "al_create_hash");
  return sb == 4 ? finish<float, true>(ws, n, rec, n, nullptr, table_size, cell_start, entries, capacity, s,
                                       "al_create_hash")
                 : finish<double, true>(ws, n, rec, n, nullptr, table_size, cell_start, entries, capacity, s,
                                        "al_create_hash");
}

int ipc_al_merge(ipc_active_set* ws, int entry_bytes, const int* old_cell_start, const void* old_entries,
                 int old_table_size, int n_old, const void* new_entries, int n_new, int table_size, int* cell_start,
                 void* entries, int capacity, void* stream) {
  const int sb = scalar_of(entry_bytes, IPC_ALCONSTRAINT_BYTES_F32, IPC_ALCONSTRAINT_BYTES_F64);
  if (!ws) return as_fail(IPC_EINVAL, "al_merge: null workspace");
  if (!sb) return as_fail(IPC_EINVAL, "al_merge: entry_bytes must be 104 (fp32 scalars) or 184 (fp64 scalars)");
  if (n_old < 0 || n_new < 0 || table_size < 1 || capacity < 0 || (n_old > 0 && old_table_size < 1))
    return as_fail(IPC_EINVAL, "al_merge: negative count or capacity, or a table_size < 1");
  if ((long long)n_old + n_new > capacity)
    return as_fail(IPC_EINVAL, "al_merge: capacity below the old plus the new entries");
  if (!cell_start || (capacity > 0 && !entries) || (n_old > 0 && (!old_cell_start || !old_entries)) ||
      (n_new > 0 && !new_entries))
    return as_fail(IPC_EINVAL, "al_merge: null pointer");
  if (!aligned(old_entries, sb) || !aligned(new_entries, sb) || !aligned(entries, sb) || !aligned(cell_start, 4) ||
      !aligned(old_cell_start, 4))
    return as_fail(IPC_EINVAL, "al_merge: misaligned pointer");
  if ((entries && (entries == old_entries || entries == new_entries)) || cell_start == old_cell_start)
    return as_fail(IPC_EINVAL, "al_merge: the destination table must not be one of the inputs");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int m = n_old + n_new;
  hipError_t e = reserve(ws, table_size, m);
  if (e == hipSuccess) e = fill_i32(ws->cnt.as<int>(), table_size + 1, 0, s);
  const unsigned* oe = static_cast<const unsigned*>(old_entries);
  const unsigned* ne = static_cast<const unsigned*>(new_entries);
  if (e == hipSuccess && m > 0) {
    if (sb == 4)
      hipLaunchKernelGGL(count_merge_kernel<float>, dim3(nblk(m, kB)), dim3(kB), 0, s, oe, n_old, old_cell_start,
                         old_table_size, ne, n_new, table_size, ws->cnt.as<int>(), ws->bucket.as<int>());
    else
      hipLaunchKernelGGL(count_merge_kernel<double>, dim3(nblk(m, kB)), dim3(kB), 0, s, oe, n_old, old_cell_start,
                         old_table_size, ne, n_new, table_size, ws->cnt.as<int>(), ws->bucket.as<int>());
    e = hipGetLastError();
  }
  if (e != hipSuccess) return as_hip(e, "al_merge");
  return sb == 4 ? finish<float, false>(ws, m, oe, n_old, ne, table_size, cell_start, entries, capacity, s, "al_merge")
                 : finish<double, false>(ws, m, oe, n_old, ne, table_size, cell_start, entries, capacity, s,
                                         "al_merge");
}

static int update_entries(bool forces, void* entries, int entry_bytes, int n, const void* x_hat, int n_vertices,
                          double mu, void* stream) {
  const char* who = forces ? "al_update_forces" : "al_update_slack";
  const int sb = scalar_of(entry_bytes, IPC_ALCONSTRAINT_BYTES_F32, IPC_ALCONSTRAINT_BYTES_F64);
  if (!sb) return as_fail(IPC_EINVAL, forces ? "al_update_forces: entry_bytes must be 104 or 184"
                                             : "al_update_slack: entry_bytes must be 104 or 184");
  if (n < 0 || n_vertices < 0) return as_fail(IPC_EINVAL, forces ? "al_update_forces: negative size"
                                                                 : "al_update_slack: negative size");
  if (n == 0) return IPC_OK;
  if (!entries || !x_hat) return as_fail(IPC_EINVAL, forces ? "al_update_forces: null pointer"
                                                            : "al_update_slack: null pointer");
  if (!aligned(entries, sb) || !aligned(x_hat, sb))
    return as_fail(IPC_EINVAL, forces ? "al_update_forces: misaligned pointer" : "al_update_slack: misaligned pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  unsigned* e = static_cast<unsigned*>(entries);
  const dim3 g(nblk(n, kB));
  if (sb == 4) {
    const float* x = static_cast<const float*>(x_hat);
    if (forces)
      hipLaunchKernelGGL(update_forces_kernel<float>, g, dim3(kB), 0, s, e, n, x, n_vertices, float(mu));
    else
      hipLaunchKernelGGL(update_slack_kernel<float>, g, dim3(kB), 0, s, e, n, x, n_vertices, float(mu));
  } else {
    const double* x = static_cast<const double*>(x_hat);
    if (forces)
      hipLaunchKernelGGL(update_forces_kernel<double>, g, dim3(kB), 0, s, e, n, x, n_vertices, mu);
    else
      hipLaunchKernelGGL(update_slack_kernel<double>, g, dim3(kB), 0, s, e, n, x, n_vertices, mu);
  }
  return as_hip(hipGetLastError(), who);
}

int ipc_al_linearize(void* entries, int entry_bytes, int n, const void* x_safe, const void* x_rest, int n_vertices,
                     int is_point_triangle, double point_edge_threshold, void* stream) {
  const int sb = scalar_of(entry_bytes, IPC_ALCONSTRAINT_BYTES_F32, IPC_ALCONSTRAINT_BYTES_F64);
  if (!sb) return as_fail(IPC_EINVAL, "al_linearize: entry_bytes must be 104 or 184");
  if (n < 0 || n_vertices < 0) return as_fail(IPC_EINVAL, "al_linearize: negative size");
  if (n == 0) return IPC_OK;
  if (!entries || !x_safe || (!is_point_triangle && !x_rest)) return as_fail(IPC_EINVAL, "al_linearize: null pointer");
  if (!aligned(entries, sb) || !aligned(x_safe, sb) || !aligned(x_rest, sb))
    return as_fail(IPC_EINVAL, "al_linearize: misaligned pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  unsigned* e = static_cast<unsigned*>(entries);
  const dim3 g(nblk(n, kB));
  if (sb == 4)
    hipLaunchKernelGGL(linearize_kernel<float>, g, dim3(kB), 0, s, e, n, static_cast<const float*>(x_safe),
                       static_cast<const float*>(x_rest), n_vertices, is_point_triangle ? 1 : 0,
                       float(point_edge_threshold));
  else
    hipLaunchKernelGGL(linearize_kernel<double>, g, dim3(kB), 0, s, e, n, static_cast<const double*>(x_safe),
                       static_cast<const double*>(x_rest), n_vertices, is_point_triangle ? 1 : 0,
                       point_edge_threshold);
  return as_hip(hipGetLastError(), "al_linearize");
}

int ipc_al_update_slack(void* entries, int entry_bytes, int n, const void* x_hat, int n_vertices, double mu,
                        void* stream) {
  return update_entries(false, entries, entry_bytes, n, x_hat, n_vertices, mu, stream);
}

int ipc_al_update_forces(void* entries, int entry_bytes, int n, const void* x_hat, int n_vertices, double mu,
                         void* stream) {
  return update_entries(true, entries, entry_bytes, n, x_hat, n_vertices, mu, stream);
}

}  // extern "C"

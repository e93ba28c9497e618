// Fixed per-body BVHs for affine-body contact (fixed_bvh_contacts.py).
//
// Every body's edges and triangles get a tree once, in the body's rest frame
// (initialize_fixed_body_bvhs, fixed_bvh_contacts.py:511-597). The trees of
// all bodies live in one forest: leaves are body-contiguous and in Morton
// order of the rest frame; each body's implicit 8-ary hierarchy is stored
// level by level in shared node arrays and described by a TreeDesc. The
// forest is built on the host at set-up (rest geometry never changes) and
// uploaded.
//
// Per step, on the device:
//   1. body frames: A^-1, p and the padding scale |A^-1|_F in FP64, the
//      world AABB of the rest box through A (affine_body_aabbs, :47-67) or,
//      for CCD, the box over both ends of the step plus the target-motion
//      extent and the near-singular flag (affine_swept_body_aabbs,
//      _target_motion_extent, :121-173);
//   2. body pairs a < b whose boxes overlap with padding
//      (compact_overlapping_body_pairs, :83-91);
//   3. tasks (source primitive, target body): every edge of a pair's first
//      body, every vertex of either body against the other, kept only when
//      the primitive's padded world box reaches the target body's box (the
//      compaction of compact_edge_query_tasks / compact_point_query_tasks,
//      :389-419, with this extra prune, which is conservative because the
//      target's primitives lie inside its body box);
//   4. each task maps its primitive into the target's rest frame, pads the
//      box there (|A^-1|_F scaled, plus the motion extent for CCD) and walks
//      the target's tree with the eight-lane traversal of the world BVH;
//      acceptance stays in world space (batched_fixed_bvh_*_query /
//      batched_fixed_bvh_*_toi, :202-386).
// Local boxes are computed in FP64 and rounded outward to fp32, so the
// transform never makes the query less conservative than the padding says.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "forest.hpp"
#include "ipc_device.hpp"
#include "traverse.hpp"

namespace ipcd {

// ---------------------------------------------------------------- host build ----
namespace {

unsigned spread10(unsigned v) {
  v = (v * 0x00010001u) & 0xFF0000FFu;
  v = (v * 0x00000101u) & 0x0F00F00Fu;
  v = (v * 0x00000011u) & 0xC30C30C3u;
  v = (v * 0x00000005u) & 0x49249249u;
  return v;
}

unsigned morton3(const float c[3], const float lo[3], const float hi[3]) {
  unsigned q[3];
  for (int a = 0; a < 3; ++a) {
    float ext = hi[a] - lo[a];
    float t = ext > 0.0f ? (c[a] - lo[a]) / ext : 0.0f;
    t = std::min(std::max(t, 0.0f), 1.0f);
    q[a] = std::min(1023u, unsigned(t * 1024.0f));
  }
  return (spread10(q[0]) << 2) | (spread10(q[1]) << 1) | spread10(q[2]);
}

// Indices of `items` sorted by the Morton code of their centres (ties by id).
std::vector<int> morton_order(const std::vector<int>& items, const std::vector<float>& cen) {
  float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (size_t i = 0; i < items.size(); ++i)
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], cen[3 * i + a]);
      hi[a] = std::max(hi[a], cen[3 * i + a]);
    }
  std::vector<unsigned long long> key(items.size());
  for (size_t i = 0; i < items.size(); ++i)
    key[i] = (static_cast<unsigned long long>(morton3(&cen[3 * i], lo, hi)) << 32) | unsigned(items[i]);
  std::vector<int> idx(items.size());
  std::iota(idx.begin(), idx.end(), 0);
  std::sort(idx.begin(), idx.end(), [&](int a, int b) { return key[a] < key[b]; });
  return idx;
}

template <typename T>
hipError_t upload(DevBuf& b, const std::vector<T>& v) {
  hipError_t e = b.ensure(sizeof(T) * std::max<size_t>(v.size(), 1));
  if (e != hipSuccess || v.empty()) return e;
  return hipMemcpy(b.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice);
}

}  // namespace

void Forest::release() {
  desc.release();
  leaf_id.release();
  node_lo.release();
  node_hi.release();
  slot_vid.release();
  n_trees = n_leaves = n_nodes = max_leaves = 0;
}

hipError_t build_forest(Forest& F, int n_bodies, int n_prims, int k, const int* prims, const int* vbody,
                        const float* rest) {
  std::vector<std::vector<int>> per(n_bodies);
  for (int p = 0; p < n_prims; ++p) {
    int b = vbody[prims[k * p]];
    if (b >= 0 && b < n_bodies) per[b].push_back(p);
  }
  std::vector<TreeDesc> D(n_bodies);
  std::vector<float4> nlo, nhi;
  std::vector<int4> vid;
  std::vector<int> lid;
  const float4 empty_lo = make_float4(INFINITY, INFINITY, INFINITY, 0.f);
  const float4 empty_hi = make_float4(-INFINITY, -INFINITY, -INFINITY, 0.f);
  const int4 no_vid = make_int4(-1, -1, -1, -1);
  int max_leaves = 0;
  for (int b = 0; b < n_bodies; ++b) {
    const std::vector<int>& L = per[b];
    TreeDesc d{};
    d.leaf_base = int(lid.size());
    d.n_leaves = int(L.size());
    max_leaves = std::max(max_leaves, d.n_leaves);
    if (L.empty()) {
      D[b] = d;
      continue;
    }
    std::vector<float> blo(3 * L.size()), bhi(3 * L.size()), cen(3 * L.size());
    for (size_t i = 0; i < L.size(); ++i) {
      for (int a = 0; a < 3; ++a) {
        float lo = rest[3 * prims[k * L[i]] + a], hi = lo;
        for (int j = 1; j < k; ++j) {
          float v = rest[3 * prims[k * L[i] + j] + a];
          lo = std::min(lo, v);
          hi = std::max(hi, v);
        }
        blo[3 * i + a] = lo;
        bhi[3 * i + a] = hi;
        cen[3 * i + a] = 0.5f * (lo + hi);
      }
    }
    std::vector<int> ord = morton_order(L, cen);
    // level sizes
    d.count[0] = d.n_leaves;
    int top = 0;
    do {
      ++top;
      d.count[top] = (d.count[top - 1] + kFan - 1) / kFan;
    } while (d.count[top] > 1 && top + 1 < kMaxLevels);
    d.top = top;
    // level 0: leaves in Morton order, then padding
    std::vector<float4> clo, chi;
    for (int i : ord) {
      clo.push_back(make_float4(blo[3 * i], blo[3 * i + 1], blo[3 * i + 2], 0.f));
      chi.push_back(make_float4(bhi[3 * i], bhi[3 * i + 1], bhi[3 * i + 2], 0.f));
      lid.push_back(L[i]);
    }
    for (int lev = 0; lev <= top; ++lev) {
      if (lev > 0) {  // merge the level below
        std::vector<float4> nxl(d.count[lev], empty_lo), nxh(d.count[lev], empty_hi);
        for (int g = 0; g < d.count[lev]; ++g)
          for (int c = 0; c < kFan && g * kFan + c < d.count[lev - 1]; ++c) {
            const float4 &l = clo[g * kFan + c], &h = chi[g * kFan + c];
            nxl[g] = make_float4(std::min(nxl[g].x, l.x), std::min(nxl[g].y, l.y), std::min(nxl[g].z, l.z), 0.f);
            nxh[g] = make_float4(std::max(nxh[g].x, h.x), std::max(nxh[g].y, h.y), std::max(nxh[g].z, h.z), 0.f);
          }
        clo.swap(nxl);
        chi.swap(nxh);
      }
      int slots = lev < top ? kFan * d.count[lev + 1] : 1;
      d.offset[lev] = int(nlo.size());
      for (int g = 0; g < slots; ++g) {
        bool real = g < d.count[lev];
        nlo.push_back(real ? clo[g] : empty_lo);
        nhi.push_back(real ? chi[g] : empty_hi);
        if (lev == 0 && real) {
          int p = L[ord[g]];
          vid.push_back(make_int4(prims[k * p], prims[k * p + 1], k == 3 ? prims[k * p + 2] : -1, p));
        } else {
          vid.push_back(no_vid);
        }
      }
    }
    D[b] = d;
  }
  F.release();
  hipError_t e = upload(F.desc, D);
  if (e == hipSuccess) e = upload(F.leaf_id, lid);
  if (e == hipSuccess) e = upload(F.node_lo, nlo);
  if (e == hipSuccess) e = upload(F.node_hi, nhi);
  if (e == hipSuccess) e = upload(F.slot_vid, vid);
  F.n_trees = n_bodies;
  F.n_leaves = int(lid.size());
  F.n_nodes = int(nlo.size());
  F.max_leaves = max_leaves;
  return e;
}

void body_vertex_lists(int n_bodies, int nv, const int* vbody, const float* rest, int* offsets, int* ids, float* lo,
                       float* hi) {
  std::vector<std::vector<int>> per(n_bodies);
  for (int v = 0; v < nv; ++v)
    if (vbody[v] >= 0 && vbody[v] < n_bodies) per[vbody[v]].push_back(v);
  int at = 0;
  for (int b = 0; b < n_bodies; ++b) {
    offsets[b] = at;
    const std::vector<int>& L = per[b];
    std::vector<float> cen(3 * L.size());
    float l[3] = {INFINITY, INFINITY, INFINITY}, h[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (size_t i = 0; i < L.size(); ++i)
      for (int a = 0; a < 3; ++a) {
        float v = rest[3 * L[i] + a];
        cen[3 * i + a] = v;
        l[a] = std::min(l[a], v);
        h[a] = std::max(h[a], v);
      }
    for (int i : morton_order(L, cen)) ids[at++] = L[i];
    for (int a = 0; a < 3; ++a) {
      lo[4 * b + a] = L.empty() ? 0.f : l[a];
      hi[4 * b + a] = L.empty() ? 0.f : h[a];
    }
    lo[4 * b + 3] = hi[4 * b + 3] = 0.f;
  }
  offsets[n_bodies] = at;
}

// ------------------------------------------------------------- body frames ----
__device__ __forceinline__ void load_affine(const float* __restrict__ q, int b, double A[9], double p[3]) {
  const float* Q = q + 12ll * b;
  for (int i = 0; i < 3; ++i) {
    p[i] = Q[i];
    for (int j = 0; j < 3; ++j) A[3 * i + j] = Q[3 + 3 * j + i];  // column j of A is Q[3 + 3j .. 3 + 3j + 2]
  }
}

__device__ __forceinline__ double det3(const double A[9]) {
  return A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) + A[2] * (A[3] * A[7] - A[4] * A[6]);
}

__device__ __forceinline__ void inverse3(const double A[9], double det, double R[9]) {
  double s = 1.0 / det;
  R[0] = (A[4] * A[8] - A[5] * A[7]) * s;
  R[1] = (A[2] * A[7] - A[1] * A[8]) * s;
  R[2] = (A[1] * A[5] - A[2] * A[4]) * s;
  R[3] = (A[5] * A[6] - A[3] * A[8]) * s;
  R[4] = (A[0] * A[8] - A[2] * A[6]) * s;
  R[5] = (A[2] * A[3] - A[0] * A[5]) * s;
  R[6] = (A[3] * A[7] - A[4] * A[6]) * s;
  R[7] = (A[1] * A[6] - A[0] * A[7]) * s;
  R[8] = (A[0] * A[4] - A[1] * A[3]) * s;
}

// World box of a local box (center c, half extent e) through (A, p).
__device__ __forceinline__ void affine_box(const double A[9], const double p[3], const double c[3], const double e[3],
                                           double lo[3], double hi[3]) {
  for (int i = 0; i < 3; ++i) {
    double w = p[i] + A[3 * i] * c[0] + A[3 * i + 1] * c[1] + A[3 * i + 2] * c[2];
    double r = fabs(A[3 * i]) * e[0] + fabs(A[3 * i + 1]) * e[1] + fabs(A[3 * i + 2]) * e[2];
    lo[i] = w - r;
    hi[i] = w + r;
  }
}

__global__ void __launch_bounds__(64) body_frames_kernel(FixedStep f) {
  run_prologue(f.pro);
  int b = blockIdx.x * 64 + threadIdx.x;
  if (b >= f.n_bodies) return;
  double A0[9], p0[3];
  load_affine(f.q, b, A0, p0);
  float4 bl = f.body_lo[b], bh = f.body_hi[b];
  double c[3] = {0.5 * (double(bl.x) + bh.x), 0.5 * (double(bl.y) + bh.y), 0.5 * (double(bl.z) + bh.z)};
  double e[3] = {0.5 * (double(bh.x) - bl.x), 0.5 * (double(bh.y) - bl.y), 0.5 * (double(bh.z) - bl.z)};
  double lo[3], hi[3];
  affine_box(A0, p0, c, e, lo, hi);
  double d0 = det3(A0);
  double mindet = fabs(d0);
  BodyFrame F;
  inverse3(A0, d0, F.inv);
  double fro = 0.0;
  for (int k = 0; k < 9; ++k) fro += F.inv[k] * F.inv[k];
  F.fro = sqrt(fro);
  for (int i = 0; i < 3; ++i) {
    F.p[i] = p0[i];
    F.motion[i] = 0.0;
  }
  if (f.dq) {
    double dA[9], dp[3], A1[9], p1[3];
    load_affine(f.dq, b, dA, dp);
    for (int k = 0; k < 9; ++k) A1[k] = A0[k] - dA[k];
    for (int i = 0; i < 3; ++i) p1[i] = p0[i] - dp[i];
    double lo1[3], hi1[3];
    affine_box(A1, p1, c, e, lo1, hi1);
    for (int i = 0; i < 3; ++i) {
      lo[i] = fmin(lo[i], lo1[i]);
      hi[i] = fmax(hi[i], hi1[i]);
    }
    mindet = fmin(mindet, fabs(det3(A1)));
    // motion of the rest box in the start frame: D = A0^-1 (A1 - A0), C = A0^-1 (p1 - p0)
    double D[9], C[3];
    for (int i = 0; i < 3; ++i) {
      for (int j = 0; j < 3; ++j)
        D[3 * i + j] = F.inv[3 * i] * (-dA[j]) + F.inv[3 * i + 1] * (-dA[3 + j]) + F.inv[3 * i + 2] * (-dA[6 + j]);
      C[i] = F.inv[3 * i] * (-dp[0]) + F.inv[3 * i + 1] * (-dp[1]) + F.inv[3 * i + 2] * (-dp[2]);
    }
    for (int i = 0; i < 3; ++i) {
      double dc = D[3 * i] * c[0] + D[3 * i + 1] * c[1] + D[3 * i + 2] * c[2] + C[i];
      F.motion[i] = fabs(dc) + fabs(D[3 * i]) * e[0] + fabs(D[3 * i + 1]) * e[1] + fabs(D[3 * i + 2]) * e[2];
    }
  }
  if (!(mindet >= 1.0e-8)) *f.singular = f.gen;
  // fp32 positions agree with (A, p) only to rounding: widen by 1e-6 relative
  for (int i = 0; i < 3; ++i) {
    double slack = 1.0e-6 * fmax(fabs(lo[i]), fabs(hi[i]));
    lo[i] -= slack;
    hi[i] += slack;
  }
  f.frames[b] = F;
  f.world_lo[b] = make_float4(__double2float_rd(lo[0]), __double2float_rd(lo[1]), __double2float_rd(lo[2]), 0.f);
  f.world_hi[b] = make_float4(__double2float_ru(hi[0]), __double2float_ru(hi[1]), __double2float_ru(hi[2]), 0.f);
}

hipError_t launch_body_frames(const FixedStep& f, hipStream_t s) {
  if (f.n_bodies <= 0) return hipSuccess;
  hipLaunchKernelGGL(body_frames_kernel, dim3((f.n_bodies + 63) / 64), dim3(64), 0, s, f);
  return hipGetLastError();
}

// -------------------------------------------------------------- body pairs ----
__device__ __forceinline__ bool boxes_touch(float4 al, float4 ah, float4 bl, float4 bh, float pad) {
  // _aabbs_overlap (fixed_bvh_contacts.py:70-80)
  return al.x - pad <= bh.x && bl.x - pad <= ah.x && al.y - pad <= bh.y && bl.y - pad <= ah.y &&
         al.z - pad <= bh.z && bl.z - pad <= ah.z;
}

__global__ void __launch_bounds__(256) body_pairs_kernel(int nb, const float4* __restrict__ lo,
                                                         const float4* __restrict__ hi, float pad,
                                                         int2* __restrict__ pairs, int* __restrict__ count, int cap) {
  long long total = (long long)nb * nb;
  long long stride = (long long)gridDim.x * blockDim.x;
  for (long long t0 = (long long)blockIdx.x * blockDim.x + (threadIdx.x & ~63); t0 < total; t0 += stride) {
    long long t = t0 + __lane_id();
    bool ok = false;
    int a = 0, b = 0;
    if (t < total) {
      a = int(t / nb);
      b = int(t % nb);
      ok = a < b && boxes_touch(lo[a], hi[a], lo[b], hi[b], pad);
    }
    int out = wave_append(ok, count);
    if (out >= 0 && out < cap) pairs[out] = make_int2(a, b);
  }
}

hipError_t launch_body_pairs(int n_bodies, const float4* lo, const float4* hi, float pad, int2* pairs, int* count,
                             int cap, hipStream_t s) {
  if (n_bodies <= 1) return hipSuccess;
  long long total = (long long)n_bodies * n_bodies;
  long long blocks = std::min<long long>((total + 255) / 256, 4096);
  hipLaunchKernelGGL(body_pairs_kernel, dim3(unsigned(blocks)), dim3(256), 0, s, n_bodies, lo, hi, pad, pairs, count,
                     cap);
  return hipGetLastError();
}

// ------------------------------------------------------------------- tasks ----
__device__ __forceinline__ bool box_reaches(V3<float> lo, V3<float> hi, float4 bl, float4 bh) {
  return !(lo.x > bh.x || lo.y > bh.y || lo.z > bh.z || bl.x > hi.x || bl.y > hi.y || bl.z > hi.z);
}

// Workgroup `bid` of `nblk` over the (pair, edge of the source body) space;
// tasks are staged in `st` (the caller flushes it).
__device__ __forceinline__ void edge_tasks_block(const TaskLaunch& T, int bid, int nblk, BlockStage& st) {
  const StageOut out{nullptr, T.tasks_ee, T.count_ee, T.cap_ee};
  if (*T.singular == T.gen) return;
  int np = *T.n_pairs;
  np = np < T.pair_cap ? np : T.pair_cap;
  long long total = (long long)np * T.max_edges;
  long long stride = (long long)nblk * blockDim.x;
  for (long long t0 = (long long)bid * blockDim.x + (threadIdx.x & ~63); t0 < total; t0 += stride) {
    long long t = t0 + __lane_id();
    bool ok = false;
    int e = 0, tgt = 0;
    if (t < total) {
      int2 pr = T.pairs[t / T.max_edges];
      int k = int(t % T.max_edges);
      int src = pr.x;
      tgt = pr.y;
      if (T.ccd && T.fixed && T.fixed[src] != 0 && T.fixed[tgt] == 0) {
        src = pr.y;
        tgt = pr.x;
      }
      const TreeDesc& d = T.edesc[src];
      if (k < d.n_leaves) {
        e = T.eleaf_id[d.leaf_base + k];
        V3<float> a = ldx(T.x0, T.edges[2 * e]), b = ldx(T.x0, T.edges[2 * e + 1]);
        V3<float> lo = vmin(a, b), hi = vmax(a, b);
        if (T.ccd) {
          V3<float> a1 = ldx(T.x1, T.edges[2 * e]), b1 = ldx(T.x1, T.edges[2 * e + 1]);
          lo = vmin(lo, vmin(a1, b1));
          hi = vmax(hi, vmax(a1, b1));
        }
        V3<float> pv = mk(T.pad, T.pad, T.pad);
        ok = box_reaches(lo - pv, hi + pv, T.world_lo[tgt], T.world_hi[tgt]);
      }
    }
    stage_append(st, out, ok, make_int4(0, 0, 0, 0), make_int2(e, tgt));
  }
}

// Workgroup `bid` of `nblk` over the (pair, direction, vertex) space.
__device__ __forceinline__ void vertex_tasks_block(const TaskLaunch& T, int bid, int nblk, BlockStage& st) {
  const StageOut out{nullptr, T.tasks_pt, T.count_pt, T.cap_pt};
  if (*T.singular == T.gen) return;
  int np = *T.n_pairs;
  np = np < T.pair_cap ? np : T.pair_cap;
  long long per = 2ll * T.max_verts;
  long long total = (long long)np * per;
  long long stride = (long long)nblk * blockDim.x;
  for (long long t0 = (long long)bid * blockDim.x + (threadIdx.x & ~63); t0 < total; t0 += stride) {
    long long t = t0 + __lane_id();
    bool ok = false;
    int v = 0, tgt = 0;
    if (t < total) {
      int2 pr = T.pairs[t / per];
      int rem = int(t % per);
      int dir = rem / T.max_verts, k = rem % T.max_verts;
      int src = dir ? pr.y : pr.x;
      tgt = dir ? pr.x : pr.y;
      int vb = T.vert_base[src];
      if (k < T.vert_base[src + 1] - vb) {
        v = T.vert_ids[vb + k];
        V3<float> lo = ldx(T.x0, v), hi = lo;
        if (T.ccd) {
          V3<float> p1 = ldx(T.x1, v);
          lo = vmin(lo, p1);
          hi = vmax(hi, p1);
        }
        V3<float> pv = mk(T.pad, T.pad, T.pad);
        ok = box_reaches(lo - pv, hi + pv, T.world_lo[tgt], T.world_hi[tgt]);
      }
    }
    stage_append(st, out, ok, make_int4(0, 0, 0, 0), make_int2(v, tgt));
  }
}

// Edge tasks in workgroups [0, blocks_ee), vertex tasks in the rest. Each
// workgroup stages its tasks in LDS and reserves its output range with one
// global atomicAdd (the two counters share a cache line).
__global__ void __launch_bounds__(256) tasks_kernel(TaskLaunch T, int blocks_ee) {
  __shared__ BlockStage st;
  stage_begin(st);
  if (int(blockIdx.x) < blocks_ee) {
    edge_tasks_block(T, blockIdx.x, blocks_ee, st);
    stage_flush(st, StageOut{nullptr, T.tasks_ee, T.count_ee, T.cap_ee});
  } else {
    vertex_tasks_block(T, blockIdx.x - blocks_ee, gridDim.x - blocks_ee, st);
    stage_flush(st, StageOut{nullptr, T.tasks_pt, T.count_pt, T.cap_pt});
  }
}

hipError_t launch_fixed_tasks(const TaskLaunch& t, hipStream_t s) {
  // The kernels read the device pair count and stride over what exists.
  // Each half is capped at 1024 workgroups (4 per CU): a grid sized by the
  // pair list's capacity is mostly workgroups without work, and dispatching
  // them cost more than the work itself (32 us against 13 us per launch at
  // 64 bodies).
  long long ee = (long long)t.pair_cap * t.max_edges, pt = 2ll * t.pair_cap * t.max_verts;
  long long blocks_ee = (ee > 0 && t.tasks_ee) ? std::min<long long>((ee + 255) / 256, 1024) : 0;
  long long blocks_pt = (pt > 0 && t.tasks_pt) ? std::min<long long>((pt + 255) / 256, 1024) : 0;
  if (blocks_ee + blocks_pt == 0) return hipSuccess;
  hipLaunchKernelGGL(tasks_kernel, dim3(unsigned(blocks_ee + blocks_pt)), dim3(256), 0, s, t, int(blocks_ee));
  return hipGetLastError();
}

// ------------------------------------------------------------------ queries ----
struct FixedArgs {
  const int2* tasks;
  const int* task_count;
  int task_cap;
  const int* edges;
  const int* tris;
  const float* x0;
  const float* x1;
  const BodyFrame* frames;
  const TreeDesc* desc;
  const float4* node_lo;
  const float4* node_hi;
  const int4* slot_vid;
  float thickness;
  float pad;
  int4* rec;
  int2* ids;
  int capacity;
  int* count;
};

__device__ __forceinline__ V3<double> to_local(const BodyFrame& F, V3<float> w) {
  double d0 = double(w.x) - F.p[0], d1 = double(w.y) - F.p[1], d2 = double(w.z) - F.p[2];
  return mk(F.inv[0] * d0 + F.inv[1] * d1 + F.inv[2] * d2, F.inv[3] * d0 + F.inv[4] * d1 + F.inv[5] * d2,
            F.inv[6] * d0 + F.inv[7] * d1 + F.inv[8] * d2);
}

// The task's own primitive: vertex ids a0, a1 (edge) or the point, and for
// DCD their positions at x0, loaded once per task; lv is the leaf's record.
struct TaskPrim {
  int src, a0, a1;
  V3<float> p0, p1;
};

template <int MODE>
__device__ __forceinline__ bool fixed_accept(const FixedArgs& A, const TaskPrim& P, int4 lv, int4& rec) {
  if (MODE == Q_DCD_EE || MODE == Q_CCD_EE) {
    int b0 = lv.x, b1 = lv.y;
    rec = make_int4(P.a0, P.a1, b0, b1);
    if (MODE == Q_CCD_EE) return true;
    EdgeClosest ec = closest_point_edge_edge(P.p0, P.p1, ldx(A.x0, b0), ldx(A.x0, b1), 1.0e-6f);
    return ec.dist < A.thickness * 2.0f;
  }
  int t0 = lv.x, t1 = lv.y, t2 = lv.z;
  if (MODE == Q_CCD_PT) {
    rec = make_int4(P.src, t0, t1, t2);
    return true;
  }
  rec = make_int4(P.src, t1, t0, t2);
  TriClosest tc = closest_point_triangle(P.p0, ldx(A.x0, t0), ldx(A.x0, t1), ldx(A.x0, t2));
  return tc.dist < A.thickness * 2.0f;
}

// One task per eight lanes, with the step of query_kernel in cull.hip: a
// stack entry (level L, mask, cbase) stands for level-L nodes of the task's
// tree whose children sit at slots cbase + 8 c + 0..7 of the forest's node
// arrays; every load of a step is issued up front (padded levels keep them
// in bounds). The grid is persistent (tasks are counted on the device) and
// each wave strides over groups of eight tasks, so every lane of a wave runs
// the same trip count.
struct FixedSmem {
  unsigned long long stack[kGroupsPerBlock][kMaskStack];
  BlockStage stage;
};

// Workgroup `bid` of the `nblk` serving A.
template <int MODE>
__device__ __forceinline__ void fixed_query_block(const FixedArgs& A, int bid, int nblk, FixedSmem& sm) {
  auto& stack = sm.stack;
  BlockStage& stage = sm.stage;
  const StageOut out{A.rec, A.ids, A.count, A.capacity};
  stage_begin(stage);
  const int tid = threadIdx.x;
  const int grp = tid / kFan, r = tid % kFan;
  const int lane = __lane_id();
  const int group_shift = lane & ~(kFan - 1);
  const int gw = lane / kFan;  // group within the wave
  const int groups_per_wave = 64 / kFan;
  int n = *A.task_count;
  n = n < A.task_cap ? n : A.task_cap;
  const int waves_total = nblk * (kQueryBlock / 64);
  for (int wbase = (bid * (kQueryBlock / 64) + tid / 64) * groups_per_wave; wbase < n;
       wbase += waves_total * groups_per_wave) {
    int slot = wbase + gw;
    bool live = slot < n;
    TaskPrim P;
    P.src = P.a0 = P.a1 = 0;
    P.p0 = P.p1 = mk(0.f, 0.f, 0.f);
    int tb = 0;
    float qlo[3] = {0.f, 0.f, 0.f}, qhi[3] = {-1.f, -1.f, -1.f};
    if (live) {
      int2 task = A.tasks[slot];
      P.src = task.x;
      tb = task.y;
      live = A.desc[tb].n_leaves > 0;
    }
    if (live) {
      const int src = P.src;
      const BodyFrame& F = A.frames[tb];
      V3<double> lo, hi;
      if (MODE == Q_DCD_EE || MODE == Q_CCD_EE) {
        P.a0 = A.edges[2 * src];
        P.a1 = A.edges[2 * src + 1];
        V3<float> w0 = ldx(A.x0, P.a0), w1 = ldx(A.x0, P.a1);
        if (MODE == Q_DCD_EE) {
          P.p0 = w0;
          P.p1 = w1;
        }
        V3<double> a = to_local(F, w0), b = to_local(F, w1);
        lo = vmin(a, b);
        hi = vmax(a, b);
        if (MODE == Q_CCD_EE) {
          V3<double> a1 = to_local(F, ldx(A.x1, P.a0)), b1 = to_local(F, ldx(A.x1, P.a1));
          lo = vmin(lo, vmin(a1, b1));
          hi = vmax(hi, vmax(a1, b1));
        }
      } else {
        V3<float> w = ldx(A.x0, src);
        if (MODE == Q_DCD_PT) P.p0 = w;
        lo = hi = to_local(F, w);
        if (MODE == Q_CCD_PT) {
          V3<double> p1 = to_local(F, ldx(A.x1, src));
          lo = vmin(lo, p1);
          hi = vmax(hi, p1);
        }
      }
      V3<double> inf;
      if (MODE == Q_DCD_EE || MODE == Q_DCD_PT) {
        double s = F.fro * (2.0 * double(A.thickness));
        inf = mk(s, s, s);
      } else {
        double s = F.fro * double(A.pad);
        inf = mk(F.motion[0] + s, F.motion[1] + s, F.motion[2] + s);
      }
      // fp32 positions agree with the body frame only to rounding: widen by 1e-6 relative
      inf = inf + mk(1.0e-6 * fmax(fabs(lo.x), fabs(hi.x)), 1.0e-6 * fmax(fabs(lo.y), fabs(hi.y)),
                     1.0e-6 * fmax(fabs(lo.z), fabs(hi.z)));
      lo = lo - inf;
      hi = hi + inf;
      qlo[0] = __double2float_rd(lo.x);
      qlo[1] = __double2float_rd(lo.y);
      qlo[2] = __double2float_rd(lo.z);
      qhi[0] = __double2float_ru(hi.x);
      qhi[1] = __double2float_ru(hi.y);
      qhi[2] = __double2float_ru(hi.z);
    }
    const TreeDesc* D = A.desc + tb;
    int sp = 0;
    if (live) {
      stack[grp][0] = mask_entry(D->top, 1u, D->offset[D->top - 1]);  // the root
      sp = 1;
    }
    while (true) {
      bool active = sp > 0;
      int L = 0, cbase = 0, c = -1;
      if (active) {
        unsigned long long e = stack[grp][sp - 1];
        L = int(e >> 40) & 0xf;
        unsigned mask = unsigned(e >> 32) & 0xffu;
        cbase = int(e & 0xffffffffu);
        c = __ffs(mask) - 1;
        mask &= mask - 1u;
        if (mask == 0u) --sp;
        else stack[grp][sp - 1] = mask_entry(L, mask, cbase);
      }
      const int at = c >= 0 ? cbase + c * kFan + r : 0;
      const float4 clo = A.node_lo[at], chi = A.node_hi[at];
      const int4 lv = A.slot_vid[L == 1 ? at : 0];
      const bool internal = L >= 2;
      const int child_off = D->offset[L >= 1 ? L - 1 : 0], child_cnt = D->count[L >= 1 ? L - 1 : 0];
      const int grand_off = D->offset[internal ? L - 2 : 0];
      // only real children (bitwise, so that no load is sunk behind a branch)
      const bool hit = (c >= 0) & (at - child_off < child_cnt) & overlaps(qlo, qhi, clo, chi);
      unsigned gmask = unsigned(__ballot(internal && hit) >> group_shift) & 0xffu;
      if (internal && gmask != 0u && sp < kMaskStack) {
        int rel = cbase - child_off + c * kFan;  // first child's index within level L - 1
        stack[grp][sp] = mask_entry(L - 1, gmask, grand_off + rel * kFan);
        ++sp;
      }
      bool want = false;
      int4 rec = make_int4(0, 0, 0, 0);
      if (L == 1 && hit) want = fixed_accept<MODE>(A, P, lv, rec);
      stage_append(stage, out, want, rec, make_int2(P.src, lv.w));
      if (!__any(sp > 0)) break;
    }
  }
  stage_flush(stage, out);
}

template <int MODE>
__global__ void __launch_bounds__(kQueryBlock) fixed_query_kernel(FixedArgs A) {
  __shared__ FixedSmem sm;
  fixed_query_block<MODE>(A, blockIdx.x, gridDim.x, sm);
}

// Workgroups [0, blocks_a) serve A, the rest B.
template <int MA, int MB>
__global__ void __launch_bounds__(kQueryBlock) fixed_query_pair_kernel(FixedArgs A, FixedArgs B, int blocks_a) {
  __shared__ FixedSmem sm;
  if (int(blockIdx.x) < blocks_a) fixed_query_block<MA>(A, blockIdx.x, blocks_a, sm);
  else fixed_query_block<MB>(B, blockIdx.x - blocks_a, gridDim.x - blocks_a, sm);
}

namespace {

FixedArgs fixed_args(const FixedQuery& q) {
  FixedArgs A;
  A.tasks = q.tasks;
  A.task_count = q.task_count;
  A.task_cap = q.task_cap;
  A.edges = q.edges;
  A.tris = q.tris;
  A.x0 = q.x0;
  A.x1 = q.x1;
  A.frames = q.frames;
  A.desc = q.forest->desc.as<TreeDesc>();
  A.node_lo = q.forest->node_lo.as<float4>();
  A.node_hi = q.forest->node_hi.as<float4>();
  A.slot_vid = q.forest->slot_vid.as<int4>();
  A.thickness = q.thickness;
  A.pad = q.pad;
  A.rec = q.rec;
  A.ids = q.ids;
  A.capacity = q.capacity;
  A.count = q.count;
  return A;
}

// Persistent grid of a query: one workgroup per 32 tasks of capacity, at
// most 2048.
int fixed_blocks(const FixedQuery& q) {
  int blocks = (q.task_cap + kGroupsPerBlock - 1) / kGroupsPerBlock;
  return blocks > 2048 ? 2048 : blocks;
}

bool fixed_live(const FixedQuery& q) { return q.task_cap > 0 && q.forest->n_leaves > 0; }

}  // namespace

hipError_t launch_fixed_query(const FixedQuery& q, hipStream_t s) {
  if (!fixed_live(q)) return hipSuccess;
  FixedArgs A = fixed_args(q);
  int blocks = fixed_blocks(q);
  switch (q.mode) {
    case Q_DCD_EE: hipLaunchKernelGGL(fixed_query_kernel<Q_DCD_EE>, dim3(blocks), dim3(kQueryBlock), 0, s, A); break;
    case Q_DCD_PT: hipLaunchKernelGGL(fixed_query_kernel<Q_DCD_PT>, dim3(blocks), dim3(kQueryBlock), 0, s, A); break;
    case Q_CCD_EE: hipLaunchKernelGGL(fixed_query_kernel<Q_CCD_EE>, dim3(blocks), dim3(kQueryBlock), 0, s, A); break;
    default: hipLaunchKernelGGL(fixed_query_kernel<Q_CCD_PT>, dim3(blocks), dim3(kQueryBlock), 0, s, A); break;
  }
  return hipGetLastError();
}

hipError_t launch_fixed_queries(const FixedQuery& ee, const FixedQuery& pt, hipStream_t s) {
  if (!fixed_live(ee)) return launch_fixed_query(pt, s);
  if (!fixed_live(pt)) return launch_fixed_query(ee, s);
  FixedArgs A = fixed_args(ee), B = fixed_args(pt);
  int blocks_a = fixed_blocks(ee), blocks_b = fixed_blocks(pt);
  dim3 grid(blocks_a + blocks_b), block(kQueryBlock);
  if (ee.mode == Q_DCD_EE && pt.mode == Q_DCD_PT)
    hipLaunchKernelGGL((fixed_query_pair_kernel<Q_DCD_EE, Q_DCD_PT>), grid, block, 0, s, A, B, blocks_a);
  else if (ee.mode == Q_CCD_EE && pt.mode == Q_CCD_PT)
    hipLaunchKernelGGL((fixed_query_pair_kernel<Q_CCD_EE, Q_CCD_PT>), grid, block, 0, s, A, B, blocks_a);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

}  // namespace ipcd

// Contact context: the device buffers ContactSolverBase owns in the
// reference (contact.py:444-521) — topology, BVHs, CCD positions and
// candidate lists — and the entry points that drive the culling and CCD
// kernels with them.
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/ipc_warp_amd.h"
#include "cull.hpp"
#include "ipc_kernels.hpp"

using namespace ipcd;

struct ipc_ctx {
  int nv = 0, ne = 0, nf = 0;
  DevBuf edges, tris, body;
  Bvh edge_bvh, tri_bvh;
  DevBuf x1;               // CCD end positions x - dx
  DevBuf cand_ee, cand_pt; // int4 candidate quadruples
  int cap_ee = 0, cap_pt = 0;
  DevBuf small;            // [0] uint64 toi bits, [8..] int counters
  void* host = nullptr;    // pinned readback area
  int last_counts[2] = {0, 0};
  // optional per-kernel timing (HIP events recorded on the launch stream)
  bool timing = false;
  struct Pending { int kind; hipEvent_t a, b; };
  std::vector<Pending> pending;
  double total_ms[IPC_TIMER_KINDS] = {0};
  long long calls[IPC_TIMER_KINDS] = {0};
};

namespace {

int ctx_fail(int code, const char* what) {
  set_error(what);
  return code;
}

int ctx_hip(hipError_t e, const char* where) {
  if (e == hipSuccess) return IPC_OK;
  char buf[512];
  std::snprintf(buf, sizeof(buf), "%s: %s", where, hipGetErrorString(e));
  set_error(buf);
  return -int(e);
}

unsigned long long* toi_bits(ipc_ctx* c) { return c->small.as<unsigned long long>(); }
int* counters(ipc_ctx* c) { return reinterpret_cast<int*>(c->small.as<char>() + 64); }

constexpr unsigned long long kOneBits = 0x3FF0000000000000ull;

// Runs `launch` between two events on `s` when the context is timing.
template <typename F>
hipError_t timed(ipc_ctx* c, int kind, hipStream_t s, F launch) {
  if (!c->timing) return launch();
  hipEvent_t a, b;
  hipError_t e = hipEventCreate(&a);
  if (e != hipSuccess) return e;
  e = hipEventCreate(&b);
  if (e != hipSuccess) { (void)hipEventDestroy(a); return e; }
  (void)hipEventRecord(a, s);
  e = launch();
  (void)hipEventRecord(b, s);
  c->pending.push_back({kind, a, b});
  return e;
}

}  // namespace

extern "C" {

int ipc_ctx_create(ipc_ctx** out, int n_vertices, const int* edges, int n_edges, const int* triangles,
                   int n_triangles, const int* body) {
  if (!out) return ctx_fail(IPC_EINVAL, "ctx_create: out is null");
  *out = nullptr;
  if (n_vertices < 0 || n_edges < 0 || n_triangles < 0)
    return ctx_fail(IPC_EINVAL, "ctx_create: negative size");
  if ((n_edges && !edges) || (n_triangles && !triangles) || (n_vertices && !body))
    return ctx_fail(IPC_EINVAL, "ctx_create: null topology pointer");
  ipc_ctx* c = new ipc_ctx();
  c->nv = n_vertices;
  c->ne = n_edges;
  c->nf = n_triangles;
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = c->edges.ensure(sizeof(int) * 2 * (size_t)n_edges + 16);
  if (e == hipSuccess) e = c->tris.ensure(sizeof(int) * 3 * (size_t)n_triangles + 16);
  if (e == hipSuccess) e = c->body.ensure(sizeof(int) * (size_t)n_vertices + 16);
  if (e == hipSuccess) e = c->small.ensure(256);
  if (e == hipSuccess && n_edges) e = hipMemcpy(c->edges.p, edges, sizeof(int) * 2 * (size_t)n_edges, hipMemcpyDefault);
  if (e == hipSuccess && n_triangles)
    e = hipMemcpy(c->tris.p, triangles, sizeof(int) * 3 * (size_t)n_triangles, hipMemcpyDefault);
  if (e == hipSuccess && n_vertices) e = hipMemcpy(c->body.p, body, sizeof(int) * (size_t)n_vertices, hipMemcpyDefault);
  if (e == hipSuccess) e = hipHostMalloc(&c->host, 256, hipHostMallocDefault);
  // Every vertex index of the topology must lie in [0, n_vertices): the
  // culling kernels gather positions through it.
  int range[4] = {0, -1, 0, -1};
  if (e == hipSuccess && n_edges) e = index_range(c->edges.as<int>(), 2ll * n_edges, c->small.as<int>(), nullptr);
  if (e == hipSuccess && n_triangles)
    e = index_range(c->tris.as<int>(), 3ll * n_triangles, c->small.as<int>() + 2, nullptr);
  if (e == hipSuccess && (n_edges || n_triangles)) e = hipMemcpy(range, c->small.p, sizeof(range), hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    int st = ctx_hip(e, "ctx_create");
    ipc_ctx_destroy(c);
    return st;
  }
  bool bad = (n_edges && (range[0] < 0 || range[1] >= n_vertices)) ||
             (n_triangles && (range[2] < 0 || range[3] >= n_vertices));
  if (bad) {
    ipc_ctx_destroy(c);
    return ctx_fail(IPC_EINVAL, "ctx_create: an edge or triangle names a vertex outside [0, n_vertices)");
  }
  *out = c;
  return IPC_OK;
}

int ipc_ctx_destroy(ipc_ctx* c) {
  if (!c) return IPC_OK;
  (void)hipDeviceSynchronize();
  for (auto& p : c->pending) { (void)hipEventDestroy(p.a); (void)hipEventDestroy(p.b); }
  c->edges.release(); c->tris.release(); c->body.release();
  c->edge_bvh.release(); c->tri_bvh.release();
  c->x1.release(); c->cand_ee.release(); c->cand_pt.release(); c->small.release();
  if (c->host) (void)hipHostFree(c->host);
  delete c;
  return IPC_OK;
}

// edge_edge_collision over a freshly built edge BVH (contact.py:672-704):
// boxes +- (thickness + buffer), accepted pairs into contacts/e0e1 up to
// capacity, total count into *count (device int).
int ipc_detect_collision_ee(ipc_ctx* c, const float* x, float thickness, float buffer, int exclude_same_body,
                            int* contacts, int* e0e1, int capacity, int* count, void* stream) {
  if (!c || !x || !count) return ctx_fail(IPC_EINVAL, "detect_collision_ee: null pointer");
  if (capacity < 0 || (capacity > 0 && !contacts)) return ctx_fail(IPC_EINVAL, "detect_collision_ee: bad capacity");
  hipStream_t s = static_cast<hipStream_t>(stream);
  int st = ctx_hip(launch_zero_i32(count, 1, s), "detect_collision_ee");
  if (st || c->ne == 0) return st;
  float pad = thickness + buffer;
  st = ctx_hip(timed(c, IPC_TIMER_BUILD_EDGE, s, [&] {
                 return build_bvh(c->edge_bvh, BOX_EDGE, c->ne, c->edges.as<int>(), c->body.as<int>(), x, nullptr,
                                  pad, s);
               }),
               "edge BVH");
  if (st) return st;
  QueryLaunch q{};
  q.mode = Q_DCD_EE;
  q.n_query = c->ne;
  q.order = c->edge_bvh.leaf_id.as<int>();
  q.edges = c->edges.as<int>();
  q.body = c->body.as<int>();
  q.x0 = x;
  q.qlo = c->edge_bvh.leaf_lo.as<float4>();
  q.qhi = c->edge_bvh.leaf_hi.as<float4>();
  q.thickness = thickness;
  q.exclude_same_body = exclude_same_body;
  q.bvh = &c->edge_bvh;
  q.rec = reinterpret_cast<int4*>(contacts);
  q.ids = reinterpret_cast<int2*>(e0e1);
  q.capacity = capacity;
  q.count = count;
  return ctx_hip(timed(c, IPC_TIMER_QUERY_DCD_EE, s, [&] { return run_query(q, s); }), "detect_collision_ee query");
}

// point_triangle_collision over a freshly built triangle BVH
// (contact.py:708-724): every vertex queries x_i +- 2 (thickness + buffer).
int ipc_detect_collision_pt(ipc_ctx* c, const float* x, float thickness, float buffer, int exclude_same_body,
                            int* contacts, int* e0e1, int capacity, int* count, void* stream) {
  if (!c || !x || !count) return ctx_fail(IPC_EINVAL, "detect_collision_pt: null pointer");
  if (capacity < 0 || (capacity > 0 && !contacts)) return ctx_fail(IPC_EINVAL, "detect_collision_pt: bad capacity");
  hipStream_t s = static_cast<hipStream_t>(stream);
  int st = ctx_hip(launch_zero_i32(count, 1, s), "detect_collision_pt");
  if (st || c->nf == 0 || c->nv == 0) return st;
  st = ctx_hip(timed(c, IPC_TIMER_BUILD_TRI, s, [&] {
                 return build_bvh(c->tri_bvh, BOX_TRI, c->nf, c->tris.as<int>(), c->body.as<int>(), x, nullptr, 0.0f,
                                  s);
               }),
               "triangle BVH");
  if (st) return st;
  QueryLaunch q{};
  q.mode = Q_DCD_PT;
  q.n_query = c->nv;
  q.tris = c->tris.as<int>();
  q.body = c->body.as<int>();
  q.x0 = x;
  q.pad = (thickness + buffer) * 2.0f;
  q.thickness = thickness;
  q.exclude_same_body = exclude_same_body;
  q.bvh = &c->tri_bvh;
  q.rec = reinterpret_cast<int4*>(contacts);
  q.ids = reinterpret_cast<int2*>(e0e1);
  q.capacity = capacity;
  q.count = count;
  return ctx_hip(timed(c, IPC_TIMER_QUERY_DCD_PT, s, [&] { return run_query(q, s); }), "detect_collision_pt query");
}

// Candidate sets of edge_edge_toi / point_triangle_toi between x0 and x1.
int ipc_ccd_candidates_ee(ipc_ctx* c, const float* x0, const float* x1, float padding, int exclude_same_body,
                          int* quads, int* ids, int capacity, int* count, void* stream) {
  if (!c || !x0 || !x1 || !count) return ctx_fail(IPC_EINVAL, "ccd_candidates_ee: null pointer");
  if (capacity < 0 || (capacity > 0 && !quads)) return ctx_fail(IPC_EINVAL, "ccd_candidates_ee: bad capacity");
  hipStream_t s = static_cast<hipStream_t>(stream);
  int st = ctx_hip(launch_zero_i32(count, 1, s), "ccd_candidates_ee");
  if (st || c->ne == 0) return st;
  st = ctx_hip(timed(c, IPC_TIMER_BUILD_EDGE, s, [&] {
                 return build_bvh(c->edge_bvh, BOX_SWEPT_EDGE, c->ne, c->edges.as<int>(), c->body.as<int>(), x0, x1,
                                  padding, s);
               }),
               "swept edge BVH");
  if (st) return st;
  QueryLaunch q{};
  q.mode = Q_CCD_EE;
  q.n_query = c->ne;
  q.order = c->edge_bvh.leaf_id.as<int>();
  q.edges = c->edges.as<int>();
  q.body = c->body.as<int>();
  q.x0 = x0;
  q.x1 = x1;
  q.qlo = c->edge_bvh.leaf_lo.as<float4>();
  q.qhi = c->edge_bvh.leaf_hi.as<float4>();
  q.exclude_same_body = exclude_same_body;
  q.bvh = &c->edge_bvh;
  q.rec = reinterpret_cast<int4*>(quads);
  q.ids = reinterpret_cast<int2*>(ids);
  q.capacity = capacity;
  q.count = count;
  return ctx_hip(timed(c, IPC_TIMER_QUERY_CCD_EE, s, [&] { return run_query(q, s); }), "ccd_candidates_ee query");
}

int ipc_ccd_candidates_pt(ipc_ctx* c, const float* x0, const float* x1, float padding, int exclude_same_body,
                          int* quads, int* ids, int capacity, int* count, void* stream) {
  if (!c || !x0 || !x1 || !count) return ctx_fail(IPC_EINVAL, "ccd_candidates_pt: null pointer");
  if (capacity < 0 || (capacity > 0 && !quads)) return ctx_fail(IPC_EINVAL, "ccd_candidates_pt: bad capacity");
  hipStream_t s = static_cast<hipStream_t>(stream);
  int st = ctx_hip(launch_zero_i32(count, 1, s), "ccd_candidates_pt");
  if (st || c->nf == 0 || c->nv == 0) return st;
  st = ctx_hip(timed(c, IPC_TIMER_BUILD_TRI, s, [&] {
                 return build_bvh(c->tri_bvh, BOX_SWEPT_TRI, c->nf, c->tris.as<int>(), c->body.as<int>(), x0, x1,
                                  padding, s);
               }),
               "swept triangle BVH");
  if (st) return st;
  QueryLaunch q{};
  q.mode = Q_CCD_PT;
  q.n_query = c->nv;
  q.tris = c->tris.as<int>();
  q.body = c->body.as<int>();
  q.x0 = x0;
  q.x1 = x1;
  q.pad = padding;
  q.exclude_same_body = exclude_same_body;
  q.bvh = &c->tri_bvh;
  q.rec = reinterpret_cast<int4*>(quads);
  q.ids = reinterpret_cast<int2*>(ids);
  q.capacity = capacity;
  q.count = count;
  return ctx_hip(timed(c, IPC_TIMER_QUERY_CCD_PT, s, [&] { return run_query(q, s); }), "ccd_candidates_pt query");
}

// collision_free_step (contact.py:539-585) before its 0.9 factor: x1 = x - dx,
// swept-box candidates for EE then PT, FP64 TOI per candidate, minimum over
// all of them. Candidate lists live in the context and grow on overflow (one
// retry, like _grow_contact_list). Synchronises `stream` once to read the
// result; writes the raw TOI (1.0 = no impact) to *toi.
int ipc_collision_free_step(ipc_ctx* c, const float* x, const float* dx, float padding, int exclude_same_body,
                            double* toi, void* stream) {
  if (!c || !x || !dx || !toi) return ctx_fail(IPC_EINVAL, "collision_free_step: null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  int st = ctx_hip(c->x1.ensure(sizeof(float) * 3 * (size_t)c->nv + 16), "collision_free_step alloc");
  if (st) return st;
  if (c->cap_ee == 0) c->cap_ee = 1 << 16;
  if (c->cap_pt == 0) c->cap_pt = 1 << 16;
  for (int attempt = 0; attempt < 2; ++attempt) {
    st = ctx_hip(c->cand_ee.ensure(sizeof(int) * 4 * (size_t)c->cap_ee), "collision_free_step alloc");
    if (!st) st = ctx_hip(c->cand_pt.ensure(sizeof(int) * 4 * (size_t)c->cap_pt), "collision_free_step alloc");
    if (st) return st;
    st = ctx_hip(launch_sub(3 * c->nv, x, dx, c->x1.as<float>(), s), "collision_free_step");
    if (!st) st = ctx_hip(launch_fill_u64(toi_bits(c), 1, kOneBits, s), "collision_free_step");
    if (st) return st;
    int* cnt = counters(c);
    st = ipc_ccd_candidates_ee(c, x, c->x1.as<float>(), padding, exclude_same_body, c->cand_ee.as<int>(), nullptr,
                               c->cap_ee, cnt + 0, stream);
    if (st) return st;
    st = ipc_ccd_candidates_pt(c, x, c->x1.as<float>(), padding, exclude_same_body, c->cand_pt.as<int>(), nullptr,
                               c->cap_pt, cnt + 1, stream);
    if (st) return st;
    // TOI over the first min(count, capacity) candidates: the kernels read the
    // device count, so no host round trip is needed between the phases.
    st = ctx_hip(timed(c, IPC_TIMER_TOI_EE, s, [&] {
                   return launch_collision_time_counted(true, x, c->x1.as<float>(), c->nv, c->cand_ee.as<int>(),
                                                        cnt + 0, c->cap_ee, toi_bits(c), s);
                 }),
                 "collision_free_step EE TOI");
    if (!st)
      st = ctx_hip(timed(c, IPC_TIMER_TOI_PT, s, [&] {
                     return launch_collision_time_counted(false, x, c->x1.as<float>(), c->nv, c->cand_pt.as<int>(),
                                                          cnt + 1, c->cap_pt, toi_bits(c), s);
                   }),
                   "collision_free_step PT TOI");
    if (st) return st;
    st = ctx_hip(hipMemcpyAsync(c->host, c->small.p, 128, hipMemcpyDeviceToHost, s), "collision_free_step readback");
    if (!st) st = ctx_hip(hipStreamSynchronize(s), "collision_free_step sync");
    if (st) return st;
    unsigned long long bits;
    std::memcpy(&bits, c->host, 8);
    int n_ee, n_pt;
    std::memcpy(&n_ee, static_cast<char*>(c->host) + 64, 4);
    std::memcpy(&n_pt, static_cast<char*>(c->host) + 68, 4);
    c->last_counts[0] = n_ee;
    c->last_counts[1] = n_pt;
    bool grow = false;
    if (n_ee > c->cap_ee) { while (c->cap_ee < n_ee) c->cap_ee *= 2; grow = true; }
    if (n_pt > c->cap_pt) { while (c->cap_pt < n_pt) c->cap_pt *= 2; grow = true; }
    if (!grow) {
      double t;
      std::memcpy(&t, &bits, 8);
      *toi = t;
      return IPC_OK;
    }
  }
  return ctx_fail(IPC_EINVAL, "collision_free_step: candidate list kept overflowing");
}

// Per-kernel timing: while enabled, builds, queries and TOI launches are
// bracketed by HIP events on their stream. ipc_ctx_kernel_times waits for
// the recorded events and returns the accumulated milliseconds and launch
// counts per kind (IPC_TIMER_*), then resets them.
int ipc_ctx_set_timing(ipc_ctx* c, int enable) {
  if (!c) return ctx_fail(IPC_EINVAL, "ctx_set_timing: null context");
  c->timing = enable != 0;
  return IPC_OK;
}

int ipc_ctx_kernel_times(ipc_ctx* c, double* total_ms, long long* calls) {
  if (!c || !total_ms || !calls) return ctx_fail(IPC_EINVAL, "ctx_kernel_times: null pointer");
  for (auto& p : c->pending) {
    float ms = 0.f;
    hipError_t e = hipEventSynchronize(p.b);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, p.a, p.b);
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
    if (e != hipSuccess) {
      c->pending.clear();
      return ctx_hip(e, "ctx_kernel_times");
    }
    c->total_ms[p.kind] += ms;
    c->calls[p.kind] += 1;
  }
  c->pending.clear();
  for (int k = 0; k < IPC_TIMER_KINDS; ++k) {
    total_ms[k] = c->total_ms[k];
    calls[k] = c->calls[k];
    c->total_ms[k] = 0.0;
    c->calls[k] = 0;
  }
  return IPC_OK;
}

// Candidate counts of the last collision_free_step (EE, PT).
int ipc_ctx_last_ccd_counts(ipc_ctx* c, int* n_ee, int* n_pt) {
  if (!c || !n_ee || !n_pt) return ctx_fail(IPC_EINVAL, "ctx_last_ccd_counts: null pointer");
  *n_ee = c->last_counts[0];
  *n_pt = c->last_counts[1];
  return IPC_OK;
}

}  // extern "C"

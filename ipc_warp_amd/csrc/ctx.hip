// Contact context: the device buffers ContactSolverBase owns in the
// reference (contact.py:444-521) — topology, BVHs, CCD positions and
// candidate lists — and the entry points that drive the culling and CCD
// kernels with them.
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/ipc_warp_amd.h"
#include "cull.hpp"
#include "forest.hpp"
#include "ipc_kernels.hpp"

using namespace ipcd;

struct ipc_ctx {
  int nv = 0, ne = 0, nf = 0;
  DevBuf edges, tris, body;
  Bvh edge_bvh, tri_bvh;
  PointOrder vorder;       // Morton order of the vertices: the PT query order
  DevBuf x1;               // CCD end positions x - dx
  DevBuf cand_ee, cand_pt; // int4 candidate quadruples
  DevBuf cand_ids_ee, cand_ids_pt;  // int2 candidate primitive ids (new_intersections)
  int cap_ee = 0, cap_pt = 0;
  // dispatch order of the paired DCD and CCD query launches (QueryOrder)
  DevBuf qcost_dcd, qperm_dcd, qcost_ccd, qperm_ccd;
  // fixed per-body BVHs (ipc_ctx_init_fixed_bodies)
  int nb = 0, max_verts = 0, cap_pairs = 0, cap_tasks_ee = 0, cap_tasks_pt = 0;
  Forest fe, ft;
  DevBuf vert_base, vert_ids, body_lo, body_hi;  // per-body vertex lists and rest boxes
  DevBuf frames, world_lo, world_hi, pairs, tasks_ee, tasks_pt;
  int fixed_gen = 0;        // generation of the last fixed-body call (singular-frame flag)
  DevBuf small;            // [0] uint64 toi bits, [64..] int counters
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

// The dispatch order of a paired query launch over n_a EE and n_b PT
// queries; the step counts start at zero with the buffers.
hipError_t query_order(DevBuf& cost, DevBuf& perm, int n_a, int n_b, hipStream_t s, QueryOrder& o) {
  o = QueryOrder{};
  const int ba = query_blocks(n_a), bb = query_blocks(n_b);
  if (ba <= 0 || bb <= 0) return hipSuccess;
  const size_t need = sizeof(int) * 4 * size_t(ba + bb);
  const bool fresh = cost.bytes < need;
  hipError_t e = cost.ensure(need);
  if (e == hipSuccess) e = perm.ensure(sizeof(int) * size_t(ba + bb));
  if (e == hipSuccess && fresh) e = launch_zero_i32(cost.as<int>(), 4 * (ba + bb), s);
  if (e == hipSuccess) o = QueryOrder{cost.as<int>(), perm.as<int>(), ba, bb};
  return e;
}

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
  if (e == hipSuccess) e = hipMemset(c->small.p, 0, 256);
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
  c->edge_bvh.release(); c->tri_bvh.release(); c->vorder.release();
  c->x1.release(); c->cand_ee.release(); c->cand_pt.release();
  c->cand_ids_ee.release(); c->cand_ids_pt.release(); c->small.release();
  c->fe.release(); c->ft.release();
  c->vert_base.release(); c->vert_ids.release(); c->body_lo.release(); c->body_hi.release();
  c->frames.release(); c->world_lo.release(); c->world_hi.release(); c->pairs.release();
  c->tasks_ee.release(); c->tasks_pt.release();
  c->qcost_dcd.release(); c->qperm_dcd.release(); c->qcost_ccd.release(); c->qperm_ccd.release();
  if (c->host) (void)hipHostFree(c->host);
  delete c;
  return IPC_OK;
}

namespace {

// The DCD queries (edge_edge_collision, point_triangle_collision,
// contact.py:672-724) over the context's freshly built trees.
QueryLaunch dcd_ee_query(ipc_ctx* c, const float* x, float thickness, int exclude_same_body, int* contacts,
                         int* e0e1, int capacity, int* count) {
  QueryLaunch q{};
  q.mode = Q_DCD_EE;
  q.n_query = c->ne;
  q.order = c->edge_bvh.leaf_id.as<int>();
  q.edges = c->edges.as<int>();
  q.body = c->body.as<int>();
  q.x0 = x;
  q.qlo = c->edge_bvh.node_lo.as<float4>();  // leaves are level 0, from slot 0
  q.qhi = c->edge_bvh.node_hi.as<float4>();
  q.thickness = thickness;
  q.exclude_same_body = exclude_same_body;
  q.self_query = 1;
  q.bvh = &c->edge_bvh;
  q.rec = reinterpret_cast<int4*>(contacts);
  q.ids = reinterpret_cast<int2*>(e0e1);
  q.capacity = capacity;
  q.count = count;
  return q;
}

QueryLaunch dcd_pt_query(ipc_ctx* c, const float* x, float thickness, float buffer, int exclude_same_body,
                         int* contacts, int* e0e1, int capacity, int* count, const int* order) {
  QueryLaunch q{};
  q.mode = Q_DCD_PT;
  q.n_query = c->nv;
  q.order = order;
  q.tris = c->tris.as<int>();
  q.body = c->body.as<int>();
  q.x0 = x;
  q.pad = thickness;  // boxes of 2 (thickness + min(buffer, slack)) in the query
  q.cap = buffer;
  q.thickness = thickness;
  q.exclude_same_body = exclude_same_body;
  q.bvh = &c->tri_bvh;
  q.rec = reinterpret_cast<int4*>(contacts);
  q.ids = reinterpret_cast<int2*>(e0e1);
  q.capacity = capacity;
  q.count = count;
  return q;
}

// The swept candidate queries of edge_edge_toi / point_triangle_toi.
QueryLaunch ccd_ee_query(ipc_ctx* c, const float* x0, const float* x1, int exclude_same_body, int* quads, int* ids,
                         int capacity, int* count) {
  QueryLaunch q{};
  q.mode = Q_CCD_EE;
  q.n_query = c->ne;
  q.order = c->edge_bvh.leaf_id.as<int>();
  q.edges = c->edges.as<int>();
  q.body = c->body.as<int>();
  q.x0 = x0;
  q.x1 = x1;
  q.qlo = c->edge_bvh.node_lo.as<float4>();
  q.qhi = c->edge_bvh.node_hi.as<float4>();
  q.exclude_same_body = exclude_same_body;
  q.self_query = 1;
  q.bvh = &c->edge_bvh;
  q.rec = reinterpret_cast<int4*>(quads);
  q.ids = reinterpret_cast<int2*>(ids);
  q.capacity = capacity;
  q.count = count;
  return q;
}

QueryLaunch ccd_pt_query(ipc_ctx* c, const float* x0, const float* x1, float padding, int exclude_same_body,
                         int* quads, int* ids, int capacity, int* count, const int* order) {
  QueryLaunch q{};
  q.mode = Q_CCD_PT;
  q.n_query = c->nv;
  q.order = order;
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
  return q;
}

// Record lists are written as int4 and id lists as int2.
bool lists_aligned(const int* rec, const int* ids) {
  return (reinterpret_cast<unsigned long long>(rec) & 15ull) == 0 &&
         (reinterpret_cast<unsigned long long>(ids) & 7ull) == 0;
}

BuildPrologue zero_prologue(int* a, int* b = nullptr) {
  BuildPrologue p;
  if (a) p.zero[p.n_zero++] = a;
  if (b) p.zero[p.n_zero++] = b;
  return p;
}

}  // namespace

// edge_edge_collision over a freshly built edge BVH (contact.py:672-704):
// boxes +- (thickness + min(buffer, slack)), accepted pairs into contacts/e0e1 up to
// capacity, total count into *count (device int).
int ipc_detect_collision_ee(ipc_ctx* c, const float* x, float thickness, float buffer, int exclude_same_body,
                            int* contacts, int* e0e1, int capacity, int* count, void* stream) {
  if (!c || !x || !count) return ctx_fail(IPC_EINVAL, "detect_collision_ee: null pointer");
  if (capacity < 0 || (capacity > 0 && !contacts)) return ctx_fail(IPC_EINVAL, "detect_collision_ee: bad capacity");
  if (!lists_aligned(contacts, e0e1))
    return ctx_fail(IPC_EINVAL, "detect_collision_ee: records must be 16-byte and ids 8-byte aligned");
  hipStream_t s = static_cast<hipStream_t>(stream);
  BuildJob job{&c->edge_bvh, BOX_EDGE, c->ne, c->edges.as<int>(), nullptr, thickness, nullptr, 0, buffer};
  int st = ctx_hip(timed(c, IPC_TIMER_BUILD_EDGE, s, [&] {
                     return build_bvhs(&job, 1, c->body.as<int>(), x, zero_prologue(count), s);
                   }),
                   "edge BVH");
  if (st || c->ne == 0) return st;
  QueryLaunch q = dcd_ee_query(c, x, thickness, exclude_same_body, contacts, e0e1, capacity, count);
  return ctx_hip(timed(c, IPC_TIMER_QUERY_DCD_EE, s, [&] { return run_query(q, s); }), "detect_collision_ee query");
}

// point_triangle_collision over a freshly built triangle BVH
// (contact.py:708-724): every vertex queries x_i +- 2 (thickness + buffer).
int ipc_detect_collision_pt(ipc_ctx* c, const float* x, float thickness, float buffer, int exclude_same_body,
                            int* contacts, int* e0e1, int capacity, int* count, void* stream) {
  if (!c || !x || !count) return ctx_fail(IPC_EINVAL, "detect_collision_pt: null pointer");
  if (capacity < 0 || (capacity > 0 && !contacts)) return ctx_fail(IPC_EINVAL, "detect_collision_pt: bad capacity");
  if (!lists_aligned(contacts, e0e1))
    return ctx_fail(IPC_EINVAL, "detect_collision_pt: records must be 16-byte and ids 8-byte aligned");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int n_tri = c->nv > 0 ? c->nf : 0;
  BuildJob job{&c->tri_bvh, BOX_TRI, n_tri, c->tris.as<int>(), nullptr, 0.0f};
  int st = ctx_hip(timed(c, IPC_TIMER_BUILD_TRI, s, [&] {
                     return build_bvhs(&job, 1, c->body.as<int>(), x, zero_prologue(count), s);
                   }),
                   "triangle BVH");
  if (st || n_tri == 0) return st;
  const int* order = nullptr;
  st = ctx_hip(point_order(c->vorder, c->nv, x, s, &order), "detect_collision_pt order");
  if (st) return st;
  QueryLaunch q = dcd_pt_query(c, x, thickness, buffer, exclude_same_body, contacts, e0e1, capacity, count, order);
  return ctx_hip(timed(c, IPC_TIMER_QUERY_DCD_PT, s, [&] { return run_query(q, s); }), "detect_collision_pt query");
}

// detect_collision (contact.py:687-725): both trees in one build launch and
// both queries in one query launch on `stream` (the EE and PT halves share
// the device instead of two streams, whose event fork and join cost more
// host time than the launches they overlap); both counts stay on the
// device. The launches are recorded under the EE timer kinds.
int ipc_detect_collision(ipc_ctx* c, const float* x, float thickness, float buffer, int exclude_same_body, int* ee,
                         int* ee_ids, int ee_capacity, int* ee_count, int* pt, int* pt_ids, int pt_capacity,
                         int* pt_count, void* stream) {
  if (!c || !x || !ee_count || !pt_count) return ctx_fail(IPC_EINVAL, "detect_collision: null pointer");
  if (ee_capacity < 0 || (ee_capacity > 0 && !ee) || pt_capacity < 0 || (pt_capacity > 0 && !pt))
    return ctx_fail(IPC_EINVAL, "detect_collision: bad capacity");
  if (!lists_aligned(ee, ee_ids) || !lists_aligned(pt, pt_ids))
    return ctx_fail(IPC_EINVAL, "detect_collision: records must be 16-byte and ids 8-byte aligned");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int n_tri = c->nv > 0 ? c->nf : 0;
  BuildJob jobs[2] = {{&c->edge_bvh, BOX_EDGE, c->ne, c->edges.as<int>(), nullptr, thickness, nullptr, 0, buffer},
                      {&c->tri_bvh, BOX_TRI, n_tri, c->tris.as<int>(), nullptr, 0.0f}};
  QueryOrder qo;
  int st = ctx_hip(query_order(c->qcost_dcd, c->qperm_dcd, c->ne, n_tri > 0 ? c->nv : 0, s, qo),
                   "detect_collision order buffers");
  if (st) return st;
  st = ctx_hip(timed(c, IPC_TIMER_BUILD_EDGE, s, [&] {
                 return build_bvhs(jobs, 2, c->body.as<int>(), x, zero_prologue(ee_count, pt_count), s, &qo);
               }),
               "detect_collision BVHs");
  if (st) return st;
  const int* order = nullptr;
  if (n_tri > 0) st = ctx_hip(point_order(c->vorder, c->nv, x, s, &order), "detect_collision order");
  if (st) return st;
  QueryLaunch q[2] = {dcd_ee_query(c, x, thickness, exclude_same_body, ee, ee_ids, ee_capacity, ee_count),
                      dcd_pt_query(c, x, thickness, buffer, exclude_same_body, pt, pt_ids, pt_capacity, pt_count,
                                   order)};
  if (n_tri == 0) q[1].n_query = 0;
  return ctx_hip(timed(c, IPC_TIMER_QUERY_DCD_EE, s, [&] { return run_queries(q, 2, s, &qo); }),
                 "detect_collision query");
}

// Edge-triangle crossings of the surface at x: every edge's bounds against
// the triangle BVH, each hit that shares no vertex tested exactly
// (segment_hits_triangle, FP64 signed volumes). A penetration-free state has
// none. (edge, triangle) ids go to `ids` up to capacity; *count (device) is
// the total.
int ipc_count_intersections(ipc_ctx* c, const float* x, int* ids, int capacity, int* count, void* stream) {
  if (!c || !x || !count) return ctx_fail(IPC_EINVAL, "count_intersections: null pointer");
  if (capacity < 0 || (capacity > 0 && !ids)) return ctx_fail(IPC_EINVAL, "count_intersections: bad capacity");
  if (!lists_aligned(nullptr, ids)) return ctx_fail(IPC_EINVAL, "count_intersections: ids must be 8-byte aligned");
  hipStream_t s = static_cast<hipStream_t>(stream);
  int st = ctx_hip(launch_zero_i32(count, 1, s), "count_intersections");
  if (st || c->ne == 0 || c->nf == 0) return st;
  st = ctx_hip(build_bvh(c->tri_bvh, BOX_TRI, c->nf, c->tris.as<int>(), c->body.as<int>(), x, nullptr, 0.0f, s),
               "count_intersections triangle BVH");
  if (st) return st;
  QueryLaunch q{};
  q.mode = Q_ISECT;
  q.n_query = c->ne;
  q.edges = c->edges.as<int>();
  q.tris = c->tris.as<int>();
  q.body = c->body.as<int>();
  q.x0 = x;
  q.bvh = &c->tri_bvh;
  q.ids = reinterpret_cast<int2*>(ids);
  q.capacity = capacity;
  q.count = count;
  return ctx_hip(run_query(q, s), "count_intersections query");
}

// Candidate sets of edge_edge_toi / point_triangle_toi between x0 and x1.
int ipc_ccd_candidates_ee(ipc_ctx* c, const float* x0, const float* x1, float padding, int exclude_same_body,
                          int* quads, int* ids, int capacity, int* count, void* stream) {
  if (!c || !x0 || !x1 || !count) return ctx_fail(IPC_EINVAL, "ccd_candidates_ee: null pointer");
  if (capacity < 0 || (capacity > 0 && !quads)) return ctx_fail(IPC_EINVAL, "ccd_candidates_ee: bad capacity");
  if (!lists_aligned(quads, ids))
    return ctx_fail(IPC_EINVAL, "ccd_candidates_ee: records must be 16-byte and ids 8-byte aligned");
  hipStream_t s = static_cast<hipStream_t>(stream);
  BuildJob job{&c->edge_bvh, BOX_SWEPT_EDGE, c->ne, c->edges.as<int>(), x1, padding};
  int st = ctx_hip(timed(c, IPC_TIMER_BUILD_EDGE, s, [&] {
                     return build_bvhs(&job, 1, c->body.as<int>(), x0, zero_prologue(count), s);
                   }),
                   "swept edge BVH");
  if (st || c->ne == 0) return st;
  QueryLaunch q = ccd_ee_query(c, x0, x1, exclude_same_body, quads, ids, capacity, count);
  return ctx_hip(timed(c, IPC_TIMER_QUERY_CCD_EE, s, [&] { return run_query(q, s); }), "ccd_candidates_ee query");
}

int ipc_ccd_candidates_pt(ipc_ctx* c, const float* x0, const float* x1, float padding, int exclude_same_body,
                          int* quads, int* ids, int capacity, int* count, void* stream) {
  if (!c || !x0 || !x1 || !count) return ctx_fail(IPC_EINVAL, "ccd_candidates_pt: null pointer");
  if (capacity < 0 || (capacity > 0 && !quads)) return ctx_fail(IPC_EINVAL, "ccd_candidates_pt: bad capacity");
  if (!lists_aligned(quads, ids))
    return ctx_fail(IPC_EINVAL, "ccd_candidates_pt: records must be 16-byte and ids 8-byte aligned");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int n_tri = c->nv > 0 ? c->nf : 0;
  BuildJob job{&c->tri_bvh, BOX_SWEPT_TRI, n_tri, c->tris.as<int>(), x1, padding};
  int st = ctx_hip(timed(c, IPC_TIMER_BUILD_TRI, s, [&] {
                     return build_bvhs(&job, 1, c->body.as<int>(), x0, zero_prologue(count), s);
                   }),
                   "swept triangle BVH");
  if (st || n_tri == 0) return st;
  const int* order = nullptr;
  st = ctx_hip(point_order(c->vorder, c->nv, x0, s, &order), "ccd_candidates_pt order");
  if (st) return st;
  QueryLaunch q = ccd_pt_query(c, x0, x1, padding, exclude_same_body, quads, ids, capacity, count, order);
  return ctx_hip(timed(c, IPC_TIMER_QUERY_CCD_PT, s, [&] { return run_query(q, s); }), "ccd_candidates_pt query");
}

// collision_free_step (contact.py:539-585) before its 0.9 factor: x1 = x - dx,
// swept-box candidates for EE and PT, FP64 TOI per candidate, minimum over
// all of them. Five launches on `stream`: x1; both swept trees (the first
// workgroup also clears the candidate counters and the TOI minimum); both
// queries; both TOI lists (the TOI kernels read the candidate counts from
// device memory, so no host round trip is needed between the phases); the
// readback. The launches are recorded under the EE timer kinds. Candidate
// lists live in the context and grow on overflow (one retry, like
// _grow_contact_list). Synchronises `stream` once to read the result;
// writes the raw TOI (1.0 = no impact) to *toi.
int ipc_collision_free_step(ipc_ctx* c, const float* x, const float* dx, float padding, int exclude_same_body,
                            double* toi, void* stream) {
  if (!c || !x || !dx || !toi) return ctx_fail(IPC_EINVAL, "collision_free_step: null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  int st = ctx_hip(c->x1.ensure(sizeof(float) * 3 * (size_t)c->nv + 16), "collision_free_step alloc");
  if (st) return st;
  if (c->cap_ee == 0) c->cap_ee = 1 << 16;
  if (c->cap_pt == 0) c->cap_pt = 1 << 16;
  const int n_tri = c->nv > 0 ? c->nf : 0;
  for (int attempt = 0; attempt < 2; ++attempt) {
    st = ctx_hip(c->cand_ee.ensure(sizeof(int) * 4 * (size_t)c->cap_ee), "collision_free_step alloc");
    if (!st) st = ctx_hip(c->cand_pt.ensure(sizeof(int) * 4 * (size_t)c->cap_pt), "collision_free_step alloc");
    if (st) return st;
    const float* x1 = c->x1.as<float>();
    int* cnt = counters(c);
    // x1 = x - dx: the build reads the end positions as x - dx and its leaf
    // launch writes x1 for the queries and the TOI launch that follow
    BuildJob jobs[2] = {{&c->edge_bvh, BOX_SWEPT_EDGE, c->ne, c->edges.as<int>(), x1, padding, dx, c->nv},
                        {&c->tri_bvh, BOX_SWEPT_TRI, n_tri, c->tris.as<int>(), x1, padding, dx, c->nv}};
    BuildPrologue pro = zero_prologue(cnt + 0, cnt + 1);
    pro.toi = toi_bits(c);
    pro.toi_init = kOneBits;
    QueryOrder qo;
    st = ctx_hip(query_order(c->qcost_ccd, c->qperm_ccd, c->ne, n_tri > 0 ? c->nv : 0, s, qo),
                 "collision_free_step order buffers");
    if (st) return st;
    st = ctx_hip(timed(c, IPC_TIMER_BUILD_EDGE, s, [&] { return build_bvhs(jobs, 2, c->body.as<int>(), x, pro, s, &qo); }),
                 "collision_free_step BVHs");
    if (st) return st;
    const int* order = nullptr;
    if (n_tri > 0) st = ctx_hip(point_order(c->vorder, c->nv, x, s, &order), "collision_free_step order");
    if (st) return st;
    QueryLaunch q[2] = {ccd_ee_query(c, x, x1, exclude_same_body, c->cand_ee.as<int>(), nullptr, c->cap_ee, cnt + 0),
                        ccd_pt_query(c, x, x1, padding, exclude_same_body, c->cand_pt.as<int>(), nullptr, c->cap_pt,
                                     cnt + 1, order)};
    if (n_tri == 0) q[1].n_query = 0;
    st = ctx_hip(timed(c, IPC_TIMER_QUERY_CCD_EE, s, [&] { return run_queries(q, 2, s, &qo); }),
                 "collision_free_step query");
    if (!st)
      st = ctx_hip(timed(c, IPC_TIMER_TOI_EE, s, [&] {
                     return launch_collision_time_pair(x, x1, c->nv, c->cand_ee.as<int>(), cnt + 0,
                                                       c->ne > 0 ? c->cap_ee : 0, c->cand_pt.as<int>(), cnt + 1,
                                                       n_tri > 0 ? c->cap_pt : 0, toi_bits(c), s);
                   }),
                   "collision_free_step TOI");
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

// new_intersections (contact.py:587-670) for one primitive kind: the swept
// candidate query of edge_edge_toi / point_triangle_toi between x0 and x1,
// FP64 TOI per candidate, and every candidate with TOI < 1 appended as
// (record, ids, alpha = TOI) (edge_edge_intersections /
// point_triangle_intersections, contact.py:370-442). The candidate list lives
// in the context and grows on overflow; *count (device) may exceed capacity,
// and then the caller grows its lists and calls again. Synchronises `stream`
// once; writes the raw minimum TOI of this kind (1.0 = none) to *toi.
static int ccd_intersections(ipc_ctx* c, bool ee, const float* x0, const float* x1, float padding,
                             int exclude_same_body, int* records, int* ids, double* alpha, int capacity, int* count,
                             double* toi, void* stream) {
  const char* what = ee ? "ccd_intersections_ee" : "ccd_intersections_pt";
  if (!c || !x0 || !x1 || !count || !toi) return ctx_fail(IPC_EINVAL, what);
  if (capacity < 0 || (capacity > 0 && (!records || !ids || !alpha)))
    return ctx_fail(IPC_EINVAL, ee ? "ccd_intersections_ee: bad capacity or null list"
                                   : "ccd_intersections_pt: bad capacity or null list");
  if ((reinterpret_cast<unsigned long long>(records) & 15ull) || (reinterpret_cast<unsigned long long>(ids) & 7ull) ||
      (reinterpret_cast<unsigned long long>(alpha) & 7ull))
    return ctx_fail(IPC_EINVAL, ee ? "ccd_intersections_ee: records must be 16-byte and ids/alpha 8-byte aligned"
                                   : "ccd_intersections_pt: records must be 16-byte and ids/alpha 8-byte aligned");
  hipStream_t s = static_cast<hipStream_t>(stream);
  int& cap = ee ? c->cap_ee : c->cap_pt;
  if (cap == 0) cap = 1 << 16;
  DevBuf& cq = ee ? c->cand_ee : c->cand_pt;
  DevBuf& ci = ee ? c->cand_ids_ee : c->cand_ids_pt;
  const int k = ee ? 0 : 1;
  for (int attempt = 0; attempt < 2; ++attempt) {
    int st = ctx_hip(cq.ensure(sizeof(int) * 4 * (size_t)cap), what);
    if (!st) st = ctx_hip(ci.ensure(sizeof(int) * 2 * (size_t)cap), what);
    if (!st) st = ctx_hip(launch_fill_u64(toi_bits(c), 1, kOneBits, s), what);
    if (!st) st = ctx_hip(launch_zero_i32(count, 1, s), what);
    if (st) return st;
    int* cnt = counters(c) + k;
    st = ee ? ipc_ccd_candidates_ee(c, x0, x1, padding, exclude_same_body, cq.as<int>(), ci.as<int>(), cap, cnt, stream)
            : ipc_ccd_candidates_pt(c, x0, x1, padding, exclude_same_body, cq.as<int>(), ci.as<int>(), cap, cnt,
                                    stream);
    if (st) return st;
    st = ctx_hip(timed(c, ee ? IPC_TIMER_TOI_EE : IPC_TIMER_TOI_PT, s, [&] {
                   return launch_record_intersections(ee, x0, x1, c->nv, cq.as<int>(), ci.as<int>(), cnt, cap, records,
                                                      ids, alpha, capacity, count, toi_bits(c), s);
                 }),
                 what);
    if (!st) st = ctx_hip(hipMemcpyAsync(c->host, c->small.p, 128, hipMemcpyDeviceToHost, s), what);
    if (!st) st = ctx_hip(hipStreamSynchronize(s), what);
    if (st) return st;
    unsigned long long bits;
    std::memcpy(&bits, c->host, 8);
    int n;
    std::memcpy(&n, static_cast<char*>(c->host) + 64 + 4 * k, 4);
    c->last_counts[k] = n;
    if (n > cap) {
      while (cap < n) cap *= 2;
      continue;
    }
    double t;
    std::memcpy(&t, &bits, 8);
    *toi = t;
    return IPC_OK;
  }
  return ctx_fail(IPC_EINVAL, ee ? "ccd_intersections_ee: candidate list kept overflowing"
                                 : "ccd_intersections_pt: candidate list kept overflowing");
}

int ipc_ccd_intersections_ee(ipc_ctx* c, const float* x0, const float* x1, float padding, int exclude_same_body,
                             int* records, int* ids, double* alpha, int capacity, int* count, double* toi,
                             void* stream) {
  return ccd_intersections(c, true, x0, x1, padding, exclude_same_body, records, ids, alpha, capacity, count, toi,
                           stream);
}

int ipc_ccd_intersections_pt(ipc_ctx* c, const float* x0, const float* x1, float padding, int exclude_same_body,
                             int* records, int* ids, double* alpha, int capacity, int* count, double* toi,
                             void* stream) {
  return ccd_intersections(c, false, x0, x1, padding, exclude_same_body, records, ids, alpha, capacity, count, toi,
                           stream);
}

// ------------------------------------------------------ fixed body BVHs ----
// Counters in the small buffer: [0], [1] CCD candidates (EE, PT), [2] body
// pairs, [3], [4] tasks (EE, PT), [5] singular-frame flag (the generation of
// the last call that met a near-singular frame).
namespace {

// A new generation for a fixed-body call (never 0, the flag's initial value).
int next_gen(ipc_ctx* c) {
  c->fixed_gen = c->fixed_gen >= 0x3fffffff ? 1 : c->fixed_gen + 1;
  return c->fixed_gen;
}

int grow_pow2(int cap, int need) {
  long long c = cap > 0 ? cap : 1;
  while (c < need) c *= 2;
  return c > (1ll << 30) ? (1 << 30) : int(c);
}

int ensure_fixed_buffers(ipc_ctx* c) {
  int st = ctx_hip(c->pairs.ensure(sizeof(int2) * (size_t)c->cap_pairs), "fixed buffers");
  if (!st) st = ctx_hip(c->tasks_ee.ensure(sizeof(int2) * (size_t)c->cap_tasks_ee), "fixed buffers");
  if (!st) st = ctx_hip(c->tasks_pt.ensure(sizeof(int2) * (size_t)c->cap_tasks_pt), "fixed buffers");
  return st;
}

// Body frames, body pairs and tasks of one fixed-body step (DCD: dq =
// NULL): three launches, the first of which runs `pro` (the caller's
// counters; the pair and task counters are added here). The singular-frame
// flag (counter 5) is set to `gen` instead of being cleared.
int fixed_cull(ipc_ctx* c, const float* x0, const float* x1, const float* q, const float* dq, const int* fixed,
               float pair_pad, float task_pad, BuildPrologue pro, int gen, hipStream_t s) {
  int* cnt = counters(c);
  return ctx_hip(timed(c, IPC_TIMER_FIXED_CULL, s, [&] {
                   FixedStep f{c->nb, q, dq, c->body_lo.as<float4>(), c->body_hi.as<float4>(), c->world_lo.as<float4>(),
                               c->world_hi.as<float4>(), c->frames.as<BodyFrame>(), cnt + 5, gen, pro};
                   hipError_t e = launch_body_frames(f, s);
                   if (e == hipSuccess)
                     e = launch_body_pairs(c->nb, c->world_lo.as<float4>(), c->world_hi.as<float4>(), pair_pad,
                                           c->pairs.as<int2>(), cnt + 2, c->cap_pairs, s);
                   if (e != hipSuccess) return e;
                   TaskLaunch t{};
                   t.ccd = dq != nullptr;
                   t.pairs = c->pairs.as<int2>();
                   t.n_pairs = cnt + 2;
                   t.pair_cap = c->cap_pairs;
                   t.singular = cnt + 5;
                   t.gen = gen;
                   t.x0 = x0;
                   t.x1 = x1;
                   t.world_lo = c->world_lo.as<float4>();
                   t.world_hi = c->world_hi.as<float4>();
                   t.pad = task_pad;
                   t.edges = c->edges.as<int>();
                   t.edesc = c->fe.desc.as<TreeDesc>();
                   t.eleaf_id = c->fe.leaf_id.as<int>();
                   t.max_edges = c->fe.max_leaves;
                   t.fixed = fixed;
                   t.vert_base = c->vert_base.as<int>();
                   t.vert_ids = c->vert_ids.as<int>();
                   t.max_verts = c->ft.n_leaves > 0 ? c->max_verts : 0;
                   t.tasks_ee = c->tasks_ee.as<int2>();
                   t.count_ee = cnt + 3;
                   t.cap_ee = c->cap_tasks_ee;
                   t.tasks_pt = c->tasks_pt.as<int2>();
                   t.count_pt = cnt + 4;
                   t.cap_pt = c->cap_tasks_pt;
                   return launch_fixed_tasks(t, s);
                 }),
                 "fixed cull");
}

FixedQuery fixed_query(ipc_ctx* c, int mode, const float* x0, const float* x1, float thickness, float pad, int4* rec,
                       int2* ids, int capacity, int* count) {
  bool ee = mode == Q_DCD_EE || mode == Q_CCD_EE;
  FixedQuery fq{};
  fq.mode = mode;
  fq.tasks = ee ? c->tasks_ee.as<int2>() : c->tasks_pt.as<int2>();
  fq.task_count = counters(c) + (ee ? 3 : 4);
  fq.task_cap = ee ? c->cap_tasks_ee : c->cap_tasks_pt;
  fq.edges = c->edges.as<int>();
  fq.tris = c->tris.as<int>();
  fq.x0 = x0;
  fq.x1 = x1;
  fq.frames = c->frames.as<BodyFrame>();
  fq.forest = ee ? &c->fe : &c->ft;
  fq.thickness = thickness;
  fq.pad = pad;
  fq.rec = rec;
  fq.ids = ids;
  fq.capacity = capacity;
  fq.count = count;
  return fq;
}

// Grows the pair and task lists after an overflow; true if anything grew.
bool grow_fixed(ipc_ctx* c, const int* host_cnt) {
  bool grow = false;
  if (host_cnt[2] > c->cap_pairs) { c->cap_pairs = grow_pow2(c->cap_pairs, host_cnt[2]); grow = true; }
  if (host_cnt[3] > c->cap_tasks_ee) { c->cap_tasks_ee = grow_pow2(c->cap_tasks_ee, host_cnt[3]); grow = true; }
  if (host_cnt[4] > c->cap_tasks_pt) { c->cap_tasks_pt = grow_pow2(c->cap_tasks_pt, host_cnt[4]); grow = true; }
  return grow;
}

}  // namespace

int ipc_ctx_init_fixed_bodies(ipc_ctx* c, int n_bodies, const float* xcs) {
  if (!c || !xcs) return ctx_fail(IPC_EINVAL, "ctx_init_fixed_bodies: null pointer");
  if (n_bodies <= 0) return ctx_fail(IPC_EINVAL, "ctx_init_fixed_bodies: n_bodies must be > 0");
  std::vector<int> body(c->nv), edges(2 * (size_t)c->ne), tris(3 * (size_t)c->nf);
  std::vector<float> rest(3 * (size_t)c->nv);
  hipError_t e = hipSuccess;
  if (c->nv) e = hipMemcpy(body.data(), c->body.p, sizeof(int) * body.size(), hipMemcpyDeviceToHost);
  if (e == hipSuccess && c->ne) e = hipMemcpy(edges.data(), c->edges.p, sizeof(int) * edges.size(), hipMemcpyDeviceToHost);
  if (e == hipSuccess && c->nf) e = hipMemcpy(tris.data(), c->tris.p, sizeof(int) * tris.size(), hipMemcpyDeviceToHost);
  if (e == hipSuccess && c->nv) e = hipMemcpy(rest.data(), xcs, sizeof(float) * rest.size(), hipMemcpyDefault);
  if (e != hipSuccess) return ctx_hip(e, "ctx_init_fixed_bodies");
  for (int b : body)
    if (b < 0 || b >= n_bodies) return ctx_fail(IPC_EINVAL, "ctx_init_fixed_bodies: a body id is outside [0, n_bodies)");
  e = build_forest(c->fe, n_bodies, c->ne, 2, edges.data(), body.data(), rest.data());
  if (e == hipSuccess) e = build_forest(c->ft, n_bodies, c->nf, 3, tris.data(), body.data(), rest.data());
  if (e != hipSuccess) return ctx_hip(e, "ctx_init_fixed_bodies: forest");
  std::vector<int> vbase(n_bodies + 1), vids(std::max(c->nv, 1));
  std::vector<float> lo(4 * (size_t)n_bodies), hi(4 * (size_t)n_bodies);
  body_vertex_lists(n_bodies, c->nv, body.data(), rest.data(), vbase.data(), vids.data(), lo.data(), hi.data());
  int max_verts = 0;
  for (int b = 0; b < n_bodies; ++b) max_verts = std::max(max_verts, vbase[b + 1] - vbase[b]);
  e = c->vert_base.ensure(sizeof(int) * vbase.size());
  if (e == hipSuccess) e = c->vert_ids.ensure(sizeof(int) * vids.size());
  if (e == hipSuccess) e = c->body_lo.ensure(sizeof(float) * lo.size());
  if (e == hipSuccess) e = c->body_hi.ensure(sizeof(float) * hi.size());
  if (e == hipSuccess) e = c->frames.ensure(sizeof(BodyFrame) * (size_t)n_bodies);
  if (e == hipSuccess) e = c->world_lo.ensure(sizeof(float4) * (size_t)n_bodies);
  if (e == hipSuccess) e = c->world_hi.ensure(sizeof(float4) * (size_t)n_bodies);
  if (e == hipSuccess) e = hipMemcpy(c->vert_base.p, vbase.data(), sizeof(int) * vbase.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(c->vert_ids.p, vids.data(), sizeof(int) * vids.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(c->body_lo.p, lo.data(), sizeof(float) * lo.size(), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(c->body_hi.p, hi.data(), sizeof(float) * hi.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) return ctx_hip(e, "ctx_init_fixed_bodies");
  c->nb = n_bodies;
  c->max_verts = max_verts;
  long long all_pairs = (long long)n_bodies * (n_bodies - 1) / 2;
  c->cap_pairs = int(std::max(1ll, std::min(all_pairs, 1ll << 20)));
  c->cap_tasks_ee = c->cap_tasks_pt = 1 << 16;
  return ensure_fixed_buffers(c);
}

// FixedBodyContactSolver.detect_collision (fixed_bvh_contacts.py:758-827).
int ipc_detect_collision_fixed(ipc_ctx* c, const float* x, const float* q, float thickness, int* ee, int* ee_ids,
                               int ee_capacity, int* ee_count, int* pt, int* pt_ids, int pt_capacity, int* pt_count,
                               void* stream) {
  if (!c || !x || !q || !ee_count || !pt_count) return ctx_fail(IPC_EINVAL, "detect_collision_fixed: null pointer");
  if (c->nb == 0) return ctx_fail(IPC_EINVAL, "detect_collision_fixed: call ipc_ctx_init_fixed_bodies first");
  if (ee_capacity < 0 || pt_capacity < 0 || (ee_capacity > 0 && !ee) || (pt_capacity > 0 && !pt))
    return ctx_fail(IPC_EINVAL, "detect_collision_fixed: bad capacity or null list");
  if ((reinterpret_cast<unsigned long long>(ee) & 15ull) || (reinterpret_cast<unsigned long long>(pt) & 15ull))
    return ctx_fail(IPC_EINVAL, "detect_collision_fixed: contact lists must be 16-byte aligned");
  if (!(thickness > 0.0f)) return ctx_fail(IPC_EINVAL, "detect_collision_fixed: thickness must be > 0");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const float pad = 2.0f * thickness;
  for (int attempt = 0; attempt < 4; ++attempt) {
    int st = ensure_fixed_buffers(c);
    if (st) return st;
    const int gen = next_gen(c);
    BuildPrologue pro;
    pro.zero[0] = ee_count;
    pro.zero[1] = pt_count;
    pro.n_zero = 2;
    pro.range = counters(c) + 2;  // pairs, tasks EE, tasks PT
    pro.n_range = 3;
    st = fixed_cull(c, x, nullptr, q, nullptr, nullptr, pad, pad, pro, gen, s);
    if (st) return st;
    st = ctx_hip(timed(c, IPC_TIMER_FIXED_QUERY_DCD, s, [&] {
                   return launch_fixed_queries(
                       fixed_query(c, Q_DCD_EE, x, nullptr, thickness, 0.0f, reinterpret_cast<int4*>(ee),
                                   reinterpret_cast<int2*>(ee_ids), ee_capacity, ee_count),
                       fixed_query(c, Q_DCD_PT, x, nullptr, thickness, 0.0f, reinterpret_cast<int4*>(pt),
                                   reinterpret_cast<int2*>(pt_ids), pt_capacity, pt_count),
                       s);
                 }),
                 "detect_collision_fixed query");
    if (!st) st = ctx_hip(hipMemcpyAsync(c->host, c->small.p, 128, hipMemcpyDeviceToHost, s), "detect_collision_fixed");
    if (!st) st = ctx_hip(hipStreamSynchronize(s), "detect_collision_fixed sync");
    if (st) return st;
    int hc[6];
    std::memcpy(hc, static_cast<char*>(c->host) + 64, sizeof(hc));
    if (hc[5] == gen) {
      // A (near-)singular body frame: world-space DCD instead (same pair set).
      st = ipc_detect_collision_ee(c, x, thickness, thickness, 1, ee, ee_ids, ee_capacity, ee_count, stream);
      if (!st) st = ipc_detect_collision_pt(c, x, thickness, thickness, 1, pt, pt_ids, pt_capacity, pt_count, stream);
      return st;
    }
    if (!grow_fixed(c, hc)) return IPC_OK;
  }
  return ctx_fail(IPC_EINVAL, "detect_collision_fixed: task lists kept overflowing");
}

// FixedBodyContactSolver.collision_free_step (fixed_bvh_contacts.py:710-756)
// before its 0.9 factor.
int ipc_collision_free_step_fixed(ipc_ctx* c, const float* x, const float* dx, const float* q, const float* dq,
                                  const int* fixed_bodies, float padding, double* toi, int* used_fallback,
                                  void* stream) {
  if (!c || !x || !dx || !q || !dq || !toi) return ctx_fail(IPC_EINVAL, "collision_free_step_fixed: null pointer");
  if (c->nb == 0) return ctx_fail(IPC_EINVAL, "collision_free_step_fixed: call ipc_ctx_init_fixed_bodies first");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (used_fallback) *used_fallback = 0;
  int st = ctx_hip(c->x1.ensure(sizeof(float) * 3 * (size_t)c->nv + 16), "collision_free_step_fixed alloc");
  if (st) return st;
  if (c->cap_ee == 0) c->cap_ee = 1 << 16;
  if (c->cap_pt == 0) c->cap_pt = 1 << 16;
  for (int attempt = 0; attempt < 4; ++attempt) {
    st = ensure_fixed_buffers(c);
    if (!st) st = ctx_hip(c->cand_ee.ensure(sizeof(int) * 4 * (size_t)c->cap_ee), "collision_free_step_fixed alloc");
    if (!st) st = ctx_hip(c->cand_pt.ensure(sizeof(int) * 4 * (size_t)c->cap_pt), "collision_free_step_fixed alloc");
    if (!st) st = ctx_hip(launch_sub(3 * c->nv, x, dx, c->x1.as<float>(), s), "collision_free_step_fixed");
    if (st) return st;
    int* cnt = counters(c);
    const float* x1 = c->x1.as<float>();
    const int gen = next_gen(c);
    BuildPrologue pro;
    pro.range = cnt;  // candidates EE, PT, pairs, tasks EE, tasks PT
    pro.n_range = 5;
    pro.toi = toi_bits(c);
    pro.toi_init = kOneBits;
    st = fixed_cull(c, x, x1, q, dq, fixed_bodies, padding, padding, pro, gen, s);
    if (st) return st;
    st = ctx_hip(timed(c, IPC_TIMER_FIXED_QUERY_CCD, s, [&] {
                   return launch_fixed_queries(
                       fixed_query(c, Q_CCD_EE, x, x1, 0.0f, padding, c->cand_ee.as<int4>(), nullptr, c->cap_ee,
                                   cnt + 0),
                       fixed_query(c, Q_CCD_PT, x, x1, 0.0f, padding, c->cand_pt.as<int4>(), nullptr, c->cap_pt,
                                   cnt + 1),
                       s);
                 }),
                 "collision_free_step_fixed query");
    if (!st)
      st = ctx_hip(timed(c, IPC_TIMER_TOI_EE, s, [&] {
                     return launch_collision_time_pair(x, x1, c->nv, c->cand_ee.as<int>(), cnt + 0, c->cap_ee,
                                                       c->cand_pt.as<int>(), cnt + 1, c->cap_pt, toi_bits(c), s);
                   }),
                   "collision_free_step_fixed TOI");
    if (!st) st = ctx_hip(hipMemcpyAsync(c->host, c->small.p, 128, hipMemcpyDeviceToHost, s), "collision_free_step_fixed");
    if (!st) st = ctx_hip(hipStreamSynchronize(s), "collision_free_step_fixed sync");
    if (st) return st;
    int hc[6];
    std::memcpy(hc, static_cast<char*>(c->host) + 64, sizeof(hc));
    if (hc[5] == gen) {
      // min |det A| < 1e-8 at either end: the world-space swept CCD
      // (fixed_bvh_contacts.py:733-735); the fixed path only pairs
      // different bodies, so the fallback excludes same-body pairs too.
      if (used_fallback) *used_fallback = 1;
      return ipc_collision_free_step(c, x, dx, padding, 1, toi, stream);
    }
    bool grow = grow_fixed(c, hc);
    if (hc[0] > c->cap_ee) { c->cap_ee = grow_pow2(c->cap_ee, hc[0]); grow = true; }
    if (hc[1] > c->cap_pt) { c->cap_pt = grow_pow2(c->cap_pt, hc[1]); grow = true; }
    if (grow) continue;
    c->last_counts[0] = hc[0];
    c->last_counts[1] = hc[1];
    unsigned long long bits;
    std::memcpy(&bits, c->host, 8);
    double t;
    std::memcpy(&t, &bits, 8);
    *toi = t;
    return IPC_OK;
  }
  return ctx_fail(IPC_EINVAL, "collision_free_step_fixed: candidate lists kept overflowing");
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

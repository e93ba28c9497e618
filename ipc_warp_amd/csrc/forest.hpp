// Fixed per-body BVHs for affine bodies (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include "cull.hpp"

namespace ipcd {

// One body's tree inside a Forest: its primitives are leaf_id[leaf_base ..
// leaf_base + n_leaves) in Morton order of the body's rest frame. Level L
// (0 = leaves) holds count[L] entries in the node arrays from slot
// offset[L], padded with empty boxes to 8 count[L+1] slots (one slot for the
// root level `top`; top >= 1 when n_leaves > 0).
struct TreeDesc {
  int leaf_base, n_leaves, top, pad;
  int count[kMaxLevels];
  int offset[kMaxLevels];
};

// All bodies' trees over one primitive kind (edges or triangles), built once
// from rest coordinates (fp32 primitive bounds, no padding).
struct Forest {
  int n_trees = 0, n_leaves = 0, n_nodes = 0, max_leaves = 0;
  DevBuf desc;              // TreeDesc per body
  DevBuf leaf_id;           // primitive id per leaf, body by body
  DevBuf node_lo, node_hi;  // float4 per slot of every level of every tree
  DevBuf slot_vid;          // int4 per slot: a leaf's vertex ids (v2 = -1 for an edge) and primitive id, else -1
  void release();
};

// Per-body frame of a step (FP64): inverse of A, translation p, Frobenius
// norm of A^-1 (the padding scale) and, for CCD, the motion extent of the
// body's rest box in its own start frame.
struct BodyFrame {
  double inv[9];
  double p[3];
  double fro;
  double motion[3];
};

// Builds F from host arrays: prims int [n_prims][k] (k = 2 edges, 3
// triangles), vertex body ids and rest positions float [V][3]. A primitive
// belongs to the body of its first vertex.
hipError_t build_forest(Forest& F, int n_bodies, int n_prims, int k, const int* prims, const int* vbody,
                        const float* rest);

// Per-body vertex lists in Morton order of the rest positions (offsets
// [n_bodies + 1], ids [V]) and rest AABBs of each body's vertices.
void body_vertex_lists(int n_bodies, int nv, const int* vbody, const float* rest, int* offsets, int* ids,
                       float* lo, float* hi);

struct FixedStep {
  int n_bodies;
  const float* q;    // float [n_bodies][4][3]: p, A[:,0], A[:,1], A[:,2]
  const float* dq;   // CCD only: q1 = q - dq
  const float4* body_lo;
  const float4* body_hi;
  float4* world_lo;  // out: world (DCD) or swept (CCD) body boxes
  float4* world_hi;
  BodyFrame* frames;
  // Set to `gen` when a body's |det A| (either end for CCD) < 1e-8. The
  // caller gives each call a new generation instead of clearing the flag,
  // which the same launch may set.
  int* singular;
  int gen;
  BuildPrologue pro;  // run by the frames launch (the call's first)
};
hipError_t launch_body_frames(const FixedStep& f, hipStream_t s);

hipError_t launch_body_pairs(int n_bodies, const float4* lo, const float4* hi, float pad, int2* pairs, int* count,
                             int cap, hipStream_t s);

struct TaskLaunch {
  bool ccd;
  const int2* pairs;
  const int* n_pairs;
  int pair_cap;
  const int* singular;  // == gen: this call fell back, no tasks
  int gen;
  const float* x0;
  const float* x1;
  const float4* world_lo;
  const float4* world_hi;
  float pad;
  // edges: per-body edge lists are the edge forest's leaves
  const int* edges;
  const TreeDesc* edesc;
  const int* eleaf_id;
  int max_edges;
  const int* fixed;  // CCD EE: a fixed source body is swapped to be the target
  // points: per-body vertex lists
  const int* vert_base;
  const int* vert_ids;
  int max_verts;
  int2* tasks_ee;
  int* count_ee;
  int cap_ee;
  int2* tasks_pt;
  int* count_pt;
  int cap_pt;
};
hipError_t launch_fixed_tasks(const TaskLaunch& t, hipStream_t s);

struct FixedQuery {
  int mode;  // QueryMode
  const int2* tasks;
  const int* task_count;
  int task_cap;
  const int* edges;
  const int* tris;
  const float* x0;
  const float* x1;
  const BodyFrame* frames;
  const Forest* forest;
  float thickness;  // DCD: acceptance dist < 2 thickness, local padding 2 thickness |A^-1|_F
  float pad;        // CCD padding
  int4* rec;
  int2* ids;
  int capacity;
  int* count;
};
hipError_t launch_fixed_query(const FixedQuery& q, hipStream_t s);
// The EE and PT queries of one call (modes Q_DCD_EE + Q_DCD_PT or
// Q_CCD_EE + Q_CCD_PT) in one launch.
hipError_t launch_fixed_queries(const FixedQuery& ee, const FixedQuery& pt, hipStream_t s);

}  // namespace ipcd

/* libipc_warp_amd — MI355X (gfx950) IPC contact inner loop, C ABI.
 *
 * Drop-in for the contact kernels of Luke-Skycrawler/ipc_warp. Every entry
 * point takes raw device pointers, sizes and scalars plus a hipStream_t
 * (passed as void*, NULL = the default stream), enqueues its work on that
 * stream and returns without synchronising, like a Warp launch.
 *
 * Data layout (the reference's own buffers):
 *   x, xcs      float32 [V][3]       vertex positions, rest positions
 *               (NewtonState.x, Soup.xcs — fem/params.py:43-50,
 *                fem/geometry.py:165-171; x aliases Soup.x_transformed)
 *   contacts    int32 [n][4]         XConstraint.a1a2b1b2 (contact.py:61-67);
 *               EE (a0, a1, b0, b1), PT (p, t1, t0, t2) as point_triangle_
 *               collision appends it (contact.py:336)
 *   rows, cols  int32 [16 * total], vals float32 [16 * total][3][3]
 *               Triplets (fem/fem.py:15-19), block (offset + i) * 16 + ii * 4 + jj
 *   b           float32 [V][3]       force accumulator (the reference's `b`
 *               stores -gradient; dynamic_contacts.py:193,231)
 * The per-contact XConstraint.l0 is always 2 * thickness (contact.py:180,
 * 195), so it is a scalar argument here. n_vertices is V: a contact or
 * candidate naming a vertex outside [0, V) is treated as inactive (zero
 * blocks, no force, no energy, TOI 1) instead of being read out of bounds.
 *
 * Status codes: 0 = success; IPC_EINVAL for a bad argument; a negative
 * hipError_t value (-code) for a HIP failure; a positive value from a
 * pair-producing entry point is the capacity it needed (grow and call again,
 * as ContactSolverBase._grow_contact_list does, contact.py:523-537).
 * ipc_last_error() describes the last failure on the calling thread.
 *
 * Alignment: contact, candidate and record lists (int32 [n][4]) and the
 * triplet arrays are read and written as 16-byte vectors and must be
 * 16-byte aligned; primitive-id lists (int32 [n][2]) 8-byte aligned. A
 * misaligned pointer is refused with IPC_EINVAL before any launch.
 */
#ifndef IPC_WARP_AMD_H
#define IPC_WARP_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

#define IPC_OK 0
#define IPC_EINVAL (-100000)

/* ABI version of this header (bumped on any signature change). */
int ipc_abi_version(void);
/* Human-readable description of the last error on this thread. */
const char* ipc_last_error(void);

/* ------------------------------------------------- barrier Hessian ---
 * Replaces the launch of contact_hessian_ee(states, soup, contacts, triplets,
 * b, kappa), dim = n_contacts (dynamic_contacts.py:151-207, launched at
 * dynamic_contacts.py:287 and affine_body_dynamics.py:470-476).
 * For every contact whose fp32 edge-edge distance is < l0: the 16 blocks of
 * the PSD-filtered 12x12 barrier Hessian go to triplet (offset + i) * 16 +
 * ii * 4 + jj and the force is atomically added into b. Inactive contacts get
 * zero blocks with row = col = 0. xcs supplies the rest edges of the
 * mollifier threshold. */
int ipc_contact_hessian_ee(const float* x, const float* xcs, int n_vertices, const int* contacts, int n_contacts, float l0,
                           double kappa, int offset, int* rows, int* cols, float* vals, float* b, void* stream);

/* ------------------------------------------------- barrier Hessian ---
 * Replaces contact_hessian_pt(states, soup, contacts, triplets, b, offset,
 * kappa), dim = n_contacts (dynamic_contacts.py:209-246, launched at
 * dynamic_contacts.py:288 with offset = n_ee). */
int ipc_contact_hessian_pt(const float* x, int n_vertices, const int* contacts, int n_contacts, float l0, double kappa, int offset,
                           int* rows, int* cols, float* vals, float* b, void* stream);

/* ------------------------------------- barrier Hessian, both lists ---
 * Replaces the contact part of RodComplexBC.compute_A
 * (dynamic_contacts.py:281-288): contact_hessian_ee over the n_ee EE
 * contacts at triplet offset 0, then contact_hessian_pt over the n_pt PT
 * contacts at offset n_ee, in one launch. rows/cols/vals hold
 * 16 (n_ee + n_pt) blocks; results equal the two calls above. */
int ipc_collision_hessian(const float* x, const float* xcs, int n_vertices, const int* ee, int n_ee, const int* pt,
                          int n_pt, float l0, double kappa, int* rows, int* cols, float* vals, float* b, void* stream);

/* ------------------------------------------------- barrier energy ---
 * Replace contact_energy_ee / contact_energy_pt(states, soup, contacts, e,
 * kappa) (dynamic_contacts.py:374-398, launched at :304-305). The sum over
 * active contacts is accumulated in FP64 and ADDED to energy[0] (a device
 * double the caller zeroes, as the reference zeroes `e`). */
int ipc_contact_energy_ee(const float* x, const float* xcs, int n_vertices, const int* contacts, int n_contacts, float l0,
                          double kappa, double* energy, void* stream);
int ipc_contact_energy_pt(const float* x, int n_vertices, const int* contacts, int n_contacts, float l0, double kappa,
                          double* energy, void* stream);

/* -------------------------------------- barrier energy, both lists ---
 * Replaces the kernels of compute_collision_energy
 * (dynamic_contacts.py:300-305): energy[0] is set to zero (the reference's
 * wp.zeros), then the EE and PT sums are added, in FP64. The caller reads
 * energy[0] (the reference's host transfer, :306-307). */
int ipc_collision_energy(const float* x, const float* xcs, int n_vertices, const int* ee, int n_ee, const int* pt,
                         int n_pt, float l0, double kappa, double* energy, void* stream);

/* ------------------------------------------------ fp64 scalars ---
 * The four contact kernels above with the reference's fp64 scalars
 * (scalar_types.py:6-28 with USE_FLOAT32 = False: vec3d positions, mat33d
 * blocks, vec3d forces): x / xcs double [V][3], vals double [16 n][3][3]
 * (16-byte aligned), b double [V][3], l0 double. The distance filter and the
 * feature dispatch run in double; the layout of the triplets and the zero
 * blocks of inactive contacts are as in fp32. */
int ipc_contact_hessian_ee_f64(const double* x, const double* xcs, int n_vertices, const int* contacts,
                               int n_contacts, double l0, double kappa, int offset, int* rows, int* cols, double* vals,
                               double* b, void* stream);
int ipc_contact_hessian_pt_f64(const double* x, int n_vertices, const int* contacts, int n_contacts, double l0,
                               double kappa, int offset, int* rows, int* cols, double* vals, double* b, void* stream);
int ipc_contact_energy_ee_f64(const double* x, const double* xcs, int n_vertices, const int* contacts, int n_contacts,
                              double l0, double kappa, double* energy, void* stream);
int ipc_contact_energy_pt_f64(const double* x, int n_vertices, const int* contacts, int n_contacts, double l0,
                              double kappa, double* energy, void* stream);
/* pt_collision_time / ee_collision_time with fp64 scalars: x0, x1 double
 * [V][3] (vec3d), the same FP64 arithmetic as ipc_collision_time_ee/pt.
 * The candidate culling keeps fp32 positions in both modes: the reference
 * casts to wp.vec3 for its boxes and distance tests (contact.py:207-229,
 * 313-336), so fp64 callers pass fp32 casts of x to the culling and DCD
 * entry points. */
int ipc_collision_time_ee_f64(const double* x0, const double* x1, int n_vertices, const int* quads, int n,
                              double* toi, unsigned long long* toi_min, void* stream);
int ipc_collision_time_pt_f64(const double* x0, const double* x1, int n_vertices, const int* quads, int n,
                              double* toi, unsigned long long* toi_min, void* stream);

/* ----------------------------------------------- contact distance ---
 * Replace get_contact_points / get_contact_points_pt (contact.py:148-161)
 * over fetch_dist_v0v1 / fetch_dist_v0v1_pt (contact.py:88-145): midpoint of
 * the two closest points (points, float [n][3]) and the fp32 distance
 * (dists, float [n]). features (int [n], optional) receives the feature the
 * Hessian kernels dispatch on: PT 0..6 as closest_point_triangle's type
 * (contact.py:230-302); EE kind * 16 + slot0 * 4 + slot1 with kind 0 PP,
 * 1 PE, 3 EE (dynamic_contacts.py:48-108; xcs may be NULL, which disables
 * the near-parallel fallback). Any output pointer may be NULL. */
int ipc_contact_points_ee(const float* x, const float* xcs, int n_vertices, const int* contacts, int n_contacts, float* points,
                          float* dists, int* features, void* stream);
int ipc_contact_points_pt(const float* x, int n_vertices, const int* contacts, int n_contacts, float* points, float* dists,
                          int* features, void* stream);

/* ------------------------------------------------ CCD time of impact ---
 * pt_collision_time / ee_collision_time (ccd/ccd.py:304-357) over a list of
 * candidate vertex quadruples (int [n][4]: EE a0, a1, b0, b1; PT p, t0, t1,
 * t2, the argument order of edge_edge_toi / point_triangle_toi,
 * ccd/ccd.py:360-427) between positions x0 (t = 0) and x1 (t = 1).
 * toi (double [n], optional) receives each candidate's TOI (1.0 = none).
 * toi_min (optional) is a device uint64 holding the bit pattern of a
 * non-negative double; every TOI < 1 is folded in with atomicMin, as the
 * reference's wp.atomic_min(toi, 0, t). Initialise it to the bits of 1.0. */
int ipc_collision_time_ee(const float* x0, const float* x1, int n_vertices, const int* quads, int n, double* toi,
                          unsigned long long* toi_min, void* stream);
int ipc_collision_time_pt(const float* x0, const float* x1, int n_vertices, const int* quads, int n, double* toi,
                          unsigned long long* toi_min, void* stream);
/* The reference's distance-based conservative advancement
 * (conservative_ee_toi / conservative_pt_toi, ccd/ccd.py:86-163; defined
 * there but not called on its path, ccd/ccd.py:388,421): per candidate, in
 * FP64, steps of 0.9 * distance / (motion bound relative to the mean
 * motion) until the distance drops below a tenth of the initial one or the
 * steps reach 1. Same arguments and outputs as ipc_collision_time_ee/pt
 * (quads EE (a0, a1, b0, b1), PT (p, t0, t1, t2)). */
int ipc_conservative_time_ee(const float* x0, const float* x1, int n_vertices, const int* quads, int n, double* toi,
                             unsigned long long* toi_min, void* stream);
int ipc_conservative_time_pt(const float* x0, const float* x1, int n_vertices, const int* quads, int n, double* toi,
                             unsigned long long* toi_min, void* stream);

/* ===================================================== contact context ===
 * The device state ContactSolverBase owns (contact.py:444-521): topology
 * (copied at creation from device or host pointers: edges int32 [E][2],
 * triangles int32 [F][3], body int32 [V]), the BVHs and the CCD buffers.
 * One context per mesh soup. A context's buffers (BVHs, leaf and query
 * orders, candidate lists) are reused from one call to the next, so calls on
 * one context must run one after another on the GPU: issue them on one
 * stream, or order calls on different streams with events; calls that
 * overlap on the GPU race on those buffers. Different contexts are
 * independent. */
typedef struct ipc_ctx ipc_ctx;
int ipc_ctx_create(ipc_ctx** ctx, int n_vertices, const int* edges, int n_edges, const int* triangles,
                   int n_triangles, const int* body);
int ipc_ctx_destroy(ipc_ctx* ctx);

/* ------------------------------------------------------ DCD ---
 * Replace edge_edge_collision / point_triangle_collision and the BVH
 * updates of detect_collision (contact.py:199-336,672-725). A BVH over the
 * current positions x is built, every edge (EE) or every vertex (PT) queries
 * it (edge boxes +- (thickness + s); vertex boxes +- 2 (thickness + s)
 * against triangle bounds, s = min(buffer, thickness 2^-10 + 2^-16 M), M
 * the primitive's largest coordinate magnitude: the reference's inflation
 * where s = buffer, and otherwise still wide enough for every pair its fp32
 * distance test accepts, so the pairs found are the reference's), and pairs
 * passing the reference's filters (EE: i < j, no shared vertex; both: not
 * the same body when exclude_same_body) and its fp32 distance test
 * (< 2 thickness) are appended: contacts int32 [capacity][4] as
 * XConstraint.a1a2b1b2 (EE (a1, a2, b1, b2), PT (p, t[1], t[0], t[2])),
 * e0e1 int32 [capacity][2] (optional) as XConstraint.e0e1. *count (device
 * int) receives the total number found; entries beyond capacity are not
 * written (read the count, grow, call again — _grow_contact_list,
 * contact.py:523-537). Order within the list is unspecified. */
int ipc_detect_collision_ee(ipc_ctx* ctx, const float* x, float thickness, float buffer, int exclude_same_body,
                            int* contacts, int* e0e1, int capacity, int* count, void* stream);
int ipc_detect_collision_pt(ipc_ctx* ctx, const float* x, float thickness, float buffer, int exclude_same_body,
                            int* contacts, int* e0e1, int capacity, int* count, void* stream);

/* Both kinds of detect_collision (contact.py:687-725) in one call: both
 * trees in one build launch and both queries in one query launch on
 * `stream` (the halves share the device within each launch); arguments and
 * results as the two calls above. No host synchronisation. */
int ipc_detect_collision(ipc_ctx* ctx, const float* x, float thickness, float buffer, int exclude_same_body, int* ee,
                         int* ee_ids, int ee_capacity, int* ee_count, int* pt, int* pt_ids, int pt_capacity,
                         int* pt_count, void* stream);

/* Penetration check (no reference counterpart; it verifies the Newton
 * loop's states): counts the edge-triangle pairs of the soup at x that
 * cross (segment through the closed triangle, no shared vertex; FP64 signed
 * volumes of the fp32 positions). (edge, triangle) ids go to ids int32
 * [capacity][2] (may be NULL with capacity 0); *count (device int) is the
 * total. A state reached by CCD-bounded moves from an intersection-free one
 * has none. */
int ipc_count_intersections(ipc_ctx* ctx, const float* x, int* ids, int capacity, int* count, void* stream);

/* --------------------------------------------- CCD candidates ---
 * The candidate sets of edge_edge_toi / point_triangle_toi
 * (ccd/ccd.py:166-210,360-427): swept boxes over x0 and x1 padded by
 * `padding` (1e-7 in the reference), filters as above without a distance
 * test. quads int32 [capacity][4] (EE a0, a1, b0, b1; PT p, t0, t1, t2),
 * ids int32 [capacity][2] (optional, primitive ids), count as above. */
int ipc_ccd_candidates_ee(ipc_ctx* ctx, const float* x0, const float* x1, float padding, int exclude_same_body,
                          int* quads, int* ids, int capacity, int* count, void* stream);
int ipc_ccd_candidates_pt(ipc_ctx* ctx, const float* x0, const float* x1, float padding, int exclude_same_body,
                          int* quads, int* ids, int capacity, int* count, void* stream);

/* ------------------------------------------- CCD step bound ---
 * Replaces collision_free_step(dx) (contact.py:539-585) up to its final
 * `0.9 * toi if toi < 1 else 1`: x1 = x - dx, EE and PT candidates, FP64
 * time of impact per candidate and their minimum. Synchronises `stream` once
 * and writes the raw minimum TOI (1.0 = no impact) to *toi (host double).
 * Candidate buffers belong to the context and grow as needed. */
int ipc_collision_free_step(ipc_ctx* ctx, const float* x, const float* dx, float padding, int exclude_same_body,
                            double* toi, void* stream);
/* Candidate counts (EE, PT) of the context's last collision_free_step. */
int ipc_ctx_last_ccd_counts(ipc_ctx* ctx, int* n_ee, int* n_pt);

/* Replaces new_intersections (contact.py:587-670), one primitive kind per
 * call: the candidates of edge_edge_intersections / point_triangle_
 * intersections (contact.py:370-442) between x0 and x1, FP64 TOI per
 * candidate, and every candidate with TOI < 1 appended (append_ccd,
 * contact.py:185-197) as records int32 [capacity][4] (EE (a0, a1, b0, b1);
 * PT (p, t0, t1, t2)), ids int32 [capacity][2] (e0e1: edge i < j; point,
 * triangle) and alpha double [capacity] (= TOI). *count (device int) is the
 * number of such pairs and may exceed capacity: grow the lists and call
 * again. Synchronises `stream` once and writes this kind's raw minimum TOI
 * (1.0 = none) to *toi (host double); new_intersections returns
 * 0.9 * min(toi_ee, toi_pt) if it is < 1, else 1. */
int ipc_ccd_intersections_ee(ipc_ctx* ctx, const float* x0, const float* x1, float padding, int exclude_same_body,
                             int* records, int* ids, double* alpha, int capacity, int* count, double* toi,
                             void* stream);
int ipc_ccd_intersections_pt(ipc_ctx* ctx, const float* x0, const float* x1, float padding, int exclude_same_body,
                             int* records, int* ids, double* alpha, int capacity, int* count, double* toi,
                             void* stream);

/* ============================================== fixed per-body BVHs ===
 * Affine bodies (ABD): vertex v sits at A_b xcs[v] + p_b for its body b, with
 * q float [n_bodies][4][3] = (p, A[:,0], A[:,1], A[:,2]) per body (the
 * reference's affine_q layout, fixed_bvh_contacts.py:32-36).
 *
 * Replaces FixedBodyContactSolver.initialize_fixed_body_bvhs
 * (fixed_bvh_contacts.py:511-597): one tree per body over its edges and one
 * over its triangles, built once from the body-local rest positions xcs
 * (float [V][3], device or host), plus per-body vertex lists and rest boxes.
 * Every vertex's body id must lie in [0, n_bodies). */
int ipc_ctx_init_fixed_bodies(ipc_ctx* ctx, int n_bodies, const float* xcs);
/* Replaces FixedBodyContactSolver.detect_collision (fixed_bvh_contacts.py:
 * 758-827): body pairs whose world boxes overlap within 2 thickness, then
 * the edges of each pair's first body against the second body's edge tree
 * and the vertices of each body against the other's triangle tree, in the
 * target body's rest frame with padding 2 thickness |A^-1|_F; accepted by
 * the world-space fp32 distance < 2 thickness. Records as in
 * ipc_detect_collision_ee / _pt: EE (source edge, target edge) vertices with
 * e0e1 = (source, target), PT (p, t[1], t[0], t[2]) with (p, triangle).
 * *ee_count / *pt_count (device) may exceed the capacities: grow and call
 * again. Synchronises `stream` once. A body with |det A| < 1e-8 makes the
 * call use the world-space DCD (same-body pairs excluded) instead. */
int ipc_detect_collision_fixed(ipc_ctx* ctx, const float* x, const float* q, float thickness, int* ee, int* ee_ids,
                               int ee_capacity, int* ee_count, int* pt, int* pt_ids, int pt_capacity, int* pt_count,
                               void* stream);
/* Replaces FixedBodyContactSolver.collision_free_step (fixed_bvh_contacts.py:
 * 710-756) up to its 0.9 factor: x1 = x - dx with dq the matching per-body
 * affine step (q1 = q - dq); swept body boxes and pairs (padding),
 * candidates from the fixed trees with query boxes widened by the target's
 * motion extent (a fixed source body of an EE pair is swapped to be the
 * target when fixed_bodies, int [n_bodies] or NULL, marks it), FP64 TOI per
 * candidate. Writes the raw minimum TOI (1.0 = none) to *toi (host). When a
 * body's |det A| < 1e-8 at either end it runs ipc_collision_free_step with
 * same-body pairs excluded instead and sets *used_fallback (may be NULL). */
int ipc_collision_free_step_fixed(ipc_ctx* ctx, const float* x, const float* dx, const float* q, const float* dq,
                                  const int* fixed_bodies, float padding, double* toi, int* used_fallback,
                                  void* stream);

/* ------------------------------------------------------ stage timing ---
 * The role of RodBCBase.profile_timer (time_integrator_base.py:245-252):
 * while enabled, each BVH build, query and TOI launch of the context is
 * bracketed by HIP events on its stream. ipc_ctx_kernel_times waits for
 * them and returns, per kind, the accumulated milliseconds and number of
 * launches since the last call (arrays of IPC_TIMER_KINDS entries). The
 * calls that serve both halves (ipc_detect_collision,
 * ipc_collision_free_step and the fixed-body CCD's TOI) launch each stage
 * once for EE and PT together and record it under the EE kind. */
#define IPC_TIMER_BUILD_EDGE 0
#define IPC_TIMER_BUILD_TRI 1
#define IPC_TIMER_QUERY_DCD_EE 2
#define IPC_TIMER_QUERY_DCD_PT 3
#define IPC_TIMER_QUERY_CCD_EE 4
#define IPC_TIMER_QUERY_CCD_PT 5
#define IPC_TIMER_TOI_EE 6
#define IPC_TIMER_TOI_PT 7
#define IPC_TIMER_FIXED_CULL 8
#define IPC_TIMER_FIXED_QUERY_DCD 9
#define IPC_TIMER_FIXED_QUERY_CCD 10
#define IPC_TIMER_KINDS 11
int ipc_ctx_set_timing(ipc_ctx* ctx, int enable);
int ipc_ctx_kernel_times(ipc_ctx* ctx, double* total_ms, long long* calls);

/* ================================================= learned warm start ===
 * No counterpart in the reference: the MLP 48 -> 256 -> 256 -> 256 -> 3
 * (GELU, fp32) that predicts a vertex's converged Newton displacement from
 * 48 features (the layout is documented in ipc_warp_amd/warmstart.py).
 * params: IPC_MLP_PARAMS floats, W1 [256][48], b1, W2 [256][256], b2,
 * W3 [256][256], b3, W4 [3][256], b4 (W row-major, out x in).
 * x float [n][48], y float [n][3]. */
#define IPC_MLP_IN 48
#define IPC_MLP_HIDDEN 256
#define IPC_MLP_OUT 3
#define IPC_MLP_PARAMS 144899
/* Forward pass. work (float [6 * 256 * n]) keeps Z1, H1, Z2, H2, Z3, H3 for
 * a backward pass; with work = NULL, scratch (float [3 * 256 * n]) holds the
 * activations. */
int ipc_mlp_forward(const float* params, int n, const float* x, float* y, float* work, float* scratch, void* stream);
/* Inference that allocates nothing: ipc_mlp_forward with work = NULL, its
 * activations held in a device array inside the library (3 * 256 *
 * IPC_MLP_NOALLOC_MAX_ROWS floats, 12 MiB, one per GPU the library runs on,
 * kept until the process ends) instead of a scratch argument. No
 * counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops). n must not exceed
 * IPC_MLP_NOALLOC_MAX_ROWS (IPC_EINVAL otherwise). Every call on one GPU
 * shares that array, so calls that overlap on one GPU (other streams,
 * other threads) race; calls on one stream are ordered and safe. */
#define IPC_MLP_NOALLOC_MAX_ROWS 4096
int ipc_mlp_forward_noalloc(const float* params, int n, const float* x, float* y, void* stream);
/* Backward pass of the forward pass that filled `work`: ADDS the gradient of
 * the loss with respect to the parameters into grad (IPC_MLP_PARAMS floats)
 * given dy = dLoss/dy; scratch float [2 * 256 * n]. */
int ipc_mlp_backward(const float* params, int n, const float* x, const float* work, const float* dy, float* grad,
                     float* scratch, void* stream);
/* Weighted mean squared error: loss[0] += sum_i w_i |y_i - t_i|^2 / weight_sum
 * (double, device) and dy_i = 2 w_i (y_i - t_i) / weight_sum; weight may be
 * NULL (all ones); dy or loss may be NULL. */
int ipc_mlp_mse(const float* y, const float* target, const float* weight, int n, float weight_sum, float* dy,
                double* loss, void* stream);
/* One Adam step (bias-corrected, step >= 1) over n parameters. */
int ipc_mlp_adam(float* params, const float* grad, float* m, float* v, long long n, float lr, float beta1, float beta2,
                 float eps, int step, void* stream);
/* p[0..n) = 0 (a fill kernel, safe inside captured graphs). */
int ipc_zero_f32(float* p, long long n, void* stream);

/* Per-vertex input features of the learned warm start (no counterpart in the
 * reference, whose Newton loop starts from the previous state,
 * time_integrator_base.py:375-401). features (float [n_vertices][48]) =
 * base (float [n_vertices][12], NULL = zeros) followed by the vertex's
 * IPC_WS_SLOTS nearest records of the contact lists ee / pt (int32 [n][4],
 * the layouts ipc_detect_collision_* writes), nearest first by fp32 distance
 * (ties by record: EE records first, then by index and vertex position), each
 * as ((partner - self) / dhat [3], dist / dhat, w, is_pt): self is the
 * closest point on the vertex's own primitive, partner the one on the other
 * primitive, w the vertex's weight in self. Unused slots are zero. slots is
 * workspace of 6 * n_vertices unsigned 64-bit words. */
#define IPC_WS_BASE 12
#define IPC_WS_SLOTS 6
#define IPC_WS_SLOT_DIM 6
int ipc_warmstart_features(const float* x, int n_vertices, const float* base, const int* ee, int n_ee, const int* pt,
                           int n_pt, float dhat, unsigned long long* slots, float* features, void* stream);

/* ================================================== contact records ===
 * The reference's Contacts.list records (XConstraint, contact.py:61-69):
 * a1a2b1b2 int32[4] at byte 0, e0e1 int32[2] at 16, then l0, alpha, lam as
 * `scalar` from byte 24: 36-byte records with fp32 scalars
 * (scalar_types.py USE_FLOAT32), 48-byte records with fp64 scalars. */
#define IPC_XCONSTRAINT_BYTES_F32 36
#define IPC_XCONSTRAINT_BYTES_F64 48
/* contacts[i] = records[i].a1a2b1b2 (int32 [n][4], 16-byte aligned), so a
 * caller's own contact lists (contacts_new.list / contacts_pt.list) feed the
 * Hessian, energy and contact-point entry points. The kernels take one l0
 * for all contacts; when l0_mismatch (device int32, may be NULL) is given,
 * the number of records whose l0 differs from `l0` is ADDED to it. */
int ipc_contact_records_pack(const void* records, int record_bytes, int n, float l0, int* contacts, int* l0_mismatch,
                             void* stream);
/* The fields append writes (contact.py:163-183): records[i].a1a2b1b2 =
 * contacts[i], .e0e1 = e0e1[i], .l0 = l0, .alpha = alpha[i] (append_ccd,
 * contact.py:185-197) or 1e-6 when alpha is NULL (append); .lam is left
 * untouched (the augmented-Lagrangian multiplier). */
int ipc_contact_records_unpack(const int* contacts, const int* e0e1, int n, float l0, const double* alpha,
                               void* records, int record_bytes, void* stream);

/* ============================================== block-sparse assembly ===
 * BSR matrices of 3x3 fp32 blocks as warp.sparse keeps them: offsets int32
 * [n_rows + 1], columns int32 [nnz] sorted and distinct within each row,
 * values float [nnz][3][3] row-major. An ipc_bsr is the workspace of these
 * calls and keeps the pattern of its last successful ipc_bsr_from_triplets;
 * calls on one workspace must run one after another on the GPU (one stream,
 * or ordered with events). */
typedef struct ipc_bsr ipc_bsr;
int ipc_bsr_create(ipc_bsr** bsr);
int ipc_bsr_destroy(ipc_bsr* bsr);
/* Replaces bsr_from_triplets(n_rows, n_cols, rows, cols, vals)
 * (dynamic_contacts.py:290-296; affine_body_dynamics.py:500-503;
 * augmented_lagrangian.py:631): triplet t = (rows[t], cols[t], vals[t]
 * float [3][3]) is dropped when its row or column is out of range or its
 * nine values are all zero (NaN is nonzero); the kept triplets of one
 * (row, column) are summed in fp32 in ascending t from +0. Writes offsets
 * always and columns/values when the block count fits `capacity`. Returns 0,
 * or, when capacity < n_triplets and the pattern needs more blocks, the
 * block count (columns/values untouched; the call synchronises `stream`
 * only in that case). */
int ipc_bsr_from_triplets(ipc_bsr* bsr, int n_rows, int n_cols, const int* rows, const int* cols, const float* vals,
                          int n_triplets, int* offsets, int* columns, float* values, int capacity, void* stream);
/* Values of the last pattern ipc_bsr_from_triplets built on this workspace
 * for new triplet values with the same rows and columns (the pattern reuse of
 * solve_ldlt, dynamic_contacts.py:314-333): every block sums its triplets in
 * the same order as the build, a triplet dropped by the build stays out and a
 * block whose triplets are now zero stays in the pattern as zeros. */
int ipc_bsr_refill(ipc_bsr* bsr, const float* vals, int n_triplets, float* values, void* stream);
/* z = alpha x + beta y over the union of the two patterns (bsr_axpy(x, y,
 * alpha, beta), dynamic_contacts.py:297; augmented_lagrangian.py:632, which
 * writes the sum into y): both inputs canonical (columns sorted and distinct
 * per row); a block present on one side only counts as zero on the other;
 * each value is alpha * xv + beta * yv in fp32. Writes z_offsets always and
 * z_columns / z_values when the union fits `capacity`. x_nnz / y_nnz are the
 * block counts of x and y or upper bounds of them (their capacities). Returns
 * 0, or, when capacity < x_nnz + y_nnz and the union needs more blocks, its
 * block count (after synchronising `stream`). z must not alias x or y. */
int ipc_bsr_axpy(ipc_bsr* bsr, int n_rows, const int* x_offsets, const int* x_columns, const float* x_values, int x_nnz,
                 float alpha, const int* y_offsets, const int* y_columns, const float* y_values, int y_nnz, float beta,
                 int* z_offsets, int* z_columns, float* z_values, int capacity, void* stream);

/* y = alpha A x + beta y for a BSR matrix of n_rows block rows (x float
 * [n_cols][3], y float [n_rows][3]; beta = 0 does not read y); warp.sparse's
 * bsr_mv, the product inside the reference's cg. */
int ipc_bsr_mv(int n_rows, const int* offsets, const int* columns, const float* values, const float* x, float* y,
               float alpha, float beta, void* stream);
/* Conjugate gradients on a symmetric positive (semi)definite BSR matrix
 * (warp.optim.linear.cg as called at time_integrator_base.py:500-507 and
 * affine_body_dynamics.py:580-592): solves A x = b from the initial guess in
 * x (float [n_rows][3]), stopping when |r| <= tol |b| (checked on the host
 * every check_every iterations, default 10, as the reference's cg does) or
 * after maxiter iterations (<= 0: 3 n_rows); iterations run in whole groups
 * of check_every rounded up to even, captured once as a HIP graph for the
 * given pointers and size and replayed. jacobi != 0 preconditions with
 * the inverse diagonal of the diagonal blocks. Vectors are fp32, dot
 * products fp64 and summed in a fixed order. Writes the iteration count of
 * the last check and |r| there. Synchronises `stream` at each check. An
 * ipc_cg is the solver's workspace (one stream at a time). */
typedef struct ipc_cg ipc_cg;
int ipc_cg_create(ipc_cg** cg);
int ipc_cg_destroy(ipc_cg* cg);
int ipc_bsr_cg(ipc_cg* cg, int n_rows, const int* offsets, const int* columns, const float* values, const float* b,
               float* x, float tol, int maxiter, int check_every, int jacobi, int* iterations, double* residual,
               void* stream);

/* ================================================= elastic tetrahedra ===
 * Stable Neo-Hookean tets (fem/snh.py:1-84): tets int32 [n][4] (16-byte
 * aligned), rest shapes Bm float [n][3][3] (the inverse of Dm = [x0 - x3,
 * x1 - x3, x2 - x3]) and volumes W float [n]. mu / lam are the Lamé
 * parameters the reference keeps as module constants (fem/params.py:7),
 * reparametrised inside as fem/snh.py:6-12 does. */
/* compute_Dm (fem/fem.py:21-34). */
int ipc_tet_rest(const float* xcs, int n_vertices, const int* tets, int n_tets, float* Bm, float* W, void* stream);
/* tet_kernel_sparse (fem/fem.py:138-203): block (offset + e) * 16 + 4 i + j
 * = d(-f_i)/dx_j of tet e (row tets[e][i], column tets[e][j]), the elastic
 * force f added into b (float [V][3], may be NULL). rows, cols and vals are
 * written as 16-byte vectors and must be 16-byte aligned. */
int ipc_tet_hessian(const float* x, int n_vertices, const int* tets, int n_tets, const float* Bm, const float* W,
                    float mu, float lam, int offset, int* rows, int* cols, float* vals, float* b, void* stream);
/* compute_Psi (time_integrator_base.py:124-137): psi[e] = W[e] psi(F_e)
 * (float [n], may be NULL); their sum ADDED to *energy (double, device,
 * may be NULL). */
int ipc_tet_energy(const float* x, int n_vertices, const int* tets, int n_tets, const float* Bm, const float* W,
                   float mu, float lam, float* psi, double* energy, void* stream);

/* ==================================================== affine bodies ===
 * Contact reduction onto affine-body degrees of freedom
 * (affine_body_dynamics.py:144-190; body b has q = [p, a0, a1, a2] and
 * vertex i of body b sits at p + r_i.x a0 + r_i.y a1 + r_i.z a2). The
 * reference writes 16 body triplets per vertex triplet and sorts them; this
 * sums them in place. An ipc_abd is the call's workspace (one stream at a
 * time). */
typedef struct ipc_abd ipc_abd;
int ipc_abd_create(ipc_abd** abd);
int ipc_abd_destroy(ipc_abd* abd);
/* Vertex triplets (rows, cols int32 [n], vals float [n][3][3], e.g. from
 * ipc_collision_hessian) reduced like _reduce_contact_hessian: triplet
 * (v_i, v_j, H) adds scale * c_a(v_i) c_c(v_j) H to block (4 b_i + a,
 * 4 b_j + c), c_0 = 1, c_k = rest[v][k-1]; triplets out of range, all zero,
 * or touching a fixed body (fixed int32 [n_bodies] != 0) add nothing. Output
 * BSR over 4 n_bodies block rows: body pair (A, B) receiving a block owns all
 * 16 blocks (4A + a, 4B + c), rows sorted by column; offsets int32
 * [4 n_bodies + 1] always written, columns / values when 16 * pairs <=
 * capacity. Blocks are summed in fp32 per lane and wave and in fp64 across
 * waves, then scaled and rounded to fp32 (the cross-wave order is that of
 * the atomics: results agree to fp64 rounding). When force (float [V][3],
 * the vertex force b) is given, _reduce_contact_force's sums are ADDED to
 * reduced_force (float [4 n_bodies][3], unscaled). Synchronises `stream`
 * once (the pattern size). Returns 0, or the block count when it exceeds
 * capacity (nothing else written). n_bodies <= 8192. */
int ipc_abd_reduce_contacts(ipc_abd* abd, int n_bodies, const int* body, const float* rest, const int* fixed,
                            int n_vertices, const int* rows, const int* cols, const float* vals, int n_triplets,
                            const float* force, double scale, int* offsets, int* columns, float* values, int capacity,
                            float* reduced_force, void* stream);

/* =========================================== augmented-Lagrangian set ===
 * The active set of the barrier-free AL solver (augmented_lagrangian.py:
 * 42-197, 506-547): a hash table (the reference's Hash) of table_size
 * buckets, cell_start int32 [table_size + 1] (bucket b holds entries
 * cell_start[b] .. cell_start[b+1]; cell_start[table_size] is the entry
 * count) and entries, ALConstraint records in the reference's layout:
 * a1a2b1b2 int32[4], e0e1 int32[2], then l0, alpha, lam, gamma, slack,
 * mollifier as `scalar`, distance_type int32, d0 `scalar`, grad `scalar`[12]
 * (104 bytes with fp32 scalars, 184 with fp64: 4 bytes of padding after
 * distance_type). Constraints come in as XConstraint records
 * (IPC_XCONSTRAINT_BYTES_*), vertex TOIs as `scalar` [n_vertices]. Within a
 * bucket, entries are ordered by decreasing input index (the reference's
 * order when its threads run one after another). An ipc_active_set is the
 * workspace of these calls (one stream at a time); tables are the
 * caller's. */
#define IPC_ALCONSTRAINT_BYTES_F32 104
#define IPC_ALCONSTRAINT_BYTES_F64 184
typedef struct ipc_active_set ipc_active_set;
int ipc_active_set_create(ipc_active_set** ws);
int ipc_active_set_destroy(ipc_active_set* ws);
/* accumulate_vertex_toi (augmented_lagrangian.py:105-113): vertex_toi[v] =
 * min(vertex_toi[v], alpha) over the constraints on v (the caller fills it
 * with 1 first, as update_active_set does at :707-712). */
int ipc_al_vertex_toi(const void* constraints, int record_bytes, int n, void* vertex_toi, int n_vertices,
                      void* stream);
/* ActiveSet.create_hash (augmented_lagrangian.py:525-535, count_constraints
 * / register_constraints :124-149): the constraints with alpha <= vertex_toi
 * + 1e-12 at one of their vertices become entries with lam 0, gamma 1,
 * slack 0, mollifier 1, distance_type 0, d0 0, grad 0; entries past the
 * count are zeroed. capacity >= n. */
int ipc_al_create_hash(ipc_active_set* ws, const void* constraints, int record_bytes, int n, const void* vertex_toi,
                       int n_vertices, int table_size, int* cell_start, void* entries, int capacity, void* stream);
/* ActiveSet.merge (augmented_lagrangian.py:537-547, count_merge /
 * register_merge :167-197): the old table's n_old entries with gamma >=
 * 0.01, then the new table's n_new entries whose e0e1 the old table does
 * not hold, into the destination table (capacity >= n_old + n_new; entries
 * past the count zeroed). The destination must not overlap the inputs. */
int ipc_al_merge(ipc_active_set* ws, int entry_bytes, const int* old_cell_start, const void* old_entries,
                 int old_table_size, int n_old, const void* new_entries, int n_new, int table_size, int* cell_start,
                 void* entries, int capacity, void* stream);
/* ActiveSet.update_slack / update_forces (augmented_lagrangian.py:564-572,
 * kernels update_slack_variables / update_active_set_forces :326-345) on
 * the n entries of a table, in place. With c = d0 + grad . x_hat (x_hat
 * `scalar` [n_vertices][3]; constraint_value :76-85) and mu the AL
 * penalty (cast to `scalar`): update_slack sets slack = max(0, c - lam/mu);
 * update_forces sets lam -= mu c and gamma = 1 where slack == 0, else
 * lam = 0 and gamma *= 0.9. An entry naming a vertex out of range is left
 * unchanged. */
/* ActiveSet.linearize (augmented_lagrangian.py:549-562, kernel
 * linearize_constraints :270-323) on the n entries, in place: at x_safe
 * (`scalar` [n_vertices][3]) each constraint's feature — PT: the region of
 * closest_point_triangle; EE: where the mollifier (rest edges x_rest, eps_x
 * = 1e-3 |ea|^2 |eb|^2) is below point_edge_threshold, the parallel
 * point-edge choice of :208-267 (distance_type 1..4), else
 * edge_edge_distance_gradient_hessian's dispatch — gives grad = grad(d^2) /
 * (2 d) and d0 = d - l0 - grad . x_safe; mollifier and distance_type are
 * stored too. Classification in fp32, the feature gradient in FP64. */
int ipc_al_linearize(void* entries, int entry_bytes, int n, const void* x_safe, const void* x_rest, int n_vertices,
                     int is_point_triangle, double point_edge_threshold, void* stream);
int ipc_al_update_slack(void* entries, int entry_bytes, int n, const void* x_hat, int n_vertices, double mu,
                        void* stream);
int ipc_al_update_forces(void* entries, int entry_bytes, int n, const void* x_hat, int n_vertices, double mu,
                         void* stream);

#ifdef __cplusplus
}
#endif

#endif /* IPC_WARP_AMD_H */

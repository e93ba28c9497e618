// extern "C" entry points of libipc_warp_amd (declared in include/ipc_warp_amd.h).
// Validation of sizes and pointers happens here on the host, before any
// launch; nothing throws across the ABI.
#include <cstdio>
#include <cstring>

#include "../../include/ipc_warp_amd.h"
#include "ipc_kernels.hpp"

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char* what) {
  ipcd::set_error(what);
  return code;
}

int hip_status(hipError_t e, const char* where) {
  if (e == hipSuccess) return IPC_OK;
  std::snprintf(g_err, sizeof(g_err), "%s: %s", where, hipGetErrorString(e));
  return -int(e);
}

bool aligned16(const void* p) { return (reinterpret_cast<unsigned long long>(p) & 15ull) == 0; }

}  // namespace

void ipcd::set_error(const char* msg) { std::snprintf(g_err, sizeof(g_err), "%s", msg); }

extern "C" {

int ipc_abi_version(void) { return 2; }

const char* ipc_last_error(void) { return g_err; }

static int check_hessian_args(const void* x, int nv, const int* contacts, int n, double l0, int offset, int* rows,
                              int* cols, const void* vals, const void* b) {
  if (n < 0 || offset < 0 || nv < 0) return fail(IPC_EINVAL, "contact_hessian: negative n_vertices, n_contacts or offset");
  if ((long long)(offset) + n > (1ll << 27)) return fail(IPC_EINVAL, "contact_hessian: offset + n_contacts exceeds 2^27");
  if (n == 0) return IPC_OK;
  if (!x || !contacts || !rows || !cols || !vals || !b) return fail(IPC_EINVAL, "contact_hessian: null pointer");
  if (!(l0 > 0.0)) return fail(IPC_EINVAL, "contact_hessian: l0 must be > 0");
  if (!aligned16(contacts) || !aligned16(rows) || !aligned16(cols) || !aligned16(vals))
    return fail(IPC_EINVAL, "contact_hessian: contacts/rows/cols/vals must be 16-byte aligned");
  return IPC_OK;
}

int ipc_contact_hessian_ee(const float* x, const float* xcs, int n_vertices, const int* contacts, int n_contacts,
                           float l0, double kappa, int offset, int* rows, int* cols, float* vals, float* b,
                           void* stream) {
  int st = check_hessian_args(x, n_vertices, contacts, n_contacts, l0, offset, rows, cols, vals, b);
  if (st != IPC_OK || n_contacts == 0) return st;
  if (!xcs) return fail(IPC_EINVAL, "contact_hessian_ee: xcs (rest positions) is required");
  return hip_status(ipcd::launch_contact_hessian(true, x, xcs, n_vertices, contacts, n_contacts, l0, kappa, offset, rows, cols,
                                                 vals, b, static_cast<hipStream_t>(stream)),
                    "contact_hessian_ee");
}

int ipc_contact_hessian_pt(const float* x, int n_vertices, const int* contacts, int n_contacts, float l0,
                           double kappa, int offset, int* rows, int* cols, float* vals, float* b, void* stream) {
  int st = check_hessian_args(x, n_vertices, contacts, n_contacts, l0, offset, rows, cols, vals, b);
  if (st != IPC_OK || n_contacts == 0) return st;
  return hip_status(ipcd::launch_contact_hessian(false, x, nullptr, n_vertices, contacts, n_contacts, l0, kappa, offset, rows,
                                                 cols, vals, b, static_cast<hipStream_t>(stream)),
                    "contact_hessian_pt");
}

int ipc_collision_hessian(const float* x, const float* xcs, int n_vertices, const int* ee, int n_ee, const int* pt,
                          int n_pt, float l0, double kappa, int* rows, int* cols, float* vals, float* b, void* stream) {
  int st = check_hessian_args(x, n_vertices, ee, n_ee, l0, 0, rows, cols, vals, b);
  if (st == IPC_OK) st = check_hessian_args(x, n_vertices, pt, n_pt, l0, n_ee, rows, cols, vals, b);
  if (st != IPC_OK || (n_ee == 0 && n_pt == 0)) return st;
  if (n_ee > 0 && !xcs) return fail(IPC_EINVAL, "collision_hessian: xcs (rest positions) is required");
  return hip_status(ipcd::launch_collision_hessian(x, xcs, n_vertices, ee, n_ee, pt, n_pt, l0, kappa, rows, cols, vals, b,
                                                   static_cast<hipStream_t>(stream)),
                    "collision_hessian");
}

static int check_energy_args(const void* x, int nv, const int* contacts, int n, double l0, double* energy) {
  if (n < 0 || nv < 0) return fail(IPC_EINVAL, "contact_energy: negative n_vertices or n_contacts");
  if (n == 0) return IPC_OK;
  if (!x || !contacts || !energy) return fail(IPC_EINVAL, "contact_energy: null pointer");
  if (!(l0 > 0.0)) return fail(IPC_EINVAL, "contact_energy: l0 must be > 0");
  if (!aligned16(contacts)) return fail(IPC_EINVAL, "contact_energy: contacts must be 16-byte aligned");
  return IPC_OK;
}

int ipc_contact_energy_ee(const float* x, const float* xcs, int n_vertices, const int* contacts, int n_contacts,
                          float l0, double kappa, double* energy, void* stream) {
  int st = check_energy_args(x, n_vertices, contacts, n_contacts, l0, energy);
  if (st != IPC_OK || n_contacts == 0) return st;
  if (!xcs) return fail(IPC_EINVAL, "contact_energy_ee: xcs (rest positions) is required");
  return hip_status(ipcd::launch_contact_energy(true, x, xcs, n_vertices, contacts, n_contacts, l0, kappa, energy,
                                                static_cast<hipStream_t>(stream)),
                    "contact_energy_ee");
}

int ipc_contact_energy_pt(const float* x, int n_vertices, const int* contacts, int n_contacts, float l0,
                          double kappa, double* energy, void* stream) {
  int st = check_energy_args(x, n_vertices, contacts, n_contacts, l0, energy);
  if (st != IPC_OK || n_contacts == 0) return st;
  return hip_status(ipcd::launch_contact_energy(false, x, nullptr, n_vertices, contacts, n_contacts, l0, kappa, energy,
                                                static_cast<hipStream_t>(stream)),
                    "contact_energy_pt");
}

int ipc_collision_energy(const float* x, const float* xcs, int n_vertices, const int* ee, int n_ee, const int* pt,
                         int n_pt, float l0, double kappa, double* energy, void* stream) {
  if (!energy) return fail(IPC_EINVAL, "collision_energy: null energy");
  int st = check_energy_args(x, n_vertices, ee, n_ee, l0, energy);
  if (st == IPC_OK) st = check_energy_args(x, n_vertices, pt, n_pt, l0, energy);
  if (st != IPC_OK) return st;
  if (n_ee > 0 && !xcs) return fail(IPC_EINVAL, "collision_energy: xcs (rest positions) is required");
  return hip_status(ipcd::launch_collision_energy(x, xcs, n_vertices, ee, n_ee, pt, n_pt, l0, kappa, energy,
                                                  static_cast<hipStream_t>(stream)),
                    "collision_energy");
}

int ipc_contact_hessian_ee_f64(const double* x, const double* xcs, int n_vertices, const int* contacts,
                               int n_contacts, double l0, double kappa, int offset, int* rows, int* cols, double* vals,
                               double* b, void* stream) {
  int st = check_hessian_args(x, n_vertices, contacts, n_contacts, l0, offset, rows, cols, vals, b);
  if (st != IPC_OK || n_contacts == 0) return st;
  if (!xcs) return fail(IPC_EINVAL, "contact_hessian_ee_f64: xcs (rest positions) is required");
  return hip_status(ipcd::launch_contact_hessian_f64(true, x, xcs, n_vertices, contacts, n_contacts, l0, kappa, offset,
                                                     rows, cols, vals, b, static_cast<hipStream_t>(stream)),
                    "contact_hessian_ee_f64");
}

int ipc_contact_hessian_pt_f64(const double* x, int n_vertices, const int* contacts, int n_contacts, double l0,
                               double kappa, int offset, int* rows, int* cols, double* vals, double* b, void* stream) {
  int st = check_hessian_args(x, n_vertices, contacts, n_contacts, l0, offset, rows, cols, vals, b);
  if (st != IPC_OK || n_contacts == 0) return st;
  return hip_status(ipcd::launch_contact_hessian_f64(false, x, nullptr, n_vertices, contacts, n_contacts, l0, kappa,
                                                     offset, rows, cols, vals, b, static_cast<hipStream_t>(stream)),
                    "contact_hessian_pt_f64");
}

int ipc_contact_energy_ee_f64(const double* x, const double* xcs, int n_vertices, const int* contacts, int n_contacts,
                              double l0, double kappa, double* energy, void* stream) {
  int st = check_energy_args(x, n_vertices, contacts, n_contacts, l0, energy);
  if (st != IPC_OK || n_contacts == 0) return st;
  if (!xcs) return fail(IPC_EINVAL, "contact_energy_ee_f64: xcs (rest positions) is required");
  return hip_status(ipcd::launch_contact_energy_f64(true, x, xcs, n_vertices, contacts, n_contacts, l0, kappa, energy,
                                                    static_cast<hipStream_t>(stream)),
                    "contact_energy_ee_f64");
}

int ipc_contact_energy_pt_f64(const double* x, int n_vertices, const int* contacts, int n_contacts, double l0,
                              double kappa, double* energy, void* stream) {
  int st = check_energy_args(x, n_vertices, contacts, n_contacts, l0, energy);
  if (st != IPC_OK || n_contacts == 0) return st;
  return hip_status(ipcd::launch_contact_energy_f64(false, x, nullptr, n_vertices, contacts, n_contacts, l0, kappa,
                                                    energy, static_cast<hipStream_t>(stream)),
                    "contact_energy_pt_f64");
}

static int check_points_args(const float* x, int nv, const int* contacts, int n) {
  if (n < 0 || nv < 0) return fail(IPC_EINVAL, "contact_points: negative n_vertices or n_contacts");
  if (n == 0) return IPC_OK;
  if (!x || !contacts) return fail(IPC_EINVAL, "contact_points: null pointer");
  if (!aligned16(contacts)) return fail(IPC_EINVAL, "contact_points: contacts must be 16-byte aligned");
  return IPC_OK;
}

int ipc_contact_points_ee(const float* x, const float* xcs, int n_vertices, const int* contacts, int n_contacts,
                          float* points, float* dists, int* features, void* stream) {
  int st = check_points_args(x, n_vertices, contacts, n_contacts);
  if (st != IPC_OK || n_contacts == 0) return st;
  return hip_status(ipcd::launch_contact_points(true, x, xcs, n_vertices, contacts, n_contacts, points, dists, features,
                                                static_cast<hipStream_t>(stream)),
                    "contact_points_ee");
}

int ipc_contact_points_pt(const float* x, int n_vertices, const int* contacts, int n_contacts, float* points,
                          float* dists, int* features, void* stream) {
  int st = check_points_args(x, n_vertices, contacts, n_contacts);
  if (st != IPC_OK || n_contacts == 0) return st;
  return hip_status(ipcd::launch_contact_points(false, x, nullptr, n_vertices, contacts, n_contacts, points, dists, features,
                                                static_cast<hipStream_t>(stream)),
                    "contact_points_pt");
}

static int collision_time(bool ee, const float* x0, const float* x1, int nv, const int* quads, int n, double* toi,
                          unsigned long long* toi_min, void* stream) {
  if (n < 0 || nv < 0) return fail(IPC_EINVAL, "collision_time: negative n_vertices or n");
  if (n == 0) return IPC_OK;
  if (!x0 || !x1 || !quads) return fail(IPC_EINVAL, "collision_time: null pointer");
  if (!aligned16(quads)) return fail(IPC_EINVAL, "collision_time: quads must be 16-byte aligned");
  return hip_status(ipcd::launch_collision_time(ee, x0, x1, nv, quads, n, toi, toi_min, static_cast<hipStream_t>(stream)),
                    ee ? "collision_time_ee" : "collision_time_pt");
}

int ipc_collision_time_ee(const float* x0, const float* x1, int n_vertices, const int* quads, int n, double* toi,
                          unsigned long long* toi_min, void* stream) {
  return collision_time(true, x0, x1, n_vertices, quads, n, toi, toi_min, stream);
}

int ipc_collision_time_pt(const float* x0, const float* x1, int n_vertices, const int* quads, int n, double* toi,
                          unsigned long long* toi_min, void* stream) {
  return collision_time(false, x0, x1, n_vertices, quads, n, toi, toi_min, stream);
}

static int collision_time_f64(bool ee, const double* x0, const double* x1, int nv, const int* quads, int n,
                              double* toi, unsigned long long* toi_min, void* stream) {
  if (n < 0 || nv < 0) return fail(IPC_EINVAL, "collision_time_f64: negative n_vertices or n");
  if (n == 0) return IPC_OK;
  if (!x0 || !x1 || !quads) return fail(IPC_EINVAL, "collision_time_f64: null pointer");
  if (!aligned16(quads)) return fail(IPC_EINVAL, "collision_time_f64: quads must be 16-byte aligned");
  return hip_status(
      ipcd::launch_collision_time_f64(ee, x0, x1, nv, quads, n, toi, toi_min, static_cast<hipStream_t>(stream)),
      ee ? "collision_time_ee_f64" : "collision_time_pt_f64");
}

int ipc_collision_time_ee_f64(const double* x0, const double* x1, int n_vertices, const int* quads, int n,
                              double* toi, unsigned long long* toi_min, void* stream) {
  return collision_time_f64(true, x0, x1, n_vertices, quads, n, toi, toi_min, stream);
}

int ipc_collision_time_pt_f64(const double* x0, const double* x1, int n_vertices, const int* quads, int n,
                              double* toi, unsigned long long* toi_min, void* stream) {
  return collision_time_f64(false, x0, x1, n_vertices, quads, n, toi, toi_min, stream);
}

int ipc_mlp_forward(const float* params, int n, const float* x, float* y, float* work, float* scratch, void* stream) {
  if (n < 0) return fail(IPC_EINVAL, "mlp_forward: negative n");
  if (n == 0) return IPC_OK;
  if (!params || !x || !y || (!work && !scratch)) return fail(IPC_EINVAL, "mlp_forward: null pointer");
  if ((long long)n * IPC_MLP_HIDDEN * 6 > (1ll << 31)) return fail(IPC_EINVAL, "mlp_forward: batch too large");
  return hip_status(ipcd::mlp_forward(params, n, x, y, work, scratch, static_cast<hipStream_t>(stream)),
                    "mlp_forward");
}

int ipc_mlp_forward_noalloc(const float* params, int n, const float* x, float* y, void* stream) {
  if (n < 0) return fail(IPC_EINVAL, "mlp_forward_noalloc: negative n");
  if (n > IPC_MLP_NOALLOC_MAX_ROWS) return fail(IPC_EINVAL, "mlp_forward_noalloc: n exceeds IPC_MLP_NOALLOC_MAX_ROWS");
  if (n == 0) return IPC_OK;
  if (!params || !x || !y) return fail(IPC_EINVAL, "mlp_forward_noalloc: null pointer");
  return hip_status(ipcd::mlp_forward_noalloc(params, n, x, y, static_cast<hipStream_t>(stream)),
                    "mlp_forward_noalloc");
}

int ipc_mlp_backward(const float* params, int n, const float* x, const float* work, const float* dy, float* grad,
                     float* scratch, void* stream) {
  if (n < 0) return fail(IPC_EINVAL, "mlp_backward: negative n");
  if (n == 0) return IPC_OK;
  if (!params || !x || !work || !dy || !grad || !scratch) return fail(IPC_EINVAL, "mlp_backward: null pointer");
  if ((long long)n * IPC_MLP_HIDDEN * 6 > (1ll << 31)) return fail(IPC_EINVAL, "mlp_backward: batch too large");
  return hip_status(ipcd::mlp_backward(params, n, x, work, dy, grad, scratch, static_cast<hipStream_t>(stream)),
                    "mlp_backward");
}

int ipc_mlp_mse(const float* y, const float* target, const float* weight, int n, float weight_sum, float* dy,
                double* loss, void* stream) {
  if (n < 0) return fail(IPC_EINVAL, "mlp_mse: negative n");
  if (n == 0) return IPC_OK;
  if (!y || !target) return fail(IPC_EINVAL, "mlp_mse: null pointer");
  if (!(weight_sum > 0.0f)) return fail(IPC_EINVAL, "mlp_mse: weight_sum must be > 0");
  return hip_status(ipcd::mlp_mse(y, target, weight, n, weight_sum, dy, loss, static_cast<hipStream_t>(stream)),
                    "mlp_mse");
}

int ipc_mlp_adam(float* params, const float* grad, float* m, float* v, long long n, float lr, float beta1, float beta2,
                 float eps, int step, void* stream) {
  if (n < 0 || step < 1) return fail(IPC_EINVAL, "mlp_adam: negative n or step < 1");
  if (n == 0) return IPC_OK;
  if (!params || !grad || !m || !v) return fail(IPC_EINVAL, "mlp_adam: null pointer");
  return hip_status(ipcd::mlp_adam(params, grad, m, v, n, lr, beta1, beta2, eps, step, static_cast<hipStream_t>(stream)),
                    "mlp_adam");
}

int ipc_zero_f32(float* p, long long n, void* stream) {
  if (n < 0) return fail(IPC_EINVAL, "zero_f32: negative n");
  if (n == 0) return IPC_OK;
  if (!p) return fail(IPC_EINVAL, "zero_f32: null pointer");
  return hip_status(ipcd::zero_f32(p, n, static_cast<hipStream_t>(stream)), "zero_f32");
}

int ipc_warmstart_features(const float* x, int n_vertices, const float* base, const int* ee, int n_ee, const int* pt,
                           int n_pt, float dhat, unsigned long long* slots, float* features, void* stream) {
  if (n_vertices < 0 || n_ee < 0 || n_pt < 0) return fail(IPC_EINVAL, "warmstart_features: negative size");
  if ((long long)n_ee + n_pt > (1ll << 27)) return fail(IPC_EINVAL, "warmstart_features: more than 2^27 records");
  if (n_vertices == 0) return IPC_OK;
  if (!x || !slots || !features) return fail(IPC_EINVAL, "warmstart_features: null pointer");
  if ((n_ee > 0 && !ee) || (n_pt > 0 && !pt)) return fail(IPC_EINVAL, "warmstart_features: null contact list");
  if (!(dhat > 0.0f)) return fail(IPC_EINVAL, "warmstart_features: dhat must be > 0");
  if ((n_ee > 0 && !aligned16(ee)) || (n_pt > 0 && !aligned16(pt)))
    return fail(IPC_EINVAL, "warmstart_features: contact lists must be 16-byte aligned");
  return hip_status(ipcd::warmstart_features(x, n_vertices, base, ee, n_ee, pt, n_pt, dhat, slots, features,
                                             static_cast<hipStream_t>(stream)),
                    "warmstart_features");
}

}  // extern "C"

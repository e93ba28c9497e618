// Internal launcher declarations of libipc_warp_amd (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

namespace ipcd {

hipError_t launch_contact_hessian(bool ee, const float* x, const float* xcs, int nv, const int* contacts, int n, float l0,
                                  double kappa, int offset, int* rows, int* cols, float* vals, float* b,
                                  hipStream_t s);
hipError_t launch_contact_energy(bool ee, const float* x, const float* xcs, int nv, const int* contacts, int n, float l0,
                                 double kappa, double* energy, hipStream_t s);
// The same with fp64 scalars (positions, blocks and forces in double).
hipError_t launch_contact_hessian_f64(bool ee, const double* x, const double* xcs, int nv, const int* contacts, int n,
                                      double l0, double kappa, int offset, int* rows, int* cols, double* vals,
                                      double* b, hipStream_t s);
hipError_t launch_contact_energy_f64(bool ee, const double* x, const double* xcs, int nv, const int* contacts, int n,
                                     double l0, double kappa, double* energy, hipStream_t s);
// Both contact lists in one launch: the Hessian with the PT triplets at
// offset n_ee; the energy after zeroing *energy.
hipError_t launch_collision_hessian(const float* x, const float* xcs, int nv, const int* ee, int n_ee, const int* pt,
                                    int n_pt, float l0, double kappa, int* rows, int* cols, float* vals, float* b,
                                    hipStream_t s);
hipError_t launch_collision_energy(const float* x, const float* xcs, int nv, const int* ee, int n_ee, const int* pt,
                                   int n_pt, float l0, double kappa, double* energy, hipStream_t s);
hipError_t launch_contact_points(bool ee, const float* x, const float* xcs, int nv, const int* contacts, int n, float* points,
                                 float* dists, int* features, hipStream_t s);
hipError_t launch_collision_time(bool ee, const float* x0, const float* x1, int nv, const int* quads, int n, double* toi_out,
                                 unsigned long long* toi_min, hipStream_t s);
hipError_t launch_collision_time_f64(bool ee, const double* x0, const double* x1, int nv, const int* quads, int n,
                                     double* toi_out, unsigned long long* toi_min, hipStream_t s);

hipError_t launch_collision_time_counted(bool ee, const float* x0, const float* x1, int nv, const int* quads,
                                         const int* n_dev, int capacity, unsigned long long* toi_min, hipStream_t s);
// Both candidate lists of a CCD call (counts on the device, capacities
// cap_ee / cap_pt) in one launch, folded into *toi_min.
hipError_t launch_collision_time_pair(const float* x0, const float* x1, int nv, const int* quads_ee, const int* n_ee,
                                      int cap_ee, const int* quads_pt, const int* n_pt, int cap_pt,
                                      unsigned long long* toi_min, hipStream_t s);

hipError_t launch_record_intersections(bool ee, const float* x0, const float* x1, int nv, const int* quads,
                                      const int* ids, const int* n_dev, int cap_in, int* rec_out, int* ids_out,
                                      double* alpha_out, int cap_out, int* count_out, unsigned long long* toi_min,
                                      hipStream_t s);

hipError_t mlp_forward(const float* prm, int n, const float* x, float* y, float* work, float* tmp, hipStream_t s);
hipError_t mlp_forward_noalloc(const float* prm, int n, const float* x, float* y, hipStream_t s);
hipError_t mlp_backward(const float* prm, int n, const float* x, const float* work, const float* dy, float* grad,
                        float* scratch, hipStream_t s);
hipError_t mlp_mse(const float* y, const float* t, const float* w, int n, float wsum, float* dy, double* loss,
                   hipStream_t s);
hipError_t mlp_adam(float* p, const float* g, float* m, float* v, long long n, float lr, float b1, float b2, float eps,
                    int step, hipStream_t s);
hipError_t zero_f32(float* p, long long n, hipStream_t s);
hipError_t warmstart_features(const float* x, int nv, const float* base, const int* ee, int n_ee, const int* pt,
                              int n_pt, float dhat, unsigned long long* slots, float* out, hipStream_t s);

// Records the calling thread's last error message (read by ipc_last_error).
void set_error(const char* msg);

}  // namespace ipcd

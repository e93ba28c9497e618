// The reference's contact records (XConstraint, contact.py:61-69: a1a2b1b2
// vec4i, e0e1 vec2i, then l0, alpha, lam as `scalar`) packed into the
// library's int32 [n][4] contact lists and written back: one copy kernel
// each way, so a caller that keeps its Contacts.list arrays can hand them to
// the Hessian and energy entry points and read the library's DCD lists in
// its own layout.
#include <cstdio>

#include "../../include/ipc_warp_amd.h"
#include "ipc_kernels.hpp"

using namespace ipcd;

namespace {

constexpr int kB = 256;

int rec_fail(int code, const char* what) {
  set_error(what);
  return code;
}

int rec_hip(hipError_t e, const char* where) {
  if (e == hipSuccess) return IPC_OK;
  char buf[512];
  std::snprintf(buf, sizeof(buf), "%s: %s", where, hipGetErrorString(e));
  set_error(buf);
  return -int(e);
}

// Field offsets: a1a2b1b2 at 0, e0e1 at 16, l0 / alpha / lam at 24 + k * s
// with s = sizeof(scalar) (4 in the 36-byte layout, 8 in the 48-byte one).
template <typename S>
__global__ void pack_kernel(const char* rec, int bytes, int n, float l0, int4* contacts, int* mismatch) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  const char* r = rec + static_cast<long long>(i) * bytes;
  const int* q = reinterpret_cast<const int*>(r);
  contacts[i] = make_int4(q[0], q[1], q[2], q[3]);
  if (mismatch) {
    const S rl0 = *reinterpret_cast<const S*>(r + 24);
    if (rl0 != static_cast<S>(l0)) atomicAdd(mismatch, 1);
  }
}

template <typename S>
__global__ void unpack_kernel(const int4* contacts, const int2* e0e1, int n, float l0, const double* alpha, char* rec,
                              int bytes) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i >= n) return;
  char* r = rec + static_cast<long long>(i) * bytes;
  const int4 c = contacts[i];
  int* q = reinterpret_cast<int*>(r);
  q[0] = c.x;
  q[1] = c.y;
  q[2] = c.z;
  q[3] = c.w;
  const int2 e = e0e1[i];
  q[4] = e.x;
  q[5] = e.y;
  S* s = reinterpret_cast<S*>(r + 24);
  s[0] = static_cast<S>(l0);
  s[1] = alpha ? static_cast<S>(alpha[i]) : static_cast<S>(1e-6);
  // lam (s[2]) is the augmented-Lagrangian multiplier: append leaves it alone
}

bool layout_ok(int bytes) { return bytes == IPC_XCONSTRAINT_BYTES_F32 || bytes == IPC_XCONSTRAINT_BYTES_F64; }

}  // namespace

extern "C" {

int ipc_contact_records_pack(const void* records, int record_bytes, int n, float l0, int* contacts, int* l0_mismatch,
                             void* stream) {
  if (n < 0) return rec_fail(IPC_EINVAL, "contact_records_pack: negative n");
  if (!layout_ok(record_bytes))
    return rec_fail(IPC_EINVAL, "contact_records_pack: record_bytes must be 36 (fp32 scalars) or 48 (fp64 scalars)");
  if (n == 0) return IPC_OK;
  if (!records || !contacts) return rec_fail(IPC_EINVAL, "contact_records_pack: null pointer");
  const unsigned long long ra = reinterpret_cast<unsigned long long>(records);
  if ((ra & (record_bytes == IPC_XCONSTRAINT_BYTES_F64 ? 7ull : 3ull)) ||
      (reinterpret_cast<unsigned long long>(contacts) & 15ull) ||
      (reinterpret_cast<unsigned long long>(l0_mismatch) & 3ull))
    return rec_fail(IPC_EINVAL, "contact_records_pack: misaligned pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int g = (n + kB - 1) / kB;
  const char* rec = static_cast<const char*>(records);
  int4* out = reinterpret_cast<int4*>(contacts);
  if (record_bytes == IPC_XCONSTRAINT_BYTES_F32)
    hipLaunchKernelGGL(pack_kernel<float>, dim3(g), dim3(kB), 0, s, rec, record_bytes, n, l0, out, l0_mismatch);
  else
    hipLaunchKernelGGL(pack_kernel<double>, dim3(g), dim3(kB), 0, s, rec, record_bytes, n, l0, out, l0_mismatch);
  return rec_hip(hipGetLastError(), "contact_records_pack");
}

int ipc_contact_records_unpack(const int* contacts, const int* e0e1, int n, float l0, const double* alpha,
                               void* records, int record_bytes, void* stream) {
  if (n < 0) return rec_fail(IPC_EINVAL, "contact_records_unpack: negative n");
  if (!layout_ok(record_bytes))
    return rec_fail(IPC_EINVAL, "contact_records_unpack: record_bytes must be 36 (fp32 scalars) or 48 (fp64 scalars)");
  if (n == 0) return IPC_OK;
  if (!contacts || !e0e1 || !records) return rec_fail(IPC_EINVAL, "contact_records_unpack: null pointer");
  if ((reinterpret_cast<unsigned long long>(contacts) & 15ull) || (reinterpret_cast<unsigned long long>(e0e1) & 7ull) ||
      (reinterpret_cast<unsigned long long>(alpha) & 7ull) ||
      (reinterpret_cast<unsigned long long>(records) & (record_bytes == IPC_XCONSTRAINT_BYTES_F64 ? 7ull : 3ull)))
    return rec_fail(IPC_EINVAL, "contact_records_unpack: misaligned pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int g = (n + kB - 1) / kB;
  const int4* c = reinterpret_cast<const int4*>(contacts);
  const int2* e = reinterpret_cast<const int2*>(e0e1);
  char* rec = static_cast<char*>(records);
  if (record_bytes == IPC_XCONSTRAINT_BYTES_F32)
    hipLaunchKernelGGL(unpack_kernel<float>, dim3(g), dim3(kB), 0, s, c, e, n, l0, alpha, rec, record_bytes);
  else
    hipLaunchKernelGGL(unpack_kernel<double>, dim3(g), dim3(kB), 0, s, c, e, n, l0, alpha, rec, record_bytes);
  return rec_hip(hipGetLastError(), "contact_records_unpack");
}

}  // extern "C"

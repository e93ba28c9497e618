// Distance-based conservative advancement (ccd/ccd.py:7-163,
// conservative_ee_toi / conservative_pt_toi): the reference's alternative
// to the cubic-root CCD, defined but not called on its path. Per candidate,
// in FP64: the motions relative to their mean bound how fast the two
// primitives can approach; the pair advances by 0.9 distance / bound at a
// time until the distance falls below a tenth of the initial one or the
// step reaches 1.
#include <cmath>

#include "../../include/ipc_warp_amd.h"
#include "ipc_device.hpp"
#include "ipc_kernels.hpp"

namespace ipcd {
namespace {

using D3 = V3<double>;

__device__ inline double clamp01(double v) { return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v); }

// Squared segment distance with the degenerate and near-parallel cases of
// the reference (ccd/ccd.py:7-38).
__device__ double segment_distance_sq(D3 a0, D3 a1, D3 b0, D3 b1) {
  const D3 da = a1 - a0, db = b1 - b0, r = a0 - b0;
  const double aa = dot(da, da), bb = dot(db, db), f = dot(db, r);
  double s = 0.0, t = 0.0;
  if (aa <= 1.0e-20 && bb <= 1.0e-20) return dot(r, r);
  if (aa <= 1.0e-20) {
    t = clamp01(f / bb);
  } else if (bb <= 1.0e-20) {
    s = clamp01(-dot(da, r) / aa);
  } else {
    const double ab = dot(da, db), c = dot(da, r);
    const double denom = aa * bb - ab * ab;
    if (denom > 1.0e-20) s = clamp01((ab * f - c * bb) / denom);
    t = (ab * s + f) / bb;
    if (t < 0.0) {
      t = 0.0;
      s = clamp01(-c / aa);
    } else if (t > 1.0) {
      t = 1.0;
      s = clamp01((ab - c) / aa);
    }
  }
  const D3 d = (a0 + s * da) - (b0 + t * db);
  return dot(d, d);
}

// Squared point-triangle distance by Voronoi region (ccd/ccd.py:41-83); a
// degenerate triangle falls back to its three edges.
__device__ double triangle_distance_sq(D3 p, D3 a, D3 b, D3 c) {
  const D3 ab = b - a, ac = c - a, ap = p - a;
  const double d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0 && d2 <= 0.0) return dot(ap, ap);
  const D3 bp = p - b;
  const double d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0 && d4 <= d3) return dot(bp, bp);
  const double vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
    const double v = d1 / (d1 - d3);
    const D3 e = p - (a + v * ab);
    return dot(e, e);
  }
  const D3 cp = p - c;
  const double d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0 && d5 <= d6) return dot(cp, cp);
  const double vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
    const double w = d2 / (d2 - d6);
    const D3 e = p - (a + w * ac);
    return dot(e, e);
  }
  const double va = d3 * d6 - d5 * d4;
  if (va <= 0.0 && d4 >= d3 && d5 >= d6) {
    const double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    const D3 e = p - (b + w * (c - b));
    return dot(e, e);
  }
  const double denom = va + vb + vc;
  if (fabs(denom) <= 1.0e-20) {
    const double dab = segment_distance_sq(p, p, a, b), dac = segment_distance_sq(p, p, a, c),
                 dbc = segment_distance_sq(p, p, b, c);
    return fmin(dab, fmin(dac, dbc));
  }
  const double inv = 1.0 / denom;
  const double v = vb * inv, w = vc * inv;
  const D3 e = p - (a + v * ab + w * ac);
  return dot(e, e);
}

template <bool EE>
__device__ double pair_distance(const D3* x) {
  return EE ? sqrt(segment_distance_sq(x[0], x[1], x[2], x[3])) : sqrt(triangle_distance_sq(x[0], x[1], x[2], x[3]));
}

// x[4] at t = 0, d[4] the motions to t = 1 (EE: a0 a1 b0 b1; PT: p a b c).
template <bool EE>
__device__ double conservative_toi(D3* x, D3* d) {
  const D3 mean = (((d[0] + d[1]) + d[2]) + d[3]) * 0.25;
#pragma unroll
  for (int k = 0; k < 4; ++k) d[k] = d[k] - mean;
  double max_disp;
  if (EE) {
    max_disp = sqrt(fmax(dot(d[0], d[0]), dot(d[1], d[1]))) + sqrt(fmax(dot(d[2], d[2]), dot(d[3], d[3])));
  } else {
    max_disp = sqrt(dot(d[0], d[0])) + sqrt(fmax(dot(d[1], d[1]), fmax(dot(d[2], d[2]), dot(d[3], d[3]))));
  }
  if (max_disp <= 1.0e-20) return 1.0;
  double distance = pair_distance<EE>(x);
  const double gap = 0.1 * distance;
  double toi = 0.0;
  for (int it = 0; it < 100; ++it) {
    const double step = 0.9 * distance / max_disp;
    if (step <= 1.0e-15) return toi;
#pragma unroll
    for (int k = 0; k < 4; ++k) x[k] = x[k] + step * d[k];
    distance = pair_distance<EE>(x);
    if (toi > 0.0 && distance < gap) return toi;
    toi += step;
    if (toi >= 1.0) return 1.0;
  }
  return fmin(toi, 1.0);
}

template <bool EE>
__global__ void __launch_bounds__(256) conservative_time_kernel(const float* __restrict__ x0, const float* __restrict__ x1,
                                                                int nv, const int4* __restrict__ quads, int n,
                                                                double* toi_out, unsigned long long* toi_min) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double t = 1.0;
  if (i < n) {
    const int4 q = quads[i];
    const int id[4] = {q.x, q.y, q.z, q.w};
    bool ok = true;
#pragma unroll
    for (int k = 0; k < 4; ++k) ok &= unsigned(id[k]) < unsigned(nv);
    if (ok) {
      D3 x[4], d[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        x[k] = to_d(ldx(x0, id[k]));
        d[k] = to_d(ldx(x1, id[k])) - x[k];
      }
      t = conservative_toi<EE>(x, d);
    }
    if (toi_out) toi_out[i] = t;
  }
  if (toi_min) {
    // per-wave minimum, then one atomic on the IEEE bits (t >= 0)
    unsigned long long bits = __double_as_longlong(t);
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned long long o = __shfl_xor(bits, off, 64);
      bits = o < bits ? o : bits;
    }
    if ((threadIdx.x & 63) == 0 && bits < 0x3FF0000000000000ull &&
        bits < __hip_atomic_load(toi_min, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
      atomicMin(toi_min, bits);
  }
}

}  // namespace
}  // namespace ipcd

using namespace ipcd;

static int conservative_time(bool ee, const float* x0, const float* x1, int nv, const int* quads, int n, double* toi,
                             unsigned long long* toi_min, void* stream) {
  if (n < 0 || nv < 0) {
    set_error("conservative_time: negative n_vertices or n");
    return IPC_EINVAL;
  }
  if (n == 0) return IPC_OK;
  if (!x0 || !x1 || !quads) {
    set_error("conservative_time: null pointer");
    return IPC_EINVAL;
  }
  if (reinterpret_cast<unsigned long long>(quads) & 15ull) {
    set_error("conservative_time: quads must be 16-byte aligned");
    return IPC_EINVAL;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int g = (n + 255) / 256;
  const int4* q4 = reinterpret_cast<const int4*>(quads);
  if (ee)
    hipLaunchKernelGGL(conservative_time_kernel<true>, dim3(g), dim3(256), 0, s, x0, x1, nv, q4, n, toi, toi_min);
  else
    hipLaunchKernelGGL(conservative_time_kernel<false>, dim3(g), dim3(256), 0, s, x0, x1, nv, q4, n, toi, toi_min);
  const hipError_t e = hipGetLastError();
  if (e == hipSuccess) return IPC_OK;
  set_error(hipGetErrorString(e));
  return -int(e);
}

extern "C" {

int ipc_conservative_time_ee(const float* x0, const float* x1, int n_vertices, const int* quads, int n, double* toi,
                             unsigned long long* toi_min, void* stream) {
  return conservative_time(true, x0, x1, n_vertices, quads, n, toi, toi_min, stream);
}

int ipc_conservative_time_pt(const float* x0, const float* x1, int n_vertices, const int* quads, int n, double* toi,
                             unsigned long long* toi_min, void* stream) {
  return conservative_time(false, x0, x1, n_vertices, quads, n, toi, toi_min, stream);
}

}  // extern "C"

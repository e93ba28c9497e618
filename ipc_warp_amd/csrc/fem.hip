// Elastic tetrahedra (fem/fem.py:21-34 compute_Dm, :138-203
// tet_kernel_sparse; time_integrator_base.py:124-137 compute_Psi) with the
// stable Neo-Hookean model of fem/snh.py:1-84.
//
// The reference launches 16 threads per tet (one per block (i, j)), and
// each recomputes F, P and the three directional derivatives dP for its
// column j. Here one lane owns a tet: F and P once, the twelve directional
// derivatives (j, k) once each, and every block (i, j) column k read from
// the same dH = -W dP Bm^T. Each value is formed by the reference's own
// operations in its order, so the blocks equal what its threads compute.
#include <cstdio>

#include "../../include/ipc_warp_amd.h"
#include "ipc_kernels.hpp"

namespace ipcd {
namespace {

// 3x3 matrix, row-major m[r][c]
struct M3 {
  float m[3][3];
};

__device__ inline M3 mzero() {
  M3 a;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) a.m[r][c] = 0.0f;
  return a;
}

__device__ inline M3 from_cols(const float* c0, const float* c1, const float* c2) {
  M3 a;
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    a.m[r][0] = c0[r];
    a.m[r][1] = c1[r];
    a.m[r][2] = c2[r];
  }
  return a;
}

__device__ inline M3 matmul(const M3& a, const M3& b) {
  M3 o;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      float s = 0.0f;
#pragma unroll
      for (int k = 0; k < 3; ++k) s += a.m[r][k] * b.m[k][c];
      o.m[r][c] = s;
    }
  return o;
}

// a @ b^T
__device__ inline M3 matmul_t(const M3& a, const M3& b) {
  M3 o;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      float s = 0.0f;
#pragma unroll
      for (int k = 0; k < 3; ++k) s += a.m[r][k] * b.m[c][k];
      o.m[r][c] = s;
    }
  return o;
}

__device__ inline M3 scale(float s, const M3& a) {
  M3 o;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) o.m[r][c] = s * a.m[r][c];
  return o;
}

__device__ inline M3 add(const M3& a, const M3& b) {
  M3 o;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) o.m[r][c] = a.m[r][c] + b.m[r][c];
  return o;
}

__device__ inline float ddot(const M3& a, const M3& b) {
  float s = 0.0f;
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) s += a.m[r][c] * b.m[r][c];
  return s;
}

__device__ inline float det3(const M3& a) {
  return a.m[0][0] * (a.m[1][1] * a.m[2][2] - a.m[1][2] * a.m[2][1]) -
         a.m[0][1] * (a.m[1][0] * a.m[2][2] - a.m[1][2] * a.m[2][0]) +
         a.m[0][2] * (a.m[1][0] * a.m[2][1] - a.m[1][1] * a.m[2][0]);
}

__device__ inline void cross3(const float* a, const float* b, float* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}

// column c of a
__device__ inline void col(const M3& a, int c, float* v) {
  v[0] = a.m[0][c];
  v[1] = a.m[1][c];
  v[2] = a.m[2][c];
}

struct Snh {
  float mu, lam, alpha;  // the reparametrised constants (fem/snh.py:6-12)
};

__device__ inline Snh reparam(float mu, float lam) {
  Snh s;
  s.mu = float(4.0 / 3.0) * mu;
  s.lam = lam + float(5.0 / 6.0) * mu;
  s.alpha = 1.0f + 0.75f * s.mu / s.lam;
  return s;
}

// dJ/dF: columns f1 x f2, f2 x f0, f0 x f1 of F's columns
__device__ inline M3 cofactor(const M3& F) {
  float f0[3], f1[3], f2[3], c0[3], c1[3], c2[3];
  col(F, 0, f0);
  col(F, 1, f1);
  col(F, 2, f2);
  cross3(f1, f2, c0);
  cross3(f2, f0, c1);
  cross3(f0, f1, c2);
  return from_cols(c0, c1, c2);
}

__device__ inline M3 pk1(const M3& F, const Snh& s) {
  const float Ic = ddot(F, F);
  const M3 cof = cofactor(F);
  const float distortional = s.mu * (1.0f - 1.0f / (Ic + 1.0f));
  const float volumetric = s.lam * (det3(F) - s.alpha);
  return add(scale(distortional, F), scale(volumetric, cof));
}

// Directional derivative of PK1 at F along dF (fem/snh.py:43-84).
__device__ inline M3 tangent(const M3& F, const M3& dF, const Snh& s) {
  const M3 cof = cofactor(F);
  const float Ic = ddot(F, F);
  const float Ic1 = Ic + 1.0f;
  const float dIc = 2.0f * ddot(F, dF);
  const float distortional = s.mu * (1.0f - 1.0f / Ic1);
  const float d_distortional = s.mu * dIc / (Ic1 * Ic1);
  const float dJ = ddot(cof, dF);
  const float volumetric = s.lam * (det3(F) - s.alpha);
  float f0[3], f1[3], f2[3], g0[3], g1[3], g2[3];
  col(F, 0, f0);
  col(F, 1, f1);
  col(F, 2, f2);
  col(dF, 0, g0);
  col(dF, 1, g1);
  col(dF, 2, g2);
  float a[3], b[3], c0[3], c1[3], c2[3];
  cross3(g1, f2, a);
  cross3(f1, g2, b);
  c0[0] = a[0] + b[0]; c0[1] = a[1] + b[1]; c0[2] = a[2] + b[2];
  cross3(g2, f0, a);
  cross3(f2, g0, b);
  c1[0] = a[0] + b[0]; c1[1] = a[1] + b[1]; c1[2] = a[2] + b[2];
  cross3(g0, f1, a);
  cross3(f0, g1, b);
  c2[0] = a[0] + b[0]; c2[1] = a[1] + b[1]; c2[2] = a[2] + b[2];
  const M3 dcof = from_cols(c0, c1, c2);
  return add(add(add(scale(distortional, dF), scale(d_distortional, F)), scale(s.lam * dJ, cof)),
             scale(volumetric, dcof));
}

__device__ inline bool tet_ok(const int4 t, int nv) {
  return unsigned(t.x) < unsigned(nv) && unsigned(t.y) < unsigned(nv) && unsigned(t.z) < unsigned(nv) &&
         unsigned(t.w) < unsigned(nv);
}

// Ds = [x0 - x3, x1 - x3, x2 - x3], F = Ds Bm
__device__ inline M3 deformation(const float* x, int4 t, const float* bm, M3& Bm) {
  float d[3][3];
  const int id[3] = {t.x, t.y, t.z};
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) d[c][r] = x[3ll * id[c] + r] - x[3ll * t.w + r];
  const M3 Ds = from_cols(d[0], d[1], d[2]);
#pragma unroll
  for (int r = 0; r < 3; ++r)
#pragma unroll
    for (int c = 0; c < 3; ++c) Bm.m[r][c] = bm[3 * r + c];
  return matmul(Ds, Bm);
}

__global__ void tet_rest_kernel(const float* __restrict__ xcs, int nv, const int4* __restrict__ T, int n,
                                float* __restrict__ Bm, float* __restrict__ W) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int4 t = T[e];
  float* bm = Bm + 9ll * e;
  if (!tet_ok(t, nv)) {
#pragma unroll
    for (int k = 0; k < 9; ++k) bm[k] = 0.0f;
    W[e] = 0.0f;
    return;
  }
  float d[3][3];
  const int id[3] = {t.x, t.y, t.z};
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int r = 0; r < 3; ++r) d[c][r] = xcs[3ll * id[c] + r] - xcs[3ll * t.w + r];
  const M3 Dm = from_cols(d[0], d[1], d[2]);
  const float det = det3(Dm);
  // inverse by the adjugate over the determinant
  const float inv = 1.0f / det;
  const float (&m)[3][3] = Dm.m;
  bm[0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) * inv;
  bm[1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * inv;
  bm[2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * inv;
  bm[3] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) * inv;
  bm[4] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * inv;
  bm[5] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * inv;
  bm[6] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) * inv;
  bm[7] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * inv;
  bm[8] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * inv;
  W[e] = fabsf(det) / 6.0f;
}

// Per wave in LDS: its 32 tets' vertex ids and flags, their forces on the
// way to b, and one block row (blocks (i, 0..3), 36 floats) of each tet's
// values on the way to memory.
struct TetStage {
  float f[4][32 * 12];
  int v[4][32 * 4];
  int ok[4][32];
  float4 row[4][32 * 9];
};

// Two lanes per tet, lanes p and p + 32 of a wave: the first takes the
// directional derivatives of columns j = 0, 1, the second those of j = 2, 3
// (six each), and the first the force. Blocks (offset + e) * 16 + 4 i + j
// with row T[e][i], column T[e][j]; the force added into b. A tet naming a
// vertex out of range writes zero blocks with row = col = 0 and no force.
//
// A tet's 16 blocks are 576 contiguous bytes and a wave's tets are
// consecutive, so the wave's output is one contiguous range. The blocks go
// out block row by block row through LDS (each lane puts its two blocks of
// row i there), and the wave stores 32 tets' row i with consecutive lanes
// on consecutive 16-byte words; a lane storing its own tet's 576 bytes
// touched 64 cache lines per store instruction. Row and column indices go
// out the same way from the staged vertex ids.
constexpr int kTetsPerWave = 32;
constexpr int kTetsPerBlock = 4 * kTetsPerWave;

__global__ void __launch_bounds__(256) tet_hessian_kernel(const float* __restrict__ x, int nv,
                                                          const int4* __restrict__ T, int n,
                                                          const float* __restrict__ Bmv, const float* __restrict__ Wv,
                                                          float mu, float lam, long long base, int* __restrict__ rows,
                                                          int* __restrict__ cols, float* __restrict__ vals, float* b) {
  __shared__ TetStage st;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int p = lane & 31, half = lane >> 5;
  const int first = blockIdx.x * kTetsPerBlock + wave * kTetsPerWave;
  if (first >= n) return;  // the whole wave is past the end
  const int nw = min(kTetsPerWave, n - first);
  const int e = first + p;
  int4 t = make_int4(0, 0, 0, 0);
  bool ok = false;
  if (e < n) {
    t = T[e];
    ok = tet_ok(t, nv);
  }
  // h[jj][k][l][i] = dH(j, k).m[l][i] with j = 2 half + jj, i < 3: entry
  // (l, k) of block (i, j)
  float h[2][3][3][3];
  float force[12];
#pragma unroll
  for (int q = 0; q < 12; ++q) force[q] = 0.0f;
  if (ok) {
    const Snh s = reparam(mu, lam);
    M3 Bm;
    const M3 F = deformation(x, t, Bmv + 9ll * e, Bm);
    const float W = Wv[e];
#pragma unroll
    for (int jj = 0; jj < 2; ++jj)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const int j = 2 * half + jj;
        M3 dDs = mzero();
        if (j < 3) {
          dDs.m[k][0] = j == 0 ? -1.0f : 0.0f;
          dDs.m[k][1] = j == 1 ? -1.0f : 0.0f;
          dDs.m[k][2] = j == 2 ? -1.0f : 0.0f;
        } else {
          dDs.m[k][0] = 1.0f;
          dDs.m[k][1] = 1.0f;
          dDs.m[k][2] = 1.0f;
        }
        const M3 dF = matmul(dDs, Bm);
        const M3 dP = tangent(F, dF, s);
        const M3 dH = matmul_t(scale(-W, dP), Bm);
#pragma unroll
        for (int l = 0; l < 3; ++l)
#pragma unroll
          for (int i = 0; i < 3; ++i) h[jj][k][l][i] = dH.m[l][i];
      }
    if (b && half == 0) {
      // columns of H = -W P Bm^T, the last the negated sum
      const M3 H = matmul_t(scale(-W, pk1(F, s)), Bm);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int l = 0; l < 3; ++l)
          force[3 * i + l] = i == 3 ? ((-H.m[l][0]) - H.m[l][1]) - H.m[l][2] : H.m[l][i];
    }
  }
  float* sf = st.f[wave];
  int* sv = st.v[wave];
  int* so = st.ok[wave];
  float* sr = reinterpret_cast<float*>(st.row[wave]);
  if (half == 0) {
#pragma unroll
    for (int q = 0; q < 12; ++q) sf[p * 12 + q] = force[q];
    sv[p * 4 + 0] = t.x;
    sv[p * 4 + 1] = t.y;
    sv[p * 4 + 2] = t.z;
    sv[p * 4 + 3] = t.w;
    so[p] = ok ? 1 : 0;
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  const long long t0 = (base + first) * 16;
  {
    int4* rows4 = reinterpret_cast<int4*>(rows + t0);
    int4* cols4 = reinterpret_cast<int4*>(cols + t0);
    for (int q = lane; q < nw * 4; q += 64) {
      const int sl = q >> 2, i = q & 3;
      int4 rv = make_int4(0, 0, 0, 0), cv = make_int4(0, 0, 0, 0);
      if (so[sl]) {
        const int vi = sv[sl * 4 + i];
        rv = make_int4(vi, vi, vi, vi);
        cv = make_int4(sv[sl * 4 + 0], sv[sl * 4 + 1], sv[sl * 4 + 2], sv[sl * 4 + 3]);
      }
      rows4[q] = rv;
      cols4[q] = cv;
    }
  }
  float4* v4 = reinterpret_cast<float4*>(vals + 9 * t0);
  const float4* sr4 = st.row[wave];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    float blk[18];  // blocks (i, 2 half), (i, 2 half + 1): entry (l, k) of block jj at 9 jj + 3 l + k
#pragma unroll
    for (int jj = 0; jj < 2; ++jj)
#pragma unroll
      for (int l = 0; l < 3; ++l)
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const float df = i == 3 ? -(h[jj][k][l][0] + h[jj][k][l][1] + h[jj][k][l][2]) : h[jj][k][l][i];
          blk[9 * jj + 3 * l + k] = ok ? 0.0f + df : 0.0f;
        }
    float2* dst = reinterpret_cast<float2*>(sr + p * 36 + 18 * half);
#pragma unroll
    for (int q = 0; q < 9; ++q) dst[q] = make_float2(blk[2 * q], blk[2 * q + 1]);
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    for (int q = lane; q < nw * 9; q += 64) {
      const int sl = q / 9, u = q - 9 * sl;
      v4[sl * 36 + i * 9 + u] = sr4[q];
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  }
  // forces into b, component q of the wave's 12 nw by lane q % 64, so that
  // one atomic instruction adds consecutive components of consecutive tets
  if (b) {
    for (int q = lane; q < 12 * nw; q += 64) {
      const float v = sf[q];
      const int sl = q / 12, comp = q - 12 * sl, vert = comp / 3;
      if (so[sl] && v != 0.0f) atomicAdd(b + 3ll * sv[sl * 4 + vert] + (comp - 3 * vert), v);
    }
  }
}

// Grid-stride over the tets: each lane sums its tets' energies in double,
// then the block's sum goes into *energy with one atomic (the atomics on
// one address serialise, so their number is kept to the grid's size).
constexpr int kEnergyGrid = 1024;

__global__ void __launch_bounds__(256) tet_energy_kernel(const float* __restrict__ x, int nv,
                                                         const int4* __restrict__ T, int n,
                                                         const float* __restrict__ Bmv, const float* __restrict__ Wv,
                                                         float mu, float lam, float* psi_out, double* energy) {
  __shared__ double partial[4];
  double ed = 0.0;
  const Snh s = reparam(mu, lam);
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
    const int4 t = T[e];
    float pe = 0.0f;
    if (tet_ok(t, nv)) {
      M3 Bm;
      const M3 F = deformation(x, t, Bmv + 9ll * e, Bm);
      const float Ic = ddot(F, F);
      const float JmA = det3(F) - s.alpha;
      const float psi = 0.5f * (s.lam * JmA * JmA + s.mu * (Ic - 3.0f) - s.mu * logf(Ic + 1.0f));
      pe = Wv[e] * psi;
    }
    if (psi_out) psi_out[e] = pe;
    ed += pe;
  }
  if (energy) {
    for (int off = 32; off > 0; off >>= 1) ed += __shfl_xor(ed, off, 64);
    if ((threadIdx.x & 63) == 0) partial[threadIdx.x >> 6] = ed;
    __syncthreads();
    if (threadIdx.x == 0) {
      const double sum = ((partial[0] + partial[1]) + partial[2]) + partial[3];
      if (sum != 0.0) atomicAdd(energy, sum);
    }
  }
}

}  // namespace
}  // namespace ipcd

using namespace ipcd;

namespace {

int fem_fail(const char* what) {
  set_error(what);
  return IPC_EINVAL;
}

int fem_hip(hipError_t e, const char* where) {
  if (e == hipSuccess) return IPC_OK;
  char buf[512];
  std::snprintf(buf, sizeof(buf), "%s: %s", where, hipGetErrorString(e));
  set_error(buf);
  return -int(e);
}

bool a16(const void* p) { return (reinterpret_cast<unsigned long long>(p) & 15ull) == 0; }

}  // namespace

extern "C" {

int ipc_tet_rest(const float* xcs, int n_vertices, const int* tets, int n_tets, float* Bm, float* W, void* stream) {
  if (n_vertices < 0 || n_tets < 0) return fem_fail("tet_rest: negative size");
  if (n_tets == 0) return IPC_OK;
  if (!xcs || !tets || !Bm || !W) return fem_fail("tet_rest: null pointer");
  if (!a16(tets)) return fem_fail("tet_rest: tets must be 16-byte aligned");
  hipLaunchKernelGGL(tet_rest_kernel, dim3((n_tets + 255) / 256), dim3(256), 0, static_cast<hipStream_t>(stream), xcs,
                     n_vertices, reinterpret_cast<const int4*>(tets), n_tets, Bm, W);
  return fem_hip(hipGetLastError(), "tet_rest");
}

int ipc_tet_hessian(const float* x, int n_vertices, const int* tets, int n_tets, const float* Bm, const float* W,
                    float mu, float lam, int offset, int* rows, int* cols, float* vals, float* b, void* stream) {
  if (n_vertices < 0 || n_tets < 0 || offset < 0) return fem_fail("tet_hessian: negative size");
  if (n_tets == 0) return IPC_OK;
  if (!x || !tets || !Bm || !W || !rows || !cols || !vals) return fem_fail("tet_hessian: null pointer");
  if (!a16(tets) || !a16(rows) || !a16(cols) || !a16(vals))
    return fem_fail("tet_hessian: tets/rows/cols/vals must be 16-byte aligned");
  hipLaunchKernelGGL(tet_hessian_kernel, dim3((n_tets + kTetsPerBlock - 1) / kTetsPerBlock), dim3(256), 0,
                     static_cast<hipStream_t>(stream), x,
                     n_vertices, reinterpret_cast<const int4*>(tets), n_tets, Bm, W, mu, lam, (long long)offset, rows,
                     cols, vals, b);
  return fem_hip(hipGetLastError(), "tet_hessian");
}

int ipc_tet_energy(const float* x, int n_vertices, const int* tets, int n_tets, const float* Bm, const float* W,
                   float mu, float lam, float* psi, double* energy, void* stream) {
  if (n_vertices < 0 || n_tets < 0) return fem_fail("tet_energy: negative size");
  if (n_tets == 0) return IPC_OK;
  if (!x || !tets || !Bm || !W) return fem_fail("tet_energy: null pointer");
  if (!a16(tets)) return fem_fail("tet_energy: tets must be 16-byte aligned");
  const int grid = (n_tets + 255) / 256 < kEnergyGrid ? (n_tets + 255) / 256 : kEnergyGrid;
  hipLaunchKernelGGL(tet_energy_kernel, dim3(grid), dim3(256), 0, static_cast<hipStream_t>(stream), x,
                     n_vertices, reinterpret_cast<const int4*>(tets), n_tets, Bm, W, mu, lam, psi, energy);
  return fem_hip(hipGetLastError(), "tet_energy");
}

}  // extern "C"

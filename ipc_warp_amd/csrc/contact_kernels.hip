// Barrier gradient / PSD Hessian / energy kernels and CCD time-of-impact
// kernels of libipc_warp_amd (gfx950, wave64).
//
// One lane owns one contact. Contacts are a stream of int4 vertex quadruples
// (16 B, one 128-bit load per lane, coalesced across the wave). Vertex
// positions are gathered from the float32 [V, 3] array (L2/MALL resident at
// every configuration this library targets).
#include "ipc_device.hpp"
#include "ipc_kernels.hpp"

namespace ipcd {

// --------------------------------------------------------------- per pair ----
// Indices outside [0, n_vertices) make a contact inactive instead of faulting.
__device__ __forceinline__ bool valid4(int4 c, int nv) {
  return (unsigned)c.x < (unsigned)nv && (unsigned)c.y < (unsigned)nv && (unsigned)c.z < (unsigned)nv &&
         (unsigned)c.w < (unsigned)nv;
}

// contact_hessian_ee (dynamic_contacts.py:151-207). Returns false for an
// inactive contact (fp32 distance >= l0). On an active contact, fills the
// compressed Hessian and the force (-gradient) contribution of each slot.
// P: the positions' (and classification's) type; R: the outputs' type.
template <typename P, typename R>
__device__ __forceinline__ bool pair_ee(const P* __restrict__ x, const P* __restrict__ xcs, int4 c, P l0,
                                        double kappa, bool want_hessian, CompressedHessianT<R>& C, R force[12],
                                        double* energy) {
  V3<P> xf[4] = {ldx(x, c.x), ldx(x, c.y), ldx(x, c.z), ldx(x, c.w)};
  EdgeClosestT<P> ec = closest_point_edge_edge(xf[0], xf[1], xf[2], xf[3], lit<P>(1.0e-6f, 1.0e-6));
  if (!(ec.dist < l0)) return false;
  V3<P> ra = ldx(xcs, c.y) - ldx(xcs, c.x), rb = ldx(xcs, c.w) - ldx(xcs, c.z);
  P eps_x = lit<P>(1.0e-3f, 1.0e-3) * dot(ra, ra) * dot(rb, rb);
  V3<P> cr = cross(xf[1] - xf[0], xf[3] - xf[2]);
  bool near_parallel = eps_x > P(0) && dot(cr, cr) < eps_x;
  Feature f = ee_feature(ec, cr, eps_x);
  V3<double> xo = to_d(xf[0]);
  V3<double> xr[4] = {mk(0.0, 0.0, 0.0), to_d(xf[1]) - xo, to_d(xf[2]) - xo, to_d(xf[3]) - xo};
  FeatureGeom g = feature_geom(f, xr);
  double d2 = dot(g.r, g.r);
  Barrier B = barrier(d2, double(l0) * double(l0), kappa);
  double m = 1.0;
  // grad m over the slots is (-ga, ga, -gb, gb)
  V3<double> ga = mk(0.0, 0.0, 0.0), gb = mk(0.0, 0.0, 0.0);
  if (near_parallel) {
    // IPC edge-edge mollifier, m(c) = (c/eps_x)(2 - c/eps_x), c = |ea x eb|^2.
    // c is a cancellation of nearly parallel edges: eb is formed from the
    // absolute coordinates (one rounding of x3 - x2), not as the difference
    // of the two offsets from slot 0, which round at the scale of |x - x0|.
    V3<double> ea = xr[1], eb = to_d(xf[3]) - to_d(xf[2]);
    V3<double> n = cross(ea, eb);
    double ex = double(eps_x);
    double q = dot(n, n) / ex;
    m = q * (2.0 - q);
    double dm = 2.0 / ex * (1.0 - q);
    ga = cross(eb, n) * (2.0 * dm);
    gb = cross(n, ea) * (2.0 * dm);
  }
  if (energy) *energy = m * B.b;
  if (!want_hessian) return true;
  // force = -(m b' grad d^2 + b grad m), grad m zero outside the band. It is
  // formed first, so that the feature geometry is dead before the spectral
  // projection, the kernel's register peak.
  double bm = near_parallel ? B.b : 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    double s = -2.0 * m * B.db * g.w[i];
    V3<double> gmi = i == 0 ? mk(-ga.x, -ga.y, -ga.z) : (i == 1 ? ga : (i == 2 ? mk(-gb.x, -gb.y, -gb.z) : gb));
    force[3 * i + 0] = R(s * g.r.x - bm * gmi.x);
    force[3 * i + 1] = R(s * g.r.y - bm * gmi.y);
    force[3 * i + 2] = R(s * g.r.z - bm * gmi.z);
  }
  if (near_parallel) {
    // Near-parallel edges keep only the Gauss-Newton term m b'' g g^T.
    C.cn = R(0);
    C.cg = R(4.0 * m * B.d2b * d2);
  } else {
    // Outside the mollified band m = 1 and grad m = 0, so the reference's
    // b * H_m and PSD(sym(grad m, b' g)) terms are exactly zero.
    C.cn = R(B.db);
    C.cg = R(4.0 * B.d2b * d2);
  }
  nsd_compressed(g, !near_parallel, C);
  return true;
}

// contact_hessian_pt (dynamic_contacts.py:209-246). The record is
// (p, t1, t0, t2) and slots 1..3 are the triangle (a, b, c) of
// fetch_dist_v0v1_pt (contact.py:116-145).
template <typename P, typename R>
__device__ __forceinline__ bool pair_pt(const P* __restrict__ x, int4 c, P l0, double kappa, bool want_hessian,
                                        CompressedHessianT<R>& C, R force[12], double* energy) {
  V3<P> xf[4] = {ldx(x, c.x), ldx(x, c.y), ldx(x, c.z), ldx(x, c.w)};
  TriClosestT<P> tc = closest_point_triangle(xf[0], xf[1], xf[2], xf[3]);
  if (!(tc.dist < l0)) return false;
  Feature f = pt_feature(tc.type);
  V3<double> xo = to_d(xf[0]);
  V3<double> xr[4] = {mk(0.0, 0.0, 0.0), to_d(xf[1]) - xo, to_d(xf[2]) - xo, to_d(xf[3]) - xo};
  FeatureGeom g = feature_geom(f, xr);
  double d2 = dot(g.r, g.r);
  Barrier B = barrier(d2, double(l0) * double(l0), kappa);
  if (energy) *energy = B.b;
  if (!want_hessian) return true;
  C.cn = R(B.db);
  C.cg = R(4.0 * B.d2b * d2);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    double s = -2.0 * B.db * g.w[i];
    force[3 * i + 0] = R(s * g.r.x);
    force[3 * i + 1] = R(s * g.r.y);
    force[3 * i + 2] = R(s * g.r.z);
  }
  nsd_compressed(g, true, C);
  return true;
}

// ----------------------------------------------------------- hessian kernel ----
// Writes the 16 triplets of contact i at (offset + i) * 16 + ii * 4 + jj
// (dynamic_contacts.py:196-207,236-246); inactive contacts get zero blocks
// with row = col = 0, which is what the reference's zero-initialised
// triplet arrays hold for them. The force is atomically added into b.
// Contact of this lane when each wave takes `lpw` contacts (lanes lpw..63
// idle; -1 there). A short list is then spread over more SIMDs, and each
// wave meets fewer of the feature dispatch's branch paths, which a wave runs
// one after another.
// `block` is the workgroup's index within its half of a launch.
__device__ __forceinline__ int contact_slot(int lpw, int block) {
  int lane = threadIdx.x & 63;
  int wave = int((block * blockDim.x + threadIdx.x) >> 6);
  return lane < lpw ? wave * lpw + lane : -1;
}

// One contact list's Hessian launch (or half of a paired launch). P is the
// positions' type, R the outputs' (vals, b): float and float, or double and
// double with fp64 scalars.
template <typename P, typename R>
struct HessianArgsT {
  const P* x;
  const P* xcs;
  int nv;
  const int4* contacts;
  int n;
  P l0;
  double kappa;
  int offset;
  int* rows;
  int* cols;
  R* vals;
  R* b;
  int lpw;
};
using HessianArgs = HessianArgsT<float, float>;

// Per wave in LDS: the forces on their way to b (12 per contact slot), the
// slots' 4 vertex ids and active flags, and 144 bytes of every slot's
// triplet values on their way to memory (one block row in fp32, half of one
// in fp64).
template <typename R>
struct HessianStageT {
  R f[4][64 * 12];
  int v[4][64 * 4];
  int act[4][64];
  float4 row[4][64 * 9];
};

// The wave's triplets are one contiguous range of memory (16 blocks per
// contact, its contacts consecutive). A lane computes its own contact, but
// the stores are issued for consecutive addresses across the wave: a lane
// writing its own contact's 576 contiguous bytes made every store
// instruction touch 64 cache lines with 16 bytes each, and those stores
// took about half of the kernel's time. Row and column indices are stored
// from the staged vertex ids; the values go block row by block row through
// LDS (each lane expands its contact's row q, then the wave stores the 64
// contacts' rows q, 144 contiguous bytes per contact; in fp64 a row goes in
// two halves of 144 bytes).
template <bool EE, typename P, typename R>
__device__ __forceinline__ void hessian_block(const HessianArgsT<P, R>& A, int block, HessianStageT<R>& st) {
  const P* __restrict__ x = A.x;
  const P* __restrict__ xcs = A.xcs;
  const int nv = A.nv, n = A.n, offset = A.offset, lpw = A.lpw;
  const P l0 = A.l0;
  const double kappa = A.kappa;
  R* __restrict__ b = A.b;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i = contact_slot(lpw, block);
  R force[12];
#pragma unroll
  for (int k = 0; k < 12; ++k) force[k] = R(0);
  int4 c = make_int4(0, 0, 0, 0);
  CompressedHessianT<R> C;
  bool active = false;
  if (i >= 0 && i < n) {
    c = A.contacts[i];
    active = valid4(c, nv) && (EE ? pair_ee<P, R>(x, xcs, c, l0, kappa, true, C, force, nullptr)
                                  : pair_pt<P, R>(x, c, l0, kappa, true, C, force, nullptr));
  }
  if (!active) {
#pragma unroll
    for (int k = 0; k < 12; ++k) force[k] = R(0);
  }
  R* sf = st.f[wave];
  int* sv = st.v[wave];
  int* sa = st.act[wave];
  float4* sr = st.row[wave];
  if (lane < lpw) {
#pragma unroll
    for (int k = 0; k < 12; ++k) sf[lane * 12 + k] = force[k];
    sv[lane * 4 + 0] = c.x;
    sv[lane * 4 + 1] = c.y;
    sv[lane * 4 + 2] = c.z;
    sv[lane * 4 + 3] = c.w;
    sa[lane] = active ? 1 : 0;
  }
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  // the wave's contacts [first, first + nw) own triplets [offset + first, ...) * 16
  const int first = int((block * blockDim.x + threadIdx.x) >> 6) * lpw;
  const int nw = first < n ? min(lpw, n - first) : 0;
  const long long tbase = (long long)(offset + first) * 16;
  // rows and columns: 4 int4 per contact each (block ii * 4 + jj: row c[ii], column c[jj])
  {
    int4* rows4 = reinterpret_cast<int4*>(A.rows + tbase);
    int4* cols4 = reinterpret_cast<int4*>(A.cols + tbase);
    for (int t = lane; t < nw * 4; t += 64) {
      const int s = t >> 2, ii = t & 3;
      int4 rv = make_int4(0, 0, 0, 0), cv = make_int4(0, 0, 0, 0);
      if (sa[s]) {
        const int vi = sv[s * 4 + ii];
        rv = make_int4(vi, vi, vi, vi);
        cv = make_int4(sv[s * 4 + 0], sv[s * 4 + 1], sv[s * 4 + 2], sv[s * 4 + 3]);
      }
      rows4[t] = rv;
      cols4[t] = cv;
    }
  }
  // values, 144 bytes at a time: kPer consecutive 3x3 blocks of a block row
  // (4 in fp32, 2 in fp64) = 9 16-byte units; a contact holds 36 kHalves
  // units, kHalves per block row.
  constexpr int kHalves = int(sizeof(R)) / 4;
  constexpr int kPer = 4 / kHalves;
  float4* v4 = reinterpret_cast<float4*>(A.vals + tbase * 9);
#pragma unroll
  for (int q = 0; q < 4 * kHalves; ++q) {
    if (lane < lpw) {
      R blk[9 * kPer];
      if (active) {
#pragma unroll
        for (int r = 0; r < kPer; ++r) expand_block(C, q / kHalves, (q % kHalves) * kPer + r, blk + 9 * r);
      } else {
#pragma unroll
        for (int k = 0; k < 9 * kPer; ++k) blk[k] = R(0);
      }
#pragma unroll
      for (int k = 0; k < 9; ++k) __builtin_memcpy(&sr[lane * 9 + k], blk + k * (16 / int(sizeof(R))), 16);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    // slot s's part q: unit s * 36 kHalves + q * 9 + k of the wave's range
    for (int t = lane; t < nw * 9; t += 64) {
      const int s = t / 9, k = t - 9 * s;
      v4[s * 36 * kHalves + q * 9 + k] = sr[t];
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  }
  // Forces into b. Device-scope float atomics cost a memory transaction per
  // cache line an atomic instruction touches. The wave's 12 lpw force
  // components (contact slot j / 12, component j % 12) are added transposed
  // from LDS, value j by lane j % 64: an instruction then adds consecutive
  // components of consecutive contacts, a few lines when their vertices are
  // adjacent instead of one line per lane. A zero force (a vertex with zero
  // weight in the contact's feature, or no active contact) is not added.
  for (int j = lane; j < 12 * lpw; j += 64) {
    const R v = sf[j];
    if (v != R(0)) {
      const int slot = j / 12, comp = j - 12 * slot, vert = comp / 3;
      unsafeAtomicAdd(b + 3 * (long long)sv[slot * 4 + vert] + (comp - 3 * vert), v);
    }
  }
}

template <bool EE, typename P = float, typename R = float>
__global__ void __launch_bounds__(256) contact_hessian_kernel(HessianArgsT<P, R> A) {
  __shared__ HessianStageT<R> st;
  hessian_block<EE>(A, blockIdx.x, st);
}

// Both halves of the contact Hessian in one launch: workgroups [0, blocks_e)
// take the EE list, the rest the PT list.
template <typename P = float, typename R = float>
__global__ void __launch_bounds__(256) contact_hessian_pair_kernel(HessianArgsT<P, R> E, HessianArgsT<P, R> T,
                                                                   int blocks_e) {
  __shared__ HessianStageT<R> st;
  if (int(blockIdx.x) < blocks_e) hessian_block<true>(E, blockIdx.x, st);
  else hessian_block<false>(T, blockIdx.x - blocks_e, st);
}

// ------------------------------------------------------------ energy kernel ----
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// contact_energy_ee / contact_energy_pt (dynamic_contacts.py:374-398): the
// per-contact energies are summed in FP64 per wave, per block through LDS,
// then one atomicAdd(double) per block into energy[0].
template <typename P>
struct EnergyArgsT {
  const P* x;
  const P* xcs;
  int nv;
  const int4* contacts;
  int n;
  P l0;
  double kappa;
  double* energy;
  int lpw;
};
using EnergyArgs = EnergyArgsT<float>;

// Energy of the contact this lane takes in (virtual) workgroup `block`.
template <bool EE, typename P>
__device__ __forceinline__ double energy_of(const EnergyArgsT<P>& A, int block) {
  const int i = contact_slot(A.lpw, block);
  double e = 0.0;
  if (i >= 0 && i < A.n) {
    CompressedHessianT<P> C;
    P force[12];
    double ei = 0.0;
    const int4 c = A.contacts[i];
    const bool active = valid4(c, A.nv) && (EE ? pair_ee<P, P>(A.x, A.xcs, c, A.l0, A.kappa, false, C, force, &ei)
                                               : pair_pt<P, P>(A.x, c, A.l0, A.kappa, false, C, force, &ei));
    if (active) e = ei;
  }
  return e;
}

// The workgroup's sum, added into *energy with one atomic: atomics on one
// address serialise, so a launch makes at most its grid's worth of them.
__device__ __forceinline__ void energy_flush(double e, double* partial, double* energy) {
  e = wave_sum(e);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) partial[wave] = e;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < int(blockDim.x >> 6); ++w) s += partial[w];
    if (s != 0.0) unsafeAtomicAdd(energy, s);
  }
}

// A launch covers `blocks` virtual workgroups (blocks_spread's count) with
// at most kEnergyGrid real ones, each walking every gridDim-th.
constexpr int kEnergyGrid = 512;
static inline int energy_grid(int blocks) { return blocks < kEnergyGrid ? blocks : kEnergyGrid; }

template <bool EE, typename P = float>
__global__ void __launch_bounds__(256) contact_energy_kernel(EnergyArgsT<P> A, int blocks) {
  __shared__ double partial[4];
  double e = 0.0;
  for (int vb = blockIdx.x; vb < blocks; vb += gridDim.x) e += energy_of<EE>(A, vb);
  energy_flush(e, partial, A.energy);
}

// Both halves of the barrier energy in one launch: workgroups [0, grid_e)
// walk the blocks_e virtual workgroups of the EE list, the rest the
// blocks_t of the PT list; both sums go into E.energy.
template <typename P = float>
__global__ void __launch_bounds__(256) contact_energy_pair_kernel(EnergyArgsT<P> E, EnergyArgsT<P> T, int blocks_e,
                                                                  int grid_e, int blocks_t) {
  __shared__ double partial[4];
  double e = 0.0;
  if (int(blockIdx.x) < grid_e) {
    for (int vb = blockIdx.x; vb < blocks_e; vb += grid_e) e += energy_of<true>(E, vb);
  } else {
    const int grid_t = int(gridDim.x) - grid_e;
    for (int vb = int(blockIdx.x) - grid_e; vb < blocks_t; vb += grid_t) e += energy_of<false>(T, vb);
  }
  energy_flush(e, partial, E.energy);
}

// ----------------------------------------------------- contact points kernel ----
// get_contact_points / get_contact_points_pt (contact.py:148-161) over
// fetch_dist_v0v1 / fetch_dist_v0v1_pt (contact.py:88-145): midpoint of the
// two closest points, the fp32 distance, and (additionally) the feature the
// Hessian kernels dispatch on.
template <bool EE>
__global__ void __launch_bounds__(256) contact_points_kernel(const float* __restrict__ x, const float* __restrict__ xcs,
                                                             int nv, const int4* __restrict__ contacts, int n,
                                                             float* __restrict__ points, float* __restrict__ dists,
                                                             int* __restrict__ features) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int4 c = contacts[i];
  if (!valid4(c, nv)) {
    if (points) { points[3 * i] = points[3 * i + 1] = points[3 * i + 2] = __builtin_nanf(""); }
    if (dists) dists[i] = __builtin_nanf("");
    if (features) features[i] = -1;
    return;
  }
  V3<float> x0 = ldx(x, c.x), x1 = ldx(x, c.y), x2 = ldx(x, c.z), x3 = ldx(x, c.w);
  V3<float> v0, v1;
  float dist;
  int feat;
  if (EE) {
    EdgeClosest ec = closest_point_edge_edge(x0, x1, x2, x3, 1.0e-6f);
    // wp.lerp(a, b, t) = a * (1 - t) + b * t
    v0 = x0 * (1.0f - ec.s) + x1 * ec.s;
    v1 = x2 * (1.0f - ec.t) + x3 * ec.t;
    dist = ec.dist;
    float eps_x = 0.0f;
    if (xcs) {
      V3<float> ra = ldx(xcs, c.y) - ldx(xcs, c.x), rb = ldx(xcs, c.w) - ldx(xcs, c.z);
      eps_x = 1.0e-3f * dot(ra, ra) * dot(rb, rb);
    }
    Feature f = ee_feature(ec, cross(x1 - x0, x3 - x2), eps_x);
    feat = f.kind * 16 + f.s0 * 4 + (f.s1 < 0 ? 0 : f.s1);
  } else {
    TriClosest tc = closest_point_triangle(x0, x1, x2, x3);
    v0 = x0;
    v1 = tc.wa * x1 + tc.wb * x2 + (1.0f - tc.wa - tc.wb) * x3;
    dist = tc.dist;
    feat = tc.type;
  }
  V3<float> mid = (v0 + v1) * 0.5f;
  if (points) { points[3 * i] = mid.x; points[3 * i + 1] = mid.y; points[3 * i + 2] = mid.z; }
  if (dists) dists[i] = dist;
  if (features) features[i] = feat;
}

// -------------------------------------------------------------- CCD kernel ----
// Folds t (>= 0) into the running minimum *toi_min (IEEE bits, which order
// like the values): the atomic only when t is below what is there already,
// so the many waves that finish after the minimum has dropped make no
// atomic on the one address (those serialise).
__device__ __forceinline__ void toi_min_add(unsigned long long* toi_min, double t) {
  const unsigned long long bits = (unsigned long long)__double_as_longlong(t);
  if (bits < __hip_atomic_load(toi_min, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) atomicMin(toi_min, bits);
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
  return v;
}

// pt_collision_time / ee_collision_time (ccd/ccd.py:304-357) for a list of
// candidate quadruples, in FP64. The number of candidates is n, or
// min(*n_dev, n) when n_dev is given (a count produced on the device; n is
// then the list's capacity), walked by each wave 64 candidates at a time.
// Writes per-candidate TOI (optional) and folds TOI < 1 into *toi_min with
// one 64-bit atomicMin per wave (TOI >= 0, so the IEEE bit pattern orders
// like an unsigned integer).
//
// Two phases with compaction: most candidates' coplanarity cubic has no
// bracketed root in [0, 1] and their TOI is 1 after the coefficients and
// the case analysis. The others need the iterative root search, about 30
// rounds of FP64 Newton / bisection, which a wave ran for its slowest lane
// with the other lanes idle. Here a workgroup first classifies a range of
// candidates, 256 at a time, and queues the indices of those with a
// bracketed root in LDS (ballot and prefix count per wave, one LDS atomic per
// wave); when the queue holds kToiDense or more, or at the end, its four
// waves run the full collision_time on the queued candidates, 64 at a time,
// one per lane. The second phase recomputes a candidate from its positions
// with the single-pass code, so the values are the single pass's.
constexpr int kToiBlock = 256;
constexpr int kToiQueue = 1024;
constexpr int kToiDense = kToiQueue - kToiBlock;  // run the queue before it could overflow
struct ToiQueue {
  int idx[kToiQueue];
  int n;
};

template <bool EE, typename P = float>
__device__ __forceinline__ double toi_of(const P* __restrict__ x0, const P* __restrict__ x1, int nv, int4 q,
                                         bool full) {
  if (!valid4(q, nv)) return 1.0;
  const int id[4] = {q.x, q.y, q.z, q.w};
  V3<double> p0[4], p1[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) { p0[k] = to_d(ldx(x0, id[k])); p1[k] = to_d(ldx(x1, id[k])); }
  if (!full) return collision_time_may_hit(p0, p1) ? -1.0 : 1.0;
  return collision_time<EE, double>(p0, p1);
}

// Root isolation of the queued candidates by the whole workgroup; empties
// the queue. Called by every thread of the workgroup.
template <bool EE, typename P = float>
__device__ __forceinline__ void toi_drain(const P* __restrict__ x0, const P* __restrict__ x1, int nv,
                                          const int4* __restrict__ quads, double* __restrict__ toi_out,
                                          ToiQueue& tq, double& tmin) {
  __syncthreads();
  const int qn = tq.n;
  for (int k = threadIdx.x; k < qn; k += kToiBlock) {
    const int j = tq.idx[k];
    const double t = toi_of<EE, P>(x0, x1, nv, quads[j], true);
    if (toi_out) toi_out[j] = t;
    tmin = fmin(tmin, t);
  }
  __syncthreads();
  if (threadIdx.x == 0) tq.n = 0;
  __syncthreads();
}

// Workgroup `bid` of `nblk` walking one candidate list (every thread of the
// workgroup takes the same number of trips).
template <bool EE, typename P = float>
__device__ __forceinline__ void collision_time_block(const P* __restrict__ x0, const P* __restrict__ x1, int nv,
                                                     const int4* __restrict__ quads, int n,
                                                     const int* __restrict__ n_dev, double* __restrict__ toi_out,
                                                     unsigned long long* __restrict__ toi_min, int bid, int nblk,
                                                     ToiQueue& tq) {
  int count = n;
  if (n_dev) {
    int d = *n_dev;
    count = d < n ? d : n;
  }
  const int lane = threadIdx.x & 63;
  if (threadIdx.x == 0) tq.n = 0;
  __syncthreads();
  double tmin = 1.0;
  const int stride = nblk * kToiBlock;
  for (int i0 = bid * kToiBlock; i0 < count; i0 += stride) {
    const int i = i0 + threadIdx.x;
    bool need = false;
    if (i < count) {
      need = toi_of<EE, P>(x0, x1, nv, quads[i], false) < 0.0;
      if (!need && toi_out) toi_out[i] = 1.0;
    }
    const unsigned long long m = __ballot(need);
    int base = 0;
    if (lane == 0 && m) base = atomicAdd(&tq.n, __popcll(m));
    base = __shfl(base, 0, 64);
    if (need) tq.idx[base + __popcll(m & ((1ull << lane) - 1ull))] = i;
    __syncthreads();
    if (tq.n >= kToiDense) toi_drain<EE, P>(x0, x1, nv, quads, toi_out, tq, tmin);
  }
  toi_drain<EE, P>(x0, x1, nv, quads, toi_out, tq, tmin);
  if (toi_min) {
    double m = wave_min(tmin);
    if (lane == 0 && m < 1.0) toi_min_add(toi_min, m);
  }
}

template <bool EE, typename P = float>
__global__ void __launch_bounds__(256) collision_time_kernel(const P* __restrict__ x0, const P* __restrict__ x1,
                                                             int nv, const int4* __restrict__ quads, int n,
                                                             const int* __restrict__ n_dev,
                                                             double* __restrict__ toi_out,
                                                             unsigned long long* __restrict__ toi_min) {
  __shared__ ToiQueue tq;
  collision_time_block<EE, P>(x0, x1, nv, quads, n, n_dev, toi_out, toi_min, blockIdx.x, gridDim.x, tq);
}

// The EE and PT candidate lists of one CCD call in one launch: workgroups
// [0, blocks_ee) walk the EE list, the rest the PT list; both fold into the
// same minimum.
__global__ void __launch_bounds__(256) collision_time_pair_kernel(
    const float* __restrict__ x0, const float* __restrict__ x1, int nv, const int4* __restrict__ quads_ee,
    const int* __restrict__ n_ee, int cap_ee, int blocks_ee, const int4* __restrict__ quads_pt,
    const int* __restrict__ n_pt, int cap_pt, unsigned long long* __restrict__ toi_min) {
  __shared__ ToiQueue tq;
  if (int(blockIdx.x) < blocks_ee)
    collision_time_block<true>(x0, x1, nv, quads_ee, cap_ee, n_ee, nullptr, toi_min, blockIdx.x, blocks_ee, tq);
  else
    collision_time_block<false>(x0, x1, nv, quads_pt, cap_pt, n_pt, nullptr, toi_min, blockIdx.x - blocks_ee,
                                gridDim.x - blocks_ee, tq);
}

// edge_edge_intersections / point_triangle_intersections (contact.py:370-442)
// after the candidate query: FP64 TOI of the first min(*n_dev, cap_in)
// candidates; each candidate with t < 1 is appended (append_ccd,
// contact.py:185-197) as its record, its primitive ids and alpha = t, and t
// is folded into *toi_min. The loop runs per wave (every lane of a wave takes
// the same trip count), so the append is one atomic per wave: ballot,
// leader's atomicAdd, prefix popcount. Slots at or beyond cap_out are counted
// but not written, so the caller can grow the list and run again.
template <bool EE>
__global__ void __launch_bounds__(256) record_intersections_kernel(
    const float* __restrict__ x0, const float* __restrict__ x1, int nv, const int4* __restrict__ quads,
    const int2* __restrict__ ids, const int* __restrict__ n_dev, int cap_in, int4* __restrict__ rec_out,
    int2* __restrict__ ids_out, double* __restrict__ alpha_out, int cap_out, int* __restrict__ count_out,
    unsigned long long* __restrict__ toi_min) {
  int count = *n_dev;
  count = count < cap_in ? count : cap_in;
  const int lane = threadIdx.x & 63;
  const int stride = gridDim.x * blockDim.x;
  double tmin = 1.0;
  for (int i0 = blockIdx.x * blockDim.x + (threadIdx.x & ~63); i0 < count; i0 += stride) {
    int i = i0 + lane;
    double t = 1.0;
    int4 q = make_int4(0, 0, 0, 0);
    if (i < count) {
      q = quads[i];
      if (valid4(q, nv)) {
        const int id[4] = {q.x, q.y, q.z, q.w};
        V3<double> p0[4], p1[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) { p0[k] = to_d(ldx(x0, id[k])); p1[k] = to_d(ldx(x1, id[k])); }
        t = collision_time<EE, double>(p0, p1);
      }
    }
    bool hit = t < 1.0;
    tmin = fmin(tmin, t);
    unsigned long long m = __ballot(hit);
    if (m) {
      int leader = __builtin_ctzll(m);
      int base = 0;
      if (lane == leader) base = atomicAdd(count_out, __popcll(m));
      base = __shfl(base, leader, 64);
      if (hit) {
        int slot = base + __popcll(m & ((1ull << lane) - 1ull));
        if (slot < cap_out) {
          rec_out[slot] = q;
          ids_out[slot] = ids[i];
          alpha_out[slot] = t;
        }
      }
    }
  }
  if (toi_min) {
    double mn = wave_min(tmin);
    if (lane == 0 && mn < 1.0) toi_min_add(toi_min, mn);
  }
}

// --------------------------------------------------------------- launchers ----
static inline int blocks_for(long long n, int bs) { return int((n + bs - 1) / bs); }

// Contacts per wave: a list too short to fill the device several times over
// is spread thinly (contact_slot); long lists run one contact per lane. The
// EE and PT halves of one launch share the device, so they are spread by
// their total length (a PT list of 237k beside 697k EE contacts spread over
// 16 lanes a wave issued four times the instructions it needed).
constexpr int kSpreadBelow = 1 << 18;
constexpr int kSpreadLanes = 16;        // Hessian (measured best of 4..64 at ~5e3 contacts)
constexpr int kSpreadLanesEnergy = 32;  // energy (no Hessian, shorter chain)
static inline int lanes_per_wave(int n, int spread = kSpreadLanes) { return n < kSpreadBelow ? spread : 64; }
static inline int blocks_spread(int n, int lpw) { return blocks_for(blocks_for(n, lpw), 4); }  // 4 waves per block

static HessianArgs hessian_args(const float* x, const float* xcs, int nv, const int* contacts, int n, float l0,
                                double kappa, int offset, int* rows, int* cols, float* vals, float* b,
                                int n_launch) {
  return HessianArgs{x, xcs, nv, reinterpret_cast<const int4*>(contacts), n, l0, kappa, offset, rows, cols, vals, b,
                     lanes_per_wave(n_launch)};
}

static EnergyArgs energy_args(const float* x, const float* xcs, int nv, const int* contacts, int n, float l0,
                              double kappa, double* energy, int n_launch) {
  return EnergyArgs{x, xcs, nv, reinterpret_cast<const int4*>(contacts), n, l0, kappa, energy,
                    lanes_per_wave(n_launch, kSpreadLanesEnergy)};
}

hipError_t launch_contact_hessian(bool ee, const float* x, const float* xcs, int nv, const int* contacts, int n, float l0,
                                  double kappa, int offset, int* rows, int* cols, float* vals, float* b,
                                  hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const HessianArgs A = hessian_args(x, xcs, nv, contacts, n, l0, kappa, offset, rows, cols, vals, b, n);
  if (ee)
    hipLaunchKernelGGL(contact_hessian_kernel<true>, dim3(blocks_spread(n, A.lpw)), dim3(256), 0, s, A);
  else
    hipLaunchKernelGGL(contact_hessian_kernel<false>, dim3(blocks_spread(n, A.lpw)), dim3(256), 0, s, A);
  return hipGetLastError();
}

hipError_t launch_collision_hessian(const float* x, const float* xcs, int nv, const int* ee, int n_ee, const int* pt,
                                    int n_pt, float l0, double kappa, int* rows, int* cols, float* vals, float* b,
                                    hipStream_t s) {
  if (n_ee <= 0 || n_pt <= 0) {
    hipError_t e = launch_contact_hessian(true, x, xcs, nv, ee, n_ee, l0, kappa, 0, rows, cols, vals, b, s);
    if (e != hipSuccess) return e;
    return launch_contact_hessian(false, x, nullptr, nv, pt, n_pt, l0, kappa, n_ee, rows, cols, vals, b, s);
  }
  const HessianArgs E = hessian_args(x, xcs, nv, ee, n_ee, l0, kappa, 0, rows, cols, vals, b, n_ee + n_pt);
  const HessianArgs P = hessian_args(x, nullptr, nv, pt, n_pt, l0, kappa, n_ee, rows, cols, vals, b, n_ee + n_pt);
  const int be = blocks_spread(n_ee, E.lpw), bp = blocks_spread(n_pt, P.lpw);
  hipLaunchKernelGGL(contact_hessian_pair_kernel<>, dim3(be + bp), dim3(256), 0, s, E, P, be);
  return hipGetLastError();
}

// fp64 scalars: positions, blocks and forces in double.
hipError_t launch_contact_hessian_f64(bool ee, const double* x, const double* xcs, int nv, const int* contacts, int n,
                                      double l0, double kappa, int offset, int* rows, int* cols, double* vals,
                                      double* b, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const HessianArgsT<double, double> A{x, xcs, nv, reinterpret_cast<const int4*>(contacts), n, l0, kappa, offset,
                                       rows, cols, vals, b, lanes_per_wave(n)};
  if (ee)
    hipLaunchKernelGGL((contact_hessian_kernel<true, double, double>), dim3(blocks_spread(n, A.lpw)), dim3(256), 0, s,
                       A);
  else
    hipLaunchKernelGGL((contact_hessian_kernel<false, double, double>), dim3(blocks_spread(n, A.lpw)), dim3(256), 0,
                       s, A);
  return hipGetLastError();
}

hipError_t launch_contact_energy_f64(bool ee, const double* x, const double* xcs, int nv, const int* contacts, int n,
                                     double l0, double kappa, double* energy, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const EnergyArgsT<double> A{x, xcs, nv, reinterpret_cast<const int4*>(contacts), n, l0, kappa, energy,
                              lanes_per_wave(n, kSpreadLanesEnergy)};
  const int blocks = blocks_spread(n, A.lpw);
  if (ee)
    hipLaunchKernelGGL((contact_energy_kernel<true, double>), dim3(energy_grid(blocks)), dim3(256), 0, s, A, blocks);
  else
    hipLaunchKernelGGL((contact_energy_kernel<false, double>), dim3(energy_grid(blocks)), dim3(256), 0, s, A, blocks);
  return hipGetLastError();
}

hipError_t launch_contact_energy(bool ee, const float* x, const float* xcs, int nv, const int* contacts, int n, float l0,
                                 double kappa, double* energy, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const EnergyArgs A = energy_args(x, xcs, nv, contacts, n, l0, kappa, energy, n);
  const int blocks = blocks_spread(n, A.lpw);
  if (ee)
    hipLaunchKernelGGL(contact_energy_kernel<true>, dim3(energy_grid(blocks)), dim3(256), 0, s, A, blocks);
  else
    hipLaunchKernelGGL(contact_energy_kernel<false>, dim3(energy_grid(blocks)), dim3(256), 0, s, A, blocks);
  return hipGetLastError();
}

__global__ void zero_f64_kernel(double* p) { *p = 0.0; }

hipError_t launch_collision_energy(const float* x, const float* xcs, int nv, const int* ee, int n_ee, const int* pt,
                                   int n_pt, float l0, double kappa, double* energy, hipStream_t s) {
  hipLaunchKernelGGL(zero_f64_kernel, dim3(1), dim3(1), 0, s, energy);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || (n_ee <= 0 && n_pt <= 0)) return e;
  if (n_ee <= 0 || n_pt <= 0) {
    e = launch_contact_energy(true, x, xcs, nv, ee, n_ee, l0, kappa, energy, s);
    if (e != hipSuccess) return e;
    return launch_contact_energy(false, x, nullptr, nv, pt, n_pt, l0, kappa, energy, s);
  }
  const EnergyArgs E = energy_args(x, xcs, nv, ee, n_ee, l0, kappa, energy, n_ee + n_pt);
  const EnergyArgs P = energy_args(x, nullptr, nv, pt, n_pt, l0, kappa, energy, n_ee + n_pt);
  const int be = blocks_spread(n_ee, E.lpw), bp = blocks_spread(n_pt, P.lpw);
  const int ge = energy_grid(be), gp = energy_grid(bp);
  hipLaunchKernelGGL(contact_energy_pair_kernel<>, dim3(ge + gp), dim3(256), 0, s, E, P, be, ge, bp);
  return hipGetLastError();
}

hipError_t launch_contact_points(bool ee, const float* x, const float* xcs, int nv, const int* contacts, int n, float* points,
                                 float* dists, int* features, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int4* c4 = reinterpret_cast<const int4*>(contacts);
  if (ee)
    hipLaunchKernelGGL(contact_points_kernel<true>, dim3(blocks_for(n, 256)), dim3(256), 0, s, x, xcs, nv, c4, n,
                       points, dists, features);
  else
    hipLaunchKernelGGL(contact_points_kernel<false>, dim3(blocks_for(n, 256)), dim3(256), 0, s, x, xcs, nv, c4, n,
                       points, dists, features);
  return hipGetLastError();
}

// Workgroups of a TOI launch over (a capacity of) n candidates: one 256-
// candidate range each for short lists; long lists give each workgroup
// about four ranges (so its queue gathers enough root searches to fill its
// waves) while still keeping at least 1536 workgroups (6 per CU) in flight.
static inline int toi_blocks(int n) {
  const int one = blocks_for(n, kToiBlock), four = blocks_for(n, 4 * kToiBlock);
  const int g = one < 1536 ? one : (four > 1536 ? four : 1536);
  return g > 8192 ? 8192 : g;
}

static hipError_t collision_time_grid(bool ee, const float* x0, const float* x1, int nv, const int* quads, int n,
                                      const int* n_dev, double* toi_out, unsigned long long* toi_min, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int4* q4 = reinterpret_cast<const int4*>(quads);
  const int grid = toi_blocks(n);
  if (ee)
    hipLaunchKernelGGL(collision_time_kernel<true>, dim3(grid), dim3(256), 0, s, x0, x1, nv, q4, n, n_dev, toi_out,
                       toi_min);
  else
    hipLaunchKernelGGL(collision_time_kernel<false>, dim3(grid), dim3(256), 0, s, x0, x1, nv, q4, n, n_dev, toi_out,
                       toi_min);
  return hipGetLastError();
}

// fp64 scalars: vec3d positions, the same FP64 arithmetic.
hipError_t launch_collision_time_f64(bool ee, const double* x0, const double* x1, int nv, const int* quads, int n,
                                     double* toi_out, unsigned long long* toi_min, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  const int4* q4 = reinterpret_cast<const int4*>(quads);
  const int grid = toi_blocks(n);
  if (ee)
    hipLaunchKernelGGL((collision_time_kernel<true, double>), dim3(grid), dim3(256), 0, s, x0, x1, nv, q4, n, nullptr,
                       toi_out, toi_min);
  else
    hipLaunchKernelGGL((collision_time_kernel<false, double>), dim3(grid), dim3(256), 0, s, x0, x1, nv, q4, n, nullptr,
                       toi_out, toi_min);
  return hipGetLastError();
}

hipError_t launch_collision_time(bool ee, const float* x0, const float* x1, int nv, const int* quads, int n,
                                 double* toi_out, unsigned long long* toi_min, hipStream_t s) {
  return collision_time_grid(ee, x0, x1, nv, quads, n, nullptr, toi_out, toi_min, s);
}

hipError_t launch_collision_time_counted(bool ee, const float* x0, const float* x1, int nv, const int* quads,
                                         const int* n_dev, int capacity, unsigned long long* toi_min, hipStream_t s) {
  return collision_time_grid(ee, x0, x1, nv, quads, capacity, n_dev, nullptr, toi_min, s);
}

hipError_t launch_collision_time_pair(const float* x0, const float* x1, int nv, const int* quads_ee, const int* n_ee,
                                      int cap_ee, const int* quads_pt, const int* n_pt, int cap_pt,
                                      unsigned long long* toi_min, hipStream_t s) {
  if (cap_ee <= 0) return launch_collision_time_counted(false, x0, x1, nv, quads_pt, n_pt, cap_pt, toi_min, s);
  if (cap_pt <= 0) return launch_collision_time_counted(true, x0, x1, nv, quads_ee, n_ee, cap_ee, toi_min, s);
  const int blocks_ee = toi_blocks(cap_ee), blocks_pt = toi_blocks(cap_pt);
  hipLaunchKernelGGL(collision_time_pair_kernel, dim3(blocks_ee + blocks_pt), dim3(256), 0, s, x0, x1, nv,
                     reinterpret_cast<const int4*>(quads_ee), n_ee, cap_ee, blocks_ee,
                     reinterpret_cast<const int4*>(quads_pt), n_pt, cap_pt, toi_min);
  return hipGetLastError();
}

hipError_t launch_record_intersections(bool ee, const float* x0, const float* x1, int nv, const int* quads,
                                      const int* ids, const int* n_dev, int cap_in, int* rec_out, int* ids_out,
                                      double* alpha_out, int cap_out, int* count_out, unsigned long long* toi_min,
                                      hipStream_t s) {
  if (cap_in <= 0) return hipSuccess;
  int grid = blocks_for(cap_in, 256);
  if (grid > 8192) grid = 8192;
  const int4* q4 = reinterpret_cast<const int4*>(quads);
  const int2* i2 = reinterpret_cast<const int2*>(ids);
  int4* r4 = reinterpret_cast<int4*>(rec_out);
  int2* o2 = reinterpret_cast<int2*>(ids_out);
  if (ee)
    hipLaunchKernelGGL(record_intersections_kernel<true>, dim3(grid), dim3(256), 0, s, x0, x1, nv, q4, i2, n_dev,
                       cap_in, r4, o2, alpha_out, cap_out, count_out, toi_min);
  else
    hipLaunchKernelGGL(record_intersections_kernel<false>, dim3(grid), dim3(256), 0, s, x0, x1, nv, q4, i2, n_dev,
                       cap_in, r4, o2, alpha_out, cap_out, count_out, toi_min);
  return hipGetLastError();
}

}  // namespace ipcd

// Device math shared by the contact kernels of libipc_warp_amd (gfx950).
//
// Precision contract (the same for every kernel in this library):
//  * the distance filter and the feature classification run in FP32, in the
//    operation order of the reference's Warp functions, so pair sets and
//    active sets agree bit for bit with the reference's FP32 mode;
//  * the squared distance of the classified feature, its gradient, the
//    negative spectral part of its Hessian and the barrier run in FP64
//    (positions are fp32, so every difference of two of them is exact in
//    fp64);
//  * the 16 output blocks are expanded and stored in FP32.
// The library is compiled with -ffp-contract=off, so no mul/add pair is fused
// behind the source's back.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>

// Every helper is host+device so the same source can be exercised on the CPU.
#define IPCD_FN __host__ __device__ __forceinline__

namespace ipcd {

using std::fabs;
using std::log;
using std::sqrt;

template <typename T> struct V3 { T x, y, z; };
template <typename T> IPCD_FN V3<T> mk(T x, T y, T z) { return V3<T>{x, y, z}; }
template <typename T> IPCD_FN V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename T> IPCD_FN V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T> IPCD_FN V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename T> IPCD_FN V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename T> IPCD_FN T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T> IPCD_FN V3<T> cross(V3<T> a, V3<T> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename T> IPCD_FN T length(V3<T> a) { return sqrt(dot(a, a)); }
template <typename T> IPCD_FN T smin(T a, T b) { return a < b ? a : b; }
template <typename T> IPCD_FN T smax(T a, T b) { return a > b ? a : b; }
template <typename T> IPCD_FN V3<T> vmin(V3<T> a, V3<T> b) { return {smin(a.x, b.x), smin(a.y, b.y), smin(a.z, b.z)}; }
template <typename T> IPCD_FN V3<T> vmax(V3<T> a, V3<T> b) { return {smax(a.x, b.x), smax(a.y, b.y), smax(a.z, b.z)}; }
IPCD_FN V3<double> to_d(V3<float> a) { return {double(a.x), double(a.y), double(a.z)}; }
IPCD_FN V3<double> to_d(V3<double> a) { return a; }

// Vertex positions are the reference's vec3 array: float32 [V, 3], AoS.
IPCD_FN V3<float> ldx(const float* __restrict__ x, int i) {
  const float* p = x + 3 * (long long)i;
  return {p[0], p[1], p[2]};
}
// With fp64 scalars (scalar_types.py USE_FLOAT32 = False) they are vec3d.
IPCD_FN V3<double> ldx(const double* __restrict__ x, int i) {
  const double* p = x + 3 * (long long)i;
  return {p[0], p[1], p[2]};
}

// A constant of the reference's `scalar` type: the float literal in fp32,
// the double literal in fp64 (not the float value widened).
template <typename T> IPCD_FN T lit(float f, double d);
template <> IPCD_FN float lit<float>(float f, double) { return f; }
template <> IPCD_FN double lit<double>(float, double d) { return d; }

// ------------------------------------------------- closest point, triangle ----
// Voronoi-region closest point (contact.py:230-302). Feature: 0/1/2 vertex
// a/b/c, 3 edge ab, 4 edge ac, 5 edge bc, 6 face. Returns the distance.
template <typename T> struct TriClosestT { T wa, wb, dist; int type; };
using TriClosest = TriClosestT<float>;
template <typename T>
IPCD_FN TriClosestT<T> closest_point_triangle(V3<T> p, V3<T> a, V3<T> b, V3<T> c) {
  const T zero = T(0), one = T(1);
  V3<T> ab = b - a, ac = c - a, ap = p - a;
  T d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= zero && d2 <= zero) return {one, zero, length(p - a), 0};
  V3<T> bp = p - b;
  T d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= zero && d4 <= d3) return {zero, one, length(p - b), 1};
  T vc = d1 * d4 - d3 * d2;
  if (vc <= zero && d1 >= zero && d3 <= zero) {
    T v = d1 / (d1 - d3);
    V3<T> point = a + v * ab;
    return {one - v, v, length(p - point), 3};
  }
  V3<T> cp = p - c;
  T d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= zero && d5 <= d6) return {zero, zero, length(p - c), 2};
  T vb = d5 * d2 - d1 * d6;
  if (vb <= zero && d2 >= zero && d6 <= zero) {
    T w = d2 / (d2 - d6);
    V3<T> point = a + w * ac;
    return {one - w, zero, length(p - point), 4};
  }
  T va = d3 * d6 - d5 * d4;
  if (va <= zero && (d4 - d3) >= zero && (d5 - d6) >= zero) {
    T w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    V3<T> point = b + w * (c - b);
    return {zero, one - w, length(p - point), 5};
  }
  T denom = one / (va + vb + vc);
  T v = vb * denom, w = vc * denom;
  T u = one - v - w;
  V3<T> point = a + ab * v + ac * w;
  return {u, v, length(p - point), 6};
}

// ------------------------------------------------- closest points, segments ----
// Segment-segment closest points (Ericson, Real-Time Collision Detection
// 5.1.9), the role of Warp's closest_point_edge_edge(p1, q1, p2, q2, eps).
template <typename T> struct EdgeClosestT { T s, t, dist; };
using EdgeClosest = EdgeClosestT<float>;
template <typename T> IPCD_FN T clamp_unit(T v) { return v < T(0) ? T(0) : (v > T(1) ? T(1) : v); }
template <typename T>
IPCD_FN EdgeClosestT<T> closest_point_edge_edge(V3<T> p1, V3<T> q1, V3<T> p2, V3<T> q2, T eps) {
  V3<T> d1 = q1 - p1, d2 = q2 - p2, r = p1 - p2;
  T a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
  T s, t;
  if (a <= eps && e <= eps) {
    s = T(0); t = T(0);
  } else if (a <= eps) {
    s = T(0); t = clamp_unit(f / e);
  } else {
    T c = dot(d1, r);
    if (e <= eps) {
      t = T(0); s = clamp_unit(-c / a);
    } else {
      T b = dot(d1, d2);
      T denom = a * e - b * b;
      s = denom != T(0) ? clamp_unit((b * f - c * e) / denom) : T(0);
      t = (b * s + f) / e;
      if (t < T(0)) { t = T(0); s = clamp_unit(-c / a); }
      else if (t > T(1)) { t = T(1); s = clamp_unit((b - c) / a); }
    }
  }
  V3<T> c1 = p1 + d1 * s, c2 = p2 + d2 * t;
  return {s, t, length(c1 - c2)};
}

// ------------------------------------------------------------------ barrier ----
// IPC log barrier in squared-distance space, x = d^2, D = dhat^2:
// b(x) = -kappa (x - D)^2 ln(x / D) for 0 < x < D, and its x-derivatives.
struct Barrier { double b, db, d2b; };
IPCD_FN Barrier barrier(double x, double D, double kappa) {
  if (!(x < D)) return {0.0, 0.0, 0.0};
  double lg = log(x / D);
  double dm = x - D;
  double r = D / x;
  Barrier B;
  B.b = -kappa * dm * dm * lg;
  B.db = kappa * (D - x) * (2.0 * lg - r + 1.0);
  B.d2b = kappa * ((r + 2.0) * r - 2.0 * lg - 3.0);
  return B;
}

// ------------------------------------------------------------ features ----
enum { F_PP = 0, F_PE = 1, F_PT = 2, F_EE = 3 };
struct Feature { int kind; int s0, s1, s2, s3; };

// edge_edge_distance_gradient_hessian's dispatch (dynamic_contacts.py:48-108)
// on the closest points, computed in the reference's `scalar` type.
template <typename T>
IPCD_FN Feature ee_feature(EdgeClosestT<T> ec, V3<T> cr, T parallel_eps) {
  const T fe = lit<T>(1.0e-6f, 1.0e-6), one = T(1);
  T s = ec.s, t = ec.t;
  bool sb = s <= fe || s >= one - fe;
  bool tb = t <= fe || t >= one - fe;
  if (sb && tb) return {F_PP, (s >= one - fe) ? 1 : 0, (t >= one - fe) ? 3 : 2, -1, -1};
  if (sb) return {F_PE, (s >= one - fe) ? 1 : 0, 2, 3, -1};
  if (tb) return {F_PE, (t >= one - fe) ? 3 : 2, 0, 1, -1};
  if (parallel_eps > T(0) && dot(cr, cr) < parallel_eps) {
    T ds0 = s, ds1 = one - s, dt0 = t, dt1 = one - t;
    T m = smin(smin(ds0, ds1), smin(dt0, dt1));
    if (m == ds0) return {F_PE, 0, 2, 3, -1};
    if (m == ds1) return {F_PE, 1, 2, 3, -1};
    if (m == dt0) return {F_PE, 2, 0, 1, -1};
    return {F_PE, 3, 0, 1, -1};
  }
  return {F_EE, 0, 1, 2, 3};
}

// point_triangle_distance_gradient_hessian's dispatch (dynamic_contacts.py:110-125).
IPCD_FN Feature pt_feature(int type) {
  switch (type) {
    case 0: return {F_PP, 0, 1, -1, -1};
    case 1: return {F_PP, 0, 2, -1, -1};
    case 2: return {F_PP, 0, 3, -1, -1};
    case 3: return {F_PE, 0, 1, 2, -1};
    case 4: return {F_PE, 0, 1, 3, -1};
    case 5: return {F_PE, 0, 2, 3, -1};
    default: return {F_PT, 0, 1, 2, 3};
  }
}

// Squared distance of a feature in the minimisation form
//     d^2(x) = min_u | sum_k w_k(u) x_k |^2,
// w affine in u (k = 0 parameters for PP, 1 for PE, 2 for PT and EE). At the
// minimiser r = sum w_k x_k is the closest-point offset, grad d^2 = 2 w (x) r,
// and the Hessian is 2 (w w^T) (x) I3 - 2 P G^-1 P^T with
// P_j = w (x) e_j + D_j (x) r, e_j = sum_k D_jk x_k, G = E^T E.
struct FeatureGeom {
  double w[4];      // weights over the 4 slots
  double D[2][4];   // d w / d u_j
  V3<double> e[2];  // e_j
  V3<double> r;     // closest-point offset
  int k;
};

// Slot selection without a runtime-indexed local array (the compiler would
// place it in scratch memory): two levels of selects (v_cndmask on each
// 32-bit half), exact, and far cheaper than a 0/1-weighted blend in FP64.
IPCD_FN V3<double> pick(const V3<double> x[4], int s) {
  const V3<double> lo = s == 1 ? x[1] : x[0];
  const V3<double> hi = s == 3 ? x[3] : x[2];
  return s >= 2 ? hi : lo;
}

// Positions are given relative to slot 0 (x[0] = 0) in double.
IPCD_FN FeatureGeom feature_geom(const Feature& f, const V3<double> x[4]) {
#pragma clang fp contract(fast)
  FeatureGeom g;
  g.e[0] = g.e[1] = mk(0.0, 0.0, 0.0);
  if (f.kind == F_PP) {
    g.k = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      g.w[i] = i == f.s0 ? 1.0 : (i == f.s1 ? -1.0 : 0.0);
      g.D[0][i] = 0.0;
      g.D[1][i] = 0.0;
    }
    g.r = pick(x, f.s0) - pick(x, f.s1);
  } else if (f.kind == F_PE) {
    g.k = 1;
    V3<double> xp = pick(x, f.s0), x0 = pick(x, f.s1), x1 = pick(x, f.s2);
    V3<double> e = x1 - x0;
    V3<double> q = xp - x0;
    double lam = dot(q, e) / dot(e, e);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      g.w[i] = i == f.s0 ? 1.0 : (i == f.s1 ? -(1.0 - lam) : (i == f.s2 ? -lam : 0.0));
      g.D[0][i] = i == f.s1 ? 1.0 : (i == f.s2 ? -1.0 : 0.0);
      g.D[1][i] = 0.0;
    }
    g.e[0] = mk(-e.x, -e.y, -e.z);
    g.r = q - e * lam;
  } else if (f.kind == F_PT) {
    g.k = 2;
    V3<double> e1 = x[2] - x[1], e2 = x[3] - x[1], q = x[0] - x[1];
    double a11 = dot(e1, e1), a12 = dot(e1, e2), a22 = dot(e2, e2);
    double b1 = dot(e1, q), b2 = dot(e2, q);
    double idet = 1.0 / (a11 * a22 - a12 * a12);
    double be = (b1 * a22 - b2 * a12) * idet;
    double ga = (a11 * b2 - a12 * b1) * idet;
    g.w[0] = 1.0; g.w[1] = -(1.0 - be - ga); g.w[2] = -be; g.w[3] = -ga;
    g.D[0][0] = 0.0; g.D[0][1] = 1.0; g.D[0][2] = -1.0; g.D[0][3] = 0.0;
    g.D[1][0] = 0.0; g.D[1][1] = 1.0; g.D[1][2] = 0.0; g.D[1][3] = -1.0;
    g.e[0] = mk(-e1.x, -e1.y, -e1.z);
    g.e[1] = mk(-e2.x, -e2.y, -e2.z);
    g.r = q - e1 * be - e2 * ga;
  } else {
    g.k = 2;
    V3<double> ea = x[1] - x[0], eb = x[3] - x[2], q = x[0] - x[2];
    double a11 = dot(ea, ea), a12 = -dot(ea, eb), a22 = dot(eb, eb);
    double b1 = -dot(ea, q), b2 = dot(eb, q);
    double idet = 1.0 / (a11 * a22 - a12 * a12);
    double s = (b1 * a22 - b2 * a12) * idet;
    double t = (a11 * b2 - a12 * b1) * idet;
    g.w[0] = 1.0 - s; g.w[1] = s; g.w[2] = -(1.0 - t); g.w[3] = -t;
    g.D[0][0] = -1.0; g.D[0][1] = 1.0; g.D[0][2] = 0.0; g.D[0][3] = 0.0;
    g.D[1][0] = 0.0; g.D[1][1] = 0.0; g.D[1][2] = 1.0; g.D[1][3] = -1.0;
    g.e[0] = ea;
    g.e[1] = mk(-eb.x, -eb.y, -eb.z);
    g.r = q + ea * s - eb * t;
  }
  return g;
}

// Whether segment pq crosses triangle abc, by signed volumes in FP64 of the
// fp32 inputs: p and q strictly on opposite sides of the plane or one of them
// on it, and the line pq inside the closed triangle (the three volumes
// det[q - p, a - p, b - p], det[q - p, b - p, c - p], det[q - p, c - p, a - p]
// do not take both signs). A segment lying in the triangle's plane, or a
// degenerate triangle, does not count: that is contact, not a crossing.
// Each volume's sign is certain or taken as zero: a value within the forward
// error bound of the FP64 evaluation (Shewchuk's orient3d bound,
// (7 + 56 eps) eps times the permanent, eps = 2^-53) counts as zero, so
// exactly coplanar inputs (planar mesh regions) never produce a crossing
// from rounding noise.
IPCD_FN int orient3_sign(V3<double> a, V3<double> b, V3<double> c, V3<double> d) {
  const V3<double> u = b - a, v = c - a, w = d - a;
  const double m0 = u.y * v.z, m1 = u.z * v.y, m2 = u.z * v.x, m3 = u.x * v.z, m4 = u.x * v.y, m5 = u.y * v.x;
  const double det = w.x * (m0 - m1) + w.y * (m2 - m3) + w.z * (m4 - m5);
  const double perm = fabs(w.x) * (fabs(m0) + fabs(m1)) + fabs(w.y) * (fabs(m2) + fabs(m3)) +
                      fabs(w.z) * (fabs(m4) + fabs(m5));
  const double bound = (7.0 + 56.0 * 0x1p-53) * 0x1p-53 * perm;
  return det > bound ? 1 : (det < -bound ? -1 : 0);
}
IPCD_FN bool segment_hits_triangle(V3<double> p, V3<double> q, V3<double> a, V3<double> b, V3<double> c) {
  const int sp = orient3_sign(a, b, c, p), sq = orient3_sign(a, b, c, q);
  if (sp * sq > 0 || (sp == 0 && sq == 0)) return false;
  const int u = orient3_sign(p, q, a, b), v = orient3_sign(p, q, b, c), w = orient3_sign(p, q, c, a);
  if (u == 0 && v == 0 && w == 0) return false;
  const bool neg = u < 0 || v < 0 || w < 0, pos = u > 0 || v > 0 || w > 0;
  return !(neg && pos);
}

// 1 / sqrt(x): the device's reciprocal square root (the host pass, used
// only to check these functions on a CPU, divides).
IPCD_FN double inv_sqrt(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return rsqrt(x);
#else
  return 1.0 / sqrt(x);
#endif
}
IPCD_FN float inv_sqrt(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return rsqrtf(x);
#else
  return 1.0f / sqrtf(x);
#endif
}

// Index of (i, j) in the packed upper triangle of a symmetric n x n matrix.
template <int n> IPCD_FN constexpr int tri_idx(int i, int j) {
  return i <= j ? i * n - i * (i - 1) / 2 + (j - i) : j * n - j * (j - 1) / 2 + (i - j);
}

// Cyclic Jacobi on a symmetric n x n (row-major) in precision T. Returns
// eigenvalues in ev and eigenvectors in the columns of V. Sweeps stop once
// the off-diagonal mass is below tol2 of the diagonal's. The matrix is kept
// as its packed upper triangle and each rotation uses the symmetric update
// (a_pp -= t a_pq, a_qq += t a_pq, a_pq = 0, the other rows rotated); a
// rotation with a_pq = 0 runs with t = 0, the identity, so that the
// unrolled sweep has no branches (and no duplicated live matrices).
template <int n, typename T>
IPCD_FN void jacobi_sym(const T A[n * n], T ev[n], T V[n * n], T tol2) {
#pragma clang fp contract(fast)
  T a[n * (n + 1) / 2];
#pragma unroll
  for (int i = 0; i < n; ++i)
#pragma unroll
    for (int j = i; j < n; ++j) a[tri_idx<n>(i, j)] = A[i * n + j];
#pragma unroll
  for (int i = 0; i < n * n; ++i) V[i] = (i % (n + 1) == 0) ? T(1) : T(0);
  for (int sweep = 0; sweep < 12; ++sweep) {
    T off = T(0), diag = T(0);
#pragma unroll
    for (int i = 0; i < n; ++i) {
      diag += a[tri_idx<n>(i, i)] * a[tri_idx<n>(i, i)];
#pragma unroll
      for (int j = i + 1; j < n; ++j) off += a[tri_idx<n>(i, j)] * a[tri_idx<n>(i, j)];
    }
    if (!(off > tol2 * diag)) break;
#pragma unroll
    for (int p = 0; p < n - 1; ++p) {
#pragma unroll
      for (int q = p + 1; q < n; ++q) {
        const T apq = a[tri_idx<n>(p, q)];
        // tan of the rotation angle, t = sgn(d) e / (|d| + sqrt(d^2 + e^2))
        // with d = a_qq - a_pp, e = 2 a_pq (the textbook theta = d / e form
        // with its two divisions folded into one)
        const T d = a[tri_idx<n>(q, q)] - a[tri_idx<n>(p, p)], e = T(2) * apq;
        T t = (d >= T(0) ? e : -e) / (fabs(d) + sqrt(d * d + e * e));
        t = apq != T(0) ? t : T(0);
        const T c = inv_sqrt(t * t + T(1)), s = t * c;
        a[tri_idx<n>(p, p)] -= t * apq;
        a[tri_idx<n>(q, q)] += t * apq;
        a[tri_idx<n>(p, q)] = T(0);
#pragma unroll
        for (int r = 0; r < n; ++r) {
          if (r == p || r == q) continue;
          const T arp = a[tri_idx<n>(r, p)], arq = a[tri_idx<n>(r, q)];
          a[tri_idx<n>(r, p)] = c * arp - s * arq;
          a[tri_idx<n>(r, q)] = s * arp + c * arq;
        }
#pragma unroll
        for (int k = 0; k < n; ++k) {
          const T vkp = V[k * n + p], vkq = V[k * n + q];
          V[k * n + p] = c * vkp - s * vkq;
          V[k * n + q] = s * vkp + c * vkq;
        }
      }
    }
  }
#pragma unroll
  for (int i = 0; i < n; ++i) ev[i] = a[tri_idx<n>(i, i)];
}

// Compressed form of the 16 blocks of a contact Hessian
//     H_ij = c_n (w^_i w^_j Ntt + w^_i u_j r^T + w^_j r u_i^T) + (cg w_i w_j + c_n s_ij) r r^T
// with r the unit closest-point direction (see feature_geom). Stored in the
// output's precision R (fp32, or fp64 with fp64 scalars).
template <typename R>
struct CompressedHessianT {
  R what[4], w[4];
  R rh[3];
  R Ntt[6];     // xx, xy, xz, yy, yz, zz of the NSD part in the w^ (x) I3 block
  R u[4][3];    // Ntz z_i
  R s[10];      // z_i^T Nzz z_j, packed upper triangle (i <= j)
  R cn, cg;
};
using CompressedHessian = CompressedHessianT<float>;

IPCD_FN int sidx(int i, int j) {
  // packed index of (min, max) in a 4x4 upper triangle
  int a = i < j ? i : j, b = i < j ? j : i;
  return a * 4 - (a * (a - 1)) / 2 + (b - a);
}

// Negative spectral part of the Hessian of d^2 for a feature, reduced to the
// (3 + k)-dimensional subspace span{w^ (x) R^3, Z (x) r^}, Z an orthonormal
// basis of the D_j projected orthogonal to w. Fills Ntt, u, s and the
// geometry fields of `C`; returns |r|^2 (the feature's squared distance).
template <typename R>
IPCD_FN double nsd_compressed(const FeatureGeom& g, bool want_nsd, CompressedHessianT<R>& C) {
#pragma clang fp contract(fast)
  double d2 = dot(g.r, g.r);
  double rn = sqrt(d2);
  V3<double> rh = rn > 0.0 ? g.r * (1.0 / rn) : mk(0.0, 0.0, 0.0);
  double wn2 = g.w[0] * g.w[0] + g.w[1] * g.w[1] + g.w[2] * g.w[2] + g.w[3] * g.w[3];
  double wn = sqrt(wn2), iwn = inv_sqrt(wn2);
  double wh[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) { wh[i] = g.w[i] * iwn; C.what[i] = R(wh[i]); C.w[i] = R(g.w[i]); }
  C.rh[0] = R(rh.x); C.rh[1] = R(rh.y); C.rh[2] = R(rh.z);
#pragma unroll
  for (int i = 0; i < 6; ++i) C.Ntt[i] = R(0);
#pragma unroll
  for (int i = 0; i < 4; ++i) { C.u[i][0] = C.u[i][1] = C.u[i][2] = R(0); }
#pragma unroll
  for (int i = 0; i < 10; ++i) C.s[i] = R(0);
  if (!want_nsd || g.k == 0) return d2;

  // Z: Gram-Schmidt of the D_j orthogonal to w^.
  double Z[4][2];
  double wd[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    wd[j] = wh[0] * g.D[j][0] + wh[1] * g.D[j][1] + wh[2] * g.D[j][2] + wh[3] * g.D[j][3];
#pragma unroll
    for (int i = 0; i < 4; ++i) Z[i][j] = g.D[j][i] - wd[j] * wh[i];
  }
  {
    double in0 = inv_sqrt(Z[0][0] * Z[0][0] + Z[1][0] * Z[1][0] + Z[2][0] * Z[2][0] + Z[3][0] * Z[3][0]);
#pragma unroll
    for (int i = 0; i < 4; ++i) Z[i][0] *= in0;
    if (g.k == 2) {
      double p = Z[0][0] * Z[0][1] + Z[1][0] * Z[1][1] + Z[2][0] * Z[2][1] + Z[3][0] * Z[3][1];
#pragma unroll
      for (int i = 0; i < 4; ++i) Z[i][1] -= p * Z[i][0];
      double in1 = inv_sqrt(Z[0][1] * Z[0][1] + Z[1][1] * Z[1][1] + Z[2][1] * Z[2][1] + Z[3][1] * Z[3][1]);
#pragma unroll
      for (int i = 0; i < 4; ++i) Z[i][1] *= in1;
    } else {
#pragma unroll
      for (int i = 0; i < 4; ++i) Z[i][1] = 0.0;
    }
  }
  // M = Q^T P: rows 0..2 = |w| e_j + (w^.D_j) r, rows 3..4 = (z_l . D_j) |r|.
  double M[5][2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    M[0][j] = wn * g.e[j].x + wd[j] * g.r.x;
    M[1][j] = wn * g.e[j].y + wd[j] * g.r.y;
    M[2][j] = wn * g.e[j].z + wd[j] * g.r.z;
#pragma unroll
    for (int l = 0; l < 2; ++l)
      M[3 + l][j] = (Z[0][l] * g.D[j][0] + Z[1][l] * g.D[j][1] + Z[2][l] * g.D[j][2] + Z[3][l] * g.D[j][3]) * rn;
  }
  // G^-1 (k x k).
  double Gi[2][2];
  if (g.k == 2) {
    double g11 = dot(g.e[0], g.e[0]), g12 = dot(g.e[0], g.e[1]), g22 = dot(g.e[1], g.e[1]);
    double idet = 1.0 / (g11 * g22 - g12 * g12);
    Gi[0][0] = g22 * idet; Gi[0][1] = -g12 * idet; Gi[1][0] = -g12 * idet; Gi[1][1] = g11 * idet;
  } else {
    Gi[0][0] = 1.0 / dot(g.e[0], g.e[0]); Gi[0][1] = Gi[1][0] = Gi[1][1] = 0.0;
    for (int a = 0; a < 5; ++a) M[a][1] = 0.0;
  }
  // K = 2 diag(|w|^2 I3, 0) - 2 M G^-1 M^T. The unit normal of the column
  // span of M's rows 0..2 (padded with zeros) is an eigenvector of K with
  // eigenvalue 2|w|^2 > 0, so the negative part lives in the 4-dimensional
  // span of Q = [(q1, 0), (q2, 0), (0, e1), (0, e2)], q1, q2 an orthonormal
  // basis of that column span: K is reduced to K4 = Q^T K Q.
  double q1[3], q2[3];
  {
    double n0 = M[0][0] * M[0][0] + M[1][0] * M[1][0] + M[2][0] * M[2][0];
    if (n0 > 0.0) {
      double in0 = inv_sqrt(n0);
#pragma unroll
      for (int a = 0; a < 3; ++a) q1[a] = M[a][0] * in0;
    } else {
      q1[0] = 1.0; q1[1] = 0.0; q1[2] = 0.0;
    }
    double p = q1[0] * M[0][1] + q1[1] * M[1][1] + q1[2] * M[2][1];
    double b[3] = {M[0][1] - p * q1[0], M[1][1] - p * q1[1], M[2][1] - p * q1[2]};
    double nb = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
    double m1 = M[0][1] * M[0][1] + M[1][1] * M[1][1] + M[2][1] * M[2][1];
    if (!(nb > 1.0e-24 * m1) || !(nb > 0.0)) {
      // second column absent or parallel: any unit vector normal to q1
      int ax = fabs(q1[0]) < 0.577 ? 0 : (fabs(q1[1]) < 0.577 ? 1 : 2);
      double ex[3] = {ax == 0 ? 1.0 : 0.0, ax == 1 ? 1.0 : 0.0, ax == 2 ? 1.0 : 0.0};
      double pe = q1[ax];
#pragma unroll
      for (int a = 0; a < 3; ++a) b[a] = ex[a] - pe * q1[a];
      nb = b[0] * b[0] + b[1] * b[1] + b[2] * b[2];
    }
    double inb = inv_sqrt(nb);
#pragma unroll
    for (int a = 0; a < 3; ++a) q2[a] = b[a] * inb;
  }
  double M4[4][2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    M4[0][j] = q1[0] * M[0][j] + q1[1] * M[1][j] + q1[2] * M[2][j];
    M4[1][j] = q2[0] * M[0][j] + q2[1] * M[1][j] + q2[2] * M[2][j];
    M4[2][j] = M[3][j];
    M4[3][j] = M[4][j];
  }
  double K[16];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    double MG0 = M4[a][0] * Gi[0][0] + M4[a][1] * Gi[1][0];
    double MG1 = M4[a][0] * Gi[0][1] + M4[a][1] * Gi[1][1];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      double v = -2.0 * (MG0 * M4[b][0] + MG1 * M4[b][1]);
      if (a == b && a < 2) v += 2.0 * wn2;
      K[a * 4 + b] = v;
    }
  }
  // The 4 x 4 eigenproblem runs in the output's precision. In fp32 (native
  // square roots and reciprocals; in FP64 their multi-instruction sequences
  // were most of the kernel's VALU work) its projection's errors stay near
  // fp32 rounding of the largest entry, below the 1e-5 bound of the blocks,
  // while every input of K (feature geometry, distances) stays FP64. K is
  // scaled by its largest diagonal magnitude first so that fp32's range
  // cannot matter.
  double kmax = fmax(fmax(fabs(K[0]), fabs(K[5])), fmax(fabs(K[10]), fabs(K[15])));
  const double ks = kmax > 0.0 ? 1.0 / kmax : 1.0;
  R Kf[16], evf[4], V[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) Kf[i] = R(K[i] * ks);
  jacobi_sym<4, R>(Kf, evf, V, lit<R>(1.0e-14f, 1.0e-28));
  double ev[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) ev[i] = double(evf[i]) * kmax;
  // N = Q N4 Q^T with N4 = sum over negative eigenvalues of lambda v v^T,
  // accumulated straight into the compressed fields: for an eigenvector v,
  // y = Q v (y0..2 = v0 q1 + v1 q2, y3 = v2, y4 = v3) and zeta_i = Z_i.(y3, y4);
  // then Ntt += lambda y y^T (3 x 3 block), u_i += lambda zeta_i y0..2 and
  // s_ij += lambda zeta_i zeta_j.
  R ntt[6] = {R(0), R(0), R(0), R(0), R(0), R(0)}, uu[4][3], ss[10];
#pragma unroll
  for (int i = 0; i < 4; ++i) uu[i][0] = uu[i][1] = uu[i][2] = R(0);
#pragma unroll
  for (int i = 0; i < 10; ++i) ss[i] = R(0);
#pragma unroll
  for (int m = 0; m < 4; ++m) {
    double lam = ev[m];
    if (lam < 0.0) {
      double y[3], zeta[4];
#pragma unroll
      for (int a = 0; a < 3; ++a) y[a] = V[0 * 4 + m] * q1[a] + V[1 * 4 + m] * q2[a];
#pragma unroll
      for (int i = 0; i < 4; ++i) zeta[i] = Z[i][0] * V[2 * 4 + m] + Z[i][1] * V[3 * 4 + m];
      ntt[0] += R(lam * y[0] * y[0]);
      ntt[1] += R(lam * y[0] * y[1]);
      ntt[2] += R(lam * y[0] * y[2]);
      ntt[3] += R(lam * y[1] * y[1]);
      ntt[4] += R(lam * y[1] * y[2]);
      ntt[5] += R(lam * y[2] * y[2]);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int a = 0; a < 3; ++a) uu[i][a] += R(lam * zeta[i] * y[a]);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = i; j < 4; ++j) ss[sidx(i, j)] += R(lam * zeta[i] * zeta[j]);
    }
  }
#pragma unroll
  for (int i = 0; i < 6; ++i) C.Ntt[i] = ntt[i];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int a = 0; a < 3; ++a) C.u[i][a] = uu[i][a];
#pragma unroll
  for (int i = 0; i < 10; ++i) C.s[i] = ss[i];
  return d2;
}

// Block (i, j) of the compressed Hessian, row-major 3x3.
template <typename R>
IPCD_FN void expand_block(const CompressedHessianT<R>& C, int i, int j, R out[9]) {
#pragma clang fp contract(fast)
  // H_ij[a][b] = c_n w^_i w^_j Ntt[a][b] + r[b] (c_n w^_i u_j[a])
  //            + r[a] (c_n w^_j u_i[b] + rr r[b]): per block three 3-vectors
  //            and the six distinct scaled Ntt entries, then two FMAs an entry.
  const R rr = C.cg * C.w[i] * C.w[j] + C.cn * C.s[sidx(i, j)];
  const R* r = C.rh;
  const R pi = C.cn * C.what[i], pj = C.cn * C.what[j];
  const R cw = pi * C.what[j];
  R A[3], B[3], N[6];
#pragma unroll
  for (int a = 0; a < 3; ++a) A[a] = pi * C.u[j][a];
#pragma unroll
  for (int b = 0; b < 3; ++b) B[b] = pj * C.u[i][b] + rr * r[b];
#pragma unroll
  for (int k = 0; k < 6; ++k) N[k] = cw * C.Ntt[k];
  const R Nt[9] = {N[0], N[1], N[2], N[1], N[3], N[4], N[2], N[4], N[5]};
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) out[a * 3 + b] = Nt[a * 3 + b] + r[b] * A[a] + r[a] * B[b];
}

}  // namespace ipcd

// ================================================================ CCD math ====
// Coplanarity cubic and root isolation (ccd/ccd.py:226-357,
// ccd/cubic_roots.py:1-193) in the precision the caller instantiates (the
// kernels use double). The cubic det[p1 - p0, p2 - p0, p3 - p0](t) is expanded
// in coordinates relative to point 0, which is mathematically the reference's
// sum of four absolute 3x3 determinants with far less cancellation.
namespace ipcd {

template <typename T> IPCD_FN T cevaluate(const T c[4], T x) { return c[0] + x * (c[1] + x * (c[2] + x * c[3])); }
template <typename T> IPCD_FN T multi_sign(T a, T b) { return b < T(0) ? -a : a; }
template <typename T> IPCD_FN bool diff_sign(T y0, T yr) { return y0 * yr < T(0); }
template <typename T> IPCD_FN void deflate(const T coef[4], T root, T out[3]) {
  out[0] = T(0); out[1] = T(0);
  out[2] = coef[3];
  out[1] = coef[2] + root * out[2];
  out[0] = coef[1] + root * out[1];
}
// Safeguarded Newton / bisection of cubic_roots.py:29-62. The reference loop
// is unbounded; this one stops after kFindClosedMaxIter rounds so every wave
// terminates. A round takes the Newton step when it stays inside the
// bracket and the bracket's midpoint otherwise; both are formed and one is
// selected, so lanes of a wave taking different branches do not run the two
// branches one after the other. The values are those of the branched loop:
// the stop tests do not depend on the new y, which is only evaluated to
// continue.
constexpr int kFindClosedMaxIter = 200;
template <typename T>
IPCD_FN T find_closed(const T coef[4], const T deriv[4], T x0, T x1, T y0) {
  const T x_error = T(1e-10);
  T xr = (x0 + x1) / T(2);
  T yr = cevaluate(coef, xr);
  T xb0 = x0, xb1 = x1;
  if (x1 - x0 <= x_error * T(2)) return xr;
  for (int it = 0; it < kFindClosedMaxIter; ++it) {
    bool side = diff_sign(y0, yr);
    if (side) xb1 = xr; else xb0 = xr;
    T dy = cevaluate(deriv, xr);
    T dx = yr / dy;
    T xn = xr - dx;
    bool newton = xn > xb0 && xn < xb1;
    T mid = (xb0 + xb1) / T(2);
    bool done = newton ? !(fabs(xr - xn) > x_error) : (xb0 == mid || xb1 == mid || xb1 - xb0 <= T(2) * x_error);
    xr = newton ? xn : mid;
    if (done) break;
    yr = cevaluate(coef, xr);
  }
  return xr;
}
template <typename T>
IPCD_FN int quadratic_roots(const T coef[3], T x0, T x1, T roots[2]) {
  roots[0] = T(0); roots[1] = T(0);
  T c = coef[0], b = coef[1], a = coef[2];
  T delta = b * b - T(4) * a * c;
  int ret = 0;
  if (delta > T(0)) {
    T d = sqrt(delta);
    T q = -(b + multi_sign(d, b)) * T(0.5);
    T rv0 = q / a, rv1 = c / q;
    T xa = smin(rv0, rv1), xb = smax(rv0, rv1);
    if (xa >= x0 && xa <= x1) roots[ret++] = xa;
    if (xb >= x0 && xb <= x1) roots[ret++] = xb;
  } else if (!(delta < T(0))) {
    T r0 = T(-0.5) * b / a;
    roots[0] = r0;
    if (r0 >= x0 && r0 <= x1) ret = 1;
  }
  return ret;
}
// The case analysis of cubic_roots.py picks one bracket for the iterative
// root and, in some cases, an interval for the deflated quadratic's roots;
// neither choice depends on the root's value, so the cases are decided
// first (cubic_case) and find_closed runs at a single call site. Lanes of a
// wave that take different cases then share one loop instead of running one
// inlined loop per case after another.
template <typename T> struct CubicCase {
  bool search, deflated;
  T lo, hi, ylo, dlo;  // bracket of the root; lower end of the quadratic's interval
  T deriv[4];
};
template <typename T>
IPCD_FN CubicCase<T> cubic_case(const T coef[4], T x0, T x1) {
  CubicCase<T> K;
  T y0 = cevaluate(coef, x0), y1 = cevaluate(coef, x1);
  T a = coef[3] * T(3), b_2 = coef[2], c = coef[1];
  K.deriv[0] = c; K.deriv[1] = T(2) * b_2; K.deriv[2] = a; K.deriv[3] = T(0);
  T delta_4 = b_2 * b_2 - a * c;
  bool search = false, deflated = false;
  T lo = x0, hi = x1, ylo = y0, dlo = x0;
  if (delta_4 > T(0)) {
    T d_2 = sqrt(delta_4);
    T q = -(b_2 + multi_sign(d_2, b_2));
    T rv0 = q / a, rv1 = c / q;
    T xa = smin(rv0, rv1), xb = smax(rv0, rv1);
    bool outside = xa >= x1 || xb <= x0 || (xa <= x0 && xb >= x1);
    if (outside) {
      search = diff_sign(y0, y1);  // bracket [x0, x1]
    } else if (xa > x0) {
      T ya = cevaluate(coef, xa);
      if (diff_sign(y0, ya)) {
        search = true;
        hi = xa;
        deflated = diff_sign(ya, y1) || (xb < x1 && diff_sign(ya, cevaluate(coef, xb)));
        dlo = xa;
      } else if (xb < x1) {
        T yb = cevaluate(coef, xb);
        if (diff_sign(ya, yb)) {
          search = true;
          lo = xa;
          hi = xb;
          ylo = ya;
          deflated = diff_sign(yb, y1);
          dlo = xb;
        } else if (diff_sign(yb, y1)) {
          search = true;
          lo = xb;
          ylo = yb;
        }
      } else if (diff_sign(ya, y1)) {
        search = true;
        lo = xa;
        ylo = ya;
      }
    } else {
      T yb = cevaluate(coef, xb);
      if (diff_sign(y0, yb)) {
        search = true;
        hi = xb;
        deflated = diff_sign(yb, y1);
        dlo = xb;
      } else if (diff_sign(yb, y1)) {
        search = true;
        lo = xb;
        ylo = yb;
      }
    }
  } else {
    search = diff_sign(y0, y1);
  }
  K.search = search;
  K.deflated = deflated;
  K.lo = lo; K.hi = hi; K.ylo = ylo; K.dlo = dlo;
  return K;
}
template <typename T>
IPCD_FN int cubic_roots(const T coef[4], T x0, T x1, T roots[3]) {
  roots[0] = roots[1] = roots[2] = T(0);
  const CubicCase<T> K = cubic_case(coef, x0, x1);
  if (!K.search) return 0;
  roots[0] = find_closed(coef, K.deriv, K.lo, K.hi, K.ylo);
  if (!K.deflated) return 1;
  T defpoly[3], qroots[2];
  deflate(coef, roots[0], defpoly);
  int r = quadratic_roots(defpoly, K.dlo, x1, qroots);
  roots[1] = qroots[0];
  roots[2] = qroots[1];
  return r + 1;
}

template <typename T> IPCD_FN T det3(V3<T> a, V3<T> b, V3<T> c) {
  return a.x * (b.y * c.z - b.z * c.y) + a.y * (b.z * c.x - b.x * c.z) + a.z * (b.x * c.y - b.y * c.x);
}

// Inside tests at a root (ccd/ccd.py:212-224) with the 2x2 least-squares
// barycentric coordinates (beta_gamma_pt / beta_gamma_ee).
template <typename T> IPCD_FN void ls2(V3<T> e1, V3<T> e2, V3<T> q, T* beta, T* gamma) {
  T a11 = dot(e1, e1), a12 = dot(e1, e2), a22 = dot(e2, e2);
  T b1 = dot(e1, q), b2 = dot(e2, q);
  T det = a11 * a22 - a12 * a12;
  *beta = (b1 * a22 - b2 * a12) / det;
  *gamma = (a11 * b2 - a12 * b1) / det;
}

// Coplanarity cubic of the 4 points at t = 0 (p0) and t = 1 (p1), relative
// to point 0: u_k, v_k with point k + 1 at u_k + t v_k.
template <typename T>
IPCD_FN void coplanarity(const V3<T> p0[4], const V3<T> p1[4], V3<T> u[3], V3<T> v[3], T coef[4]) {
  V3<T> d0 = p1[0] - p0[0];
  for (int k = 0; k < 3; ++k) { u[k] = p0[k + 1] - p0[0]; v[k] = (p1[k + 1] - p0[k + 1]) - d0; }
  coef[0] = det3(u[0], u[1], u[2]);
  coef[1] = det3(v[0], u[1], u[2]) + det3(u[0], v[1], u[2]) + det3(u[0], u[1], v[2]);
  coef[2] = det3(u[0], v[1], v[2]) + det3(v[0], u[1], v[2]) + det3(v[0], v[1], u[2]);
  coef[3] = det3(v[0], v[1], v[2]);
}

// Whether the coplanarity cubic has a root in [0, 1] at all (cubic_roots'
// case analysis finds a bracket): without one, collision_time returns 1.
template <typename T>
IPCD_FN bool collision_time_may_hit(const V3<T> p0[4], const V3<T> p1[4]) {
  V3<T> u[3], v[3];
  T coef[4];
  coplanarity(p0, p1, u, v, coef);
  return cubic_case(coef, T(0), T(1)).search;
}

// pt_collision_time / ee_collision_time (ccd/ccd.py:304-357) on the 4 points
// at t = 0 (p0) and t = 1 (p1). PT: point, t0, t1, t2; EE: a0, a1, b0, b1.
template <bool EE, typename T>
IPCD_FN T collision_time(const V3<T> p0[4], const V3<T> p1[4]) {
  V3<T> u[3], v[3];
  T coef[4];
  coplanarity(p0, p1, u, v, coef);
  T roots[3];
  int n = cubic_roots(coef, T(0), T(1), roots);
  const T tol = T(1e-8);
  for (int i = 0; i < n; ++i) {
    T t = roots[i];
    V3<T> q1 = u[0] + v[0] * t, q2 = u[1] + v[1] * t, q3 = u[2] + v[2] * t;
    T beta, gamma;
    bool ok;
    if (EE) {
      // point 0 sits at the origin: ea = q1, eb = q3 - q2, q = q2.
      ls2(q1, q3 - q2, q2, &beta, &gamma);
      ok = beta >= -tol && beta <= T(1) + tol && -gamma >= -tol && -gamma <= T(1) + tol;
    } else {
      // triangle (q1, q2, q3), point at the origin: q = -q1.
      ls2(q2 - q1, q3 - q1, mk(T(0), T(0), T(0)) - q1, &beta, &gamma);
      ok = beta >= -tol && gamma >= -tol && beta + gamma <= T(1) + tol;
    }
    if (ok) return t;
  }
  return T(1);
}

}  // namespace ipcd

// TEST INFRASTRUCTURE ONLY. This is the CPU oracle that the tests, the smoke
// check and bench.py's cpu_baseline leg use to check and time the HIP
// library. Nothing in the product (ipc_warp_amd/) includes, links or calls it.
//
// A C++ restatement of the IPC contact inner loop of Luke-Skycrawler/ipc_warp,
// written from the reference's source text. Every function cites the
// reference file:line it follows. Parts that live in the absent `ipctkwp`
// submodule (barrier, mollifier, NSD distance Hessians, beta/gamma) and in
// the absent Warp builtin `closest_point_edge_edge` are restated from their
// published algorithms and are pinned only by closed forms, brute force and
// invariants (see the tests under tests/test_oracle_*.py).
//
// Precision: `TC` is the precision that feature classification and the
// distance filter run in (float reproduces the reference's FP32 mode, which
// casts every distance query to wp.vec3); derivative math is always double.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

namespace ipco {

// ---------------------------------------------------------------- vec3 ----
template <typename T> struct V3 { T x, y, z; };
template <typename T> inline V3<T> mk(T x, T y, T z) { return V3<T>{x, y, z}; }
template <typename T> inline V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename T> inline V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T> inline V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <typename T> inline V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <typename T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T> inline V3<T> cross(V3<T> a, V3<T> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <typename T> inline T length(V3<T> a) { return std::sqrt(dot(a, a)); }
template <typename T> inline V3<T> vmin(V3<T> a, V3<T> b) { return {std::min(a.x, b.x), std::min(a.y, b.y), std::min(a.z, b.z)}; }
template <typename T> inline V3<T> vmax(V3<T> a, V3<T> b) { return {std::max(a.x, b.x), std::max(a.y, b.y), std::max(a.z, b.z)}; }
template <typename T, typename S> inline V3<T> cast3(V3<S> a) { return {T(a.x), T(a.y), T(a.z)}; }
template <typename T> inline V3<T> load3(const float* p, int i) { return {T(p[3 * i]), T(p[3 * i + 1]), T(p[3 * i + 2])}; }
template <typename T> inline V3<T> load3(const double* p, int i) { return {T(p[3 * i]), T(p[3 * i + 1]), T(p[3 * i + 2])}; }

// --------------------------------------------- closest point on triangle ----
// contact.py:230-302 (Ericson, Real-Time Collision Detection 5.1.5).
// Returns (w_a, w_b, dist) and the feature: 0/1/2 vertex a/b/c, 3 edge ab,
// 4 edge ac, 5 edge bc, 6 face. The weight of c is 1 - w_a - w_b.
template <typename T> struct TriClosest { T wa, wb, dist; int type; };

template <typename T>
TriClosest<T> closest_point_triangle(V3<T> p, V3<T> a, V3<T> b, V3<T> c) {
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

// -------------------------------------------- closest points of segments ----
// Stands in for Warp's builtin wp.closest_point_edge_edge(p1, q1, p2, q2, eps)
// (call sites contact.py:104,226; dynamic_contacts.py:52). Warp is absent, so
// this restates Ericson, Real-Time Collision Detection 5.1.9
// (ClosestPtSegmentSegment) with `eps` as the squared-length degeneracy test.
// Returns (s, t, dist).
template <typename T> struct EdgeClosest { T s, t, dist; };

template <typename T> inline T clamp01(T v) { return v < T(0) ? T(0) : (v > T(1) ? T(1) : v); }

template <typename T>
EdgeClosest<T> closest_point_edge_edge(V3<T> p1, V3<T> q1, V3<T> p2, V3<T> q2, T eps) {
  V3<T> d1 = q1 - p1, d2 = q2 - p2, r = p1 - p2;
  T a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
  T s, t;
  if (a <= eps && e <= eps) {
    s = T(0); t = T(0);
  } else if (a <= eps) {
    s = T(0); t = clamp01(f / e);
  } else {
    T c = dot(d1, r);
    if (e <= eps) {
      t = T(0); s = clamp01(-c / a);
    } else {
      T b = dot(d1, d2);
      T denom = a * e - b * b;
      s = denom != T(0) ? clamp01((b * f - c * e) / denom) : T(0);
      t = (b * s + f) / e;
      if (t < T(0)) { t = T(0); s = clamp01(-c / a); }
      else if (t > T(1)) { t = T(1); s = clamp01((b - c) / a); }
    }
  }
  V3<T> c1 = p1 + d1 * s, c2 = p2 + d2 * t;
  return {s, t, length(c1 - c2)};
}

// ------------------------------------------------------------- barrier ----
// ipc_barrier / _derivative / _derivative2 (ipctkwp, absent; call sites
// dynamic_contacts.py:176-177,182,228-229,387,397). Restated from IPC
// (Li et al. 2020): b(x) = -kappa (x - D)^2 ln(x / D) for 0 < x < D with
// x = d^2 and D = dhat^2, derivatives taken with respect to x.
struct Barrier { double b, db, d2b; };
inline Barrier barrier(double x, double D, double kappa) {
  if (!(x < D)) return {0.0, 0.0, 0.0};
  double lg = std::log(x / D);
  double dm = x - D;
  double b = -kappa * dm * dm * lg;
  double db = kappa * (D - x) * (2.0 * lg - D / x + 1.0);
  double r = D / x;
  double d2b = kappa * ((r + 2.0) * r - 2.0 * lg - 3.0);
  return {b, db, d2b};
}

// --------------------------------------------- second-order 12-DOF jets ----
// Exact value/gradient/Hessian arithmetic used to differentiate the explicit
// squared-distance formulas. It is independent of the envelope-theorem form
// the HIP kernel uses, so agreement pins both.
struct Jet {
  double v;
  double g[12];
  double h[12][12];
};
inline Jet jconst(double c) { Jet r; std::memset(&r, 0, sizeof(r)); r.v = c; return r; }
inline Jet jvar(double c, int k) { Jet r = jconst(c); r.g[k] = 1.0; return r; }
inline Jet jadd(const Jet& a, const Jet& b) {
  Jet r; r.v = a.v + b.v;
  for (int i = 0; i < 12; ++i) r.g[i] = a.g[i] + b.g[i];
  for (int i = 0; i < 12; ++i) for (int j = 0; j < 12; ++j) r.h[i][j] = a.h[i][j] + b.h[i][j];
  return r;
}
inline Jet jsub(const Jet& a, const Jet& b) {
  Jet r; r.v = a.v - b.v;
  for (int i = 0; i < 12; ++i) r.g[i] = a.g[i] - b.g[i];
  for (int i = 0; i < 12; ++i) for (int j = 0; j < 12; ++j) r.h[i][j] = a.h[i][j] - b.h[i][j];
  return r;
}
inline Jet jmul(const Jet& a, const Jet& b) {
  Jet r; r.v = a.v * b.v;
  for (int i = 0; i < 12; ++i) r.g[i] = a.v * b.g[i] + b.v * a.g[i];
  for (int i = 0; i < 12; ++i)
    for (int j = 0; j < 12; ++j)
      r.h[i][j] = a.v * b.h[i][j] + b.v * a.h[i][j] + a.g[i] * b.g[j] + b.g[i] * a.g[j];
  return r;
}
inline Jet jdiv(const Jet& a, const Jet& b) {
  // a / b = a * (1/b); (1/b)' = -b'/b^2, (1/b)'' = 2 b' b'^T / b^3 - b'' / b^2.
  Jet inv; double iv = 1.0 / b.v;
  inv.v = iv;
  for (int i = 0; i < 12; ++i) inv.g[i] = -b.g[i] * iv * iv;
  for (int i = 0; i < 12; ++i)
    for (int j = 0; j < 12; ++j) inv.h[i][j] = 2.0 * b.g[i] * b.g[j] * iv * iv * iv - b.h[i][j] * iv * iv;
  return jmul(a, inv);
}
struct JV3 { Jet x, y, z; };
inline JV3 jv_seed(V3<double> p, int slot) { return {jvar(p.x, 3 * slot), jvar(p.y, 3 * slot + 1), jvar(p.z, 3 * slot + 2)}; }
inline JV3 jv_sub(const JV3& a, const JV3& b) { return {jsub(a.x, b.x), jsub(a.y, b.y), jsub(a.z, b.z)}; }
inline Jet jv_dot(const JV3& a, const JV3& b) { return jadd(jadd(jmul(a.x, b.x), jmul(a.y, b.y)), jmul(a.z, b.z)); }
inline JV3 jv_cross(const JV3& a, const JV3& b) {
  return {jsub(jmul(a.y, b.z), jmul(a.z, b.y)), jsub(jmul(a.z, b.x), jmul(a.x, b.z)), jsub(jmul(a.x, b.y), jmul(a.y, b.x))};
}

// ------------------------------------------------------ Jacobi eigensolver ----
// Cyclic Jacobi for a symmetric n x n matrix (row-major, n <= 12), double.
inline void jacobi_eigen(int n, double* A, double* evals, double* evecs) {
  for (int i = 0; i < n; ++i) for (int j = 0; j < n; ++j) evecs[i * n + j] = (i == j) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) { tot += A[i * n + j] * A[i * n + j]; if (i != j) off += A[i * n + j] * A[i * n + j]; }
    if (off <= 1e-30 * tot || off == 0.0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        double apq = A[p * n + q];
        if (apq == 0.0) continue;
        double app = A[p * n + p], aqq = A[q * n + q];
        double theta = (aqq - app) / (2.0 * apq);
        double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {
          double akp = A[k * n + p], akq = A[k * n + q];
          A[k * n + p] = c * akp - s * akq;
          A[k * n + q] = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {
          double apk = A[p * n + k], aqk = A[q * n + k];
          A[p * n + k] = c * apk - s * aqk;
          A[q * n + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          double vkp = evecs[k * n + p], vkq = evecs[k * n + q];
          evecs[k * n + p] = c * vkp - s * vkq;
          evecs[k * n + q] = s * vkp + c * vkq;
        }
      }
  }
  for (int i = 0; i < n; ++i) evals[i] = A[i * n + i];
}

// Negative (sign=-1) or positive (sign=+1) spectral part of a symmetric 12x12.
inline void spectral_part_12(const double* H, int sign, double* out) {
  double A[144], ev[12], V[144];
  std::memcpy(A, H, sizeof(A));
  // Symmetrise against round-off from the jets.
  for (int i = 0; i < 12; ++i) for (int j = i + 1; j < 12; ++j) { double m = 0.5 * (A[i * 12 + j] + A[j * 12 + i]); A[i * 12 + j] = A[j * 12 + i] = m; }
  jacobi_eigen(12, A, ev, V);
  std::memset(out, 0, sizeof(double) * 144);
  for (int k = 0; k < 12; ++k) {
    double lam = ev[k];
    if ((sign < 0 && lam < 0.0) || (sign > 0 && lam > 0.0))
      for (int i = 0; i < 12; ++i) for (int j = 0; j < 12; ++j) out[i * 12 + j] += lam * V[i * 12 + k] * V[j * 12 + k];
  }
}

// ---------------------------------------- distance features and their NSD ----
// Feature kinds produced by the dispatch in dynamic_contacts.py:48-125.
enum FeatureKind { F_PP = 0, F_PE = 1, F_PT = 2, F_EE = 3 };
struct Feature { int kind; int slot[4]; };

// Squared distance of a feature as an exact jet over the 12 slot coordinates.
//  PP (dynamic_contacts.py:19-29): |xa - xb|^2.
//  PE (dynamic_contacts.py:31-46, ipctkwp point_edge_distance_gradient_nsd_hessian):
//     point-line distance |(p - e0) x (e1 - e0)|^2 / |e1 - e0|^2.
//  PT (ipctkwp x_to_grad_nsd_hess_pt): point-plane ((p - a).n)^2 / |n|^2.
//  EE (ipctkwp x_to_grad_nsd_hess_ee): line-line ((x2 - x0).n)^2 / |n|^2.
inline Jet feature_d2_jet(const Feature& f, const V3<double> x[4]) {
  JV3 s[4];
  for (int k = 0; k < 4; ++k) s[k] = jv_seed(x[k], k);
  const int* o = f.slot;
  if (f.kind == F_PP) {
    JV3 d = jv_sub(s[o[0]], s[o[1]]);
    return jv_dot(d, d);
  }
  if (f.kind == F_PE) {
    JV3 u = jv_sub(s[o[0]], s[o[1]]);
    JV3 e = jv_sub(s[o[2]], s[o[1]]);
    JV3 n = jv_cross(u, e);
    return jdiv(jv_dot(n, n), jv_dot(e, e));
  }
  if (f.kind == F_PT) {
    JV3 n = jv_cross(jv_sub(s[o[2]], s[o[1]]), jv_sub(s[o[3]], s[o[1]]));
    Jet sd = jv_dot(jv_sub(s[o[0]], s[o[1]]), n);
    return jdiv(jmul(sd, sd), jv_dot(n, n));
  }
  JV3 n = jv_cross(jv_sub(s[o[1]], s[o[0]]), jv_sub(s[o[3]], s[o[2]]));
  Jet sd = jv_dot(jv_sub(s[o[2]], s[o[0]]), n);
  return jdiv(jmul(sd, sd), jv_dot(n, n));
}

// Gradient of d^2 and the negative-semidefinite spectral part of its Hessian
// (the "grad_nsd_hess" contract of the ipctkwp call sites). For PP the
// reference leaves the Hessian zero (dynamic_contacts.py:25-27).
inline double feature_grad_nsd(const Feature& f, const V3<double> x[4], double g[12], double H[144]) {
  Jet d2 = feature_d2_jet(f, x);
  for (int i = 0; i < 12; ++i) g[i] = d2.g[i];
  if (f.kind == F_PP) {
    std::memset(H, 0, sizeof(double) * 144);
  } else {
    double full[144];
    for (int i = 0; i < 12; ++i) for (int j = 0; j < 12; ++j) full[i * 12 + j] = d2.h[i][j];
    spectral_part_12(full, -1, H);
  }
  return d2.v;
}

// Full (unfiltered) Hessian of d^2 for a feature, row-major 12x12.
inline double feature_hessian(const Feature& f, const V3<double> x[4], double g[12], double H[144]) {
  Jet d2 = feature_d2_jet(f, x);
  for (int i = 0; i < 12; ++i) g[i] = d2.g[i];
  for (int i = 0; i < 12; ++i) for (int j = 0; j < 12; ++j) H[i * 12 + j] = d2.h[i][j];
  return d2.v;
}

// edge_edge_distance_gradient_hessian dispatch (dynamic_contacts.py:48-108),
// classification in precision TC.
template <typename TC>
Feature ee_feature(V3<TC> x0, V3<TC> x1, V3<TC> x2, V3<TC> x3, TC parallel_eps) {
  EdgeClosest<TC> dab = closest_point_edge_edge(x0, x1, x2, x3, TC(1.0e-6));
  TC s = dab.s, t = dab.t, fe = TC(1.0e-6), one = TC(1);
  bool sb = s <= fe || s >= one - fe;
  bool tb = t <= fe || t >= one - fe;
  if (sb && tb) {
    int oa = (s >= one - fe) ? 1 : 0;
    int ob = (t >= one - fe) ? 3 : 2;
    return {F_PP, {oa, ob, -1, -1}};
  }
  if (sb) {
    int oa = (s >= one - fe) ? 1 : 0;
    return {F_PE, {oa, 2, 3, -1}};
  }
  if (tb) {
    int ob = (t >= one - fe) ? 3 : 2;
    return {F_PE, {ob, 0, 1, -1}};
  }
  V3<TC> cr = cross(x1 - x0, x3 - x2);
  if (parallel_eps > TC(0) && dot(cr, cr) < parallel_eps) {
    TC ds0 = s, ds1 = one - s, dt0 = t, dt1 = one - t;
    TC m = std::min(std::min(ds0, ds1), std::min(dt0, dt1));
    if (m == ds0) return {F_PE, {0, 2, 3, -1}};
    if (m == ds1) return {F_PE, {1, 2, 3, -1}};
    if (m == dt0) return {F_PE, {2, 0, 1, -1}};
    return {F_PE, {3, 0, 1, -1}};
  }
  return {F_EE, {0, 1, 2, 3}};
}

// point_triangle_distance_gradient_hessian dispatch (dynamic_contacts.py:110-125).
inline Feature pt_feature_from_type(int type) {
  switch (type) {
    case 0: return {F_PP, {0, 1, -1, -1}};
    case 1: return {F_PP, {0, 2, -1, -1}};
    case 2: return {F_PP, {0, 3, -1, -1}};
    case 3: return {F_PE, {0, 1, 2, -1}};
    case 4: return {F_PE, {0, 1, 3, -1}};
    case 5: return {F_PE, {0, 2, 3, -1}};
    default: return {F_PT, {0, 1, 2, 3}};
  }
}

// ------------------------------------------------------------ mollifier ----
// ee_mollifier_threshold / _value / _gradient_hessian_psd (ipctkwp, absent;
// call sites dynamic_contacts.py:128-134,178-181,383-386). Restated from IPC:
// eps_x = 1e-3 |ea_rest|^2 |eb_rest|^2, c = |ea x eb|^2,
// m(c) = (c/eps_x)(2 - c/eps_x) for c < eps_x, else 1.
template <typename T>
inline T ee_mollifier_threshold(V3<T> r0, V3<T> r1, V3<T> r2, V3<T> r3) {
  V3<T> ea = r1 - r0, eb = r3 - r2;
  return T(1.0e-3) * dot(ea, ea) * dot(eb, eb);
}

struct Mollifier { double m; double g[12]; double H[144]; };
// `active` is the branch c < eps_x as decided by the caller (the reference
// decides it in the pipeline precision).
inline Mollifier ee_mollifier(const V3<double> x[4], double eps_x, bool active) {
  Mollifier M;
  std::memset(&M, 0, sizeof(M));
  if (!active) { M.m = 1.0; return M; }
  JV3 s[4];
  for (int k = 0; k < 4; ++k) s[k] = jv_seed(x[k], k);
  JV3 n = jv_cross(jv_sub(s[1], s[0]), jv_sub(s[3], s[2]));
  Jet c = jv_dot(n, n);
  double q = c.v / eps_x;
  M.m = q * (2.0 - q);
  double dm = 2.0 / eps_x * (1.0 - q);  // dm/dc
  double d2m = -2.0 / (eps_x * eps_x);   // d2m/dc2
  double full[144];
  for (int i = 0; i < 12; ++i) M.g[i] = dm * c.g[i];
  for (int i = 0; i < 12; ++i) for (int j = 0; j < 12; ++j) full[i * 12 + j] = d2m * c.g[i] * c.g[j] + dm * c.h[i][j];
  spectral_part_12(full, +1, M.H);
  return M;
}

// symmetric_outer_product_psd_12 (dynamic_contacts.py:136-149).
inline void symmetric_outer_product_psd_12(const double u[12], const double v[12], double out[144]) {
  std::memset(out, 0, sizeof(double) * 144);
  double un = 0, vn = 0;
  for (int i = 0; i < 12; ++i) { un += u[i] * u[i]; vn += v[i] * v[i]; }
  un = std::sqrt(un); vn = std::sqrt(vn);
  double np = un * vn;
  if (np > 1.0e-30) {
    double w[12], scale = 0.5 / np;
    for (int i = 0; i < 12; ++i) w[i] = vn * u[i] + un * v[i];
    for (int i = 0; i < 12; ++i) for (int j = 0; j < 12; ++j) out[i * 12 + j] = scale * w[i] * w[j];
  }
}

// ---------------------------------------------------- contact kernels ----
// d2 policy: 0 = the reference's FP32 mode, d2 = dist*dist of the TC closest
// point; 1 = the squared distance of the classified feature recomputed in
// double (the HIP library's policy).
struct PairResult {
  bool active;
  Feature feature;        // the dispatched feature (valid when active)
  double grad_force[12];  // the value atomically added into `b` (force = -grad E)
  double H[144];
  double energy;
};

template <typename TC> inline V3<TC> cast3(const V3<double>& v) { return V3<TC>{TC(v.x), TC(v.y), TC(v.z)}; }

// contact_hessian_ee (dynamic_contacts.py:151-207) + contact_energy_ee (:374-388)
// at positions xd (rest positions rd for the mollifier threshold). The
// reference's buffers are fp32: a caller passing fp32 values widened to
// double gets them back exactly in the TC = float classification. l0 is the
// `scalar` 2 * thickness (a float widened to double in fp32).
template <typename TC>
PairResult contact_ee_at(const V3<double> xd[4], const V3<double> rd[4], double l0f, double kappa, int d2_policy) {
  PairResult R;
  std::memset(&R, 0, sizeof(R));
  V3<TC> xc[4];
  for (int k = 0; k < 4; ++k) xc[k] = cast3<TC>(xd[k]);
  EdgeClosest<TC> ec = closest_point_edge_edge(xc[0], xc[1], xc[2], xc[3], TC(1.0e-6));
  TC l0 = TC(l0f);
  if (!(ec.dist < l0)) return R;
  R.active = true;
  TC eps_x = ee_mollifier_threshold(cast3<TC>(rd[0]), cast3<TC>(rd[1]), cast3<TC>(rd[2]), cast3<TC>(rd[3]));
  V3<TC> cr = cross(xc[1] - xc[0], xc[3] - xc[2]);
  bool near_parallel = eps_x > TC(0) && dot(cr, cr) < eps_x;
  Feature f = ee_feature(xc[0], xc[1], xc[2], xc[3], eps_x);
  R.feature = f;
  double g[12], Hn[144];
  double d2m = feature_grad_nsd(f, xd, g, Hn);
  double d2 = d2_policy == 0 ? double(ec.dist * ec.dist) : d2m;
  double d02 = d2_policy == 0 ? double(l0 * l0) : double(l0f) * double(l0f);
  Barrier B = barrier(d2, d02, kappa);
  Mollifier M = ee_mollifier(xd, double(eps_x), near_parallel);
  double bg[12];
  for (int i = 0; i < 12; ++i) bg[i] = B.db * g[i];
  if (near_parallel) {
    for (int i = 0; i < 12; ++i) for (int j = 0; j < 12; ++j) R.H[i * 12 + j] = M.m * B.d2b * g[i] * g[j];
  } else {
    double P[144];
    symmetric_outer_product_psd_12(M.g, bg, P);
    for (int i = 0; i < 12; ++i)
      for (int j = 0; j < 12; ++j)
        R.H[i * 12 + j] = M.m * (B.d2b * g[i] * g[j] + B.db * Hn[i * 12 + j]) + B.b * M.H[i * 12 + j] + P[i * 12 + j];
  }
  for (int i = 0; i < 12; ++i) R.grad_force[i] = -(M.m * bg[i] + B.b * M.g[i]);
  R.energy = M.m * B.b;
  return R;
}

template <typename TC, typename XT>
PairResult contact_ee(const XT* x, const XT* xcs, const int c[4], double l0f, double kappa, int d2_policy) {
  V3<double> xd[4], rd[4];
  for (int k = 0; k < 4; ++k) { xd[k] = load3<double>(x, c[k]); rd[k] = load3<double>(xcs, c[k]); }
  return contact_ee_at<TC>(xd, rd, l0f, kappa, d2_policy);
}

// contact_hessian_pt (dynamic_contacts.py:209-246) + contact_energy_pt (:390-398).
// Slot order is the contact record (p, t1, t0, t2) written by
// point_triangle_collision (contact.py:336); fetch_dist_v0v1_pt
// (contact.py:116-145) treats slots 1..3 as the triangle (a, b, c).
template <typename TC>
PairResult contact_pt_at(const V3<double> xd[4], double l0f, double kappa, int d2_policy) {
  PairResult R;
  std::memset(&R, 0, sizeof(R));
  V3<TC> xc[4];
  for (int k = 0; k < 4; ++k) xc[k] = cast3<TC>(xd[k]);
  TriClosest<TC> tc = closest_point_triangle(xc[0], xc[1], xc[2], xc[3]);
  TC l0 = TC(l0f);
  if (!(tc.dist < l0)) return R;
  R.active = true;
  Feature f = pt_feature_from_type(tc.type);
  R.feature = f;
  double g[12], Hn[144];
  double d2m = feature_grad_nsd(f, xd, g, Hn);
  double d2 = d2_policy == 0 ? double(tc.dist * tc.dist) : d2m;
  double d02 = d2_policy == 0 ? double(l0 * l0) : double(l0f) * double(l0f);
  Barrier B = barrier(d2, d02, kappa);
  for (int i = 0; i < 12; ++i) for (int j = 0; j < 12; ++j) R.H[i * 12 + j] = B.d2b * g[i] * g[j] + B.db * Hn[i * 12 + j];
  for (int i = 0; i < 12; ++i) R.grad_force[i] = -B.db * g[i];
  R.energy = B.b;
  return R;
}

template <typename TC, typename XT>
PairResult contact_pt(const XT* x, const int c[4], double l0f, double kappa, int d2_policy) {
  V3<double> xd[4];
  for (int k = 0; k < 4; ++k) xd[k] = load3<double>(x, c[k]);
  return contact_pt_at<TC>(xd, l0f, kappa, d2_policy);
}

// --------------------------------------------------------------- CCD ----
// ccd/cubic_roots.py restated. x_error = 1e-10 (:4).
template <typename T> inline T evaluate(const T c[4], T x) { return c[0] + x * (c[1] + x * (c[2] + x * c[3])); }  // :6-8
template <typename T> inline T multi_sign(T a, T b) { return b < T(0) ? -a : a; }                               // :10-15
template <typename T> inline bool is_different_sign(T y0, T yr) { return y0 * yr < T(0); }                      // :17-19
template <typename T> inline void deflate(const T coef[4], T root, T out[3]) {                                  // :21-27
  out[0] = T(0); out[1] = T(0);
  out[2] = coef[3];
  for (int i = 2; i > 0; --i) out[i - 1] = coef[i] + root * out[i];
}
// The reference loop has no iteration cap; this restatement stops after
// kFindClosedMaxIter rounds and counts the event in *capped (tests assert it
// never happens on their inputs). The HIP kernel uses the same cap.
constexpr int kFindClosedMaxIter = 200;
template <typename T>
T find_closed(const T coef[4], const T deriv[4], T x0, T x1, T y0, T /*y1*/, int* capped) {  // :29-62
  const T x_error = T(1e-10);
  T xr = (x0 + x1) / T(2);
  T yr = evaluate(coef, xr);
  T xb0 = x0, xb1 = x1;
  if (x1 - x0 <= x_error * T(2)) return xr;
  for (int it = 0;; ++it) {
    if (it >= kFindClosedMaxIter) { if (capped) ++*capped; break; }
    bool side = is_different_sign(y0, yr);
    if (side) xb1 = xr; else xb0 = xr;
    T dy = evaluate(deriv, xr);
    T dx = yr / dy;
    T xn = xr - dx;
    if (xn > xb0 && xn < xb1) {
      T stepsize = std::fabs(xr - xn);
      xr = xn;
      if (stepsize > x_error) yr = evaluate(coef, xr);
      else break;
    } else {
      xr = (xb0 + xb1) / T(2);
      yr = evaluate(coef, xr);
      if (xb0 == xr || xb1 == xr || xb1 - xb0 <= T(2) * x_error) break;
    }
  }
  return xr;
}
template <typename T>
int quadratic_roots(const T coef[3], T x0, T x1, T roots[2]) {  // :63-93
  roots[0] = T(0); roots[1] = T(0);
  T c = coef[0], b = coef[1], a = coef[2];
  T delta = b * b - T(4) * a * c;
  int ret = 0;
  if (delta > T(0)) {
    T d = std::sqrt(delta);
    T q = -(b + multi_sign(d, b)) * T(0.5);
    T rv0 = q / a, rv1 = c / q;
    T xa = std::min(rv0, rv1), xb = std::max(rv0, rv1);
    if (xa >= x0 && xa <= x1) roots[ret++] = xa;
    if (xb >= x0 && xb <= x1) roots[ret++] = xb;
  } else if (delta < T(0)) {
    ret = 0;
  } else {
    T r0 = T(-0.5) * b / a;
    roots[0] = r0;
    if (r0 >= x0 && r0 <= x1) ret = 1;
  }
  return ret;
}
// wp.min / wp.max of two scalars (NaN handling matches a < b ? a : b).
template <typename T> inline T smin(T a, T b) { return a < b ? a : b; }
template <typename T> inline T smax(T a, T b) { return a > b ? a : b; }

template <typename T>
int cubic_roots(const T coef[4], T x0, T x1, T roots[3], int* capped) {  // :95-193
  roots[0] = roots[1] = roots[2] = T(0);
  T y0 = evaluate(coef, x0), y1 = evaluate(coef, x1);
  T a = coef[3] * T(3), b_2 = coef[2], c = coef[1];
  T deriv[4] = {c, T(2) * b_2, a, T(0)};
  T delta_4 = b_2 * b_2 - a * c;
  T defpoly[3], qroots[2];
  if (delta_4 > T(0)) {
    T d_2 = std::sqrt(delta_4);
    T q = -(b_2 + multi_sign(d_2, b_2));
    T rv0 = q / a, rv1 = c / q;
    T xa = smin(rv0, rv1), xb = smax(rv0, rv1);
    if (is_different_sign(y0, y1)) {
      if (xa >= x1 || xb <= x0 || (xa <= x0 && xb >= x1)) {
        roots[0] = find_closed(coef, deriv, x0, x1, y0, y1, capped);
        return 1;
      }
    } else {
      if (xa >= x1 || xb <= x0 || (xa <= x0 && xb >= x1)) return 0;
    }
    if (xa > x0) {
      T ya = evaluate(coef, xa);
      if (is_different_sign(y0, ya)) {
        roots[0] = find_closed(coef, deriv, x0, xa, y0, ya, capped);
        if (is_different_sign(ya, y1) || (xb < x1 && is_different_sign(ya, evaluate(coef, xb)))) {
          deflate(coef, roots[0], defpoly);
          int r = quadratic_roots(defpoly, xa, x1, qroots);
          roots[1] = qroots[0]; roots[2] = qroots[1];
          return r + 1;
        }
        return 1;
      }
      if (xb < x1) {
        T yb = evaluate(coef, xb);
        if (is_different_sign(ya, yb)) {
          roots[0] = find_closed(coef, deriv, xa, xb, ya, yb, capped);
          if (is_different_sign(yb, y1)) {
            deflate(coef, roots[0], defpoly);
            int r = quadratic_roots(defpoly, xb, x1, qroots);
            roots[1] = qroots[0]; roots[2] = qroots[1];
            return r + 1;
          }
          return 1;
        }
        if (is_different_sign(yb, y1)) {
          roots[0] = find_closed(coef, deriv, xb, x1, yb, y1, capped);
          return 1;
        }
      } else if (is_different_sign(ya, y1)) {
        roots[0] = find_closed(coef, deriv, xa, x1, ya, y1, capped);
        return 1;
      }
    } else {
      T yb = evaluate(coef, xb);
      if (is_different_sign(y0, yb)) {
        roots[0] = find_closed(coef, deriv, x0, xb, y0, yb, capped);
        if (is_different_sign(yb, y1)) {
          deflate(coef, roots[0], defpoly);
          int r = quadratic_roots(defpoly, xb, x1, qroots);
          roots[1] = qroots[0]; roots[2] = qroots[1];
          return r + 1;
        }
        return 1;
      }
      if (is_different_sign(yb, y1)) {
        roots[0] = find_closed(coef, deriv, xb, x1, yb, y1, capped);
        return 1;
      }
    }
    return 0;  // the reference's `while True: ... break` falls through with ret = 0
  }
  if (is_different_sign(y0, y1)) {
    roots[0] = find_closed(coef, deriv, x0, x1, y0, y1, capped);
    return 1;
  }
  return 0;
}

// cubic_binomial (ccd/ccd.py:293-300): coefficients of prod_i (b_i + t a_i).
template <typename T> inline void cubic_binomial(const T a[3], const T b[3], T out[4]) {
  out[0] = b[0] * b[1] * b[2];
  out[1] = a[0] * b[1] * b[2] + b[0] * b[1] * a[2] + b[0] * a[1] * b[2];
  out[2] = a[0] * a[1] * b[2] + a[0] * b[1] * a[2] + b[0] * a[1] * a[2];
  out[3] = a[0] * a[1] * a[2];
}
// det_polynomial (ccd/ccd.py:252-290): det(B + t A) as ascending coefficients;
// A and B are 3x3 with columns given as V3.
template <typename T> inline void det_polynomial(const V3<T> A[3], const V3<T> B[3], T out[4]) {
  auto el = [](const V3<T> M[3], int r, int c) -> T { return r == 0 ? M[c].x : (r == 1 ? M[c].y : M[c].z); };
  const int perm_pos[3][3][2] = {{{0, 0}, {1, 1}, {2, 2}}, {{0, 1}, {1, 2}, {2, 0}}, {{0, 2}, {1, 0}, {2, 1}}};
  const int perm_neg[3][3][2] = {{{0, 0}, {1, 2}, {2, 1}}, {{0, 1}, {1, 0}, {2, 2}}, {{0, 2}, {1, 1}, {2, 0}}};
  T pos[4] = {0, 0, 0, 0}, neg[4] = {0, 0, 0, 0}, tmp[4];
  for (int p = 0; p < 3; ++p) {
    T a[3], b[3];
    for (int k = 0; k < 3; ++k) { a[k] = el(A, perm_pos[p][k][0], perm_pos[p][k][1]); b[k] = el(B, perm_pos[p][k][0], perm_pos[p][k][1]); }
    cubic_binomial(a, b, tmp);
    for (int k = 0; k < 4; ++k) pos[k] += tmp[k];
  }
  for (int p = 0; p < 3; ++p) {
    T a[3], b[3];
    for (int k = 0; k < 3; ++k) { a[k] = el(A, perm_neg[p][k][0], perm_neg[p][k][1]); b[k] = el(B, perm_neg[p][k][0], perm_neg[p][k][1]); }
    cubic_binomial(a, b, tmp);
    for (int k = 0; k < 4; ++k) neg[k] += tmp[k];
  }
  for (int k = 0; k < 4; ++k) out[k] = pos[k] - neg[k];
}
// build_and_solve_4_points_coplanar (ccd/ccd.py:226-249), absolute coordinates.
template <typename T>
void coplanarity_cubic_absolute(const V3<T> p0[4], const V3<T> p1[4], T coef[4]) {
  const int cols[4][3] = {{1, 2, 3}, {0, 2, 3}, {0, 1, 3}, {0, 1, 2}};
  T acc[4] = {0, 0, 0, 0};
  for (int k = 0; k < 4; ++k) {
    V3<T> A[3], B[3];
    for (int j = 0; j < 3; ++j) { B[j] = p0[cols[k][j]]; A[j] = p1[cols[k][j]] - p0[cols[k][j]]; }
    T d[4];
    det_polynomial(A, B, d);
    T sgn = (k % 2 == 0) ? T(1) : T(-1);
    for (int i = 0; i < 4; ++i) acc[i] = acc[i] + sgn * d[i];
  }
  // The reference sums t = det1 - det2 + det3 - det4 left to right.
  for (int i = 0; i < 4; ++i) coef[i] = acc[i];
}
template <typename T> inline T det3(V3<T> a, V3<T> b, V3<T> c) {
  return a.x * (b.y * c.z - b.z * c.y) + a.y * (b.z * c.x - b.x * c.z) + a.z * (b.x * c.y - b.y * c.x);
}
// The same cubic, det[p1 - p0, p2 - p0, p3 - p0](t), expanded in coordinates
// relative to point 0: u_k = p_k(0) - p_0(0), v_k = (p_k(1) - p_k(0)) - (p_0(1) - p_0(0)).
// This is the HIP library's deliberate deviation (less cancellation).
template <typename T>
void coplanarity_cubic_relative(const V3<T> p0[4], const V3<T> p1[4], T coef[4], V3<T> u[3], V3<T> v[3]) {
  V3<T> d0 = p1[0] - p0[0];
  for (int k = 0; k < 3; ++k) { u[k] = p0[k + 1] - p0[0]; v[k] = (p1[k + 1] - p0[k + 1]) - d0; }
  coef[0] = det3(u[0], u[1], u[2]);
  coef[1] = det3(v[0], u[1], u[2]) + det3(u[0], v[1], u[2]) + det3(u[0], u[1], v[2]);
  coef[2] = det3(u[0], v[1], v[2]) + det3(v[0], u[1], v[2]) + det3(v[0], v[1], u[2]);
  coef[3] = det3(v[0], v[1], v[2]);
}

// beta_gamma_pt / beta_gamma_ee (ipctkwp, absent; call sites ccd/ccd.py:212-224).
// Restated as the 2x2 least-squares coordinates the inside tests imply:
//  PT: x0 - x1 = beta (x2 - x1) + gamma (x3 - x1);
//  EE: x2 - x0 = beta (x1 - x0) + gamma (x3 - x2), so the hit is -gamma on edge b.
template <typename T>
inline void solve_2x2_ls(V3<T> e1, V3<T> e2, V3<T> q, T* beta, T* gamma) {
  T a11 = dot(e1, e1), a12 = dot(e1, e2), a22 = dot(e2, e2);
  T b1 = dot(e1, q), b2 = dot(e2, q);
  T det = a11 * a22 - a12 * a12;
  *beta = (b1 * a22 - b2 * a12) / det;
  *gamma = (a11 * b2 - a12 * b1) / det;
}
template <typename T> inline bool verify_root_pt(V3<T> x0, V3<T> x1, V3<T> x2, V3<T> x3) {  // ccd/ccd.py:212-216
  T beta, gamma;
  solve_2x2_ls(x2 - x1, x3 - x1, x0 - x1, &beta, &gamma);
  T tol = T(1e-8);
  return beta >= -tol && gamma >= -tol && beta + gamma <= T(1) + tol;
}
template <typename T> inline bool verify_root_ee(V3<T> x0, V3<T> x1, V3<T> x2, V3<T> x3) {  // ccd/ccd.py:218-224
  T beta, gamma;
  solve_2x2_ls(x1 - x0, x3 - x2, x2 - x0, &beta, &gamma);
  T tol = T(1e-8);
  return beta >= -tol && beta <= T(1) + tol && -gamma >= -tol && -gamma <= T(1) + tol;
}
template <typename T> inline V3<T> lerp(V3<T> a, V3<T> b, T t) { return a * (T(1) - t) + b * t; }

// pt_collision_time / ee_collision_time (ccd/ccd.py:304-357). relative=false
// restates the reference exactly (absolute coordinates, lerp); relative=true
// is the HIP library's coordinate form.
template <typename T>
T collision_time(bool is_ee, bool relative, const V3<T> p0[4], const V3<T> p1[4], int* capped) {
  T coef[4], roots[3];
  V3<T> u[3], v[3];
  if (relative) coplanarity_cubic_relative(p0, p1, coef, u, v);
  else coplanarity_cubic_absolute(p0, p1, coef);
  int n = cubic_roots(coef, T(0), T(1), roots, capped);
  for (int i = 0; i < n; ++i) {
    T t = roots[i];
    V3<T> q[4];
    if (relative) {
      q[0] = mk(T(0), T(0), T(0));
      for (int k = 0; k < 3; ++k) q[k + 1] = u[k] + v[k] * t;
    } else {
      for (int k = 0; k < 4; ++k) q[k] = lerp(p0[k], p1[k], t);
    }
    bool ok = is_ee ? verify_root_ee(q[0], q[1], q[2], q[3]) : verify_root_pt(q[0], q[1], q[2], q[3]);
    if (ok) return t;
  }
  return T(1);
}


// ccd/ccd.py:7-163: the distance-based conservative advancement alternates
// (restated from the reference's text; T = float reproduces its fp32 default,
// T = double the library's FP64 evaluation).
template <typename T>
T ca_clamp01(T v) { return v < T(0) ? T(0) : (v > T(1) ? T(1) : v); }

template <typename T>
T ca_segment_distance_sq(V3<T> a0, V3<T> a1, V3<T> b0, V3<T> b1) {
  const V3<T> da = a1 - a0, db = b1 - b0, r = a0 - b0;
  const T aa = dot(da, da), bb = dot(db, db), f = dot(db, r);
  T s = T(0), t = T(0);
  if (aa <= T(1e-20) && bb <= T(1e-20)) return dot(r, r);
  if (aa <= T(1e-20)) {
    t = ca_clamp01(f / bb);
  } else if (bb <= T(1e-20)) {
    s = ca_clamp01(-dot(da, r) / aa);
  } else {
    const T ab = dot(da, db), c = dot(da, r);
    const T denom = aa * bb - ab * ab;
    if (denom > T(1e-20)) s = ca_clamp01((ab * f - c * bb) / denom);
    t = (ab * s + f) / bb;
    if (t < T(0)) {
      t = T(0);
      s = ca_clamp01(-c / aa);
    } else if (t > T(1)) {
      t = T(1);
      s = ca_clamp01((ab - c) / aa);
    }
  }
  const V3<T> d = (a0 + s * da) - (b0 + t * db);
  return dot(d, d);
}

template <typename T>
T ca_triangle_distance_sq(V3<T> p, V3<T> a, V3<T> b, V3<T> c) {
  const V3<T> ab = b - a, ac = c - a, ap = p - a;
  const T d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= T(0) && d2 <= T(0)) return dot(ap, ap);
  const V3<T> bp = p - b;
  const T d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= T(0) && d4 <= d3) return dot(bp, bp);
  const T vc = d1 * d4 - d3 * d2;
  if (vc <= T(0) && d1 >= T(0) && d3 <= T(0)) {
    const T v = d1 / (d1 - d3);
    const V3<T> e = p - (a + v * ab);
    return dot(e, e);
  }
  const V3<T> cp = p - c;
  const T d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= T(0) && d5 <= d6) return dot(cp, cp);
  const T vb = d5 * d2 - d1 * d6;
  if (vb <= T(0) && d2 >= T(0) && d6 <= T(0)) {
    const T w = d2 / (d2 - d6);
    const V3<T> e = p - (a + w * ac);
    return dot(e, e);
  }
  const T va = d3 * d6 - d5 * d4;
  if (va <= T(0) && d4 >= d3 && d5 >= d6) {
    const T w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    const V3<T> e = p - (b + w * (c - b));
    return dot(e, e);
  }
  const T denom = va + vb + vc;
  if (std::fabs(denom) <= T(1e-20)) {
    const T dab = ca_segment_distance_sq(p, p, a, b), dac = ca_segment_distance_sq(p, p, a, c),
            dbc = ca_segment_distance_sq(p, p, b, c);
    return std::min(dab, std::min(dac, dbc));
  }
  const T inv = T(1) / denom;
  const T v = vb * inv, w = vc * inv;
  const V3<T> e = p - (a + v * ab + w * ac);
  return dot(e, e);
}

template <typename T>
T ca_distance(bool ee, const V3<T>* x) {
  return ee ? std::sqrt(ca_segment_distance_sq(x[0], x[1], x[2], x[3]))
            : std::sqrt(ca_triangle_distance_sq(x[0], x[1], x[2], x[3]));
}

template <typename T>
T conservative_toi(bool ee, V3<T> x[4], V3<T> d[4]) {
  const V3<T> mean = (((d[0] + d[1]) + d[2]) + d[3]) * T(0.25);
  for (int k = 0; k < 4; ++k) d[k] = d[k] - mean;
  T max_disp;
  if (ee)
    max_disp = std::sqrt(std::max(dot(d[0], d[0]), dot(d[1], d[1]))) + std::sqrt(std::max(dot(d[2], d[2]), dot(d[3], d[3])));
  else
    max_disp = std::sqrt(dot(d[0], d[0])) + std::sqrt(std::max(dot(d[1], d[1]), std::max(dot(d[2], d[2]), dot(d[3], d[3]))));
  if (max_disp <= T(1e-20)) return T(1);
  T distance = ca_distance(ee, x);
  const T gap = T(0.1) * distance;
  T toi = T(0);
  for (int it = 0; it < 100; ++it) {
    const T step = T(0.9) * distance / max_disp;
    if (step <= T(1e-15)) return toi;
    for (int k = 0; k < 4; ++k) x[k] = x[k] + step * d[k];
    distance = ca_distance(ee, x);
    if (toi > T(0) && distance < gap) return toi;
    toi += step;
    if (toi >= T(1)) return T(1);
  }
  return std::min(toi, T(1));
}

}  // namespace ipco

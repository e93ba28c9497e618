// TEST INFRASTRUCTURE ONLY: C entry points of the CPU oracle (see
// ipc_oracle.hpp). Called through ctypes by tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg; never by the product.
#include "ipc_oracle.hpp"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <map>
#include <unordered_map>
#include <utility>
#include <vector>

using namespace ipco;

namespace {

struct Box { float lo[3], hi[3]; int id; };

inline bool overlap(const Box& a, const Box& b) {
  // Inclusive AABB test (the reference's padded-overlap predicate,
  // fixed_bvh_contacts.py:70-80 with padding already folded into the boxes).
  for (int k = 0; k < 3; ++k)
    if (a.lo[k] > b.hi[k] || b.lo[k] > a.hi[k]) return false;
  return true;
}

// Calls fn(a_id, b_id) once for every overlapping pair with one box from A and
// one from B (every unordered pair when self = true), found through a uniform
// hash grid; every candidate is confirmed with `overlap`, so the enumeration
// is exact and independent of the grid's parameters. Cell size is twice the
// median box extent. Boxes spanning more than kBigCells cells on an axis
// ("big") stay out of the grid: each one visits the grid cells it covers
// (or, when those exceed kMaxVisit, scans every small box), and big boxes
// are tested against each other directly. A pair found through the grid is
// reported only from the cell holding the lower corner of the two boxes'
// intersection, so each pair is reported once.
template <typename Fn>
void grid_pairs(const std::vector<Box>& A, const std::vector<Box>& B, bool self, Fn fn) {
  const std::vector<Box>& Bs = self ? A : B;
  if (A.empty() || Bs.empty()) return;
  std::vector<float> ext;
  ext.reserve(A.size() + Bs.size());
  for (const Box& b : A) ext.push_back(std::max(b.hi[0] - b.lo[0], std::max(b.hi[1] - b.lo[1], b.hi[2] - b.lo[2])));
  if (!self)
    for (const Box& b : Bs) ext.push_back(std::max(b.hi[0] - b.lo[0], std::max(b.hi[1] - b.lo[1], b.hi[2] - b.lo[2])));
  std::nth_element(ext.begin(), ext.begin() + ext.size() / 2, ext.end());
  double h = std::max(2.0 * double(ext[ext.size() / 2]), 1e-6);
  const int kBigCells = 16;
  const long long kMaxVisit = 1 << 16;
  auto cell = [h](float v) { return (long long)std::floor(double(v) / h); };
  auto is_big = [&](const Box& b) {
    for (int k = 0; k < 3; ++k)
      if (cell(b.hi[k]) - cell(b.lo[k]) > kBigCells) return true;
    return false;
  };
  auto key = [](long long i, long long j, long long k) {
    return (unsigned long long)((i & 0x1FFFFF) | ((j & 0x1FFFFF) << 21) | ((k & 0x1FFFFF) << 42));
  };
  using Grid = std::unordered_map<unsigned long long, std::vector<int>>;
  auto build = [&](const std::vector<Box>& S, const std::vector<int>& ids) {
    Grid g;
    g.reserve(ids.size() * 2);
    for (int b : ids) {
      const Box& x = S[b];
      for (long long i = cell(x.lo[0]); i <= cell(x.hi[0]); ++i)
        for (long long j = cell(x.lo[1]); j <= cell(x.hi[1]); ++j)
          for (long long k = cell(x.lo[2]); k <= cell(x.hi[2]); ++k) g[key(i, j, k)].push_back(b);
    }
    return g;
  };
  // visit(x, S, g, ids, skip, emit): every box y of S in grid g (built from
  // `ids`) overlapping x, reported once per pair, skip(y) excluded.
  auto visit = [&](const Box& x, const std::vector<Box>& S, const Grid& g, const std::vector<int>& ids, auto skip,
                   auto emit) {
    long long n = 1;
    for (int k = 0; k < 3; ++k) n *= cell(x.hi[k]) - cell(x.lo[k]) + 1;
    if (n > kMaxVisit) {
      for (int b : ids)
        if (!skip(b) && overlap(x, S[b])) emit(b);
      return;
    }
    for (long long i = cell(x.lo[0]); i <= cell(x.hi[0]); ++i)
      for (long long j = cell(x.lo[1]); j <= cell(x.hi[1]); ++j)
        for (long long k = cell(x.lo[2]); k <= cell(x.hi[2]); ++k) {
          auto it = g.find(key(i, j, k));
          if (it == g.end()) continue;
          for (int b : it->second) {
            if (skip(b)) continue;
            const Box& y = S[b];
            if (!overlap(x, y)) continue;
            if (cell(std::max(x.lo[0], y.lo[0])) != i || cell(std::max(x.lo[1], y.lo[1])) != j ||
                cell(std::max(x.lo[2], y.lo[2])) != k)
              continue;
            emit(b);
          }
        }
  };
  std::vector<int> bigA, smallA, bigB, smallB;
  for (size_t i = 0; i < A.size(); ++i) (is_big(A[i]) ? bigA : smallA).push_back(int(i));
  if (self) { bigB = bigA; smallB = smallA; }
  else for (size_t i = 0; i < Bs.size(); ++i) (is_big(Bs[i]) ? bigB : smallB).push_back(int(i));
  const Grid gridB = build(Bs, smallB);
  // small A against small B
  for (int a : smallA)
    visit(A[a], Bs, gridB, smallB, [&](int b) { return self && b <= a; }, [&](int b) { fn(A[a].id, Bs[b].id); });
  // big A against small B (self: this also covers small-big)
  for (int a : bigA)
    visit(A[a], Bs, gridB, smallB, [&](int) { return false; }, [&](int b) { fn(A[a].id, Bs[b].id); });
  // big A against big B
  for (int a : bigA)
    for (int b : bigB) {
      if (self && b <= a) continue;
      if (overlap(A[a], Bs[b])) fn(A[a].id, Bs[b].id);
    }
  // small A against big B
  if (!self && !bigB.empty()) {
    const Grid gridA = build(A, smallA);
    for (int b : bigB)
      visit(Bs[b], A, gridA, smallA, [&](int) { return false; }, [&](int a) { fn(A[a].id, Bs[b].id); });
  }
}

inline V3<float> ld(const float* x, int i) { return load3<float>(x, i); }

inline Box box_of(const V3<float>* pts, int n, float pad, int id) {
  Box b;
  V3<float> lo = pts[0], hi = pts[0];
  for (int k = 1; k < n; ++k) { lo = vmin(lo, pts[k]); hi = vmax(hi, pts[k]); }
  b.lo[0] = lo.x - pad; b.lo[1] = lo.y - pad; b.lo[2] = lo.z - pad;
  b.hi[0] = hi.x + pad; b.hi[1] = hi.y + pad; b.hi[2] = hi.z + pad;
  b.id = id;
  return b;
}

struct Out {
  int* rec; int* ee; int64_t cap; int64_t n;
  void push(int a, int b, int c, int d, int e0, int e1) {
    if (n < cap) {
      rec[4 * n] = a; rec[4 * n + 1] = b; rec[4 * n + 2] = c; rec[4 * n + 3] = d;
      if (ee) { ee[2 * n] = e0; ee[2 * n + 1] = e1; }
    }
    ++n;
  }
};

// get_contact_points / get_contact_points_pt (contact.py:148-161) over
// fetch_dist_v0v1 / fetch_dist_v0v1_pt (contact.py:88-145), in precision
// T: midpoint (v0 + v1) / 2 of the closest points (wp.lerp(a, b, t) =
// a (1 - t) + b t; PT v1 = alpha t0 + beta t1 + (1 - alpha - beta) t2 over
// slots 1..3), the distance, and the feature code the dispatch picks (PT:
// closest_point_triangle's type 0..6; EE: kind * 16 + slot0 * 4 + slot1,
// the layout of ipc_contact_points_ee; xcs NULL disables the near-parallel
// fallback).
template <typename T>
static void contact_points_t(int is_ee, const float* x, const float* xcs, const int* contacts, int n, double* points,
                             double* dists, int* features) {
  for (int i = 0; i < n; ++i) {
    const int* c = contacts + 4 * i;
    V3<T> p[4];
    for (int k = 0; k < 4; ++k) p[k] = load3<T>(x, c[k]);
    V3<T> v0, v1;
    T dist;
    int feat;
    if (is_ee) {
      EdgeClosest<T> ec = closest_point_edge_edge(p[0], p[1], p[2], p[3], T(1.0e-6));
      v0 = p[0] * (T(1) - ec.s) + p[1] * ec.s;
      v1 = p[2] * (T(1) - ec.t) + p[3] * ec.t;
      dist = ec.dist;
      T eps_x = xcs ? ee_mollifier_threshold(load3<T>(xcs, c[0]), load3<T>(xcs, c[1]), load3<T>(xcs, c[2]),
                                             load3<T>(xcs, c[3]))
                    : T(0);
      Feature f = ee_feature(p[0], p[1], p[2], p[3], eps_x);
      feat = f.kind * 16 + f.slot[0] * 4 + (f.slot[1] < 0 ? 0 : f.slot[1]);
    } else {
      TriClosest<T> tc = closest_point_triangle(p[0], p[1], p[2], p[3]);
      v0 = p[0];
      v1 = tc.wa * p[1] + tc.wb * p[2] + (T(1) - tc.wa - tc.wb) * p[3];
      dist = tc.dist;
      feat = tc.type;
    }
    V3<T> mid = (v0 + v1) * T(0.5);
    if (points) { points[3 * i] = mid.x; points[3 * i + 1] = mid.y; points[3 * i + 2] = mid.z; }
    if (dists) dists[i] = dist;
    if (features) features[i] = feat;
  }
}

}  // namespace

extern "C" {

// ------------------------------------------------------------ primitives ----
void oracle_closest_point_triangle(int use_double, int n, const double* p, const double* a, const double* b,
                                   const double* c, double* out, int* type) {
  for (int i = 0; i < n; ++i) {
    if (use_double) {
      auto r = closest_point_triangle(mk(p[3 * i], p[3 * i + 1], p[3 * i + 2]), mk(a[3 * i], a[3 * i + 1], a[3 * i + 2]),
                                      mk(b[3 * i], b[3 * i + 1], b[3 * i + 2]), mk(c[3 * i], c[3 * i + 1], c[3 * i + 2]));
      out[3 * i] = r.wa; out[3 * i + 1] = r.wb; out[3 * i + 2] = r.dist; type[i] = r.type;
    } else {
      auto f = [](const double* v, int i) { return mk(float(v[3 * i]), float(v[3 * i + 1]), float(v[3 * i + 2])); };
      auto r = closest_point_triangle(f(p, i), f(a, i), f(b, i), f(c, i));
      out[3 * i] = r.wa; out[3 * i + 1] = r.wb; out[3 * i + 2] = r.dist; type[i] = r.type;
    }
  }
}

void oracle_closest_point_edge_edge(int use_double, int n, const double* p1, const double* q1, const double* p2,
                                    const double* q2, double eps, double* out) {
  for (int i = 0; i < n; ++i) {
    if (use_double) {
      auto g = [](const double* v, int i) { return mk(v[3 * i], v[3 * i + 1], v[3 * i + 2]); };
      auto r = closest_point_edge_edge(g(p1, i), g(q1, i), g(p2, i), g(q2, i), eps);
      out[3 * i] = r.s; out[3 * i + 1] = r.t; out[3 * i + 2] = r.dist;
    } else {
      auto f = [](const double* v, int i) { return mk(float(v[3 * i]), float(v[3 * i + 1]), float(v[3 * i + 2])); };
      auto r = closest_point_edge_edge(f(p1, i), f(q1, i), f(p2, i), f(q2, i), float(eps));
      out[3 * i] = r.s; out[3 * i + 1] = r.t; out[3 * i + 2] = r.dist;
    }
  }
}

void oracle_barrier(int n, const double* x, double D, double kappa, double* out) {
  for (int i = 0; i < n; ++i) {
    Barrier b = barrier(x[i], D, kappa);
    out[3 * i] = b.b; out[3 * i + 1] = b.db; out[3 * i + 2] = b.d2b;
  }
}

// kind: 0 PP, 1 PE, 2 PT, 3 EE; nsd: 0 full Hessian of d^2, 1 its NSD part
// (PP: zero). Returns the feature's squared distance.
double oracle_feature_hessian(int kind, const int* slot, const double* x12, int nsd, double* g, double* H) {
  Feature f{kind, {slot[0], slot[1], slot[2], slot[3]}};
  V3<double> x[4];
  for (int k = 0; k < 4; ++k) x[k] = mk(x12[3 * k], x12[3 * k + 1], x12[3 * k + 2]);
  return nsd ? feature_grad_nsd(f, x, g, H) : feature_hessian(f, x, g, H);
}

// Feature chosen by the EE / PT dispatch, in float (prec=0) or double.
// Writes kind and slots; returns nothing.
void oracle_feature_ee(int use_double, const double* x12, double parallel_eps, int* kind, int* slot) {
  Feature f;
  if (use_double) {
    auto g = [&](int k) { return mk(x12[3 * k], x12[3 * k + 1], x12[3 * k + 2]); };
    f = ee_feature(g(0), g(1), g(2), g(3), parallel_eps);
  } else {
    auto g = [&](int k) { return mk(float(x12[3 * k]), float(x12[3 * k + 1]), float(x12[3 * k + 2])); };
    f = ee_feature(g(0), g(1), g(2), g(3), float(parallel_eps));
  }
  *kind = f.kind;
  for (int k = 0; k < 4; ++k) slot[k] = f.slot[k];
}

// Mollifier value, gradient and PSD Hessian with the branch decided in double.
double oracle_ee_mollifier(const double* x12, double eps_x, double* g, double* H) {
  V3<double> x[4];
  for (int k = 0; k < 4; ++k) x[k] = mk(x12[3 * k], x12[3 * k + 1], x12[3 * k + 2]);
  V3<double> cr = cross(x[1] - x[0], x[3] - x[2]);
  bool active = eps_x > 0 && dot(cr, cr) < eps_x;
  Mollifier m = ee_mollifier(x, eps_x, active);
  std::memcpy(g, m.g, sizeof(m.g));
  std::memcpy(H, m.H, sizeof(m.H));
  return m.m;
}

// --------------------------------------------------------- contact kernels ----
// contact_hessian_ee / contact_hessian_pt (dynamic_contacts.py:151-246).
// Writes the 16 triplets of contact i at (offset + i) * 16 + ii * 4 + jj,
// accumulates the force into b (double, V x 3) and sets active[i].
// prec: 0 = classify in float (the reference's FP32 mode), 1 = double.
void oracle_contact_hessian(int is_ee, int prec, int d2_policy, const float* x, const float* xcs, const int* contacts,
                            int n, float l0, double kappa, int offset, int* rows, int* cols, double* vals, double* b,
                            unsigned char* active, int threads) {
  std::vector<PairResult> res(n > 0 ? n : 0);
#pragma omp parallel for num_threads(threads > 0 ? threads : 1) schedule(dynamic, 64)
  for (int i = 0; i < n; ++i) {
    const int* c = contacts + 4 * i;
    if (is_ee) res[i] = prec ? contact_ee<double>(x, xcs, c, l0, kappa, d2_policy) : contact_ee<float>(x, xcs, c, l0, kappa, d2_policy);
    else res[i] = prec ? contact_pt<double>(x, c, l0, kappa, d2_policy) : contact_pt<float>(x, c, l0, kappa, d2_policy);
  }
  for (int i = 0; i < n; ++i) {
    const int* c = contacts + 4 * i;
    const PairResult& r = res[i];
    if (active) active[i] = r.active ? 1 : 0;
    if (!r.active) continue;
    for (int ii = 0; ii < 4; ++ii) {
      for (int k = 0; k < 3; ++k) b[3 * c[ii] + k] += r.grad_force[3 * ii + k];
      for (int jj = 0; jj < 4; ++jj) {
        int64_t idx = int64_t(offset + i) * 16 + ii * 4 + jj;
        if (rows) rows[idx] = c[ii];
        if (cols) cols[idx] = c[jj];
        if (vals)
          for (int r3 = 0; r3 < 3; ++r3)
            for (int c3 = 0; c3 < 3; ++c3) vals[idx * 9 + r3 * 3 + c3] = r.H[(ii * 3 + r3) * 12 + jj * 3 + c3];
      }
    }
  }
}

// The same with fp64 scalars (scalar_types.py with USE_FLOAT32 = False):
// double positions, classification and barrier argument in double.
void oracle_contact_hessian_f64(int is_ee, const double* x, const double* xcs, const int* contacts, int n, double l0,
                                double kappa, int offset, int* rows, int* cols, double* vals, double* b,
                                unsigned char* active, int threads) {
  std::vector<PairResult> res(n > 0 ? n : 0);
#pragma omp parallel for num_threads(threads > 0 ? threads : 1) schedule(dynamic, 64)
  for (int i = 0; i < n; ++i) {
    const int* c = contacts + 4 * i;
    res[i] = is_ee ? contact_ee<double>(x, xcs, c, l0, kappa, 1) : contact_pt<double>(x, c, l0, kappa, 1);
  }
  for (int i = 0; i < n; ++i) {
    const int* c = contacts + 4 * i;
    const PairResult& r = res[i];
    if (active) active[i] = r.active ? 1 : 0;
    if (!r.active) continue;
    for (int ii = 0; ii < 4; ++ii) {
      for (int k = 0; k < 3; ++k) b[3 * c[ii] + k] += r.grad_force[3 * ii + k];
      for (int jj = 0; jj < 4; ++jj) {
        int64_t idx = int64_t(offset + i) * 16 + ii * 4 + jj;
        rows[idx] = c[ii];
        cols[idx] = c[jj];
        for (int r3 = 0; r3 < 3; ++r3)
          for (int c3 = 0; c3 < 3; ++c3) vals[idx * 9 + r3 * 3 + c3] = r.H[(ii * 3 + r3) * 12 + jj * 3 + c3];
      }
    }
  }
}

double oracle_contact_energy_f64(int is_ee, const double* x, const double* xcs, const int* contacts, int n, double l0,
                                 double kappa) {
  double s = 0.0;
  for (int i = 0; i < n; ++i) {
    const int* c = contacts + 4 * i;
    const PairResult r = is_ee ? contact_ee<double>(x, xcs, c, l0, kappa, 1) : contact_pt<double>(x, c, l0, kappa, 1);
    if (r.active) s += r.energy;
  }
  return s;
}

// One contact at double positions x12 (rest12 for the EE mollifier
// threshold): energy (returned), force (-gradient, 12), 12x12 Hessian and
// the dispatched feature (kind, 4 slots; kind -1 when inactive). prec and
// d2_policy as in oracle_contact_hessian. Lets tests differentiate the
// pair's energy in double.
double oracle_contact_pair(int is_ee, int prec, int d2_policy, const double* x12, const double* rest12, float l0,
                           double kappa, double* force, double* H, int* feature) {
  V3<double> xd[4], rd[4];
  for (int k = 0; k < 4; ++k) {
    xd[k] = mk(x12[3 * k], x12[3 * k + 1], x12[3 * k + 2]);
    rd[k] = rest12 ? mk(rest12[3 * k], rest12[3 * k + 1], rest12[3 * k + 2]) : xd[k];
  }
  PairResult r;
  if (is_ee) r = prec ? contact_ee_at<double>(xd, rd, l0, kappa, d2_policy) : contact_ee_at<float>(xd, rd, l0, kappa, d2_policy);
  else r = prec ? contact_pt_at<double>(xd, l0, kappa, d2_policy) : contact_pt_at<float>(xd, l0, kappa, d2_policy);
  if (force) std::memcpy(force, r.grad_force, sizeof(r.grad_force));
  if (H) std::memcpy(H, r.H, sizeof(r.H));
  if (feature) {
    feature[0] = r.active ? r.feature.kind : -1;
    for (int k = 0; k < 4; ++k) feature[1 + k] = r.active ? r.feature.slot[k] : -1;
  }
  return r.active ? r.energy : 0.0;
}

// Midpoints, distances and feature codes of a contact list (contact_points_t).
void oracle_contact_points(int is_ee, int use_double, const float* x, const float* xcs, const int* contacts, int n,
                           double* points, double* dists, int* features) {
  if (use_double) contact_points_t<double>(is_ee, x, xcs, contacts, n, points, dists, features);
  else contact_points_t<float>(is_ee, x, xcs, contacts, n, points, dists, features);
}

// contact_energy_ee / contact_energy_pt (dynamic_contacts.py:374-398):
// sum over active contacts of m * b (EE) or b (PT), accumulated in double
// with Kahan summation in contact order.
double oracle_contact_energy(int is_ee, int prec, int d2_policy, const float* x, const float* xcs, const int* contacts,
                             int n, float l0, double kappa, int threads) {
  std::vector<double> e(n > 0 ? n : 0, 0.0);
#pragma omp parallel for num_threads(threads > 0 ? threads : 1) schedule(dynamic, 64)
  for (int i = 0; i < n; ++i) {
    const int* c = contacts + 4 * i;
    PairResult r;
    if (is_ee) r = prec ? contact_ee<double>(x, xcs, c, l0, kappa, d2_policy) : contact_ee<float>(x, xcs, c, l0, kappa, d2_policy);
    else r = prec ? contact_pt<double>(x, c, l0, kappa, d2_policy) : contact_pt<float>(x, c, l0, kappa, d2_policy);
    e[i] = r.active ? r.energy : 0.0;
  }
  double s = 0.0, comp = 0.0;
  for (int i = 0; i < n; ++i) {
    double y = e[i] - comp, t = s + y;
    comp = (t - s) - y;
    s = t;
  }
  return s;
}

// ------------------------------------------------------------------- CCD ----
int oracle_cubic_roots(int use_double, const double* coef, double x0, double x1, double* roots, int* capped) {
  if (use_double) return cubic_roots(coef, x0, x1, roots, capped);
  float cf[4] = {float(coef[0]), float(coef[1]), float(coef[2]), float(coef[3])}, rf[3];
  int n = cubic_roots(cf, float(x0), float(x1), rf, capped);
  for (int k = 0; k < 3; ++k) roots[k] = rf[k];
  return n;
}

void oracle_coplanarity_cubic(int relative, const double* p0_12, const double* p1_12, double* coef) {
  V3<double> a[4], b[4], u[3], v[3];
  for (int k = 0; k < 4; ++k) { a[k] = mk(p0_12[3 * k], p0_12[3 * k + 1], p0_12[3 * k + 2]); b[k] = mk(p1_12[3 * k], p1_12[3 * k + 1], p1_12[3 * k + 2]); }
  if (relative) coplanarity_cubic_relative(a, b, coef, u, v);
  else coplanarity_cubic_absolute(a, b, coef);
}

// pt_collision_time / ee_collision_time over a candidate list of vertex
// quadruples (PT: point, t0, t1, t2; EE: a0, a1, b0, b1). Returns the number
// of find_closed iteration caps hit.
int oracle_collision_time(int is_ee, int relative, int use_double, int n, const float* x0, const float* x1,
                          const int* quads, double* toi, int threads) {
  int capped_total = 0;
#pragma omp parallel for num_threads(threads > 0 ? threads : 1) schedule(dynamic, 64) reduction(+ : capped_total)
  for (int i = 0; i < n; ++i) {
    const int* q = quads + 4 * i;
    int capped = 0;
    if (use_double) {
      V3<double> a[4], b[4];
      for (int k = 0; k < 4; ++k) { a[k] = load3<double>(x0, q[k]); b[k] = load3<double>(x1, q[k]); }
      toi[i] = collision_time<double>(is_ee != 0, relative != 0, a, b, &capped);
    } else {
      V3<float> a[4], b[4];
      for (int k = 0; k < 4; ++k) { a[k] = load3<float>(x0, q[k]); b[k] = load3<float>(x1, q[k]); }
      toi[i] = collision_time<float>(is_ee != 0, relative != 0, a, b, &capped);
    }
    capped_total += capped;
  }
  return capped_total;
}

// The same with fp64 scalars: double positions (vec3d), FP64 arithmetic.
int oracle_collision_time_f64(int is_ee, int relative, int n, const double* x0, const double* x1, const int* quads,
                              double* toi, int threads) {
  int capped_total = 0;
#pragma omp parallel for num_threads(threads > 0 ? threads : 1) schedule(dynamic, 64) reduction(+ : capped_total)
  for (int i = 0; i < n; ++i) {
    const int* q = quads + 4 * i;
    int capped = 0;
    V3<double> a[4], b[4];
    for (int k = 0; k < 4; ++k) { a[k] = load3<double>(x0, q[k]); b[k] = load3<double>(x1, q[k]); }
    toi[i] = collision_time<double>(is_ee != 0, relative != 0, a, b, &capped);
    capped_total += capped;
  }
  return capped_total;
}

// ------------------------------------------------------------------- DCD ----
// edge_edge_collision (contact.py:199-229) with the BVH replaced by an exact
// overlap enumeration of the same inflated boxes (edge_aabb, contact.py:30-37).
// Records (edges[i], edges[j]) with i < j and e0e1 = (i, j). Returns the
// number of contacts (may exceed cap; only cap are written).
int64_t oracle_dcd_ee(const float* x, const int* edges, int n_edges, const int* body, float thickness, float buffer,
                      int exclude_same_body, int* rec, int* e0e1, int64_t cap) {
  float pad = thickness + buffer;
  std::vector<Box> boxes(n_edges);
  for (int i = 0; i < n_edges; ++i) {
    V3<float> p[2] = {ld(x, edges[2 * i]), ld(x, edges[2 * i + 1])};
    boxes[i] = box_of(p, 2, pad, i);
  }
  std::vector<std::pair<int, int>> hits;
  grid_pairs(boxes, {}, true, [&](int a, int b) { hits.emplace_back(std::min(a, b), std::max(a, b)); });
  std::sort(hits.begin(), hits.end());
  Out out{rec, e0e1, cap, 0};
  for (auto& h : hits) {
    int i = h.first, j = h.second;
    int a1 = edges[2 * i], a2 = edges[2 * i + 1], b1 = edges[2 * j], b2 = edges[2 * j + 1];
    bool connected = a1 == b1 || a1 == b2 || a2 == b1 || a2 == b2;
    bool same = exclude_same_body && body[a1] == body[b1];
    if (connected || same) continue;
    auto r = closest_point_edge_edge(ld(x, a1), ld(x, a2), ld(x, b1), ld(x, b2), 1.0e-6f);
    if (r.dist < thickness * 2.0f) out.push(a1, a2, b1, b2, i, j);
  }
  return out.n;
}

// point_triangle_collision (contact.py:304-336): every node against every
// triangle whose bounds meet the node's box +-2(thickness+buffer). Records
// (p, tri[1], tri[0], tri[2]) with e0e1 = (p, triangle), ordered by (p, triangle).
int64_t oracle_dcd_pt(const float* x, int n_nodes, const int* tris, int n_tris, const int* body, float thickness,
                      float buffer, int exclude_same_body, int* rec, int* e0e1, int64_t cap) {
  float pad = (thickness + buffer) * 2.0f;
  std::vector<Box> pb(n_nodes), tb(n_tris);
  for (int i = 0; i < n_nodes; ++i) { V3<float> p = ld(x, i); pb[i] = box_of(&p, 1, pad, i); }
  for (int j = 0; j < n_tris; ++j) {
    V3<float> p[3] = {ld(x, tris[3 * j]), ld(x, tris[3 * j + 1]), ld(x, tris[3 * j + 2])};
    tb[j] = box_of(p, 3, 0.0f, j);
  }
  std::vector<std::pair<int, int>> hits;
  grid_pairs(pb, tb, false, [&](int a, int b) { hits.emplace_back(a, b); });
  std::sort(hits.begin(), hits.end());
  Out out{rec, e0e1, cap, 0};
  for (auto& h : hits) {
    int i = h.first, j = h.second;
    int t1 = tris[3 * j], t2 = tris[3 * j + 1], t3 = tris[3 * j + 2];
    if (t1 == i || t2 == i || t3 == i) continue;
    if (exclude_same_body && body[i] == body[t1]) continue;
    auto r = closest_point_triangle(ld(x, i), ld(x, t1), ld(x, t2), ld(x, t3));
    if (r.dist < thickness * 2.0f) out.push(i, t2, t1, t3, i, j);
  }
  return out.n;
}

// ------------------------------------------------------ CCD candidate sets ----
// edge_edge_toi's candidates (ccd/ccd.py:166-186,397-427): swept edge boxes
// padded by `padding`, i < j, not connected, not same body. Records
// (a0, a1, b0, b1) ordered by (i, j).
int64_t oracle_ccd_candidates_ee(const float* x0, const float* x1, const int* edges, int n_edges, const int* body,
                                 float padding, int exclude_same_body, int* rec, int* e0e1, int64_t cap) {
  std::vector<Box> boxes(n_edges);
  for (int i = 0; i < n_edges; ++i) {
    int a = edges[2 * i], b = edges[2 * i + 1];
    V3<float> p[4] = {ld(x0, a), ld(x0, b), ld(x1, a), ld(x1, b)};
    boxes[i] = box_of(p, 4, padding, i);
  }
  std::vector<std::pair<int, int>> hits;
  grid_pairs(boxes, {}, true, [&](int a, int b) { hits.emplace_back(std::min(a, b), std::max(a, b)); });
  std::sort(hits.begin(), hits.end());
  Out out{rec, e0e1, cap, 0};
  for (auto& h : hits) {
    int i = h.first, j = h.second;
    int a0 = edges[2 * i], a1 = edges[2 * i + 1], b0 = edges[2 * j], b1 = edges[2 * j + 1];
    if (a0 == b0 || a0 == b1 || a1 == b0 || a1 == b1) continue;
    if (exclude_same_body && body[a0] == body[b0]) continue;
    out.push(a0, a1, b0, b1, i, j);
  }
  return out.n;
}

// point_triangle_toi's candidates (ccd/ccd.py:188-210,360-394): swept point
// box +-padding against padded swept triangle boxes, not incident, not same
// body. Records (p, t0, t1, t2) ordered by (p, triangle).
int64_t oracle_ccd_candidates_pt(const float* x0, const float* x1, int n_nodes, const int* tris, int n_tris,
                                 const int* body, float padding, int exclude_same_body, int* rec, int* e0e1,
                                 int64_t cap) {
  std::vector<Box> pb(n_nodes), tb(n_tris);
  for (int i = 0; i < n_nodes; ++i) { V3<float> p[2] = {ld(x0, i), ld(x1, i)}; pb[i] = box_of(p, 2, padding, i); }
  for (int j = 0; j < n_tris; ++j) {
    V3<float> p[6];
    for (int k = 0; k < 3; ++k) { p[k] = ld(x0, tris[3 * j + k]); p[3 + k] = ld(x1, tris[3 * j + k]); }
    tb[j] = box_of(p, 6, padding, j);
  }
  std::vector<std::pair<int, int>> hits;
  grid_pairs(pb, tb, false, [&](int a, int b) { hits.emplace_back(a, b); });
  std::sort(hits.begin(), hits.end());
  Out out{rec, e0e1, cap, 0};
  for (auto& h : hits) {
    int i = h.first, j = h.second;
    int t0 = tris[3 * j], t1 = tris[3 * j + 1], t2 = tris[3 * j + 2];
    if (t0 == i || t1 == i || t2 == i) continue;
    if (exclude_same_body && body[i] == body[t0]) continue;
    out.push(i, t0, t1, t2, i, j);
  }
  return out.n;
}

// Distance of point o from the plane of (v0, v0 + e1, v0 + e2), scaled by
// |e1 x e2|, is certainly nonzero: |n . (o - v0)| above the forward error
// bound of its double evaluation, 8 * 2^-53 times the sum of the products'
// magnitudes (Shewchuk's orient3d filter, rounded up).
static bool off_plane(const V3<double>& n, const V3<double>& e1, const V3<double>& e2, const V3<double>& r) {
  double perm = std::fabs(r.x) * (std::fabs(e1.y * e2.z) + std::fabs(e1.z * e2.y)) +
                std::fabs(r.y) * (std::fabs(e1.z * e2.x) + std::fabs(e1.x * e2.z)) +
                std::fabs(r.z) * (std::fabs(e1.x * e2.y) + std::fabs(e1.y * e2.x));
  return std::fabs(dot(n, r)) > 8.0 * 0x1p-53 * perm;
}

// Edge-triangle crossings (layout of ipc_count_intersections; no reference
// counterpart). Written independently of the GPU's signed-volume test:
// Moller-Trumbore ray-triangle intersection in double, the segment
// parameter in [0, 1] and the barycentrics in the closed triangle; a segment
// parallel to the triangle's plane (|det| <= 1e-300), one whose end points
// both lie in the plane within rounding (coplanar: contact, not a
// crossing), or one sharing a vertex does not count. Pairs come from the
// hash-grid box enumeration. Writes the (edge, triangle) ids, sorted, up to
// cap; returns the count.
int64_t oracle_count_intersections(const float* x, const int* edges, int n_edges, const int* tris, int n_tris,
                                   int* ids, int64_t cap) {
  std::vector<Box> eb(n_edges), tb(n_tris);
  for (int i = 0; i < n_edges; ++i) {
    V3<float> p[2] = {ld(x, edges[2 * i]), ld(x, edges[2 * i + 1])};
    eb[i] = box_of(p, 2, 0.0f, i);
  }
  for (int j = 0; j < n_tris; ++j) {
    V3<float> p[3] = {ld(x, tris[3 * j]), ld(x, tris[3 * j + 1]), ld(x, tris[3 * j + 2])};
    tb[j] = box_of(p, 3, 0.0f, j);
  }
  std::vector<std::pair<int, int>> hits;
  grid_pairs(eb, tb, false, [&](int a, int b) { hits.emplace_back(a, b); });
  std::sort(hits.begin(), hits.end());
  int64_t n = 0;
  for (auto& h : hits) {
    int i = h.first, j = h.second;
    int a0 = edges[2 * i], a1 = edges[2 * i + 1];
    int t0 = tris[3 * j], t1 = tris[3 * j + 1], t2 = tris[3 * j + 2];
    if (a0 == t0 || a0 == t1 || a0 == t2 || a1 == t0 || a1 == t1 || a1 == t2) continue;
    V3<double> o = load3<double>(x, a0), d = load3<double>(x, a1) - o;
    V3<double> v0 = load3<double>(x, t0), e1 = load3<double>(x, t1) - v0, e2 = load3<double>(x, t2) - v0;
    V3<double> nrm = cross(e1, e2);
    if (!off_plane(nrm, e1, e2, o - v0) && !off_plane(nrm, e1, e2, o + d - v0)) continue;
    V3<double> pv = cross(d, e2);
    double det = dot(e1, pv);
    if (std::fabs(det) <= 1.0e-300) continue;
    double inv = 1.0 / det;
    V3<double> tv = o - v0;
    double u = dot(tv, pv) * inv;
    if (u < 0.0 || u > 1.0) continue;
    V3<double> qv = cross(tv, e1);
    double v = dot(d, qv) * inv;
    if (v < 0.0 || u + v > 1.0) continue;
    double t = dot(e2, qv) * inv;
    if (t < 0.0 || t > 1.0) continue;
    if (n < cap) {
      ids[2 * n] = i;
      ids[2 * n + 1] = j;
    }
    ++n;
  }
  return n;
}

// Warm-start features (layout of ipc_warmstart_features; no reference
// counterpart). Written independently of the GPU's atomic insertion: every
// (record, vertex) pair is listed, the list is sorted by (vertex, fp32
// distance, record id) and the first 6 entries of each vertex are kept.
void oracle_warmstart_features(const float* x, int nv, const float* base, const int* ee, int n_ee, const int* pt,
                               int n_pt, float dhat, float* out) {
  const float inv = 1.0f / dhat;
  struct Entry { int v; float d; uint32_t rid; };
  std::vector<Entry> L;
  auto ok = [&](const int* c) {
    for (int k = 0; k < 4; ++k)
      if (c[k] < 0 || c[k] >= nv) return false;
    return true;
  };
  auto dist_of = [&](bool is_ee, const int* c) {
    V3<float> a = load3<float>(x, c[0]), b = load3<float>(x, c[1]), cc = load3<float>(x, c[2]), d = load3<float>(x, c[3]);
    return is_ee ? closest_point_edge_edge<float>(a, b, cc, d, 1.0e-6f).dist : closest_point_triangle<float>(a, b, cc, d).dist;
  };
  for (int pass = 0; pass < 2; ++pass) {
    bool is_ee = pass == 0;
    const int* C = is_ee ? ee : pt;
    int n = is_ee ? n_ee : n_pt;
    uint32_t base_id = is_ee ? 0u : 4u * uint32_t(n_ee);
    for (int i = 0; i < n; ++i) {
      const int* c = C + 4 * i;
      if (!ok(c)) continue;
      float d = dist_of(is_ee, c);
      if (!(d >= 0.0f)) continue;
      for (int k = 0; k < 4; ++k) L.push_back({c[k], d, base_id + 4u * uint32_t(i) + uint32_t(k)});
    }
  }
  std::sort(L.begin(), L.end(), [](const Entry& a, const Entry& b) {
    if (a.v != b.v) return a.v < b.v;
    if (a.d != b.d) return a.d < b.d;
    return a.rid < b.rid;
  });
  for (int v = 0; v < nv; ++v)
    for (int q = 0; q < 48; ++q) out[48ll * v + q] = q < 12 && base ? base[12ll * v + q] : 0.0f;
  size_t i = 0;
  while (i < L.size()) {
    int v = L[i].v;
    size_t j = i;
    while (j < L.size() && L[j].v == v) ++j;
    for (size_t e = i; e < j && e < i + 6; ++e) {
      uint32_t r = L[e].rid;
      bool is_ee = r < 4u * uint32_t(n_ee);
      uint32_t rr = is_ee ? r : r - 4u * uint32_t(n_ee);
      int k = int(rr & 3u);
      const int* c = (is_ee ? ee : pt) + 4 * (rr >> 2);
      V3<float> x0 = load3<float>(x, c[0]), x1 = load3<float>(x, c[1]), x2 = load3<float>(x, c[2]),
                x3 = load3<float>(x, c[3]);
      V3<float> self, partner;
      float w, dist;
      if (is_ee) {
        EdgeClosest<float> ec = closest_point_edge_edge<float>(x0, x1, x2, x3, 1.0e-6f);
        V3<float> pa = x0 * (1.0f - ec.s) + x1 * ec.s;
        V3<float> pb = x2 * (1.0f - ec.t) + x3 * ec.t;
        self = k < 2 ? pa : pb;
        partner = k < 2 ? pb : pa;
        const float ws[4] = {1.0f - ec.s, ec.s, 1.0f - ec.t, ec.t};
        w = ws[k];
        dist = ec.dist;
      } else {
        TriClosest<float> tc = closest_point_triangle<float>(x0, x1, x2, x3);
        float wc = 1.0f - tc.wa - tc.wb;
        V3<float> q = tc.wa * x1 + tc.wb * x2 + wc * x3;
        self = k == 0 ? x0 : q;
        partner = k == 0 ? q : x0;
        const float ws[4] = {1.0f, tc.wa, tc.wb, wc};
        w = ws[k];
        dist = tc.dist;
      }
      float* o = out + 48ll * v + 12 + 6 * (e - i);
      V3<float> off = partner - self;
      o[0] = off.x * inv;
      o[1] = off.y * inv;
      o[2] = off.z * inv;
      o[3] = dist * inv;
      o[4] = w;
      o[5] = is_ee ? 0.0f : 1.0f;
    }
    i = j;
  }
}

// warp.sparse.bsr_from_triplets as called at dynamic_contacts.py:290-296
// (Warp is absent; its published contract): triplets with a row or column
// out of range, or nine zero values, are dropped; the rest are ordered by
// (row, column) keeping triplet order within a key and summed per key in
// fp32 from +0. values holds 9 * n floats (the pattern has at most n blocks).
int oracle_bsr_from_triplets(int nr, int nc, const int* rows, const int* cols, const float* vals, int n, int* offsets,
                             int* columns, float* values) {
  std::vector<int> keep;
  keep.reserve(n);
  for (int t = 0; t < n; ++t) {
    if (rows[t] < 0 || rows[t] >= nr || cols[t] < 0 || cols[t] >= nc) continue;
    bool nz = false;
    for (int k = 0; k < 9; ++k) nz |= !(vals[9ll * t + k] == 0.0f);
    if (nz) keep.push_back(t);
  }
  std::stable_sort(keep.begin(), keep.end(), [&](int a, int b) {
    return rows[a] != rows[b] ? rows[a] < rows[b] : cols[a] < cols[b];
  });
  for (int r = 0; r <= nr; ++r) offsets[r] = 0;
  int nnz = 0;
  for (size_t i = 0; i < keep.size();) {
    const int r = rows[keep[i]], c = cols[keep[i]];
    float acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    size_t j = i;
    for (; j < keep.size() && rows[keep[j]] == r && cols[keep[j]] == c; ++j)
      for (int k = 0; k < 9; ++k) acc[k] += vals[9ll * keep[j] + k];
    columns[nnz] = c;
    for (int k = 0; k < 9; ++k) values[9ll * nnz + k] = acc[k];
    ++offsets[r + 1];
    ++nnz;
    i = j;
  }
  for (int r = 0; r < nr; ++r) offsets[r + 1] += offsets[r];
  return nnz;
}

// warp.sparse.bsr_axpy(x, y, alpha, beta) (dynamic_contacts.py:297): the
// union of two canonical patterns, a block missing on one side read as zero
// there, each value alpha * xv + beta * yv in fp32. zc / zv hold
// x_nnz + y_nnz blocks.
int oracle_bsr_axpy(int nr, const int* xo, const int* xc, const float* xv, float alpha, const int* yo, const int* yc,
                    const float* yv, float beta, int* zo, int* zc, float* zv) {
  int o = 0;
  zo[0] = 0;
  for (int r = 0; r < nr; ++r) {
    int i = xo[r], j = yo[r];
    while (i < xo[r + 1] || j < yo[r + 1]) {
      const int a = i < xo[r + 1] ? xc[i] : INT32_MAX, b = j < yo[r + 1] ? yc[j] : INT32_MAX;
      const bool tx = a <= b, ty = b <= a;
      for (int k = 0; k < 9; ++k) {
        const float xs = tx ? xv[9ll * i + k] : 0.0f, ys = ty ? yv[9ll * j + k] : 0.0f;
        zv[9ll * o + k] = alpha * xs + beta * ys;
      }
      zc[o++] = tx ? a : b;
      i += tx;
      j += ty;
    }
    zo[r + 1] = o;
  }
  return o;
}

// affine_body_dynamics.py:144-190: _reduce_contact_hessian's 16 body blocks
// per vertex triplet (scale * c_a(v_i) c_c(v_j) H, c_0 = 1, c_k = rest[v][k-1],
// zero when a body is fixed) summed in fp64 per block over the structural
// pattern the library uses (every body pair receiving a nonzero block owns
// its 16 DOF blocks), plus the sum of |terms| per value (the summation error
// bound), and _reduce_contact_force into rforce [4 nb][3] (fp64). Outputs
// sized by the caller: offsets [4 nb + 1], columns / values / abssum for up to
// cap_pairs body pairs. Returns the body-pair count (or -1 if above cap).
int oracle_abd_reduce(int nb, const int* body, const float* rest, const int* fixed, int nv, const int* rows,
                      const int* cols, const float* vals, int m, const float* force, double scale, int cap_pairs,
                      int* offsets, int* columns, double* values, double* abssum, double* rforce) {
  std::map<std::pair<int, int>, std::vector<double>> acc;
  for (int t = 0; t < m; ++t) {
    const int r = rows[t], c = cols[t];
    if (r < 0 || r >= nv || c < 0 || c >= nv) continue;
    const int bi = body[r], bj = body[c];
    if (bi < 0 || bi >= nb || bj < 0 || bj >= nb || fixed[bi] || fixed[bj]) continue;
    bool nz = false;
    for (int k = 0; k < 9; ++k) nz |= !(vals[9ll * t + k] == 0.0f);
    if (!nz) continue;
    auto& v = acc[{bi, bj}];
    if (v.empty()) v.assign(288, 0.0);
    const double ca[4] = {1.0, rest[3ll * r], rest[3ll * r + 1], rest[3ll * r + 2]};
    const double cc[4] = {1.0, rest[3ll * c], rest[3ll * c + 1], rest[3ll * c + 2]};
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b)
        for (int k = 0; k < 9; ++k) {
          const double term = scale * ca[a] * cc[b] * double(vals[9ll * t + k]);
          v[(a * 4 + b) * 9 + k] += term;
          v[144 + (a * 4 + b) * 9 + k] += std::fabs(term);
        }
  }
  const int pairs = int(acc.size());
  if (rforce) {
    for (int k = 0; k < 12 * nb; ++k) rforce[k] = 0.0;
    for (int i = 0; i < nv; ++i) {
      const int b = body[i];
      if (b < 0 || b >= nb || fixed[b]) continue;
      const double cf[4] = {1.0, rest[3ll * i], rest[3ll * i + 1], rest[3ll * i + 2]};
      for (int a = 0; a < 4; ++a)
        for (int k = 0; k < 3; ++k) rforce[12ll * b + 3 * a + k] += cf[a] * double(force[3ll * i + k]);
    }
  }
  if (pairs > cap_pairs) return -1;
  // body-level rows (std::map iterates keys in (A, B) order)
  std::vector<int> poff(nb + 1, 0), pcol;
  std::vector<const std::vector<double>*> pv;
  for (auto& kv : acc) {
    ++poff[kv.first.first + 1];
    pcol.push_back(kv.first.second);
    pv.push_back(&kv.second);
  }
  for (int A = 0; A < nb; ++A) poff[A + 1] += poff[A];
  for (int A = 0; A < nb; ++A) {
    const int p0 = poff[A], len = poff[A + 1] - p0;
    for (int a = 0; a < 4; ++a) {
      const long long o = 16ll * p0 + 4ll * a * len;
      offsets[4 * A + a] = int(o);
      for (int j = 0; j < len; ++j)
        for (int c = 0; c < 4; ++c) {
          const long long blk = o + 4 * j + c;
          columns[blk] = 4 * pcol[p0 + j] + c;
          for (int k = 0; k < 9; ++k) {
            values[9 * blk + k] = (*pv[p0 + j])[(a * 4 + c) * 9 + k];
            abssum[9 * blk + k] = (*pv[p0 + j])[144 + (a * 4 + c) * 9 + k];
          }
        }
    }
  }
  offsets[4 * nb] = 16 * pairs;
  return pairs;
}

// ccd/ccd.py:86-163 per candidate (quads EE (a0, a1, b0, b1), PT (p, t0, t1, t2)).
void oracle_conservative_time(int is_ee, int use_double, int n, const float* x0, const float* x1, const int* quads,
                              double* out) {
  for (int i = 0; i < n; ++i) {
    const int* q = quads + 4ll * i;
    if (use_double) {
      V3<double> x[4], d[4];
      for (int k = 0; k < 4; ++k) {
        x[k] = load3<double>(x0, q[k]);
        d[k] = load3<double>(x1, q[k]) - x[k];
      }
      out[i] = conservative_toi<double>(is_ee != 0, x, d);
    } else {
      V3<float> x[4], d[4];
      for (int k = 0; k < 4; ++k) {
        x[k] = load3<float>(x0, q[k]);
        d[k] = load3<float>(x1, q[k]) - x[k];
      }
      out[i] = conservative_toi<float>(is_ee != 0, x, d);
    }
  }
}

double oracle_segment_distance_sq(const double* a0, const double* a1, const double* b0, const double* b1) {
  return ca_segment_distance_sq(mk(a0[0], a0[1], a0[2]), mk(a1[0], a1[1], a1[2]), mk(b0[0], b0[1], b0[2]),
                                mk(b1[0], b1[1], b1[2]));
}

double oracle_triangle_distance_sq(const double* p, const double* a, const double* b, const double* c) {
  return ca_triangle_distance_sq(mk(p[0], p[1], p[2]), mk(a[0], a[1], a[2]), mk(b[0], b[1], b[2]),
                                 mk(c[0], c[1], c[2]));
}

// ---------------------------------------------------------------- FEM ----
// fem/fem.py:21-34 (compute_Dm), :138-203 (tet_kernel_sparse, one thread per
// block (e, i, j) as the reference launches it) and
// time_integrator_base.py:124-137 (compute_Psi), stable Neo-Hookean
// (fem/snh.py:1-84), fp32 in the reference's operation order.
namespace {
struct M3f { float m[3][3]; };
M3f m3_zero() { M3f a{}; return a; }
M3f m3_cols(const float* c0, const float* c1, const float* c2) {
  M3f a;
  for (int r = 0; r < 3; ++r) { a.m[r][0] = c0[r]; a.m[r][1] = c1[r]; a.m[r][2] = c2[r]; }
  return a;
}
M3f m3_mul(const M3f& a, const M3f& b) {
  M3f o;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      float s = 0.0f;
      for (int k = 0; k < 3; ++k) s += a.m[r][k] * b.m[k][c];
      o.m[r][c] = s;
    }
  return o;
}
M3f m3_transpose(const M3f& a) {
  M3f o;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) o.m[r][c] = a.m[c][r];
  return o;
}
M3f m3_scale(float s, const M3f& a) {
  M3f o;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) o.m[r][c] = s * a.m[r][c];
  return o;
}
M3f m3_add(const M3f& a, const M3f& b) {
  M3f o;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) o.m[r][c] = a.m[r][c] + b.m[r][c];
  return o;
}
float m3_ddot(const M3f& a, const M3f& b) {
  float s = 0.0f;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) s += a.m[r][c] * b.m[r][c];
  return s;
}
float m3_det(const M3f& a) {
  return a.m[0][0] * (a.m[1][1] * a.m[2][2] - a.m[1][2] * a.m[2][1]) -
         a.m[0][1] * (a.m[1][0] * a.m[2][2] - a.m[1][2] * a.m[2][0]) +
         a.m[0][2] * (a.m[1][0] * a.m[2][1] - a.m[1][1] * a.m[2][0]);
}
void v3_cross(const float* a, const float* b, float* o) {
  o[0] = a[1] * b[2] - a[2] * b[1];
  o[1] = a[2] * b[0] - a[0] * b[2];
  o[2] = a[0] * b[1] - a[1] * b[0];
}
void m3_col(const M3f& a, int c, float* v) { v[0] = a.m[0][c]; v[1] = a.m[1][c]; v[2] = a.m[2][c]; }
struct SnhP { float mu, lam, alpha; };
SnhP snh_reparam(float mu, float lam) {
  SnhP s;
  s.mu = float(4.0 / 3.0) * mu;
  s.lam = lam + float(5.0 / 6.0) * mu;
  s.alpha = 1.0f + 0.75f * s.mu / s.lam;
  return s;
}
M3f snh_cofactor(const M3f& F) {
  // partialJpartialF: the rows of F^T are F's columns
  float f0[3], f1[3], f2[3], c0[3], c1[3], c2[3];
  m3_col(F, 0, f0); m3_col(F, 1, f1); m3_col(F, 2, f2);
  v3_cross(f1, f2, c0); v3_cross(f2, f0, c1); v3_cross(f0, f1, c2);
  return m3_cols(c0, c1, c2);
}
M3f snh_pk1(const M3f& F, const SnhP& s) {
  const float Ic = m3_ddot(F, F);
  const M3f cof = snh_cofactor(F);
  const float dc = s.mu * (1.0f - 1.0f / (Ic + 1.0f));
  const float vc = s.lam * (m3_det(F) - s.alpha);
  return m3_add(m3_scale(dc, F), m3_scale(vc, cof));
}
M3f snh_tangent(const M3f& F, const M3f& dF, const SnhP& s) {
  const M3f cof = snh_cofactor(F);
  const float Ic = m3_ddot(F, F);
  const float Ic1 = Ic + 1.0f;
  const float dIc = 2.0f * m3_ddot(F, dF);
  const float dc = s.mu * (1.0f - 1.0f / Ic1);
  const float ddc = s.mu * dIc / (Ic1 * Ic1);
  const float dJ = m3_ddot(cof, dF);
  const float vc = s.lam * (m3_det(F) - s.alpha);
  float f0[3], f1[3], f2[3], g0[3], g1[3], g2[3], a[3], b[3], c0[3], c1[3], c2[3];
  m3_col(F, 0, f0); m3_col(F, 1, f1); m3_col(F, 2, f2);
  m3_col(dF, 0, g0); m3_col(dF, 1, g1); m3_col(dF, 2, g2);
  v3_cross(g1, f2, a); v3_cross(f1, g2, b);
  for (int k = 0; k < 3; ++k) c0[k] = a[k] + b[k];
  v3_cross(g2, f0, a); v3_cross(f2, g0, b);
  for (int k = 0; k < 3; ++k) c1[k] = a[k] + b[k];
  v3_cross(g0, f1, a); v3_cross(f0, g1, b);
  for (int k = 0; k < 3; ++k) c2[k] = a[k] + b[k];
  const M3f dcof = m3_cols(c0, c1, c2);
  return m3_add(m3_add(m3_add(m3_scale(dc, dF), m3_scale(ddc, F)), m3_scale(s.lam * dJ, cof)), m3_scale(vc, dcof));
}
M3f tet_F(const float* x, const int* t, const float* bm, M3f& Bm) {
  float d[3][3];
  for (int c = 0; c < 3; ++c)
    for (int r = 0; r < 3; ++r) d[c][r] = x[3ll * t[c] + r] - x[3ll * t[3] + r];
  const M3f Ds = m3_cols(d[0], d[1], d[2]);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) Bm.m[r][c] = bm[3 * r + c];
  return m3_mul(Ds, Bm);
}
}  // namespace

void oracle_tet_rest(const float* xcs, const int* T, int n, float* Bm, float* W) {
  for (int e = 0; e < n; ++e) {
    const int* t = T + 4ll * e;
    float d[3][3];
    for (int c = 0; c < 3; ++c)
      for (int r = 0; r < 3; ++r) d[c][r] = xcs[3ll * t[c] + r] - xcs[3ll * t[3] + r];
    const M3f Dm = m3_cols(d[0], d[1], d[2]);
    const float det = m3_det(Dm);
    const float inv = 1.0f / det;
    const auto& m = Dm.m;
    float* bm = Bm + 9ll * e;
    bm[0] = (m[1][1] * m[2][2] - m[1][2] * m[2][1]) * inv;
    bm[1] = (m[0][2] * m[2][1] - m[0][1] * m[2][2]) * inv;
    bm[2] = (m[0][1] * m[1][2] - m[0][2] * m[1][1]) * inv;
    bm[3] = (m[1][2] * m[2][0] - m[1][0] * m[2][2]) * inv;
    bm[4] = (m[0][0] * m[2][2] - m[0][2] * m[2][0]) * inv;
    bm[5] = (m[0][2] * m[1][0] - m[0][0] * m[1][2]) * inv;
    bm[6] = (m[1][0] * m[2][1] - m[1][1] * m[2][0]) * inv;
    bm[7] = (m[0][1] * m[2][0] - m[0][0] * m[2][1]) * inv;
    bm[8] = (m[0][0] * m[1][1] - m[0][1] * m[1][0]) * inv;
    W[e] = std::fabs(det) / 6.0f;
  }
}

// One iteration per reference thread ej = 16 e + 4 i + j.
void oracle_tet_hessian(const float* x, const int* T, int n, const float* Bmv, const float* Wv, float mu, float lam,
                        int* rows, int* cols, float* vals, float* b) {
  const SnhP s = snh_reparam(mu, lam);
  for (long long ej = 0; ej < 16ll * n; ++ej) {
    const long long e = ej / 16;
    const int jj = int(ej % 4), ii = int((ej / 4) % 4);
    const int* t = T + 4 * e;
    M3f Bm;
    const M3f F = tet_F(x, t, Bmv + 9 * e, Bm);
    M3f a = m3_zero();
    const M3f P = snh_pk1(F, s);
    const M3f H = m3_transpose(m3_mul(m3_scale(-Wv[e], P), m3_transpose(Bm)));
    for (int k = 0; k < 3; ++k) {
      M3f dDs = m3_zero();
      if (jj < 3) {
        dDs.m[k][jj] = -1.0f;
      } else {
        dDs.m[k][0] = 1.0f;
        dDs.m[k][1] = 1.0f;
        dDs.m[k][2] = 1.0f;
      }
      const M3f dF = m3_mul(dDs, Bm);
      const M3f dP = snh_tangent(F, dF, s);
      const M3f dH = m3_mul(m3_scale(-Wv[e], dP), m3_transpose(Bm));
      float df[3];
      for (int l = 0; l < 3; ++l) df[l] = ii == 3 ? -(dH.m[l][0] + dH.m[l][1] + dH.m[l][2]) : dH.m[l][ii];
      for (int l = 0; l < 3; ++l) a.m[l][k] += df[l];
    }
    rows[ej] = t[ii];
    cols[ej] = t[jj];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) vals[9 * ej + 3 * r + c] = a.m[r][c];
    if (b && jj == 0) {
      for (int l = 0; l < 3; ++l) {
        const float fi = ii == 3 ? ((-H.m[0][l]) - H.m[1][l]) - H.m[2][l] : H.m[ii][l];
        b[3ll * t[ii] + l] += fi;
      }
    }
  }
}

void oracle_tet_energy(const float* x, const int* T, int n, const float* Bmv, const float* Wv, float mu, float lam,
                       float* psi) {
  const SnhP s = snh_reparam(mu, lam);
  for (int e = 0; e < n; ++e) {
    M3f Bm;
    const M3f F = tet_F(x, T + 4ll * e, Bmv + 9ll * e, Bm);
    const float Ic = m3_ddot(F, F);
    const float JmA = m3_det(F) - s.alpha;
    const float p = 0.5f * (s.lam * JmA * JmA + s.mu * (Ic - 3.0f) - s.mu * std::log(Ic + 1.0f));
    psi[e] = Wv[e] * p;
  }
}

int oracle_max_threads(void) { return omp_get_max_threads(); }

}  // extern "C"

// ------------------------------------------- augmented-Lagrangian set ---
// The active-set hash kernels of augmented_lagrangian.py run with their
// threads one after another in index order (as Warp's CPU device runs a
// launch), on records in the reference's layout seen as int32 words.
namespace {

template <typename S> struct AlLayout {
  static constexpr int sw = sizeof(S) / 4;
  static constexpr int xwords = 6 + 3 * sw;                  // XConstraint
  static constexpr int awords = (sw == 1 ? 13 : 20) + 13 * sw;  // ALConstraint
  static constexpr int alpha = 6 + sw, lam = 6 + 2 * sw, gamma = 6 + 3 * sw, slack = 6 + 4 * sw,
                       mollifier = 6 + 5 * sw, type = 6 + 6 * sw;
};

template <typename S> S get_s(const uint32_t* r, int w) {
  S v;
  std::memcpy(&v, r + w, sizeof(S));
  return v;
}
template <typename S> void put_s(uint32_t* r, int w, S v) { std::memcpy(r + w, &v, sizeof(S)); }

// hash_pos (augmented_lagrangian.py:68-73): bit_xor of the int32 products,
// negated when negative, modulo table_size. The products wrap; the one XOR
// whose negation int32 cannot hold (INT_MIN) is taken as 2^31.
int al_hash(int e0, int e1, int table_size) {
  const int32_t v = int32_t(uint32_t(e0) * 73856093u ^ uint32_t(e1) * 19349663u);
  const int64_t mag = v < 0 ? -int64_t(v) : int64_t(v);
  return int(mag % table_size);
}

// filter_accept (augmented_lagrangian.py:116-121).
template <typename S> bool al_filter(const uint32_t* c, const S* vt, int nv) {
  const S alpha = get_s<S>(c, AlLayout<S>::alpha);
  for (int k = 0; k < 4; ++k) {
    const int v = int(c[k]);
    if (v >= 0 && v < nv && alpha <= S(vt[v] + S(1.0e-12))) return true;
  }
  return false;
}

// to_al_constraint (augmented_lagrangian.py:88-102).
template <typename S> void al_fresh(const uint32_t* c, uint32_t* out) {
  using L = AlLayout<S>;
  std::memset(out, 0, sizeof(uint32_t) * L::awords);
  std::memcpy(out, c, sizeof(uint32_t) * 6);
  put_s<S>(out, 6, get_s<S>(c, 6));
  put_s<S>(out, L::alpha, get_s<S>(c, L::alpha));
  put_s<S>(out, L::gamma, S(1));
  put_s<S>(out, L::mollifier, S(1));
}

// ActiveSet.create_hash (augmented_lagrangian.py:525-535) with
// count_constraints, array_scan(inclusive) and register_constraints
// (:124-149).
template <typename S>
int al_create(const uint32_t* rec, int n, const S* vt, int nv, int T, int* cs, uint32_t* entries, int cap) {
  using L = AlLayout<S>;
  std::fill(cs, cs + T + 1, 0);
  std::fill(entries, entries + size_t(cap) * L::awords, 0u);
  for (int i = 0; i < n; ++i) {
    const uint32_t* c = rec + size_t(i) * L::xwords;
    if (al_filter<S>(c, vt, nv)) ++cs[al_hash(int(c[4]), int(c[5]), T)];
  }
  for (int b = 1; b <= T; ++b) cs[b] += cs[b - 1];
  for (int i = 0; i < n; ++i) {
    const uint32_t* c = rec + size_t(i) * L::xwords;
    if (!al_filter<S>(c, vt, nv)) continue;
    const int out = --cs[al_hash(int(c[4]), int(c[5]), T)];
    al_fresh<S>(c, entries + size_t(out) * L::awords);
  }
  return cs[T];
}

// is_new_constraint (augmented_lagrangian.py:157-164).
template <typename S> bool al_is_new(const int* ocs, const uint32_t* oe, int oT, const uint32_t* c) {
  const int b = al_hash(int(c[4]), int(c[5]), oT);
  for (int k = ocs[b]; k < ocs[b + 1]; ++k) {
    const uint32_t* o = oe + size_t(k) * AlLayout<S>::awords;
    if (o[4] == c[4] && o[5] == c[5]) return false;
  }
  return true;
}

// ActiveSet.merge (augmented_lagrangian.py:537-547) with count_merge,
// array_scan(inclusive) and register_merge (:167-197); keep_old :152-154.
template <typename S>
int al_merge(const int* ocs, const uint32_t* oe, int oT, int n_old, const uint32_t* ne, int n_new, int T, int* cs,
             uint32_t* entries, int cap) {
  using L = AlLayout<S>;
  std::fill(cs, cs + T + 1, 0);
  std::fill(entries, entries + size_t(cap) * L::awords, 0u);
  auto source = [&](int i, bool& accept) -> const uint32_t* {
    if (i < n_old) {
      const uint32_t* c = oe + size_t(i) * L::awords;
      accept = get_s<S>(c, L::gamma) >= S(0.01);
      return c;
    }
    const uint32_t* c = ne + size_t(i - n_old) * L::awords;
    accept = al_is_new<S>(ocs, oe, oT, c);
    return c;
  };
  for (int i = 0; i < n_old + n_new; ++i) {
    bool accept;
    const uint32_t* c = source(i, accept);
    if (accept) ++cs[al_hash(int(c[4]), int(c[5]), T)];
  }
  for (int b = 1; b <= T; ++b) cs[b] += cs[b - 1];
  for (int i = 0; i < n_old + n_new; ++i) {
    bool accept;
    const uint32_t* c = source(i, accept);
    if (!accept) continue;
    const int out = --cs[al_hash(int(c[4]), int(c[5]), T)];
    std::memcpy(entries + size_t(out) * L::awords, c, sizeof(uint32_t) * L::awords);
  }
  return cs[T];
}

// constraint_value (augmented_lagrangian.py:76-85); false when a vertex is
// out of range.
template <typename S> bool al_value(const uint32_t* e, const S* x, int nv, S& c) {
  using L = AlLayout<S>;
  const int d0 = L::sw == 1 ? 13 : 20;
  c = get_s<S>(e, d0);
  for (int k = 0; k < 4; ++k) {
    const int v = int(e[k]);
    if (v < 0 || v >= nv) return false;
    const int g = d0 + L::sw * (1 + 3 * k);
    const S d = get_s<S>(e, g) * x[3ll * v] + get_s<S>(e, g + L::sw) * x[3ll * v + 1] +
                get_s<S>(e, g + 2 * L::sw) * x[3ll * v + 2];
    c = c + d;
  }
  return true;
}

// update_slack_variables (:326-331) and update_active_set_forces (:334-345).
template <typename S> void al_update(bool forces, uint32_t* entries, int n, const S* x, int nv, S mu) {
  using L = AlLayout<S>;
  for (int i = 0; i < n; ++i) {
    uint32_t* e = entries + size_t(i) * L::awords;
    S c;
    if (!al_value<S>(e, x, nv, c)) continue;
    const S lam = get_s<S>(e, L::lam);
    if (!forces) {
      const S v = c - lam / mu;
      put_s<S>(e, L::slack, v > S(0) ? v : S(0));
    } else if (get_s<S>(e, L::slack) == S(0)) {
      put_s<S>(e, L::lam, lam - mu * c);
      put_s<S>(e, L::gamma, S(1));
    } else {
      put_s<S>(e, L::lam, S(0));
      put_s<S>(e, L::gamma, get_s<S>(e, L::gamma) * S(0.9));
    }
  }
}

// linearize_constraints (augmented_lagrangian.py:270-323) with the helpers
// it calls: closest_point_triangle / wp.closest_point_edge_edge on fp32
// casts of x_safe, point_line_distance_squared and
// parallel_edge_point_edge_distance (:199-267) in S, the mollifier
// m(c) = (c/eps)(2 - c/eps) for c = |ea x eb|^2 < eps = 1e-3 |ea0|^2 |eb0|^2
// (ipctkwp's ee_mollifier_value / _threshold, restated), the dispatch of
// point_triangle_distance_gradient_hessian / edge_edge_distance_gradient_
// hessian (dynamic_contacts.py:48-125), and the gradient of the feature's
// squared distance in double (feature_grad_nsd).
template <typename S> S al_point_line(V3<S> p, V3<S> e0, V3<S> e1, S& alpha) {
  const V3<S> edge = e1 - e0;
  const S ee = dot(edge, edge);
  alpha = dot(p - e0, edge) / (ee > S(1.0e-30) ? ee : S(1.0e-30));
  const V3<S> off = p - (e0 + edge * alpha);
  return dot(off, off);
}

template <typename S>
void al_linearize(uint32_t* entries, int n, const S* xs, const S* xrest, int nv, bool is_pt, S threshold) {
  using L = AlLayout<S>;
  const int d0w = L::sw == 1 ? 13 : 20;
  for (int i = 0; i < n; ++i) {
    uint32_t* e = entries + size_t(i) * L::awords;
    int id[4];
    bool ok = true;
    for (int k = 0; k < 4; ++k) {
      id[k] = int(e[k]);
      ok = ok && id[k] >= 0 && id[k] < nv;
    }
    if (!ok) continue;
    V3<S> x[4];
    V3<float> xf[4];
    V3<double> xd[4];
    for (int k = 0; k < 4; ++k) {
      x[k] = V3<S>{xs[3 * id[k]], xs[3 * id[k] + 1], xs[3 * id[k] + 2]};
      xf[k] = V3<float>{float(x[k].x), float(x[k].y), float(x[k].z)};
      xd[k] = V3<double>{double(x[k].x), double(x[k].y), double(x[k].z)};
    }
    Feature f;
    S d2;
    S mollifier = S(1);
    int type = 0;
    if (is_pt) {
      const TriClosest<float> tc = closest_point_triangle(xf[0], xf[1], xf[2], xf[3]);
      f = pt_feature_from_type(tc.type);
      d2 = S(tc.dist) * S(tc.dist);
    } else {
      auto rest = [&](int k) { return V3<S>{xrest[3 * id[k]], xrest[3 * id[k] + 1], xrest[3 * id[k] + 2]}; };
      const V3<S> ra = rest(1) - rest(0), rb = rest(3) - rest(2);
      const S eps_x = S(1.0e-3) * dot(ra, ra) * dot(rb, rb);
      const V3<S> cr = cross(x[1] - x[0], x[3] - x[2]);
      const S c = dot(cr, cr);
      if (c < eps_x) mollifier = (c / eps_x) * (S(2) - c / eps_x);
      if (mollifier < threshold) {
        // parallel_edge_point_edge_distance (:208-267)
        S a[4], dist[4];
        dist[0] = al_point_line(x[0], x[2], x[3], a[0]);
        dist[1] = al_point_line(x[1], x[2], x[3], a[1]);
        dist[2] = al_point_line(x[2], x[0], x[1], a[2]);
        dist[3] = al_point_line(x[3], x[0], x[1], a[3]);
        S score[4];
        const S large = S(1.0e30);
        for (int k = 0; k < 4; ++k) score[k] = (a[k] < S(0) || a[k] > S(1)) ? large : dist[k];
        if (std::min(std::min(score[0], score[1]), std::min(score[2], score[3])) == large)
          for (int k = 0; k < 4; ++k) score[k] = dist[k];
        int best = 0;
        for (int k = 1; k < 4; ++k)
          if (score[k] < score[best]) best = k;
        type = best + 1;
        d2 = dist[best];
        static const int slots[4][3] = {{0, 2, 3}, {1, 2, 3}, {2, 0, 1}, {3, 0, 1}};
        f = Feature{F_PE, {slots[best][0], slots[best][1], slots[best][2], -1}};
      } else {
        const EdgeClosest<float> ec = closest_point_edge_edge(xf[0], xf[1], xf[2], xf[3], 1.0e-6f);
        const S sp = S(ec.s), tp = S(ec.t), fe = S(1.0e-6), one = S(1);
        const bool sb = sp <= fe || sp >= one - fe, tb = tp <= fe || tp >= one - fe;
        if (sb && tb) {
          f = Feature{F_PP, {sp >= one - fe ? 1 : 0, tp >= one - fe ? 3 : 2, -1, -1}};
        } else if (sb) {
          f = Feature{F_PE, {sp >= one - fe ? 1 : 0, 2, 3, -1}};
        } else if (tb) {
          f = Feature{F_PE, {tp >= one - fe ? 3 : 2, 0, 1, -1}};
        } else if (eps_x > S(0) && c < eps_x) {
          const S m = std::min(std::min(sp, one - sp), std::min(tp, one - tp));
          if (m == sp) f = Feature{F_PE, {0, 2, 3, -1}};
          else if (m == one - sp) f = Feature{F_PE, {1, 2, 3, -1}};
          else if (m == tp) f = Feature{F_PE, {2, 0, 1, -1}};
          else f = Feature{F_PE, {3, 0, 1, -1}};
        } else {
          f = Feature{F_EE, {0, 1, 2, 3}};
        }
        d2 = S(ec.dist) * S(ec.dist);
      }
    }
    double g[12], H[144];
    feature_grad_nsd(f, xd, g, H);
    const S dist = std::sqrt(d2 > S(1.0e-30) ? d2 : S(1.0e-30));
    S grad[12];
    S off = dist - get_s<S>(e, 6);
    for (int k = 0; k < 4; ++k) {
      for (int a3 = 0; a3 < 3; ++a3) grad[3 * k + a3] = S(g[3 * k + a3] / (2.0 * double(dist)));
      off = off - (grad[3 * k] * x[k].x + grad[3 * k + 1] * x[k].y + grad[3 * k + 2] * x[k].z);
    }
    put_s<S>(e, d0w, off);
    for (int j = 0; j < 12; ++j) put_s<S>(e, d0w + L::sw * (1 + j), grad[j]);
    put_s<S>(e, L::mollifier, mollifier);
    e[L::type] = uint32_t(type);
  }
}

}  // namespace

extern "C" {

void oracle_al_linearize(int use_double, uint32_t* entries, int n, const void* xs, const void* xrest, int nv,
                         int is_pt, double threshold) {
  if (use_double)
    al_linearize<double>(entries, n, static_cast<const double*>(xs), static_cast<const double*>(xrest), nv, is_pt != 0,
                         threshold);
  else
    al_linearize<float>(entries, n, static_cast<const float*>(xs), static_cast<const float*>(xrest), nv, is_pt != 0,
                        float(threshold));
}

void oracle_al_update(int forces, int use_double, uint32_t* entries, int n, const void* x, int nv, double mu) {
  if (use_double) al_update<double>(forces != 0, entries, n, static_cast<const double*>(x), nv, mu);
  else al_update<float>(forces != 0, entries, n, static_cast<const float*>(x), nv, float(mu));
}

int oracle_al_hash_pos(int e0, int e1, int table_size) { return al_hash(e0, e1, table_size); }

// accumulate_vertex_toi (augmented_lagrangian.py:105-113).
void oracle_al_vertex_toi(int use_double, const uint32_t* rec, int n, void* vt, int nv) {
  for (int i = 0; i < n; ++i) {
    if (use_double) {
      const uint32_t* c = rec + size_t(i) * AlLayout<double>::xwords;
      const double a = get_s<double>(c, AlLayout<double>::alpha);
      for (int k = 0; k < 4; ++k) {
        const int v = int(c[k]);
        if (v >= 0 && v < nv) static_cast<double*>(vt)[v] = std::min(static_cast<double*>(vt)[v], a);
      }
    } else {
      const uint32_t* c = rec + size_t(i) * AlLayout<float>::xwords;
      const float a = get_s<float>(c, AlLayout<float>::alpha);
      for (int k = 0; k < 4; ++k) {
        const int v = int(c[k]);
        if (v >= 0 && v < nv) static_cast<float*>(vt)[v] = std::min(static_cast<float*>(vt)[v], a);
      }
    }
  }
}

int oracle_al_create_hash(int use_double, const uint32_t* rec, int n, const void* vt, int nv, int table_size,
                          int* cell_start, uint32_t* entries, int capacity) {
  return use_double ? al_create<double>(rec, n, static_cast<const double*>(vt), nv, table_size, cell_start, entries,
                                        capacity)
                    : al_create<float>(rec, n, static_cast<const float*>(vt), nv, table_size, cell_start, entries,
                                       capacity);
}

int oracle_al_merge(int use_double, const int* old_cs, const uint32_t* old_e, int old_table_size, int n_old,
                    const uint32_t* new_e, int n_new, int table_size, int* cell_start, uint32_t* entries,
                    int capacity) {
  return use_double ? al_merge<double>(old_cs, old_e, old_table_size, n_old, new_e, n_new, table_size, cell_start,
                                       entries, capacity)
                    : al_merge<float>(old_cs, old_e, old_table_size, n_old, new_e, n_new, table_size, cell_start,
                                      entries, capacity);
}

}  // extern "C"

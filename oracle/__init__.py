"""CPU oracle for the IPC contact inner loop — TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import this package, and only as the checker (or the
timed CPU baseline), never as the thing the product runs.

It wraps ``libipc_oracle.so`` (``ipc_oracle.hpp``/``ipc_oracle.cpp``), a C++
restatement of the reference's algorithms:

* ``contact.py:230-302`` closest point on a triangle, ``contact.py:199-336``
  DCD pair sets, ``ccd/ccd.py:166-427`` CCD candidates and times of impact,
  ``ccd/cubic_roots.py:1-193`` root isolation,
  ``dynamic_contacts.py:19-246,374-398`` barrier gradient/Hessian/energy.
* Absent dependencies restated from their published algorithms:
  Warp's ``closest_point_edge_edge`` (Ericson 5.1.9) and the ``ipctkwp``
  barrier, mollifier, NSD distance Hessians and beta/gamma (IPC, Li et al.
  2020). Parity of those against the originals is unpinned; the tests pin
  them by closed forms, brute force and invariants instead.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libipc_oracle.so")
_lib = None

F32P = ctypes.POINTER(ctypes.c_float)
F64P = ctypes.POINTER(ctypes.c_double)
I32P = ctypes.POINTER(ctypes.c_int)
U8P = ctypes.POINTER(ctypes.c_uint8)


def build() -> str:
    """Compile the oracle with its Makefile (g++, no FMA contraction)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        c_int, c_dbl, c_flt, c_i64 = ctypes.c_int, ctypes.c_double, ctypes.c_float, ctypes.c_int64
        L.oracle_closest_point_triangle.argtypes = [c_int, c_int, F64P, F64P, F64P, F64P, F64P, I32P]
        L.oracle_closest_point_edge_edge.argtypes = [c_int, c_int, F64P, F64P, F64P, F64P, c_dbl, F64P]
        L.oracle_barrier.argtypes = [c_int, F64P, c_dbl, c_dbl, F64P]
        L.oracle_feature_hessian.argtypes = [c_int, I32P, F64P, c_int, F64P, F64P]
        L.oracle_feature_hessian.restype = c_dbl
        L.oracle_feature_ee.argtypes = [c_int, F64P, c_dbl, I32P, I32P]
        L.oracle_ee_mollifier.argtypes = [F64P, c_dbl, F64P, F64P]
        L.oracle_ee_mollifier.restype = c_dbl
        L.oracle_contact_hessian.argtypes = [c_int, c_int, c_int, F32P, F32P, I32P, c_int, c_flt, c_dbl, c_int,
                                             I32P, I32P, F64P, F64P, U8P, c_int]
        L.oracle_contact_energy.argtypes = [c_int, c_int, c_int, F32P, F32P, I32P, c_int, c_flt, c_dbl, c_int]
        L.oracle_contact_energy.restype = c_dbl
        L.oracle_contact_pair.argtypes = [c_int, c_int, c_int, F64P, F64P, c_flt, c_dbl, F64P, F64P, I32P]
        L.oracle_contact_pair.restype = c_dbl
        L.oracle_contact_points.argtypes = [c_int, c_int, F32P, F32P, I32P, c_int, F64P, F64P, I32P]
        L.oracle_cubic_roots.argtypes = [c_int, F64P, c_dbl, c_dbl, F64P, I32P]
        L.oracle_cubic_roots.restype = c_int
        L.oracle_coplanarity_cubic.argtypes = [c_int, F64P, F64P, F64P]
        L.oracle_collision_time.argtypes = [c_int, c_int, c_int, c_int, F32P, F32P, I32P, F64P, c_int]
        L.oracle_collision_time.restype = c_int
        L.oracle_dcd_ee.argtypes = [F32P, I32P, c_int, I32P, c_flt, c_flt, c_int, I32P, I32P, c_i64]
        L.oracle_dcd_ee.restype = c_i64
        L.oracle_dcd_pt.argtypes = [F32P, c_int, I32P, c_int, I32P, c_flt, c_flt, c_int, I32P, I32P, c_i64]
        L.oracle_dcd_pt.restype = c_i64
        L.oracle_ccd_candidates_ee.argtypes = [F32P, F32P, I32P, c_int, I32P, c_flt, c_int, I32P, I32P, c_i64]
        L.oracle_ccd_candidates_ee.restype = c_i64
        L.oracle_ccd_candidates_pt.argtypes = [F32P, F32P, c_int, I32P, c_int, I32P, c_flt, c_int, I32P, I32P, c_i64]
        L.oracle_ccd_candidates_pt.restype = c_i64
        L.oracle_max_threads.restype = c_int
        L.oracle_count_intersections.argtypes = [F32P, I32P, c_int, I32P, c_int, I32P, c_i64]
        L.oracle_count_intersections.restype = c_i64
        L.oracle_warmstart_features.argtypes = [F32P, c_int, F32P, I32P, c_int, I32P, c_int, c_flt, F32P]
        L.oracle_bsr_from_triplets.argtypes = [c_int, c_int, I32P, I32P, F32P, c_int, I32P, I32P, F32P]
        L.oracle_bsr_from_triplets.restype = c_int
        L.oracle_bsr_axpy.argtypes = [c_int, I32P, I32P, F32P, c_flt, I32P, I32P, F32P, c_flt, I32P, I32P, F32P]
        L.oracle_bsr_axpy.restype = c_int
        L.oracle_abd_reduce.argtypes = [c_int, I32P, F32P, I32P, c_int, I32P, I32P, F32P, c_int, F32P, c_dbl, c_int,
                                        I32P, I32P, F64P, F64P, F64P]
        L.oracle_abd_reduce.restype = c_int
        L.oracle_conservative_time.argtypes = [c_int, c_int, c_int, F32P, F32P, I32P, F64P]
        L.oracle_tet_rest.argtypes = [F32P, I32P, c_int, F32P, F32P]
        L.oracle_tet_hessian.argtypes = [F32P, I32P, c_int, F32P, F32P, c_flt, c_flt, I32P, I32P, F32P, F32P]
        L.oracle_tet_energy.argtypes = [F32P, I32P, c_int, F32P, F32P, c_flt, c_flt, F32P]
        L.oracle_contact_hessian_f64.argtypes = [c_int, F64P, F64P, I32P, c_int, c_dbl, c_dbl, c_int, I32P, I32P,
                                                 F64P, F64P, U8P, c_int]
        L.oracle_contact_energy_f64.argtypes = [c_int, F64P, F64P, I32P, c_int, c_dbl, c_dbl]
        L.oracle_contact_energy_f64.restype = c_dbl
        L.oracle_collision_time_f64.argtypes = [c_int, c_int, c_int, F64P, F64P, I32P, F64P, c_int]
        L.oracle_al_hash_pos.argtypes = [c_int, c_int, c_int]
        L.oracle_al_update.argtypes = [c_int, c_int, I32P, c_int, ctypes.c_void_p, c_int, c_dbl]
        L.oracle_al_linearize.argtypes = [c_int, I32P, c_int, ctypes.c_void_p, ctypes.c_void_p, c_int, c_int, c_dbl]
        L.oracle_al_hash_pos.restype = c_int
        L.oracle_al_vertex_toi.argtypes = [c_int, I32P, c_int, ctypes.c_void_p, c_int]
        L.oracle_al_create_hash.argtypes = [c_int, I32P, c_int, ctypes.c_void_p, c_int, c_int, I32P, I32P, c_int]
        L.oracle_al_merge.argtypes = [c_int, I32P, I32P, c_int, c_int, I32P, c_int, c_int, I32P, I32P, c_int]
        L.oracle_segment_distance_sq.argtypes = [F64P, F64P, F64P, F64P]
        L.oracle_segment_distance_sq.restype = c_dbl
        L.oracle_triangle_distance_sq.argtypes = [F64P, F64P, F64P, F64P]
        L.oracle_triangle_distance_sq.restype = c_dbl
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def max_threads() -> int:
    return int(lib().oracle_max_threads())


# ------------------------------------------------------------- primitives ----
def closest_point_triangle(p, a, b, c, double=False):
    """contact.py:230-302 -> (weights_and_dist [n,3], feature [n])."""
    p, a, b, c = (_f64(np.atleast_2d(v)) for v in (p, a, b, c))
    n = p.shape[0]
    out = np.zeros((n, 3))
    typ = np.zeros(n, dtype=np.int32)
    lib().oracle_closest_point_triangle(int(double), n, _p(p, F64P), _p(a, F64P), _p(b, F64P), _p(c, F64P),
                                        _p(out, F64P), _p(typ, I32P))
    return out, typ


def closest_point_edge_edge(p1, q1, p2, q2, eps=1e-6, double=False):
    """Warp closest_point_edge_edge stand-in -> [n,3] (s, t, dist)."""
    p1, q1, p2, q2 = (_f64(np.atleast_2d(v)) for v in (p1, q1, p2, q2))
    n = p1.shape[0]
    out = np.zeros((n, 3))
    lib().oracle_closest_point_edge_edge(int(double), n, _p(p1, F64P), _p(q1, F64P), _p(p2, F64P), _p(q2, F64P),
                                         float(eps), _p(out, F64P))
    return out


def barrier(x, D, kappa):
    """(b, db/dx, d2b/dx2) with x = d^2, D = dhat^2."""
    x = _f64(np.atleast_1d(x))
    out = np.zeros((x.shape[0], 3))
    lib().oracle_barrier(x.shape[0], _p(x, F64P), float(D), float(kappa), _p(out, F64P))
    return out


FEATURE = {"PP": 0, "PE": 1, "PT": 2, "EE": 3}


def feature_hessian(kind, slots, x12, nsd=False):
    """Squared distance of a feature, its gradient and (NSD part of) Hessian."""
    slots = _i32(list(slots) + [-1] * (4 - len(slots)))
    x12 = _f64(np.asarray(x12).reshape(12))
    g = np.zeros(12)
    H = np.zeros((12, 12))
    k = FEATURE[kind] if isinstance(kind, str) else int(kind)
    d2 = lib().oracle_feature_hessian(k, _p(slots, I32P), _p(x12, F64P), int(nsd), _p(g, F64P), _p(H, F64P))
    return d2, g, H


def feature_ee(x12, parallel_eps, double=False):
    kind = ctypes.c_int()
    slots = np.zeros(4, dtype=np.int32)
    x12 = _f64(np.asarray(x12).reshape(12))
    lib().oracle_feature_ee(int(double), _p(x12, F64P), float(parallel_eps), ctypes.byref(kind), _p(slots, I32P))
    return kind.value, slots


def ee_mollifier(x12, eps_x):
    x12 = _f64(np.asarray(x12).reshape(12))
    g = np.zeros(12)
    H = np.zeros((12, 12))
    m = lib().oracle_ee_mollifier(_p(x12, F64P), float(eps_x), _p(g, F64P), _p(H, F64P))
    return m, g, H


# ---------------------------------------------------------- contact kernels ----
# Modes: "mixed" classifies in float like the reference's FP32 mode and takes
# derivatives of the classified feature in double (the HIP library's policy);
# "ref32" also uses the float distance for the barrier argument, as the
# reference does; "fp64" classifies in double too (the all-FP64 oracle).
_MODES = {"mixed": (0, 1), "ref32": (0, 0), "fp64": (1, 1)}


def contact_hessian(kind, x, contacts, l0, kappa, xcs=None, offset=0, n_total=None, mode="mixed", threads=1):
    """contact_hessian_ee / _pt (dynamic_contacts.py:151-246).

    Returns (rows, cols, vals [16*n_total,3,3] float64, b [V,3] force, active [n]).
    """
    x = _f32(x).reshape(-1, 3)
    xcs = x if xcs is None else _f32(xcs).reshape(-1, 3)
    contacts = _i32(contacts).reshape(-1, 4)
    n = contacts.shape[0]
    n_total = offset + n if n_total is None else n_total
    rows = np.zeros(16 * n_total, dtype=np.int32)
    cols = np.zeros(16 * n_total, dtype=np.int32)
    vals = np.zeros((16 * n_total, 3, 3))
    b = np.zeros((x.shape[0], 3))
    active = np.zeros(n, dtype=np.uint8)
    prec, pol = _MODES[mode]
    lib().oracle_contact_hessian(int(kind == "ee"), prec, pol, _p(x, F32P), _p(xcs, F32P), _p(contacts, I32P), n,
                                 float(l0), float(kappa), int(offset), _p(rows, I32P), _p(cols, I32P), _p(vals, F64P),
                                 _p(b, F64P), _p(active, U8P), int(threads))
    return rows, cols, vals, b, active.astype(bool)


def contact_hessian_f64(kind, x, contacts, l0, kappa, xcs=None, offset=0, n_total=None, threads=1):
    """contact_hessian_ee / _pt with fp64 scalars (double positions and l0).

    Returns (rows, cols, vals [16*n_total,3,3], b [V,3], active [n])."""
    x = _f64(x).reshape(-1, 3)
    xcs = x if xcs is None else _f64(xcs).reshape(-1, 3)
    contacts = _i32(contacts).reshape(-1, 4)
    n = contacts.shape[0]
    n_total = offset + n if n_total is None else n_total
    rows = np.zeros(16 * n_total, dtype=np.int32)
    cols = np.zeros(16 * n_total, dtype=np.int32)
    vals = np.zeros((16 * n_total, 3, 3))
    b = np.zeros((x.shape[0], 3))
    active = np.zeros(max(n, 1), dtype=np.uint8)
    lib().oracle_contact_hessian_f64(int(kind == "ee"), _p(x, F64P), _p(xcs, F64P), _p(contacts, I32P), n, float(l0),
                                     float(kappa), int(offset), _p(rows, I32P), _p(cols, I32P), _p(vals, F64P),
                                     _p(b, F64P), _p(active, U8P), int(threads))
    return rows, cols, vals, b, active[:n].astype(bool)


def contact_energy_f64(kind, x, contacts, l0, kappa, xcs=None):
    """contact_energy_ee / _pt with fp64 scalars."""
    x = _f64(x).reshape(-1, 3)
    xcs = x if xcs is None else _f64(xcs).reshape(-1, 3)
    contacts = _i32(contacts).reshape(-1, 4)
    return float(lib().oracle_contact_energy_f64(int(kind == "ee"), _p(x, F64P), _p(xcs, F64P), _p(contacts, I32P),
                                                 contacts.shape[0], float(l0), float(kappa)))


def contact_energy(kind, x, contacts, l0, kappa, xcs=None, mode="mixed", threads=1):
    """contact_energy_ee / _pt (dynamic_contacts.py:374-398), double sum."""
    x = _f32(x).reshape(-1, 3)
    xcs = x if xcs is None else _f32(xcs).reshape(-1, 3)
    contacts = _i32(contacts).reshape(-1, 4)
    prec, pol = _MODES[mode]
    return float(lib().oracle_contact_energy(int(kind == "ee"), prec, pol, _p(x, F32P), _p(xcs, F32P),
                                             _p(contacts, I32P), contacts.shape[0], float(l0), float(kappa),
                                             int(threads)))


def contact_pair(kind, x12, l0, kappa, rest12=None, mode="fp64"):
    """One contact at double positions x12 [12] (EE: rest12 for the mollifier
    threshold; default x12): (energy, force [12] = -gradient, H [12, 12],
    feature (kind, 4 slots) with kind -1 when inactive)."""
    x12 = _f64(np.asarray(x12).reshape(12))
    r12 = None if rest12 is None else _f64(np.asarray(rest12).reshape(12))
    force = np.zeros(12)
    H = np.zeros((12, 12))
    feat = np.zeros(5, dtype=np.int32)
    prec, pol = _MODES[mode]
    e = lib().oracle_contact_pair(int(kind == "ee"), prec, pol, _p(x12, F64P), None if r12 is None else _p(r12, F64P),
                                  float(l0), float(kappa), _p(force, F64P), _p(H, F64P), _p(feat, I32P))
    return float(e), force, H, (int(feat[0]), [int(v) for v in feat[1:]])


def contact_points(kind, x, contacts, xcs=None, double=False):
    """get_contact_points(_pt) (contact.py:148-161): midpoints [n, 3], distances
    [n] and dispatched feature codes [n] (the layout of ipc_contact_points_*)."""
    x = _f32(x).reshape(-1, 3)
    xcs = None if xcs is None else _f32(xcs).reshape(-1, 3)
    contacts = _i32(contacts).reshape(-1, 4)
    n = contacts.shape[0]
    pts = np.zeros((n, 3))
    dist = np.zeros(n)
    feat = np.zeros(n, dtype=np.int32)
    lib().oracle_contact_points(int(kind == "ee"), int(double), _p(x, F32P), None if xcs is None else _p(xcs, F32P),
                                _p(contacts, I32P), n, _p(pts, F64P), _p(dist, F64P), _p(feat, I32P))
    return pts, dist, feat


# --------------------------------------------------------------------- CCD ----
def cubic_roots(coef, x0=0.0, x1=1.0, double=True):
    """ccd/cubic_roots.py:95-193 -> (n_roots, roots[3], capped)."""
    coef = _f64(coef)
    roots = np.zeros(3)
    capped = ctypes.c_int(0)
    n = lib().oracle_cubic_roots(int(double), _p(coef, F64P), float(x0), float(x1), _p(roots, F64P),
                                 ctypes.byref(capped))
    return n, roots, capped.value


def coplanarity_cubic(p0, p1, relative=False):
    p0 = _f64(np.asarray(p0).reshape(12))
    p1 = _f64(np.asarray(p1).reshape(12))
    coef = np.zeros(4)
    lib().oracle_coplanarity_cubic(int(relative), _p(p0, F64P), _p(p1, F64P), _p(coef, F64P))
    return coef


def collision_time_f64(kind, x0, x1, quads, relative=True, threads=1):
    """ccd/ccd.py:304-357 with fp64 scalars (double positions) -> toi float64 [n]."""
    a, b = _f64(x0).reshape(-1), _f64(x1).reshape(-1)
    q = _i32(quads).reshape(-1)
    n = q.size // 4
    toi = np.zeros(n)
    lib().oracle_collision_time_f64(int(kind == "ee"), int(relative), n, _p(a, F64P), _p(b, F64P), _p(q, I32P),
                                    _p(toi, F64P), int(threads))
    return toi


def collision_time(kind, x0, x1, quads, relative=False, double=True, threads=1):
    """pt_collision_time / ee_collision_time per candidate -> (toi [n], n_capped)."""
    x0 = _f32(x0).reshape(-1, 3)
    x1 = _f32(x1).reshape(-1, 3)
    quads = _i32(quads).reshape(-1, 4)
    toi = np.zeros(quads.shape[0])
    capped = lib().oracle_collision_time(int(kind == "ee"), int(relative), int(double), quads.shape[0],
                                         _p(x0, F32P), _p(x1, F32P), _p(quads, I32P), _p(toi, F64P), int(threads))
    return toi, capped


# --------------------------------------------------------------- pair sets ----
def _grow_call(fn, *args):
    cap = 1 << 12
    while True:
        rec = np.zeros((cap, 4), dtype=np.int32)
        e = np.zeros((cap, 2), dtype=np.int32)
        n = fn(*args, _p(rec, I32P), _p(e, I32P), cap)
        if n <= cap:
            return rec[:n].copy(), e[:n].copy()
        cap = 1 << int(n - 1).bit_length()


def dcd_ee(x, edges, body, thickness, buffer, exclude_same_body=False):
    """Exact EE contact set of edge_edge_collision (contact.py:199-229)."""
    x = _f32(x).reshape(-1, 3)
    edges = _i32(edges).reshape(-1)
    body = _i32(body)
    return _grow_call(lib().oracle_dcd_ee, _p(x, F32P), _p(edges, I32P), edges.size // 2, _p(body, I32P),
                      float(thickness), float(buffer), int(exclude_same_body))


def dcd_pt(x, triangles, body, thickness, buffer, exclude_same_body=False):
    """Exact PT contact set of point_triangle_collision (contact.py:304-336)."""
    x = _f32(x).reshape(-1, 3)
    tris = _i32(triangles).reshape(-1)
    body = _i32(body)
    return _grow_call(lib().oracle_dcd_pt, _p(x, F32P), x.shape[0], _p(tris, I32P), tris.size // 3, _p(body, I32P),
                      float(thickness), float(buffer), int(exclude_same_body))


def ccd_candidates_ee(x0, x1, edges, body, padding=1e-7, exclude_same_body=False):
    x0 = _f32(x0).reshape(-1, 3)
    x1 = _f32(x1).reshape(-1, 3)
    edges = _i32(edges).reshape(-1)
    body = _i32(body)
    return _grow_call(lib().oracle_ccd_candidates_ee, _p(x0, F32P), _p(x1, F32P), _p(edges, I32P), edges.size // 2,
                      _p(body, I32P), float(padding), int(exclude_same_body))


def ccd_candidates_pt(x0, x1, triangles, body, padding=1e-7, exclude_same_body=False):
    x0 = _f32(x0).reshape(-1, 3)
    x1 = _f32(x1).reshape(-1, 3)
    tris = _i32(triangles).reshape(-1)
    body = _i32(body)
    return _grow_call(lib().oracle_ccd_candidates_pt, _p(x0, F32P), _p(x1, F32P), x0.shape[0], _p(tris, I32P),
                      tris.size // 3, _p(body, I32P), float(padding), int(exclude_same_body))


def ccd_toi(x0, x1, edges, triangles, body, padding=1e-7, exclude_same_body=False, relative=False, threads=1):
    """Raw minimum TOI over EE and PT candidates (collision_free_step before the 0.9 factor)."""
    ee, _ = ccd_candidates_ee(x0, x1, edges, body, padding, exclude_same_body)
    pt, _ = ccd_candidates_pt(x0, x1, triangles, body, padding, exclude_same_body)
    t = 1.0
    if len(ee):
        te, _ = collision_time("ee", x0, x1, ee, relative=relative, threads=threads)
        t = min(t, float(te.min()))
    if len(pt):
        tp, _ = collision_time("pt", x0, x1, pt, relative=relative, threads=threads)
        t = min(t, float(tp.min()))
    return t


def warmstart_features(x, base, ee, pt, dhat):
    """float32 [V, 48] warm-start features (layout of ipc_warmstart_features)."""
    x = _f32(x)
    V = x.shape[0]
    base = None if base is None else _f32(base).reshape(V, 12)
    ee = _i32(ee).reshape(-1, 4)
    pt = _i32(pt).reshape(-1, 4)
    out = np.zeros((V, 48), np.float32)
    lib().oracle_warmstart_features(_p(x, F32P), V, None if base is None else _p(base, F32P), _p(ee, I32P), len(ee),
                                    _p(pt, I32P), len(pt), float(dhat), _p(out, F32P))
    return out


def count_intersections(x, edges, triangles):
    """(count, sorted (edge, triangle) ids) of the crossing edge-triangle pairs."""
    x = _f32(x).reshape(-1, 3)
    edges = _i32(edges).reshape(-1)
    tris = _i32(triangles).reshape(-1)
    cap = 1 << 10
    while True:
        ids = np.zeros((cap, 2), np.int32)
        n = lib().oracle_count_intersections(_p(x, F32P), _p(edges, I32P), edges.size // 2, _p(tris, I32P),
                                             tris.size // 3, _p(ids, I32P), cap)
        if n <= cap:
            return int(n), ids[:n].copy()
        cap = 1 << int(n - 1).bit_length()


def bsr_from_triplets(n_rows, n_cols, rows, cols, vals):
    """(offsets, columns, values float32 [nnz, 3, 3]) of the triplets, summed
    in fp32 in triplet order (warp.sparse.bsr_from_triplets's contract)."""
    rows, cols = _i32(rows).reshape(-1), _i32(cols).reshape(-1)
    vals = _f32(vals).reshape(-1)
    n = rows.size
    offsets = np.zeros(n_rows + 1, np.int32)
    columns = np.zeros(max(n, 1), np.int32)
    values = np.zeros(max(n, 1) * 9, np.float32)
    nnz = lib().oracle_bsr_from_triplets(n_rows, n_cols, _p(rows, I32P), _p(cols, I32P), _p(vals, F32P), n,
                                         _p(offsets, I32P), _p(columns, I32P), _p(values, F32P))
    return offsets, columns[:nnz].copy(), values[:9 * nnz].reshape(nnz, 3, 3).copy()


def bsr_axpy(x, y, alpha=1.0, beta=1.0):
    """alpha x + beta y of two (offsets, columns, values) BSR triples."""
    xo, xc, xv = _i32(x[0]), _i32(x[1]), _f32(x[2]).reshape(-1)
    yo, yc, yv = _i32(y[0]), _i32(y[1]), _f32(y[2]).reshape(-1)
    nr = xo.size - 1
    cap = max(xc.size + yc.size, 1)
    zo = np.zeros(nr + 1, np.int32)
    zc = np.zeros(cap, np.int32)
    zv = np.zeros(cap * 9, np.float32)
    nnz = lib().oracle_bsr_axpy(nr, _p(xo, I32P), _p(xc, I32P), _p(xv, F32P), alpha, _p(yo, I32P), _p(yc, I32P),
                                _p(yv, F32P), beta, _p(zo, I32P), _p(zc, I32P), _p(zv, F32P))
    return zo, zc[:nnz].copy(), zv[:9 * nnz].reshape(nnz, 3, 3).copy()


def abd_reduce(n_bodies, body, rest, fixed, rows, cols, vals, force=None, scale=1.0):
    """affine_body_dynamics.py:144-190 in fp64: (offsets, columns, values
    [nnz,3,3], abssum [nnz,3,3], reduced force [4 n_bodies, 3] or None) over
    the structural body-pair pattern (16 DOF blocks per body pair)."""
    body, fixed = _i32(body), _i32(fixed)
    rest = _f32(rest).reshape(-1)
    rows, cols = _i32(rows).reshape(-1), _i32(cols).reshape(-1)
    vals = _f32(vals).reshape(-1)
    nv, m = body.size, rows.size
    cap = max(1, min(n_bodies * n_bodies, m))
    offsets = np.zeros(4 * n_bodies + 1, np.int32)
    columns = np.zeros(16 * cap, np.int32)
    values = np.zeros(16 * cap * 9)
    abssum = np.zeros(16 * cap * 9)
    rforce = np.zeros(12 * n_bodies)
    f = None if force is None else _f32(force).reshape(-1)
    pairs = lib().oracle_abd_reduce(n_bodies, _p(body, I32P), _p(rest, F32P), _p(fixed, I32P), nv, _p(rows, I32P),
                                    _p(cols, I32P), _p(vals, F32P), m, None if f is None else _p(f, F32P),
                                    float(scale), cap, _p(offsets, I32P), _p(columns, I32P), _p(values, F64P),
                                    _p(abssum, F64P), None if f is None else _p(rforce, F64P))
    assert pairs >= 0
    n = 16 * pairs
    return (offsets, columns[:n].copy(), values[:9 * n].reshape(n, 3, 3), abssum[:9 * n].reshape(n, 3, 3),
            None if f is None else rforce.reshape(-1, 3))


def conservative_time(kind, x0, x1, quads, double=True):
    """ccd/ccd.py:86-163 per candidate quadruple (float64 array)."""
    x0, x1 = _f32(x0).reshape(-1), _f32(x1).reshape(-1)
    q = _i32(quads).reshape(-1)
    n = q.size // 4
    out = np.zeros(n)
    lib().oracle_conservative_time(int(kind == "ee"), int(double), n, _p(x0, F32P), _p(x1, F32P), _p(q, I32P),
                                   _p(out, F64P))
    return out


def segment_distance_sq(a0, a1, b0, b1):
    a = [_f64(v) for v in (a0, a1, b0, b1)]
    return lib().oracle_segment_distance_sq(*[_p(v, F64P) for v in a])


def triangle_distance_sq(p, a, b, c):
    a = [_f64(v) for v in (p, a, b, c)]
    return lib().oracle_triangle_distance_sq(*[_p(v, F64P) for v in a])


def tet_rest(xcs, tets):
    """fem/fem.py:21-34 -> (Bm float32 [n,3,3], W float32 [n])."""
    x = _f32(xcs).reshape(-1)
    T = _i32(tets).reshape(-1)
    n = T.size // 4
    Bm = np.zeros(9 * n, np.float32)
    W = np.zeros(n, np.float32)
    lib().oracle_tet_rest(_p(x, F32P), _p(T, I32P), n, _p(Bm, F32P), _p(W, F32P))
    return Bm.reshape(n, 3, 3), W


def tet_hessian(x, tets, Bm, W, mu, lam, with_force=True):
    """fem/fem.py:138-203 per reference thread -> (rows, cols, vals [16n,3,3], force [V,3])."""
    xf = _f32(x).reshape(-1)
    T = _i32(tets).reshape(-1)
    n = T.size // 4
    rows = np.zeros(16 * n, np.int32)
    cols = np.zeros(16 * n, np.int32)
    vals = np.zeros(16 * n * 9, np.float32)
    b = np.zeros(xf.size, np.float32)
    lib().oracle_tet_hessian(_p(xf, F32P), _p(T, I32P), n, _p(_f32(Bm).reshape(-1), F32P), _p(_f32(W), F32P),
                             float(mu), float(lam), _p(rows, I32P), _p(cols, I32P), _p(vals, F32P),
                             _p(b, F32P) if with_force else None)
    return rows, cols, vals.reshape(-1, 3, 3), b.reshape(-1, 3)


def tet_energy(x, tets, Bm, W, mu, lam):
    """time_integrator_base.py:124-137 -> per-tet W psi(F) (float32 [n])."""
    xf = _f32(x).reshape(-1)
    T = _i32(tets).reshape(-1)
    n = T.size // 4
    psi = np.zeros(n, np.float32)
    lib().oracle_tet_energy(_p(xf, F32P), _p(T, I32P), n, _p(_f32(Bm).reshape(-1), F32P), _p(_f32(W), F32P),
                            float(mu), float(lam), _p(psi, F32P))
    return psi


# ------------------------------------------------ augmented-Lagrangian set ---
XCONSTRAINT_WORDS = {np.float32: 9, np.float64: 12}
ALCONSTRAINT_WORDS = {np.float32: 26, np.float64: 46}


def al_hash_pos(e0, e1, table_size):
    """augmented_lagrangian.py:68-73."""
    return int(lib().oracle_al_hash_pos(int(e0), int(e1), int(table_size)))


def al_vertex_toi(records, vertex_toi):
    """augmented_lagrangian.py:105-113 on XConstraint records (int32 [n, 9]
    fp32 / [n, 12] fp64); returns the updated copy of vertex_toi."""
    r = _i32(records)
    use_double = r.shape[1] == 12
    vt = np.array(vertex_toi, dtype=np.float64 if use_double else np.float32)
    lib().oracle_al_vertex_toi(int(use_double), _p(r.reshape(-1), I32P), r.shape[0], vt.ctypes.data, vt.size)
    return vt


def al_create_hash(records, vertex_toi, table_size, capacity):
    """ActiveSet.create_hash (augmented_lagrangian.py:525-535) -> (count,
    cell_start int32 [T+1], entries int32 [capacity, 26 | 46])."""
    r = _i32(records)
    use_double = r.shape[1] == 12
    vt = np.ascontiguousarray(vertex_toi, dtype=np.float64 if use_double else np.float32)
    cs = np.zeros(table_size + 1, np.int32)
    ent = np.zeros((capacity, 46 if use_double else 26), np.int32)
    cnt = lib().oracle_al_create_hash(int(use_double), _p(r.reshape(-1), I32P), r.shape[0], vt.ctypes.data, vt.size,
                                      int(table_size), _p(cs, I32P), _p(ent, I32P), int(capacity))
    return cnt, cs, ent


def al_merge(old_cell_start, old_entries, old_table_size, n_old, new_entries, n_new, table_size, capacity):
    """ActiveSet.merge (augmented_lagrangian.py:537-547) -> (count,
    cell_start, entries)."""
    oe = _i32(old_entries)
    ne = _i32(new_entries)
    words = oe.shape[1]
    use_double = words == 46
    ocs = _i32(old_cell_start)
    cs = np.zeros(table_size + 1, np.int32)
    ent = np.zeros((capacity, words), np.int32)
    cnt = lib().oracle_al_merge(int(use_double), _p(ocs, I32P), _p(oe.reshape(-1), I32P), int(old_table_size),
                                int(n_old), _p(ne.reshape(-1), I32P), int(n_new), int(table_size), _p(cs, I32P),
                                _p(ent, I32P), int(capacity))
    return cnt, cs, ent


def al_update(entries, x, mu, forces):
    """update_slack_variables (forces=False) or update_active_set_forces
    (forces=True), augmented_lagrangian.py:326-345, on a copy of the
    ALConstraint records (int32 [n, 26 | 46]); x `scalar` [V, 3]."""
    e = np.array(entries, dtype=np.int32, copy=True)
    use_double = e.shape[1] == 46
    xs = np.ascontiguousarray(x, dtype=np.float64 if use_double else np.float32)
    lib().oracle_al_update(int(bool(forces)), int(use_double), _p(e.reshape(-1), I32P), e.shape[0], xs.ctypes.data,
                           xs.shape[0], float(mu))
    return e


def al_linearize(entries, x_safe, x_rest, is_point_triangle, threshold=1.0):
    """linearize_constraints (augmented_lagrangian.py:270-323) on a copy of
    the ALConstraint records; x_safe / x_rest `scalar` [V, 3]."""
    e = np.array(entries, dtype=np.int32, copy=True)
    use_double = e.shape[1] == 46
    dt = np.float64 if use_double else np.float32
    xs = np.ascontiguousarray(x_safe, dtype=dt)
    xr = np.ascontiguousarray(x_rest, dtype=dt)
    lib().oracle_al_linearize(int(use_double), _p(e.reshape(-1), I32P), e.shape[0], xs.ctypes.data, xr.ctypes.data,
                              xs.shape[0], int(bool(is_point_triangle)), float(threshold))
    return e

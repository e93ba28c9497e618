"""Pins of the oracle's warm-start features (layout of ipc_warmstart_features).

The checks use what the layout fixes, not the oracle's own formulas:
  * a single record seen from its four vertices: the two closest points
    rebuilt from the weights (self = sum w_k x_k over the vertex's primitive)
    must be joined by the stored offset, the offsets seen from the two sides
    are opposite, weights of a primitive sum to 1, |offset| dhat = dist, and
    dist equals the (separately pinned) closest-point distance;
  * a point above a stack of triangles at known depths keeps the 6 smallest
    depths in ascending order, ties resolved by record index;
  * base features pass through; empty and out-of-range records leave zeros.
"""
import numpy as np
import pytest

DHAT = 1e-3


def _slots(f, v):
    return f[v, 12:].reshape(6, 6)


def test_single_pt_record_geometry(oracle):
    rng = np.random.default_rng(3)
    for trial in range(50):
        tri = rng.normal(size=(3, 3)).astype(np.float32)
        p = (tri.mean(0) + rng.normal(size=3) * 0.7).astype(np.float32)
        x = np.vstack([p[None], tri]).astype(np.float32)
        # PT record layout (p, t[1], t[0], t[2])
        pt = np.array([[0, 2, 1, 3]], np.int32)
        f = oracle.warmstart_features(x, None, np.zeros((0, 4), np.int32), pt, DHAT)
        sp = _slots(f, 0)
        assert sp[0, 4] == 1.0 and sp[0, 5] == 1.0 and (sp[1:] == 0).all()
        q_from_p = x[0].astype(np.float64) + sp[0, :3].astype(np.float64) * DHAT
        q_from_t = np.zeros(3)
        wsum = 0.0
        for v in (1, 2, 3):
            s = _slots(f, v)
            assert s[0, 5] == 1.0 and (s[1:] == 0).all()
            np.testing.assert_allclose(s[0, :3], -sp[0, :3], rtol=0, atol=1e-6 * np.abs(sp[0, :3]).max())
            assert s[0, 3] == sp[0, 3]
            q_from_t += s[0, 4] * x[v]
            wsum += s[0, 4]
        assert abs(wsum - 1.0) < 1e-6
        assert 0.0 <= _slots(f, 1)[0, 4] <= 1.0
        dist = np.linalg.norm(sp[0, :3].astype(np.float64)) * DHAT
        np.testing.assert_allclose(q_from_p, q_from_t, atol=2e-6 * max(1.0, np.abs(x).max()))
        # fp32 closest point of unit-size coordinates: absolute error of a few ulp of |x|
        np.testing.assert_allclose(sp[0, 3] * DHAT, dist, rtol=0, atol=4e-7 * max(1.0, np.abs(x).max()))
        wd, _ = oracle.closest_point_triangle(x[0:1], x[2:3], x[1:2], x[3:4])
        np.testing.assert_allclose(sp[0, 3] * DHAT, wd[0, 2], rtol=1e-6)


def test_single_ee_record_geometry(oracle):
    rng = np.random.default_rng(4)
    for trial in range(50):
        x = rng.normal(size=(4, 3)).astype(np.float32)
        ee = np.array([[0, 1, 2, 3]], np.int32)
        f = oracle.warmstart_features(x, None, ee, np.zeros((0, 4), np.int32), DHAT)
        s = [_slots(f, v)[0] for v in range(4)]
        for v in range(4):
            assert s[v][5] == 0.0 and (_slots(f, v)[1:] == 0).all()
        assert abs(s[0][4] + s[1][4] - 1.0) < 1e-6 and abs(s[2][4] + s[3][4] - 1.0) < 1e-6
        pa = s[0][4] * x[0].astype(np.float64) + s[1][4] * x[1]
        pb = s[2][4] * x[2].astype(np.float64) + s[3][4] * x[3]
        tol = 2e-6 * max(1.0, np.abs(x).max())
        np.testing.assert_allclose(pa + s[0][:3] * DHAT, pb, atol=tol)
        np.testing.assert_allclose(pb + s[2][:3] * DHAT, pa, atol=tol)
        np.testing.assert_array_equal(s[0][:3], s[1][:3])
        np.testing.assert_allclose(s[0][:3], -s[2][:3], atol=1e-6 * np.abs(s[0][:3]).max())
        std = oracle.closest_point_edge_edge(x[0:1], x[1:2], x[2:3], x[3:4])
        np.testing.assert_allclose(s[0][3] * DHAT, std[0, 2], rtol=1e-6)


def test_nearest_six_of_a_stack(oracle):
    # a point at the origin above 10 triangles at depths depth[j] (exact in
    # float32), each projecting the origin to its interior; two share a depth
    depth = np.array([7, 3, 9, 1, 5, 3, 8, 2, 10, 6], np.float32) * np.float32(1e-4)
    xs = [np.zeros(3, np.float32)]
    pts = []
    for j, d in enumerate(depth):
        base = len(xs)
        xs += [np.array([-1, -d, -1], np.float32), np.array([1, -d, -1], np.float32), np.array([0, -d, 1], np.float32)]
        pts.append([0, base + 1, base, base + 2])
    x = np.array(xs, np.float32)
    pt = np.array(pts, np.int32)
    b = np.arange(x.shape[0] * 12, dtype=np.float32).reshape(-1, 12)
    f = oracle.warmstart_features(x, b, np.zeros((0, 4), np.int32), pt, DHAT)
    np.testing.assert_array_equal(f[:, :12], b)
    s = _slots(f, 0)
    order = sorted(range(10), key=lambda j: (float(depth[j]), j))[:6]
    np.testing.assert_allclose(s[:, 3], depth[order] / np.float32(DHAT), rtol=1e-6)
    # the partner (closest point on the triangle) lies straight below the point
    np.testing.assert_allclose(s[:, 1], -depth[order] / np.float32(DHAT), rtol=1e-6)
    assert np.abs(s[:, [0, 2]]).max() < 1e-6
    assert (s[:, 4] == 1).all() and (s[:, 5] == 1).all()
    # each triangle vertex sees only its own record
    for v in range(1, x.shape[0]):
        sv = _slots(f, v)
        assert sv[0, 5] == 1.0 and (sv[1:] == 0).all()


def test_empty_and_invalid_records(oracle):
    x = np.random.default_rng(0).normal(size=(8, 3)).astype(np.float32)
    z = np.zeros((0, 4), np.int32)
    f = oracle.warmstart_features(x, None, z, z, DHAT)
    assert (f == 0).all()
    bad = np.array([[0, 1, 2, 99], [-1, 1, 2, 3]], np.int32)
    f = oracle.warmstart_features(x, None, bad, bad, DHAT)
    assert (f == 0).all()

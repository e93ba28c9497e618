"""Pins the oracle's closest-point routines (contact.py:230-302 and the
segment-segment stand-in for Warp's closest_point_edge_edge) against
brute-force minimisation and textbook cases."""
import numpy as np
import pytest


def _tri_sample(a, b, c, m=40000, seed=0):
    rng = np.random.default_rng(seed)
    u = rng.random((m, 2))
    u = np.where(u.sum(1, keepdims=True) > 1, 1 - u, u)
    edges = np.linspace(0, 1, 2001)[:, None]
    pts = [a + u[:, :1] * (b - a) + u[:, 1:] * (c - a), a + edges * (b - a), a + edges * (c - a), b + edges * (c - b)]
    return np.concatenate(pts)


def test_closest_point_triangle_vs_brute_force(oracle):
    rng = np.random.default_rng(1)
    p, a, b, c = (rng.normal(size=(150, 3)) for _ in range(4))
    out, typ = oracle.closest_point_triangle(p, a, b, c, double=True)
    for i in range(150):
        pts = _tri_sample(a[i], b[i], c[i], seed=i)
        dmin = np.linalg.norm(pts - p[i], axis=1).min()
        assert out[i, 2] <= dmin + 1e-12
        assert out[i, 2] >= dmin - 2e-3  # sampling resolution
        # returned weights reproduce the closest point and its distance
        wc = 1 - out[i, 0] - out[i, 1]
        q = out[i, 0] * a[i] + out[i, 1] * b[i] + wc * c[i]
        assert abs(np.linalg.norm(p[i] - q) - out[i, 2]) < 1e-12
        assert min(out[i, 0], out[i, 1], wc) >= -1e-12
    # every feature region is exercised
    assert set(np.unique(typ)) == set(range(7))


@pytest.mark.parametrize(
    "p,feature,dist",
    [
        ((0.2, 0.2, 0.5), 6, 0.5),  # above the face
        ((-1.0, -1.0, 0.0), 0, 2 ** 0.5),  # vertex a
        ((2.0, -0.5, 0.0), 1, (1 + 0.25) ** 0.5),  # vertex b
        ((-0.5, 2.0, 0.0), 2, (0.25 + 1) ** 0.5),  # vertex c
        ((0.5, -1.0, 0.0), 3, 1.0),  # edge ab
        ((-1.0, 0.5, 0.0), 4, 1.0),  # edge ac
        ((1.0, 1.0, 0.0), 5, 0.5 ** 0.5),  # edge bc
    ],
)
def test_closest_point_triangle_regions(oracle, p, feature, dist):
    a, b, c = np.array([0.0, 0, 0]), np.array([1.0, 0, 0]), np.array([0.0, 1, 0])
    for double in (False, True):
        out, typ = oracle.closest_point_triangle(np.array(p), a, b, c, double=double)
        assert typ[0] == feature
        assert abs(out[0, 2] - dist) < 1e-6


def test_closest_point_edge_edge_vs_brute_force(oracle):
    rng = np.random.default_rng(2)
    n = 200
    p1, q1, p2, q2 = (rng.normal(size=(n, 3)) for _ in range(4))
    out = oracle.closest_point_edge_edge(p1, q1, p2, q2, double=True)
    s = np.linspace(0, 1, 801)
    for i in range(n):
        A = p1[i] + s[:, None] * (q1[i] - p1[i])
        B = p2[i] + s[:, None] * (q2[i] - p2[i])
        d = np.linalg.norm(A[:, None, :] - B[None, :, :], axis=2)
        dmin = d.min()
        assert out[i, 2] <= dmin + 1e-12
        assert out[i, 2] >= dmin - 5e-3
        c1 = p1[i] + out[i, 0] * (q1[i] - p1[i])
        c2 = p2[i] + out[i, 1] * (q2[i] - p2[i])
        assert abs(np.linalg.norm(c1 - c2) - out[i, 2]) < 1e-12
        assert 0 <= out[i, 0] <= 1 and 0 <= out[i, 1] <= 1


def test_closest_point_edge_edge_special_cases(oracle):
    # crossing segments: distance is the offset along z
    out = oracle.closest_point_edge_edge([-1, 0, 0], [1, 0, 0], [0, -1, 0.3], [0, 1, 0.3], double=True)
    assert np.allclose(out[0], [0.5, 0.5, 0.3])
    # parallel segments (denominator 0): s = 0 and t from the projection
    out = oracle.closest_point_edge_edge([0, 0, 0], [1, 0, 0], [0.5, 1, 0], [1.5, 1, 0], double=True)
    assert out[0, 2] == pytest.approx(1.0)
    # degenerate first segment (a point)
    out = oracle.closest_point_edge_edge([0, 1, 0], [0, 1, 0], [-1, 0, 0], [1, 0, 0], double=True)
    assert np.allclose(out[0], [0.0, 0.5, 1.0])
    # float and double agree on a well-conditioned case
    rng = np.random.default_rng(3)
    p1, q1, p2, q2 = (rng.normal(size=(50, 3)) for _ in range(4))
    of = oracle.closest_point_edge_edge(p1, q1, p2, q2)
    od = oracle.closest_point_edge_edge(p1, q1, p2, q2, double=True)
    assert np.allclose(of[:, 2], od[:, 2], rtol=1e-5, atol=1e-6)

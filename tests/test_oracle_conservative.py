"""Pins the oracle's conservative-advancement alternates (ccd/ccd.py:7-163)
by the textbook distances, closed-form advancement answers and the
cubic-root CCD: the squared distances equal Ericson's closest points (and
brute force), a head-on approach stops at exactly 0.9 of the impact time,
and the advancement never passes the first root of the coplanarity cubic."""
import numpy as np
import pytest

from ipc_warp_amd import scenes


def test_segment_distance_matches_closest_points_and_brute_force(oracle):
    rng = np.random.default_rng(0)
    for _ in range(200):
        a0, a1, b0, b1 = rng.normal(size=(4, 3))
        d2 = oracle.segment_distance_sq(a0, a1, b0, b1)
        _, _, dist = oracle.closest_point_edge_edge(a0[None], a1[None], b0[None], b1[None], double=True)[0]
        assert d2 == pytest.approx(dist ** 2, rel=1e-9, abs=1e-14)
    # brute force on a grid, parallel and degenerate segments included
    cases = [rng.normal(size=(4, 3)) for _ in range(5)]
    a0, a1 = np.zeros(3), np.array([1.0, 0, 0])
    cases += [np.array([a0, a1, [0.2, 1, 0], [1.5, 1, 0]]), np.array([a0, a0, [0.3, 1, 2], [0.3, 1, 2]]),
              np.array([a0, a1, [0.5, 0.5, 0.5], [0.5, 0.5, 0.5]])]
    s = np.linspace(0, 1, 401)
    for a0, a1, b0, b1 in cases:
        P = a0 + s[:, None] * (a1 - a0)
        Q = b0 + s[:, None] * (b1 - b0)
        brute = ((P[:, None] - Q[None]) ** 2).sum(-1).min()
        d2 = oracle.segment_distance_sq(a0, a1, b0, b1)
        assert d2 <= brute + 1e-12 and d2 >= brute - 1e-4 * max(1.0, brute)


def test_triangle_distance_matches_closest_points(oracle):
    rng = np.random.default_rng(1)
    for _ in range(300):
        p, a, b, c = rng.normal(size=(4, 3))
        d2 = oracle.triangle_distance_sq(p, a, b, c)
        dist = oracle.closest_point_triangle(p[None], a[None], b[None], c[None], double=True)[0][0, 2]
        assert d2 == pytest.approx(dist ** 2, rel=1e-9, abs=1e-14)
    # a degenerate (collinear) triangle: the distance to its edges
    p = np.array([0.3, 1.0, 0.0])
    a, b, c = np.zeros(3), np.array([1.0, 0, 0]), np.array([2.0, 0, 0])
    assert oracle.triangle_distance_sq(p, a, b, c) == pytest.approx(1.0)


def _run(oracle, kind, x0, x1, double=True):
    quads = np.arange(4, dtype=np.int32)[None]
    return oracle.conservative_time(kind, np.asarray(x0, np.float32), np.asarray(x1, np.float32), quads, double)[0]


def test_head_on_approach_stops_at_nine_tenths(oracle):
    # a point falling onto a fixed triangle: relative speed = motion bound,
    # one step of 0.9 d/v, then the gap test ends it
    tri = np.array([[-1, -1, 0], [2, -1, 0], [-1, 2, 0]], float)
    x0 = np.vstack([[0.1, 0.2, 0.5], tri])
    x1 = x0.copy()
    x1[0, 2] = -0.5  # crosses z = 0 at t = 0.5
    assert _run(oracle, "pt", x0, x1) == pytest.approx(0.9 * 0.5, rel=1e-12)
    # an edge moving onto a crossing edge
    x0 = np.array([[-1, 0.1, 0.25], [1, 0.1, 0.25], [0.2, -1, 0], [0.2, 1, 0]], float)
    x1 = x0.copy()
    x1[:2, 2] -= 1.0  # reaches z = 0 at t = 0.25
    assert _run(oracle, "ee", x0, x1) == pytest.approx(0.9 * 0.25, rel=1e-6)
    # no motion, and moving apart
    assert _run(oracle, "ee", x0, x0) == 1.0
    x1 = x0.copy()
    x1[:2, 2] += 1.0
    assert _run(oracle, "ee", x0, x1) == 1.0


@pytest.mark.parametrize("kind", ["ee", "pt"])
def test_never_passes_the_first_root(oracle, kind):
    x0, x1, q = scenes.ccd_stream(kind, 4000, seed=3)[:3]
    ca = oracle.conservative_time(kind, x0, x1, q)
    cubic, _ = oracle.collision_time(kind, x0, x1, q)
    hit = cubic < 1.0
    assert hit.sum() > 100
    assert (ca <= cubic + 1e-12).all()
    assert (ca[~hit] <= 1.0).all() and (ca >= 0).all()

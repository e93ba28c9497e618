"""Pins the oracle's CCD restatement (ccd/ccd.py:212-357,
ccd/cubic_roots.py:1-193) against constructed-root polynomials, the 4x4
coplanarity determinant and straight-line impacts with analytic times."""
import numpy as np
import pytest


def _poly(roots, scale):
    c = np.poly(roots)[::-1] * scale  # ascending coefficients
    return np.concatenate([c, np.zeros(4 - len(c))])


def test_cubic_roots_constructed(oracle):
    # the generator pattern of cubic_roots.py:225-269, seeded and on [0, 1]
    rng = np.random.default_rng(12)
    for _ in range(300):
        r = np.sort(rng.random(3))
        if np.min(np.diff(r)) < 1e-3:
            continue
        coef = _poly(r, rng.random() + 0.1) * (1 if rng.random() < 0.5 else -1)
        n, roots, capped = oracle.cubic_roots(coef, 0.0, 1.0)
        assert capped == 0
        assert n == 3
        assert np.allclose(roots, r, atol=1e-9)


def test_cubic_roots_partial_intervals_and_quadratics(oracle):
    rng = np.random.default_rng(13)
    for _ in range(300):
        r = rng.uniform(-1.0, 2.0, 3)
        coef = _poly(r, rng.uniform(0.1, 3.0))
        n, roots, _ = oracle.cubic_roots(coef, 0.0, 1.0)
        inside = np.sort(r[(r >= 0) & (r <= 1)])
        # distinct simple roots are all found, in ascending order
        if len(inside) and np.min(np.abs(r[:, None] - r[None, :]) + np.eye(3)) > 1e-3 and \
                np.min(np.abs(np.concatenate([r, r - 1]))) > 1e-3:
            assert n == len(inside)
            assert np.allclose(roots[:n], inside, atol=1e-9)
        # every reported root is a root
        for t in roots[:n]:
            assert abs(np.polyval(coef[::-1], t)) < 1e-8 * max(1.0, np.abs(coef).max())
    # a quadratic (c3 = 0) with roots in [0, 1]
    n, roots, _ = oracle.cubic_roots(_poly([0.25, 0.75], 2.0)[:4], 0.0, 1.0)
    assert n == 2 and np.allclose(roots[:2], [0.25, 0.75])
    # no sign change, no roots
    assert oracle.cubic_roots([1.0, 0.0, 0.0, 0.5])[0] == 0


def _det4(p):
    M = np.ones((4, 4))
    M[:3, :] = np.asarray(p).reshape(4, 3).T
    return np.linalg.det(M)


@pytest.mark.parametrize("relative", [False, True])
def test_coplanarity_cubic_is_the_4x4_determinant(oracle, relative):
    rng = np.random.default_rng(14)
    for _ in range(50):
        # dyadic-integer inputs make both sides exact
        p0 = rng.integers(-8, 8, 12) / 4.0
        p1 = rng.integers(-8, 8, 12) / 4.0
        coef = oracle.coplanarity_cubic(p0, p1, relative=relative)
        for t in (0.0, 0.25, 0.5, 1.0, 1.5):
            pt = p0 * (1 - t) + p1 * t
            # det[p1-p0, p2-p0, p3-p0] = -det [[x];[1]] with points as columns
            ref = -_det4(pt)
            val = coef[0] + t * (coef[1] + t * (coef[2] + t * coef[3]))
            assert val == pytest.approx(ref, abs=1e-9)


def _pt_impact(rng, t_star):
    """Triangle moving rigidly, point crossing its interior at t_star."""
    tri0 = rng.normal(size=(3, 3))
    vel = rng.normal(size=3) * 0.5
    w = rng.dirichlet([3, 3, 3])
    n = np.cross(tri0[1] - tri0[0], tri0[2] - tri0[0])
    n /= np.linalg.norm(n)
    hit = w @ tri0 + t_star * vel  # interior point of the moved triangle
    vp = -n * rng.uniform(0.5, 2.0) + vel + rng.normal(size=3) * 0.1 * np.cross(n, vel)
    p0 = hit - t_star * vp
    x0 = np.vstack([p0, tri0])
    x1 = np.vstack([p0 + vp, tri0 + vel])
    return x0, x1


def test_collision_time_pt_straight_line_impacts(oracle):
    rng = np.random.default_rng(15)
    n = 200
    ts = rng.uniform(0.05, 0.95, n)
    X0, X1 = [], []
    for t in ts:
        a, b = _pt_impact(rng, t)
        X0.append(a)
        X1.append(b)
    X0 = np.concatenate(X0).astype(np.float32)
    X1 = np.concatenate(X1).astype(np.float32)
    quads = np.arange(4 * n, dtype=np.int32).reshape(n, 4)
    for relative in (False, True):
        toi, capped = oracle.collision_time("pt", X0, X1, quads, relative=relative)
        assert capped == 0
        assert np.allclose(toi, ts, atol=1e-5)


def test_collision_time_ee_straight_line_impacts(oracle):
    rng = np.random.default_rng(16)
    n = 200
    ts = rng.uniform(0.05, 0.95, n)
    X0, X1 = [], []
    for t in ts:
        a0, a1 = rng.normal(size=3), rng.normal(size=3)
        b_dir = np.cross(a1 - a0, rng.normal(size=3))
        s, u = rng.uniform(0.2, 0.8, 2)
        hit = a0 + s * (a1 - a0)
        vb = rng.normal(size=3)
        b0_t = hit - u * b_dir
        b1_t = hit + (1 - u) * b_dir
        X0.append(np.vstack([a0, a1, b0_t - t * vb, b1_t - t * vb]))
        X1.append(np.vstack([a0, a1, b0_t + (1 - t) * vb, b1_t + (1 - t) * vb]))
    X0 = np.concatenate(X0).astype(np.float32)
    X1 = np.concatenate(X1).astype(np.float32)
    quads = np.arange(4 * n, dtype=np.int32).reshape(n, 4)
    for relative in (False, True):
        toi, capped = oracle.collision_time("ee", X0, X1, quads, relative=relative)
        assert capped == 0
        assert np.allclose(toi, ts, atol=1e-5)


def test_collision_time_misses_return_one(oracle):
    rng = np.random.default_rng(17)
    n = 100
    x0 = rng.normal(size=(4 * n, 3)).astype(np.float32)
    x1 = x0.copy()
    # pure translation of all four points together never changes coplanarity
    x1 += np.repeat(rng.normal(size=(n, 1, 3)), 4, axis=1).reshape(-1, 3).astype(np.float32)
    quads = np.arange(4 * n, dtype=np.int32).reshape(n, 4)
    for kind in ("pt", "ee"):
        toi, _ = oracle.collision_time(kind, x0, x1, quads)
        assert np.all(toi == 1.0)


def test_pt_collision_before_toi_is_free(oracle):
    rng = np.random.default_rng(18)
    for _ in range(50):
        x0, x1 = _pt_impact(rng, rng.uniform(0.1, 0.9))
        q = np.arange(4, dtype=np.int32).reshape(1, 4)
        toi, _ = oracle.collision_time("pt", x0.astype(np.float32), x1.astype(np.float32), q)
        for t in np.linspace(0, 0.9 * toi[0], 20):
            xt = x0 * (1 - t) + x1 * t
            _, typ = oracle.closest_point_triangle(xt[0], xt[1], xt[2], xt[3], double=True)
            d = oracle.closest_point_triangle(xt[0], xt[1], xt[2], xt[3], double=True)[0][0, 2]
            assert d > 0


@pytest.mark.parametrize("kind", ["pt", "ee"])
def test_ccd_stream_impacts_are_where_they_were_built(oracle, kind):
    """The bench's synthetic CCD stream: ~30 % of the candidates are built to
    meet at a known time, the rest move apart; the oracle finds each built
    impact near its time (fp32 rounding of coordinates ~10 moves it by
    ~1e-6) and no TOI for nearly all of the rest."""
    from ipc_warp_amd import scenes
    x0, x1, q, t = scenes.ccd_stream(kind, 4096, seed=41)
    assert x0.dtype == np.float32 and q.shape == (4096, 4)
    assert 8.0 < np.abs(x0).max() < 11.0
    rel, capped = oracle.collision_time(kind, x0, x1, q, relative=True, threads=4)
    hits = t < 1
    assert 0.25 < hits.mean() < 0.35 and capped == 0
    assert (rel[hits] < 1).all()
    assert np.median(np.abs(rel[hits] - t[hits])) < 1e-5
    assert (rel[~hits] >= 1).mean() > 0.99

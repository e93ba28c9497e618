"""Pins of the oracle's feature dispatch and composite forces that do not
restate the oracle's formulas (CPU only).

* PT dispatch (dynamic_contacts.py:110-125 over contact.py:230-302): for
  points built inside each of the seven Voronoi regions of a triangle, the
  squared distance of the feature the dispatch picks equals the squared
  distance the point was built at and the double closest-point distance
  (rel 1e-12); a wrong type -> slot table entry gives another point, edge
  or plane and fails.
* EE dispatch (dynamic_contacts.py:48-108): outside the mollified band the
  picked feature's squared distance equals closest_point_edge_edge's in
  double (rel 1e-12), for interior, point-edge and point-point cases built
  at a known distance; inside the band the picked point-edge pair is the
  one whose parameter (s, 1-s, t, 1-t) is smallest and its squared
  distance is not below the segments' distance (rel 1e-8, the rounding of
  near-parallel closest-point parameters).
* Composite forces (dynamic_contacts.py:151-246,374-398): the force the
  oracle assembles, -(m b' grad d^2 + b grad m) for EE and -b' grad d^2 for
  PT, equals minus the central difference of its own pair energy m b (b)
  in double, for pairs inside and outside the band, and the pair Hessian is
  symmetric positive semidefinite.
"""
import numpy as np
import pytest

L0 = 1e-3
KAPPA = 1e8
N_SLOTS = {0: 2, 1: 3, 2: 4, 3: 4}  # PP, PE, PT, EE


def _feature_d2(oracle, feat, x12):
    kind, slots = feat
    d2, _, _ = oracle.feature_hessian(kind, slots[:N_SLOTS[kind]], x12)
    return d2


def _frame(rng):
    a = rng.standard_normal(3)
    a /= np.linalg.norm(a)
    b = np.cross(a, rng.standard_normal(3))
    b /= np.linalg.norm(b)
    return a, b, np.cross(a, b)


def _pt_in_region(rng, region, d):
    """A triangle (A, B, C) of size ~0.05 and a point at distance d whose
    closest feature is `region` (closest_point_triangle's type: 0-2 vertex
    A/B/C, 3 edge AB, 4 edge AC, 5 edge BC, 6 face)."""
    t1, t2, n = _frame(rng)
    ang = np.array([0.0, 2.1, 4.2]) + rng.uniform(-0.3, 0.3, 3)
    rad = 0.05 * rng.uniform(0.6, 1.0, 3)
    o = rng.uniform(-1, 1, 3)
    T = o + rad[:, None] * (np.cos(ang)[:, None] * t1 + np.sin(ang)[:, None] * t2)
    A, B, C = T
    if region == 6:
        w = rng.dirichlet([3, 3, 3])
        p = w @ T + d * n * rng.choice([-1, 1])
    elif region in (3, 4, 5):
        u, v, opp = {3: (A, B, C), 4: (A, C, B), 5: (B, C, A)}[region]
        q = u + rng.uniform(0.2, 0.8) * (v - u)
        eo = np.cross(v - u, n)
        eo /= np.linalg.norm(eo)
        eo *= np.sign(eo.dot(q - opp))
        phi = rng.uniform(-1.2, 1.2)
        p = q + d * (np.cos(phi) * eo + np.sin(phi) * n)
    else:
        v, o1, o2 = {0: (A, B, C), 1: (B, A, C), 2: (C, A, B)}[region]
        u = -(o1 - v) / np.linalg.norm(o1 - v) - (o2 - v) / np.linalg.norm(o2 - v)
        u /= np.linalg.norm(u)
        u = u + rng.uniform(-0.4, 0.4) * n
        p = v + d * u / np.linalg.norm(u)
    return np.concatenate([p, A, B, C])


@pytest.mark.parametrize("region", range(7))
def test_pt_dispatch_feature_distance_is_the_closest_distance(oracle, region):
    rng = np.random.default_rng(100 + region)
    for _ in range(40):
        d = L0 * 10 ** rng.uniform(-3, -0.05)
        x12 = _pt_in_region(rng, region, d)
        e, f, H, feat = oracle.contact_pair("pt", x12, L0, KAPPA)
        assert feat[0] >= 0
        x4 = x12.reshape(4, 3)
        wd, typ = oracle.closest_point_triangle(x4[0], x4[1], x4[2], x4[3], double=True)
        assert typ[0] == region
        d2 = _feature_d2(oracle, feat, x12)
        assert d2 == pytest.approx(wd[0, 2] ** 2, rel=1e-12)
        assert d2 == pytest.approx(d * d, rel=1e-9)


def _ee_case(rng, case, d):
    """Edges a = (x0, x1), b = (x2, x3) of length ~0.05 at distance d:
    'ee' interior crossing, 'pe_b' b's end point x3 over a's interior,
    'pp' x1 and x2 the closest pair, 'band' near-parallel (cross product
    below the mollifier threshold) with interior closest points."""
    t1, t2, n = _frame(rng)
    o = rng.uniform(-1, 1, 3)
    L = 0.05 * rng.uniform(0.6, 1.2)
    a0 = o - rng.uniform(0.2, 0.8) * L * t1
    a1 = a0 + L * t1
    if case == "ee":
        ang = rng.uniform(0.4, np.pi - 0.4)
        db = np.cos(ang) * t1 + np.sin(ang) * t2
        pb = o + d * n
        sb = rng.uniform(0.2, 0.8)
        b0, b1 = pb - sb * L * db, pb + (1 - sb) * L * db
    elif case == "pe_b":
        pb = o + d * n
        rise = 0.5 * rng.standard_normal() * t2 + n
        rise /= np.linalg.norm(rise)
        b1, b0 = pb, pb + L * rise
    elif case == "pp":
        # b starts beyond a1 along a's direction, offset by d, and leaves away
        gap = d * (np.cos(0.7) * t1 + np.sin(0.7) * n)
        b0 = a1 + gap
        b1 = b0 + L * (0.6 * t1 + 0.8 * n)
    else:
        ang = rng.uniform(1e-4, 5e-3)
        db = np.cos(ang) * t1 + np.sin(ang) * t2
        # the lines cross (seen along n) at a's parameter s0 and b's t0: the
        # closest points are interior, and any of the four ends can be nearest
        s0, t0 = rng.uniform(0.05, 0.95, 2)
        pb = a0 + s0 * L * t1 + d * n
        b0 = pb - t0 * L * db
        b1 = b0 + L * db
    return np.concatenate([a0, a1, b0, b1])


@pytest.mark.parametrize("case,kind", [("ee", 3), ("pe_b", 1), ("pp", 0)])
def test_ee_dispatch_outside_the_band_is_the_segment_distance(oracle, case, kind):
    rng = np.random.default_rng({"ee": 200, "pe_b": 201, "pp": 202}[case])
    for _ in range(40):
        d = L0 * 10 ** rng.uniform(-3, -0.05)
        x12 = _ee_case(rng, case, d)
        e, f, H, feat = oracle.contact_pair("ee", x12, L0, KAPPA)
        assert feat[0] == kind
        x4 = x12.reshape(4, 3)
        st = oracle.closest_point_edge_edge(x4[0], x4[1], x4[2], x4[3], double=True)[0]
        d2 = _feature_d2(oracle, feat, x12)
        assert d2 == pytest.approx(st[2] ** 2, rel=1e-12)
        assert d2 == pytest.approx(d * d, rel=1e-9)


def test_ee_dispatch_inside_the_band_picks_the_nearest_end(oracle):
    rng = np.random.default_rng(203)
    picked = set()
    for _ in range(80):
        d = L0 * 10 ** rng.uniform(-3, -0.05)
        x12 = _ee_case(rng, "band", d)
        x4 = x12.reshape(4, 3)
        ea, eb = x4[1] - x4[0], x4[3] - x4[2]
        assert np.sum(np.cross(ea, eb) ** 2) < 1e-3 * ea.dot(ea) * eb.dot(eb)
        e, f, H, feat = oracle.contact_pair("ee", x12, L0, KAPPA)
        assert feat[0] == 1
        s, t, dist = oracle.closest_point_edge_edge(x4[0], x4[1], x4[2], x4[3], double=True)[0]
        ends = [s, 1 - s, t, 1 - t]
        point_slot = int(np.argmin(ends))
        assert feat[1][0] == point_slot
        other = [2, 3] if point_slot < 2 else [0, 1]
        assert sorted(feat[1][1:3]) == other
        picked.add(point_slot)
        # the point-line distance of the picked pair is not below the
        # segments' distance (up to the rounding of the near-parallel
        # closest-point parameters)
        assert _feature_d2(oracle, feat, x12) >= dist * dist * (1 - 1e-8)
    assert picked == {0, 1, 2, 3}


def _fd_force(oracle, kind, x12, rest12, h):
    G = np.zeros(12)
    for j in range(12):
        e = np.zeros(12)
        e[j] = h
        ep = oracle.contact_pair(kind, x12 + e, L0, KAPPA, rest12=rest12)[0]
        em = oracle.contact_pair(kind, x12 - e, L0, KAPPA, rest12=rest12)[0]
        G[j] = (ep - em) / (2 * h)
    return -G


@pytest.mark.parametrize("case", ["ee", "pe_b", "pp", "band"])
def test_ee_force_is_minus_the_energy_gradient(oracle, case):
    rng = np.random.default_rng({"ee": 300, "pe_b": 301, "pp": 302, "band": 303}[case])
    n_band = 0
    while n_band < 12:
        d = L0 * 10 ** rng.uniform(-1.5, -0.1)
        x12 = _ee_case(rng, case, d)
        rest = x12.copy()
        if case == "band":
            # the picked end must not change within the difference steps: the
            # closest-point parameters of near-parallel lines move ~1/angle
            # times faster than the points, so keep cases whose nearest end
            # leads the next by 0.1
            s, t, _ = oracle.closest_point_edge_edge(*x12.reshape(4, 3), double=True)[0]
            ends = np.sort([s, 1 - s, t, 1 - t])
            if ends[1] - ends[0] < 0.1:
                continue
        e, f, H, feat = oracle.contact_pair("ee", x12, L0, KAPPA, rest12=rest)
        x4 = x12.reshape(4, 3)
        c = np.sum(np.cross(x4[1] - x4[0], x4[3] - x4[2]) ** 2)
        eps_x = 1e-3 * np.sum((x4[1] - x4[0]) ** 2) * np.sum((x4[3] - x4[2]) ** 2)
        assert (c < eps_x) == (case == "band")
        n_band += 1
        G = _fd_force(oracle, "ee", x12, rest, h=1e-4 * d)
        assert np.allclose(f, G, rtol=1e-5, atol=1e-6 * np.abs(G).max())
        assert np.abs(H - H.T).max() <= 1e-12 * np.abs(H).max()
        assert np.linalg.eigvalsh(H).min() >= -1e-9 * np.abs(H).max()


@pytest.mark.parametrize("region", [0, 4, 6])
def test_pt_force_is_minus_the_energy_gradient(oracle, region):
    rng = np.random.default_rng(400 + region)
    for _ in range(12):
        d = L0 * 10 ** rng.uniform(-1.5, -0.1)
        x12 = _pt_in_region(rng, region, d)
        e, f, H, feat = oracle.contact_pair("pt", x12, L0, KAPPA)
        G = _fd_force(oracle, "pt", x12, None, h=1e-4 * d)
        assert np.allclose(f, G, rtol=1e-5, atol=1e-6 * np.abs(G).max())
        assert np.linalg.eigvalsh(H).min() >= -1e-9 * np.abs(H).max()

"""Pins the oracle's pair enumeration (edge_edge_collision /
point_triangle_collision, contact.py:199-336; the CCD candidate sets of
ccd/ccd.py:360-427) against O(N^2) brute force on small scenes."""
import numpy as np
import pytest

from ipc_warp_amd import scenes


def _brute_ee(x, E, body, th, buf, excl):
    out = []
    pad = np.float32(th) + np.float32(buf)
    lo = np.minimum(x[E[:, 0]], x[E[:, 1]]) - pad
    hi = np.maximum(x[E[:, 0]], x[E[:, 1]]) + pad
    for i in range(len(E)):
        ov = np.all(lo[i] <= hi, axis=1) & np.all(lo <= hi[i], axis=1)
        for j in np.nonzero(ov)[0]:
            if j <= i:
                continue
            a, b = E[i], E[j]
            if len({a[0], a[1]} & {b[0], b[1]}) or (excl and body[a[0]] == body[b[0]]):
                continue
            out.append((i, j))
    return out


def _scene_small(seed):
    soup = scenes.two_cubes_on_ground(seed=seed, n=3, gap=4e-4, jitter=2e-4)
    return soup


def test_dcd_ee_matches_brute_force(oracle):
    soup = _scene_small(0)
    th, buf = 5e-4, 5e-4
    rec, e = oracle.dcd_ee(soup.x, soup.edges, soup.body, th, buf)
    cand = _brute_ee(soup.x, soup.edges, soup.body, th, buf, False)
    d = oracle.closest_point_edge_edge(*(soup.x[soup.edges[[c[k] for c in cand], m]].astype(np.float64)
                                         for k, m in ((0, 0), (0, 1), (1, 0), (1, 1))))
    ref = [c for c, dist in zip(cand, d[:, 2]) if np.float32(dist) < np.float32(th) * np.float32(2)]
    assert [tuple(v) for v in e] == ref
    assert len(ref) > 0
    assert np.array_equal(rec[:, :2], soup.edges[e[:, 0]]) and np.array_equal(rec[:, 2:], soup.edges[e[:, 1]])


def test_dcd_pt_matches_brute_force(oracle):
    soup = _scene_small(1)
    th, buf = 5e-4, 5e-4
    rec, e = oracle.dcd_pt(soup.x, soup.triangles, soup.body, th, buf)
    x, F = soup.x, soup.triangles
    ref = []
    for i in range(len(x)):
        for j in range(len(F)):
            if i in F[j]:
                continue
            out, _ = oracle.closest_point_triangle(x[i], x[F[j, 0]], x[F[j, 1]], x[F[j, 2]])
            if np.float32(out[0, 2]) < np.float32(th) * np.float32(2):
                ref.append((i, j))
    assert [tuple(v) for v in e] == ref
    assert len(ref) > 0
    # record order (p, t1, t0, t2) of contact.py:336
    assert np.array_equal(rec[:, 1], F[e[:, 1], 1]) and np.array_equal(rec[:, 2], F[e[:, 1], 0])


@pytest.mark.parametrize("excl", [False, True])
def test_ccd_candidates_match_brute_force(oracle, excl):
    soup = _scene_small(2)
    rng = np.random.default_rng(3)
    x0 = soup.x
    x1 = (x0 + rng.normal(size=x0.shape) * 0.05).astype(np.float32)
    pad = np.float32(1e-7)
    E = soup.edges
    rec, e = oracle.ccd_candidates_ee(x0, x1, E, soup.body, 1e-7, excl)
    pts = np.stack([x0[E[:, 0]], x0[E[:, 1]], x1[E[:, 0]], x1[E[:, 1]]])
    lo, hi = pts.min(0) - pad, pts.max(0) + pad
    ref = []
    for i in range(len(E)):
        ov = np.all(lo[i] <= hi, axis=1) & np.all(lo <= hi[i], axis=1)
        for j in np.nonzero(ov)[0]:
            if j > i and not len({E[i, 0], E[i, 1]} & {E[j, 0], E[j, 1]}) and not (excl and soup.body[E[i, 0]] == soup.body[E[j, 0]]):
                ref.append((i, j))
    assert [tuple(v) for v in e] == ref
    F = soup.triangles
    rec, e = oracle.ccd_candidates_pt(x0, x1, F, soup.body, 1e-7, excl)
    tp = np.stack([x0[F[:, k]] for k in range(3)] + [x1[F[:, k]] for k in range(3)])
    tlo, thi = tp.min(0) - pad, tp.max(0) + pad
    plo, phi = np.minimum(x0, x1) - pad, np.maximum(x0, x1) + pad
    ref = []
    for i in range(len(x0)):
        ov = np.all(plo[i] <= thi, axis=1) & np.all(tlo <= phi[i], axis=1)
        for j in np.nonzero(ov)[0]:
            if i not in F[j] and not (excl and soup.body[i] == soup.body[F[j, 0]]):
                ref.append((i, j))
    assert [tuple(v) for v in e] == ref
    assert np.array_equal(rec[:, 1:], F[e[:, 1]])


def test_exclude_same_body_filters(oracle):
    soup = _scene_small(4)
    rec_all, e_all = oracle.dcd_pt(soup.x, soup.triangles, soup.body, 5e-4, 5e-4, False)
    rec_x, e_x = oracle.dcd_pt(soup.x, soup.triangles, soup.body, 5e-4, 5e-4, True)
    keep = soup.body[e_all[:, 0]] != soup.body[soup.triangles[e_all[:, 1], 0]]
    assert np.array_equal(e_x, e_all[keep])

"""Pins the oracle's barrier, distance-derivative, mollifier and assembly
restatement (dynamic_contacts.py:19-246,374-398; the absent ipctkwp
functions restated from IPC) by closed forms, finite differences and
invariants. A dropped term, wrong sign or swapped slot fails one of these."""
import numpy as np
import pytest

from ipc_warp_amd import scenes


def test_barrier_closed_form_and_derivatives(oracle):
    D, kappa = 1e-6, 1e8
    x = np.geomspace(1e-12, D * 0.999, 40)
    out = oracle.barrier(x, D, kappa)
    ref = -kappa * (x - D) ** 2 * np.log(x / D)
    assert np.allclose(out[:, 0], ref, rtol=1e-13)
    assert np.all(out[:, 0] > 0) and np.all(out[:, 1] < 0) and np.all(out[:, 2] > 0)
    h = x * 1e-5
    p, m = oracle.barrier(x + h, D, kappa), oracle.barrier(x - h, D, kappa)
    assert np.allclose((p[:, 0] - m[:, 0]) / (2 * h), out[:, 1], rtol=1e-7)
    assert np.allclose((p[:, 1] - m[:, 1]) / (2 * h), out[:, 2], rtol=1e-7)
    # zero on and beyond dhat^2, C1 at the boundary
    assert np.all(oracle.barrier([D, 2 * D], D, kappa) == 0)
    assert abs(oracle.barrier([D * (1 - 1e-9)], D, kappa)[0, 1]) < 1e-6 * kappa * D


SLOTS = {"PP": [1, 3], "PE": [2, 0, 1], "PT": [0, 1, 2, 3], "EE": [0, 1, 2, 3]}


@pytest.mark.parametrize("kind", ["PP", "PE", "PT", "EE"])
def test_feature_distance_derivatives_by_fd(oracle, kind):
    rng = np.random.default_rng(4)
    for _ in range(10):
        x = rng.normal(size=12)
        d2, g, H = oracle.feature_hessian(kind, SLOTS[kind], x)
        eps = 1e-6
        G = np.zeros(12)
        HF = np.zeros((12, 12))
        for j in range(12):
            e = np.zeros(12)
            e[j] = eps
            dp, gp, _ = oracle.feature_hessian(kind, SLOTS[kind], x + e)
            dm, gm, _ = oracle.feature_hessian(kind, SLOTS[kind], x - e)
            G[j] = (dp - dm) / (2 * eps)
            HF[:, j] = (gp - gm) / (2 * eps)
        assert np.allclose(G, g, atol=1e-7 * max(1, np.abs(g).max()))
        assert np.allclose(HF, H, atol=1e-6 * max(1, np.abs(H).max()))
        assert np.allclose(H, H.T, atol=1e-12)


def test_feature_distances_match_geometry(oracle):
    rng = np.random.default_rng(5)
    for _ in range(20):
        x = rng.normal(size=(4, 3))
        # PP
        d2, _, H = oracle.feature_hessian("PP", [1, 3], x)
        assert d2 == pytest.approx(np.sum((x[1] - x[3]) ** 2))
        ref = np.zeros((12, 12))
        ref[3:6, 3:6] = ref[9:12, 9:12] = 2 * np.eye(3)
        ref[3:6, 9:12] = ref[9:12, 3:6] = -2 * np.eye(3)
        assert np.allclose(H, ref)
        # PT: point-plane distance
        n = np.cross(x[2] - x[1], x[3] - x[1])
        d2, _, _ = oracle.feature_hessian("PT", [0, 1, 2, 3], x)
        assert d2 == pytest.approx((np.dot(x[0] - x[1], n) / np.linalg.norm(n)) ** 2)
        # EE: line-line distance
        n = np.cross(x[1] - x[0], x[3] - x[2])
        d2, _, _ = oracle.feature_hessian("EE", [0, 1, 2, 3], x)
        assert d2 == pytest.approx((np.dot(x[2] - x[0], n) / np.linalg.norm(n)) ** 2)
        # PE: point-line distance with the point in slot 2
        e = x[1] - x[0]
        q = x[2] - x[0]
        d2, _, _ = oracle.feature_hessian("PE", [2, 0, 1], x)
        assert d2 == pytest.approx(np.sum((q - np.dot(q, e) / np.dot(e, e) * e) ** 2))


@pytest.mark.parametrize("kind", ["PE", "PT", "EE"])
def test_nsd_part_is_the_negative_spectrum(oracle, kind):
    rng = np.random.default_rng(6)
    for _ in range(20):
        x = rng.normal(size=12)
        _, _, H = oracle.feature_hessian(kind, SLOTS[kind], x)
        _, _, N = oracle.feature_hessian(kind, SLOTS[kind], x, nsd=True)
        ev, V = np.linalg.eigh(H)
        ref = (V * np.minimum(ev, 0)) @ V.T
        assert np.allclose(N, ref, atol=1e-10 * np.abs(H).max())
        assert np.linalg.eigvalsh(N).max() < 1e-10 * np.abs(H).max()
        assert np.linalg.eigvalsh(H - N).min() > -1e-10 * np.abs(H).max()
    # PP: the reference leaves the NSD part zero
    _, _, N = oracle.feature_hessian("PP", [0, 1], rng.normal(size=12), nsd=True)
    assert not N.any()


def test_mollifier_value_gradient_and_psd(oracle):
    rng = np.random.default_rng(7)
    for _ in range(20):
        x = rng.normal(size=(4, 3))
        # make the edges nearly parallel so the mollifier is active
        x[3] = x[2] + (x[1] - x[0]) * 0.9 + rng.normal(size=3) * 1e-2
        ea, eb = x[1] - x[0], x[3] - x[2]
        c = np.sum(np.cross(ea, eb) ** 2)
        eps_x = 1e-3 * ea.dot(ea) * eb.dot(eb) * 10
        m, g, H = oracle.ee_mollifier(x, eps_x)
        if c >= eps_x:
            assert m == 1 and not g.any()
            continue
        q = c / eps_x
        assert m == pytest.approx(q * (2 - q))
        h = 1e-7
        G = np.zeros(12)
        for j in range(12):
            e = np.zeros(12)
            e[j] = h
            G[j] = (oracle.ee_mollifier(x.reshape(12) + e, eps_x)[0] - oracle.ee_mollifier(x.reshape(12) - e, eps_x)[0]) / (2 * h)
        assert np.allclose(G, g, atol=1e-6 * max(1, np.abs(g).max()))
        assert np.linalg.eigvalsh(H).min() > -1e-9 * max(1, np.abs(H).max())


def _pair_hessian(vals, i):
    return vals[16 * i:16 * i + 16].reshape(4, 4, 3, 3).transpose(0, 2, 1, 3).reshape(12, 12)


@pytest.mark.parametrize("kind", ["pt", "ee"])
def test_assembled_hessian_is_symmetric_psd_and_force_is_minus_gradient(oracle, kind):
    x, c = scenes.pair_stream(kind, 300, seed=8)
    l0 = np.float32(1e-3)
    kappa = 1e8
    rows, cols, vals, b, act = oracle.contact_hessian(kind, x, c, l0, kappa)
    assert act.all()
    D = float(l0) ** 2
    for i in range(0, 300, 7):
        H = _pair_hessian(vals, i)
        assert np.abs(H - H.T).max() <= 1e-12 * np.abs(H).max()
        assert np.linalg.eigvalsh(H).min() >= -1e-9 * np.abs(H).max()
        # rows/cols follow the reference's index formula
        for ii in range(4):
            for jj in range(4):
                assert rows[16 * i + 4 * ii + jj] == c[i, ii] and cols[16 * i + 4 * ii + jj] == c[i, jj]
    if kind == "pt":
        # force = -b'(d^2) grad d^2 of the dispatched feature: check by FD of
        # the energy of the same feature in double
        types = oracle.closest_point_triangle(x[c[:, 0]], x[c[:, 1]], x[c[:, 2]], x[c[:, 3]])[1]
        feats = {0: ("PP", [0, 1]), 1: ("PP", [0, 2]), 2: ("PP", [0, 3]), 3: ("PE", [0, 1, 2]),
                 4: ("PE", [0, 1, 3]), 5: ("PE", [0, 2, 3]), 6: ("PT", [0, 1, 2, 3])}
        for i in range(0, 300, 13):
            kindf, sl = feats[int(types[i])]
            x12 = x[c[i]].astype(np.float64).reshape(12)

            def energy(y):
                d2, _, _ = oracle.feature_hessian(kindf, sl, y)
                return oracle.barrier([d2], D, kappa)[0, 0]

            d2, g, Hd = oracle.feature_hessian(kindf, sl, x12)
            bb = oracle.barrier([d2], D, kappa)[0]
            h = 1e-9
            G = np.array([(energy(x12 + h * e) - energy(x12 - h * e)) / (2 * h) for e in np.eye(12)])
            f = b[c[i]].reshape(12)
            assert np.allclose(-f, G, rtol=1e-4, atol=1e-6 * np.abs(G).max())
            # unfiltered Hessian minus the assembled one is b' * PSD(H_d2) (NSD)
            H = _pair_hessian(vals, i)
            full = bb[2] * np.outer(g, g) + bb[1] * Hd
            assert np.linalg.eigvalsh(full - H).max() <= 1e-9 * np.abs(H).max()


def test_energy_matches_per_pair_barrier(oracle):
    x, c = scenes.pair_stream("pt", 200, seed=9)
    l0 = np.float32(1e-3)
    e = oracle.contact_energy("pt", x, c, l0, 1e8)
    types = oracle.closest_point_triangle(x[c[:, 0]], x[c[:, 1]], x[c[:, 2]], x[c[:, 3]])[1]
    feats = {0: ("PP", [0, 1]), 1: ("PP", [0, 2]), 2: ("PP", [0, 3]), 3: ("PE", [0, 1, 2]),
             4: ("PE", [0, 1, 3]), 5: ("PE", [0, 2, 3]), 6: ("PT", [0, 1, 2, 3])}
    tot = 0.0
    for i in range(200):
        k, sl = feats[int(types[i])]
        d2, _, _ = oracle.feature_hessian(k, sl, x[c[i]].astype(np.float64))
        tot += oracle.barrier([d2], float(l0) ** 2, 1e8)[0, 0]
    assert e == pytest.approx(tot, rel=1e-12)
    # the "ref32" mode differs only in using the fp32 distance for d^2
    e32 = oracle.contact_energy("pt", x, c, l0, 1e8, mode="ref32")
    assert e32 == pytest.approx(e, rel=1e-2)


def test_inactive_pairs_leave_zero_triplets(oracle):
    x, c = scenes.pair_stream("ee", 64, seed=10)
    x = x.copy()
    # push every odd pair's edge b far away
    for i in range(1, 64, 2):
        x[c[i, 2:]] += np.float32(1.0)
    rows, cols, vals, b, act = oracle.contact_hessian("ee", x, c, np.float32(1e-3), 1e8)
    assert act[::2].all() and not act[1::2].any()
    for i in range(1, 64, 2):
        assert not vals[16 * i:16 * i + 16].any() and not rows[16 * i:16 * i + 16].any()
        assert not b[c[i]].any()


def test_near_parallel_ee_uses_mollified_gauss_newton_term(oracle):
    x, c = scenes.pair_stream("ee", 400, seed=11)
    rows, cols, vals, b, act = oracle.contact_hessian("ee", x, c, np.float32(1e-3), 1e8)
    n_par = 0
    for i in range(400):
        xi = x[c[i]].astype(np.float64)
        ea, eb = xi[1] - xi[0], xi[3] - xi[2]
        eps_x = 1e-3 * ea.dot(ea) * eb.dot(eb)
        if np.sum(np.cross(ea, eb) ** 2) < eps_x:
            n_par += 1
            H = _pair_hessian(vals, i)
            assert np.linalg.matrix_rank(H, tol=1e-8 * np.abs(H).max()) == 1
    assert n_par > 10

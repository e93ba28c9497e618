"""Pins the oracle's elastic tets (fem/fem.py:21-34,138-203, fem/snh.py,
time_integrator_base.py:124-137): rest shapes against numpy; force =
-grad of the energy and blocks = d(-force)/dx by central differences; the
stable Neo-Hookean rest state (zero stress, and a stiffness equal to linear
elasticity's with the same Lamé parameters); rotation invariance; the
triplet index formula."""
import numpy as np
import pytest

MU, LAM = 2.0e6, 2.0e7  # fem/params.py:7


def _tets(seed, n=6, def_scale=0.05):
    rng = np.random.default_rng(seed)
    base = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    X, T = [], []
    for e in range(n):
        R, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        X.append((base @ R.T) * rng.uniform(0.5, 2.0) + rng.normal(size=3) * 3)
        T.append(np.arange(4) + 4 * e)
    X = np.concatenate(X).astype(np.float32)
    x = (X + def_scale * rng.normal(size=X.shape)).astype(np.float32)
    return X, x, np.array(T, np.int32)


def test_rest_shapes(oracle):
    X, _, T = _tets(0)
    Bm, W = oracle.tet_rest(X, T)
    for e in range(T.shape[0]):
        P = X[T[e]].astype(np.float64)
        Dm = np.stack([P[0] - P[3], P[1] - P[3], P[2] - P[3]], axis=1)
        assert np.allclose(Bm[e] @ Dm, np.eye(3), atol=1e-5)
        assert W[e] == pytest.approx(abs(np.linalg.det(Dm)) / 6, rel=1e-5)


def test_force_is_minus_energy_gradient_and_blocks_its_jacobian(oracle):
    X, x, T = _tets(1, n=3)
    Bm, W = oracle.tet_rest(X, T)
    rows, cols, vals, f = oracle.tet_hessian(x, T, Bm, W, MU, LAM)
    x64 = x.astype(np.float64)

    def energy(y):
        return float(oracle.tet_energy(y, T, Bm, W, MU, LAM).astype(np.float64).sum())

    def force(y):
        return oracle.tet_hessian(y, T, Bm, W, MU, LAM)[3].astype(np.float64)

    h = 2e-3
    G = np.zeros_like(x64)
    Kfd = np.zeros((x.size, x.size))
    for v in range(x.shape[0]):
        for c in range(3):
            e = np.zeros_like(x64)
            e[v, c] = h
            G[v, c] = (energy(x64 + e) - energy(x64 - e)) / (2 * h)
            Kfd[:, 3 * v + c] = -(force(x64 + e) - force(x64 - e)).reshape(-1) / (2 * h)
    assert np.allclose(-f, G, rtol=2e-3, atol=2e-3 * np.abs(G).max())
    K = np.zeros((x.size, x.size))
    for q in range(rows.size):
        K[3 * rows[q]:3 * rows[q] + 3, 3 * cols[q]:3 * cols[q] + 3] += vals[q]
    assert np.allclose(K, Kfd, rtol=2e-3, atol=2e-3 * np.abs(K).max())
    assert np.allclose(K, K.T, atol=1e-4 * np.abs(K).max())
    # index formula: block 16 e + 4 i + j is (T[e][i], T[e][j])
    e, i, j = np.divmod(np.arange(rows.size), 16)[0], (np.arange(rows.size) // 4) % 4, np.arange(rows.size) % 4
    assert np.array_equal(rows, T[e, i]) and np.array_equal(cols, T[e, j])


def test_rest_state_is_stress_free_linear_elasticity(oracle):
    X, _, T = _tets(2, n=2)
    Bm, W = oracle.tet_rest(X, T)
    rows, cols, vals, f = oracle.tet_hessian(X, T, Bm, W, MU, LAM)
    assert np.abs(f).max() <= 1e-6 * MU
    for e in range(T.shape[0]):
        # B: d(eps)/dx of the tet's linear shape functions; C: 2 mu I + lam 1 1^T (Voigt, engineering shear)
        G = np.vstack([Bm[e].astype(np.float64), -Bm[e].astype(np.float64).sum(0)])  # grad of shape fns, rows = vertices
        Bv = np.zeros((6, 12))
        for a in range(4):
            gx, gy, gz = G[a]
            Bv[:, 3 * a:3 * a + 3] = [[gx, 0, 0], [0, gy, 0], [0, 0, gz], [gy, gx, 0], [0, gz, gy], [gz, 0, gx]]
        C = np.zeros((6, 6))
        C[:3, :3] = LAM
        C[np.arange(3), np.arange(3)] += 2 * MU
        C[np.arange(3, 6), np.arange(3, 6)] = MU
        Kref = W[e] * Bv.T @ C @ Bv
        K = np.zeros((12, 12))
        for q in range(16 * e, 16 * e + 16):
            i, j = (q // 4) % 4, q % 4
            K[3 * i:3 * i + 3, 3 * j:3 * j + 3] = vals[q]
        assert np.allclose(K, Kref, rtol=1e-4, atol=1e-4 * np.abs(Kref).max())


def test_energy_is_rotation_invariant(oracle):
    X, x, T = _tets(3, n=4, def_scale=0.2)
    Bm, W = oracle.tet_rest(X, T)
    R, _ = np.linalg.qr(np.random.default_rng(4).normal(size=(3, 3)))
    if np.linalg.det(R) < 0:
        R[:, 0] *= -1
    e0 = oracle.tet_energy(x, T, Bm, W, MU, LAM).astype(np.float64)
    e1 = oracle.tet_energy((x.astype(np.float64) @ R.T).astype(np.float32), T, Bm, W, MU, LAM).astype(np.float64)
    assert np.allclose(e0, e1, rtol=1e-4, atol=1e-4 * np.abs(e0).max())

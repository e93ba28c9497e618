"""The Newton harness's Neo-Hookean term (ipc_warp_amd/integrator.py):
energy zero at rest, gradient by finite differences, PSD Hessian whose
unprojected part matches finite differences of the gradient."""
import numpy as np
import torch

from ipc_warp_amd import scenes
from ipc_warp_amd.integrator import NeoHookeanTets, lame


def _setup(seed=0):
    V, T = scenes.tet_grid_cube(2, 1.0)
    rest = torch.from_numpy(V)
    mu, lam = lame(1e5, 0.4)
    nh = NeoHookeanTets(rest, torch.from_numpy(T), mu, lam)
    rng = np.random.default_rng(seed)
    x = rest + torch.from_numpy(rng.uniform(-0.05, 0.05, V.shape))
    return nh, rest, x


def test_rest_state_is_stress_free():
    nh, rest, _ = _setup()
    assert abs(nh.energy(rest)) < 1e-6
    g, H = nh.gradient_hessian(rest)
    assert float(g.abs().max()) < 1e-6


def test_gradient_by_finite_differences():
    nh, rest, x = _setup(1)
    g, _ = nh.gradient_hessian(x)
    rng = np.random.default_rng(2)
    h = 1e-6
    for _ in range(10):
        d = torch.from_numpy(rng.standard_normal(x.shape))
        fd = (nh.energy(x + h * d) - nh.energy(x - h * d)) / (2 * h)
        assert abs(fd - float((g * d).sum())) <= 1e-5 * max(1.0, abs(fd))


def test_projected_hessian_is_psd_and_exact_when_convex():
    nh, rest, x = _setup(3)
    _, H = nh.gradient_hessian(x)
    ev = torch.linalg.eigvalsh(H)
    assert float(ev.min()) >= -1e-6 * float(ev.abs().max())
    # near rest NH is convex, so projection changes nothing: compare with FD of the gradient
    xs = rest + 1e-3 * (x - rest)
    g0, H0 = nh.gradient_hessian(xs)
    T = nh.T
    K = torch.zeros((xs.numel(), xs.numel()), dtype=torch.float64)
    for t in range(T.shape[0]):
        dofs = torch.stack([3 * T[t] + k for k in range(3)], 1).reshape(-1)
        K[dofs[:, None], dofs[None, :]] += H0[t]
    h = 1e-7
    for j in range(0, xs.numel(), 7):
        e = torch.zeros(xs.numel(), dtype=torch.float64)
        e[j] = h
        gp, _ = nh.gradient_hessian(xs + e.reshape(xs.shape))
        gm, _ = nh.gradient_hessian(xs - e.reshape(xs.shape))
        col = ((gp - gm) / (2 * h)).reshape(-1)
        # central differences at h = 1e-7 resolve ~1e-5 of the matrix scale
        assert float((col - K[:, j]).abs().max()) <= 1e-4 * float(K.abs().max())

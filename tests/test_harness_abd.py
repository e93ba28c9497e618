"""The affine-body harness's own terms on the CPU: mass properties of a
closed box against its closed forms, and the inertia + orthogonality
potential's gradient and Hessian (affine_body_dynamics.py:86-141,200-238)
against central differences."""
import numpy as np
import torch

from ipc_warp_amd import scenes
from ipc_warp_amd.abd import AbdIntegrator, mass_moments, mass_properties


def _box(lo, hi):
    x = np.array([[lo[0], lo[1], lo[2]], [hi[0], lo[1], lo[2]], [hi[0], hi[1], lo[2]], [lo[0], hi[1], lo[2]],
                  [lo[0], lo[1], hi[2]], [hi[0], lo[1], hi[2]], [hi[0], hi[1], hi[2]], [lo[0], hi[1], hi[2]]], float)
    f = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4], [2, 3, 7], [2, 7, 6],
                  [1, 2, 6], [1, 6, 5], [0, 4, 7], [0, 7, 3]])
    return x, f


def test_mass_properties_of_boxes():
    x1, f1 = _box((0, 0, 0), (1, 2, 3))
    x2, f2 = _box((5, 5, 5), (6, 6, 6))
    x = np.concatenate([x1, x2])
    f = np.concatenate([f1, f2 + 8])
    body = np.repeat([0, 1], 8)
    c, v, w = mass_properties(x, f, body)
    assert np.allclose(c, [[0.5, 1.0, 1.5], [5.5, 5.5, 5.5]]) and np.allclose(v, [6.0, 1.0])
    # area weights: each box's total surface area
    assert np.isclose(w[:8].sum(), 2 * (2 + 6 + 3)) and np.isclose(w[8:].sum(), 6.0)
    M = mass_moments(x - c[body], body, w, v, 1000.0)
    assert np.isclose(M[0, 0, 0], 6000.0) and np.isclose(M[1, 0, 0], 1000.0)
    assert np.allclose(M, M.transpose(0, 2, 1))
    # centred: the first moments vanish for these symmetric vertex weights
    assert np.allclose(M[:, 0, 1:], 0.0, atol=1e-9)


def _cpu_harness(nb=3, seed=0):
    rng = np.random.default_rng(seed)
    I = AbdIntegrator.__new__(AbdIntegrator)
    I.nb, I.h = nb, 1e-2
    I.project_body_hessian = False
    R = rng.normal(size=(nb, 4, 4))
    I.M = torch.from_numpy(R @ R.transpose(0, 2, 1) + 4 * np.eye(4))
    I.kv = torch.from_numpy(rng.uniform(1e5, 1e6, nb))
    I.free = torch.tensor([True] * (nb - 1) + [False])
    q = np.zeros((nb, 4, 3))
    q[:, 0] = rng.normal(size=(nb, 3))
    q[:, 1:] = np.eye(3) + 0.05 * rng.normal(size=(nb, 3, 3))
    qt = q + 0.01 * rng.normal(size=q.shape)
    return I, torch.from_numpy(q), torch.from_numpy(qt)


def test_body_terms_gradient_and_hessian_by_central_differences():
    I, q, qt = _cpu_harness()
    e, g, H = I._body_terms(q, qt, True)
    eps = 1e-6
    n = q.numel()
    G = np.zeros(n)
    HF = np.zeros((n, n))
    for k in range(n):
        d = torch.zeros(n, dtype=torch.float64)
        d[k] = eps
        ep, gp, _ = I._body_terms(q + d.reshape(q.shape), qt, False)
        em, gm, _ = I._body_terms(q - d.reshape(q.shape), qt, False)
        G[k] = (ep - em) / (2 * eps)
        HF[:, k] = ((gp - gm) / (2 * eps)).reshape(-1).numpy()
    gn = g.reshape(-1).numpy()
    assert np.allclose(G, gn, rtol=1e-6, atol=1e-6 * np.abs(gn).max())
    for b in range(I.nb - 1):
        blk = slice(12 * b, 12 * b + 12)
        assert np.allclose(H[b].numpy(), HF[blk, blk], rtol=1e-5, atol=1e-5 * np.abs(HF[blk, blk]).max())
    # the fixed body: identity system row, zero gradient
    assert np.array_equal(H[-1].numpy(), np.eye(12)) and not gn[-12:].any()


def test_projected_body_hessian_is_psd():
    I, q, qt = _cpu_harness(seed=3)
    q = q.clone()
    q[:, 1:] *= 0.8  # compressed frames: the orthogonality Hessian turns indefinite
    _, _, H = I._body_terms(q, qt, True)
    assert min(float(torch.linalg.eigvalsh(H[b]).min()) for b in range(I.nb)) < 0
    I.project_body_hessian = True
    _, _, Hp = I._body_terms(q, qt, True)
    assert min(float(torch.linalg.eigvalsh(Hp[b]).min()) for b in range(I.nb)) >= -1e-9 * float(Hp.abs().max())
    # unchanged where it was PSD already: orthonormal frames
    q2 = q.clone()
    q2[:, 1:] = torch.eye(3, dtype=q.dtype)
    I.project_body_hessian = False
    _, _, H2 = I._body_terms(q2, qt, True)
    I.project_body_hessian = True
    _, _, H2p = I._body_terms(q2, qt, True)
    assert torch.allclose(H2, H2p, rtol=1e-9, atol=1e-9 * float(H2.abs().max()))

"""GPU elastic tets (fem/fem.py:21-34,138-203; time_integrator_base.py:
124-137) against the oracle's per-thread restatement of tet_kernel_sparse:
rest shapes, triplet rows/cols and blocks bit-exact (the same operations in
the same order), per-tet energies within fp32 log rounding, forces within
the rounding of their float atomics; on the cube-drop cubes and on the
twisted rods' tets."""
import numpy as np
import pytest
import torch

from ipc_warp_amd import ops, scenes

pytestmark = pytest.mark.gpu

MU, LAM = 2.0e6, 2.0e7


def _case(name):
    if name == "cubes":
        soup = scenes.two_cubes_on_ground(seed=11, n=8, gap=4e-4, jitter=2e-4)
        rng = np.random.default_rng(5)
        x = soup.x + (1e-2 * rng.normal(size=soup.x.shape)).astype(np.float32)
        return soup.xcs, x.astype(np.float32), soup.tets
    soup, _, _ = scenes.rod_bundle(seed=0)
    T = soup.tets[:40000]
    return soup.xcs, soup.x, T


@pytest.mark.parametrize("name", ["cubes", "rods"])
def test_tet_kernels_match_oracle(dev, oracle, name):
    xcs, x, T = _case(name)
    T = np.ascontiguousarray(T, np.int32)
    n = T.shape[0]
    xcs_t, x_t, T_t = (torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in (xcs, x, T))
    Bm, W = ops.tet_rest(xcs_t, T_t)
    rBm, rW = oracle.tet_rest(xcs, T)
    torch.cuda.synchronize()
    assert np.array_equal(Bm.cpu().numpy().view(np.uint32), rBm.view(np.uint32))
    assert np.array_equal(W.cpu().numpy().view(np.uint32), rW.view(np.uint32))
    rows = torch.empty(16 * n, dtype=torch.int32, device=dev)
    cols = torch.empty_like(rows)
    vals = torch.empty((16 * n, 3, 3), dtype=torch.float32, device=dev)
    b = torch.zeros_like(x_t)
    ops.tet_hessian(x_t, T_t, Bm, W, MU, LAM, rows, cols, vals, b)
    psi = torch.empty(n, dtype=torch.float32, device=dev)
    e = torch.zeros(1, dtype=torch.float64, device=dev)
    ops.tet_energy(x_t, T_t, Bm, W, MU, LAM, psi, e)
    torch.cuda.synchronize()
    rr, rc, rv, rf = oracle.tet_hessian(x, T, rBm, rW, MU, LAM)
    assert np.array_equal(rows.cpu().numpy(), rr) and np.array_equal(cols.cpu().numpy(), rc)
    assert np.array_equal(vals.cpu().numpy().view(np.uint32), rv.view(np.uint32))
    gf = b.cpu().numpy()
    assert np.abs(gf - rf).max() <= 1e-5 * np.abs(rf).max()
    rpsi = oracle.tet_energy(x, T, rBm, rW, MU, LAM)
    assert np.allclose(psi.cpu().numpy(), rpsi, rtol=2e-6, atol=1e-6 * np.abs(rpsi).max())
    assert e.item() == pytest.approx(float(psi.cpu().numpy().astype(np.float64).sum()), rel=1e-12)


def test_out_of_range_tets_write_zeros(dev):
    x = torch.zeros((4, 3), dtype=torch.float32, device=dev)
    x[1, 0] = x[2, 1] = x[3, 2] = 1.0
    T = torch.tensor([[0, 1, 2, 3], [0, 1, 2, 9]], dtype=torch.int32, device=dev)
    Bm, W = ops.tet_rest(x, T)
    rows = torch.full((32,), -1, dtype=torch.int32, device=dev)
    cols = torch.full((32,), -1, dtype=torch.int32, device=dev)
    vals = torch.full((32, 3, 3), float("nan"), dtype=torch.float32, device=dev)
    ops.tet_hessian(x, T, Bm, W, MU, LAM, rows, cols, vals)
    torch.cuda.synchronize()
    assert W[1].item() == 0.0 and (rows[16:] == 0).all() and (vals[16:] == 0).all()
    assert torch.isfinite(vals[:16]).all()

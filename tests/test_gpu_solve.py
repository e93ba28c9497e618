"""GPU BSR product and conjugate gradients (warp.sparse.bsr_mv and
warp.optim.linear.cg as called at time_integrator_base.py:500-507) against
scipy in fp64, on the pressed two-cube contact Hessian plus a lumped mass
(an SPD Newton system). Bars: the product within fp32 rounding of the sum of
|terms| per row; CG stops with |b - A x| <= tol |b| as scipy evaluates it."""
import numpy as np
import pytest
import scipy.sparse as sp
import torch

from ipc_warp_amd import scenes, sparse
from ipc_warp_amd.contact import ContactSolverBase

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def system(dev):
    soup = scenes.two_cubes_on_ground(seed=11, n=8, gap=4e-4, jitter=2e-4)
    x = torch.from_numpy(soup.x).to(dev)
    S = ContactSolverBase(x, x.clone(), torch.from_numpy(soup.edges), torch.from_numpy(soup.triangles),
                          torch.from_numpy(soup.body), thickness=5e-4, buffer_=5e-4, kappa=1e8)
    S.detect_collision()
    trip = S.compute_collision_hessian(torch.zeros_like(x))
    V = soup.x.shape[0]
    h2, mass = 1e-4, 1.0
    idx = torch.arange(V, dtype=torch.int32, device=dev)
    rows = torch.cat([trip.rows, idx])
    cols = torch.cat([trip.cols, idx])
    vals = torch.cat([h2 * trip.vals, mass * torch.eye(3, device=dev).expand(V, 3, 3)]).contiguous()
    A = sparse.bsr_from_triplets(V, V, rows, cols, vals)
    torch.cuda.synchronize()
    nnz = A.nnz_sync()
    o, c, v = A.offsets.cpu().numpy(), A.columns[:nnz].cpu().numpy(), A.values[:nnz].cpu().numpy().astype(np.float64)
    ref = sp.bsr_matrix((v, c, o), shape=(3 * V, 3 * V)).tocsr()
    return A, ref, V


def test_bsr_mv_matches_scipy(dev, system):
    A, ref, V = system
    rng = np.random.default_rng(0)
    x = rng.normal(size=(V, 3)).astype(np.float32)
    y0 = rng.normal(size=(V, 3)).astype(np.float32)
    xt = torch.from_numpy(x).to(dev)
    absA = abs(ref)
    bound = 4e-7 * (absA @ np.abs(x.reshape(-1).astype(np.float64))) + 1e-30
    for alpha, beta in ((1.0, 0.0), (0.5, 2.0)):
        y = torch.from_numpy(y0.copy()).to(dev)
        sparse.bsr_mv(A, xt, y, alpha, beta)
        torch.cuda.synchronize()
        want = alpha * (ref @ x.reshape(-1).astype(np.float64)) + beta * y0.reshape(-1)
        err = np.abs(y.cpu().numpy().reshape(-1) - want)
        assert (err <= abs(alpha) * bound + 4e-7 * abs(beta) * np.abs(y0.reshape(-1)) + 1e-30).all()


@pytest.mark.parametrize("jacobi", [False, True])
def test_cg_solves_the_newton_system(dev, system, jacobi):
    A, ref, V = system
    rng = np.random.default_rng(1)
    b = rng.normal(size=(V, 3)).astype(np.float32)
    bt = torch.from_numpy(b).to(dev)
    x = torch.zeros_like(bt)
    it, res, goal = sparse.cg(A, bt, x, tol=1e-5, maxiter=2000, jacobi=jacobi)
    torch.cuda.synchronize()
    r = b.reshape(-1).astype(np.float64) - ref @ x.cpu().numpy().reshape(-1).astype(np.float64)
    assert 0 < it < 2000
    assert res <= goal
    # the true residual of the fp32 iterate, evaluated in fp64
    assert np.linalg.norm(r) <= 2 * 1e-5 * np.linalg.norm(b)


def test_cg_zero_rhs_and_warm_start(dev, system):
    A, ref, V = system
    b = torch.zeros((V, 3), dtype=torch.float32, device=dev)
    x = torch.zeros_like(b)
    it, res, _ = sparse.cg(A, b, x)
    assert it == 0 and res == 0.0 and not x.any()
    # started at the solution: converged before the first iteration
    rng = np.random.default_rng(2)
    xs = torch.from_numpy(rng.normal(size=(V, 3)).astype(np.float32)).to(dev)
    b = sparse.bsr_mv(A, xs)
    it, res, goal = sparse.cg(A, b, xs.clone(), tol=1e-4)
    assert it == 0 and res <= goal


def test_cg_graph_replays_are_identical(dev, system):
    """The captured iterations replayed by repeated solves, with eager work
    and synchronisations between them, give the same bits each time (the
    sums are in a fixed order) as the first solve."""
    A, ref, V = system
    rng = np.random.default_rng(3)
    b = torch.from_numpy(rng.normal(size=(V, 3)).astype(np.float32)).to(dev)
    outs = []
    for k in range(3):
        x = torch.zeros_like(b)
        it, res, goal = sparse.cg(A, b, x, tol=1e-5, maxiter=2000, jacobi=True)
        torch.cuda.synchronize()
        outs.append((it, res, x.cpu().numpy().copy()))
        # eager work between replays
        junk = torch.randn(1 << 20, device=dev).sum()
        sparse.bsr_mv(A, b)
        torch.cuda.synchronize()
        assert torch.isfinite(junk)
    for it, res, x in outs[1:]:
        assert it == outs[0][0] and res == outs[0][1]
        assert np.array_equal(x.view(np.uint32), outs[0][2].view(np.uint32))

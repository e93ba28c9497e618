"""GPU affine-body contact reduction (affine_body_dynamics.py:144-190)
against the oracle's fp64 restatement, on the contact Hessian of the
64-icosphere grid (one body fixed) and of the screw in its fixed nut.

Bars: the DOF-level pattern (offsets, columns) exact; every value within
1e-5 of the sum of the absolute values of its terms (the fp32 summation
bound: ~100 fp32 roundings per value, each 6e-8) plus fp32 rounding of the
result; the reduced force within 1e-5 of its largest entry."""
import numpy as np
import pytest
import torch

from ipc_warp_amd import scenes
from ipc_warp_amd.abd import AbdReducer
from ipc_warp_amd.contact import ContactSolverBase

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _case(dev, name):
    if name == "grid":
        soup, _ = scenes.icosphere_grid(seed=0)
        fixed = np.zeros(soup.n_bodies, np.int32)
        fixed[5] = 1
    else:
        soup = scenes.screw_and_nut(seed=0)[0]
        fixed = np.array([0, 1], np.int32)
    x = torch.from_numpy(soup.x).to(dev)
    S = ContactSolverBase(x, x.clone(), torch.from_numpy(soup.edges), torch.from_numpy(soup.triangles),
                          torch.from_numpy(soup.body), thickness=5e-4, buffer_=5e-4, kappa=1e8,
                          disable_self_collision_=True)
    S.detect_collision()
    b = torch.zeros_like(x)
    trip = S.compute_collision_hessian(b)
    # rest-frame coordinates: vertex minus its body's mean
    nb = int(soup.body.max()) + 1
    cen = np.zeros((nb, 3))
    np.add.at(cen, soup.body, soup.x.astype(np.float64))
    cen /= np.bincount(soup.body, minlength=nb)[:, None]
    rest = (soup.x - cen[soup.body]).astype(np.float32)
    return soup, nb, fixed, rest, trip, b, S


@pytest.mark.parametrize("name", ["grid", "screw"])
def test_reduction_matches_oracle(dev, oracle, name):
    soup, nb, fixed, rest, trip, b, S = _case(dev, name)
    _say(f"abd [{name}] bodies={nb} triplets={trip.rows.shape[0]}")
    scale = 1e-4
    R = AbdReducer()
    M, rf = R.reduce(nb, torch.from_numpy(soup.body).to(dev), torch.from_numpy(rest).to(dev),
                     torch.from_numpy(fixed).to(dev), trip.rows, trip.cols, trip.vals, b, scale)
    torch.cuda.synchronize()
    rows, cols, vals = trip.rows.cpu().numpy(), trip.cols.cpu().numpy(), trip.vals.cpu().numpy()
    o, c, v, a, rref = oracle.abd_reduce(nb, soup.body, rest, fixed, rows, cols, vals, b.cpu().numpy(), scale)
    nnz = int(o[-1])
    assert nnz > 0
    assert np.array_equal(M.offsets.cpu().numpy(), o)
    assert np.array_equal(M.columns[:nnz].cpu().numpy(), c)
    g = M.values[:nnz].cpu().numpy().astype(np.float64)
    err = np.abs(g - v)
    bound = 1e-5 * a + 6e-8 * np.abs(v) + 1e-30
    assert (err <= bound).all(), f"worst {float((err / bound).max())}"
    assert np.abs(rf.cpu().numpy() - rref).max() <= 1e-5 * np.abs(rref).max()
    # a second call on the same workspace gives the same pattern
    M2, _ = R.reduce(nb, torch.from_numpy(soup.body).to(dev), torch.from_numpy(rest).to(dev),
                     torch.from_numpy(fixed).to(dev), trip.rows, trip.cols, trip.vals, None, scale)
    assert torch.equal(M2.offsets, M.offsets)
    assert np.abs(M2.values[:nnz].cpu().numpy() - g).max() <= 1e-5 * a.max()


def test_capacity_is_reported_before_any_output(dev, oracle):
    from ipc_warp_amd._lib import lib

    soup, nb, fixed, rest, trip, b, S = _case(dev, "grid")
    R = AbdReducer()
    offsets = torch.empty(4 * nb + 1, dtype=torch.int32, device=dev)
    rf = torch.zeros((4 * nb, 3), dtype=torch.float32, device=dev)
    body_t = torch.from_numpy(soup.body).to(dev)
    rest_t = torch.from_numpy(rest).to(dev)
    fixed_t = torch.from_numpy(fixed).to(dev)
    st = lib().ipc_abd_reduce_contacts(R.handle, nb, body_t.data_ptr(), rest_t.data_ptr(), fixed_t.data_ptr(),
                                       soup.x.shape[0], trip.rows.data_ptr(), trip.cols.data_ptr(),
                                       trip.vals.data_ptr(), trip.rows.shape[0], b.data_ptr(), 1.0,
                                       offsets.data_ptr(), None, None, 0, rf.data_ptr(), None)
    torch.cuda.synchronize()
    assert st > 0 and st % 16 == 0 and int(offsets[-1].item()) == st
    assert not rf.any()

"""GPU fixed-body (per-body BVH) contact path (fixed_bvh_contacts.py:503-827)
against the oracle's world-space enumeration.

Both broad phases are conservative and the acceptance is the same fp32
distance test in world space, so the fixed path must return exactly the
world path's pair sets with same-body pairs excluded (bit-exact after
sorting; EE records are (source edge, target edge) with the source on the
lower body, which for body-ordered edges is the world path's i < j order),
and its CCD bound must equal the world-space minimum TOI to 1e-12."""
import numpy as np
import pytest
import torch

from ipc_warp_amd import scenes
from ipc_warp_amd._lib import IPCError
from ipc_warp_amd.contact import FixedBodyContactSolver

pytestmark = pytest.mark.gpu

TH = 5e-4


def _solver(dev, soup, q, capacity=None, fixed=None):
    x = torch.from_numpy(soup.x).to(dev)
    xcs = torch.from_numpy(soup.xcs).to(dev)
    S = FixedBodyContactSolver(x, xcs, torch.from_numpy(soup.edges), torch.from_numpy(soup.triangles),
                               torch.from_numpy(soup.body), thickness=TH, buffer_=TH, kappa=1e8,
                               disable_self_collision_=True, capacity=capacity)
    S.initialize_fixed_body_bvhs(torch.from_numpy(q).to(dev), fixed_bodies=fixed)
    return S


def _sorted(rec, ids):
    order = np.lexsort((ids[:, 1], ids[:, 0]))
    return rec[order], ids[order]


def _check_dcd(S, soup, oracle):
    S.detect_collision()
    got = {}
    for kind, C, n in (("ee", S.contacts_new, S.n_contacts), ("pt", S.contacts_pt, S.n_contacts_pt)):
        got[kind] = _sorted(C.list[:n].cpu().numpy(), C.e0e1[:n].cpu().numpy())
    r_ee, r_ee_ids = oracle.dcd_ee(soup.x, soup.edges, soup.body, TH, TH, True)
    r_pt, r_pt_ids = oracle.dcd_pt(soup.x, soup.triangles, soup.body, TH, TH, True)
    assert np.array_equal(got["ee"][1], r_ee_ids) and np.array_equal(got["ee"][0], r_ee)
    assert np.array_equal(got["pt"][1], r_pt_ids) and np.array_equal(got["pt"][0], r_pt)
    return len(r_ee), len(r_pt)


@pytest.mark.parametrize("shear,sub,spacing", [(0.0, 3, 0.995), (0.1, 2, 0.99)])
def test_fixed_dcd_matches_world_oracle(dev, oracle, shear, sub, spacing):
    soup, q = scenes.affine_icospheres(seed=1, per_axis=3, spacing=spacing, subdivisions=sub, shear=shear)
    n_ee, n_pt = _check_dcd(_solver(dev, soup, q), soup, oracle)
    assert n_ee > 0 and n_pt > 0


def test_fixed_dcd_overflow_grows(dev, oracle):
    soup, q = scenes.affine_icospheres(seed=2, per_axis=3, spacing=0.995, subdivisions=2)
    S = _solver(dev, soup, q, capacity=4)
    n_ee, n_pt = _check_dcd(S, soup, oracle)
    assert n_ee > 4 and S.contacts_new.capacity >= n_ee


def test_fixed_dcd_separated_bodies(dev, oracle):
    soup, q = scenes.affine_icospheres(seed=3, per_axis=2, spacing=1.5, subdivisions=2)
    S = _solver(dev, soup, q)
    assert _check_dcd(S, soup, oracle) == (0, 0)


def _ccd_case(seed, shear, spacing=1.01, move=0.03):
    soup, q = scenes.affine_icospheres(seed=seed, per_axis=3, spacing=spacing, subdivisions=2, shear=shear)
    dq, dx = scenes.affine_step(q, soup.xcs.astype(np.float64), soup.body, seed=seed + 1, move=move)
    return soup, q, dq, dx


@pytest.mark.parametrize("shear", [0.0, 0.1])
def test_fixed_ccd_matches_world_oracle(dev, oracle, shear):
    soup, q, dq, dx = _ccd_case(4, shear)
    S = _solver(dev, soup, q)
    dx_t, dq_t = torch.from_numpy(dx).to(dev), torch.from_numpy(dq).to(dev)
    toi, fell_back = S.raw_toi_fixed(dx_t, dq_t)
    ref = oracle.ccd_toi(soup.x, soup.x - dx, soup.edges, soup.triangles, soup.body, exclude_same_body=True,
                         relative=True, threads=8)
    assert not fell_back
    assert ref < 1.0 and abs(toi - ref) <= 1e-12
    assert abs(S.raw_toi(dx_t) - ref) <= 1e-12
    assert S.collision_free_step(dx_t, dq_t) == pytest.approx(0.9 * toi)
    zero = torch.zeros_like(dx_t)
    assert S.raw_toi_fixed(zero, torch.zeros_like(dq_t)) == (1.0, False)


def test_fixed_ccd_with_fixed_bodies(dev, oracle):
    soup, q, dq, dx = _ccd_case(6, 0.0)
    nb = q.shape[0] // 4
    fixed = (np.arange(nb) % 3 == 0).astype(np.int32)
    # fixed bodies do not move
    qv = dq.reshape(nb, 4, 3)
    qv[fixed == 1] = 0.0
    dx[fixed[soup.body] == 1] = 0.0
    S = _solver(dev, soup, q, fixed=fixed)
    toi, fell_back = S.raw_toi_fixed(torch.from_numpy(dx).to(dev), torch.from_numpy(dq).to(dev))
    ref = oracle.ccd_toi(soup.x, soup.x - dx, soup.edges, soup.triangles, soup.body, exclude_same_body=True,
                         relative=True, threads=8)
    assert not fell_back and abs(toi - ref) <= 1e-12


def test_fixed_ccd_singular_frame_falls_back(dev, oracle):
    soup, q, dq, dx = _ccd_case(8, 0.0)
    nb = q.shape[0] // 4
    qv = q.reshape(nb, 4, 3)
    dqv = dq.reshape(nb, 4, 3)
    # body 0 collapses onto its centre at the end of the step: A1 = 0, p1 = p0
    dqv[0, 0] = 0.0
    dqv[0, 1:] = qv[0, 1:]
    on0 = soup.body == 0
    dx[on0] = soup.x[on0] - qv[0, 0][None, :]
    S = _solver(dev, soup, q)
    toi, fell_back = S.raw_toi_fixed(torch.from_numpy(dx).to(dev), torch.from_numpy(dq).to(dev))
    ref = oracle.ccd_toi(soup.x, soup.x - dx, soup.edges, soup.triangles, soup.body, exclude_same_body=True,
                         relative=True, threads=8)
    assert fell_back and abs(toi - ref) <= 1e-12


def test_fixed_argument_checks(dev):
    soup, q = scenes.affine_icospheres(seed=0, per_axis=2, spacing=1.2, subdivisions=1)
    x = torch.from_numpy(soup.x).to(dev)
    S = FixedBodyContactSolver(x, x.clone(), torch.from_numpy(soup.edges), torch.from_numpy(soup.triangles),
                               torch.from_numpy(soup.body), thickness=TH, buffer_=TH, disable_self_collision_=True)
    with pytest.raises(IPCError):
        S.initialize_fixed_body_bvhs(torch.from_numpy(q[:4]).to(dev), n_bodies=1)  # body ids beyond n_bodies
    with pytest.raises(ValueError):
        S.initialize_fixed_body_bvhs(torch.from_numpy(q).to(dev).double())

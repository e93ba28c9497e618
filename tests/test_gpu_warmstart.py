"""GPU warm-start features (ipc_warmstart_features) against the oracle.

Bar: bit-exact. The features are fp32 closest-point arithmetic in the same
operation order on both sides, and the slot order is fixed by (fp32 distance,
record id), so the lock-free insertion on the GPU must select exactly what
the oracle's sort selects."""
import numpy as np
import pytest
import torch

from ipc_warp_amd import scenes, warmstart
from ipc_warp_amd.contact import ContactSolverBase

pytestmark = pytest.mark.gpu

DHAT = 1e-3


def _gpu(dev, x, base, ee, pt, dhat=DHAT):
    xt = torch.from_numpy(np.ascontiguousarray(x, np.float32)).to(dev)
    bt = None if base is None else torch.from_numpy(np.ascontiguousarray(base, np.float32)).to(dev)
    et = torch.from_numpy(np.ascontiguousarray(ee, np.int32).reshape(-1, 4)).to(dev)
    pt_ = torch.from_numpy(np.ascontiguousarray(pt, np.int32).reshape(-1, 4)).to(dev)
    out = torch.full((x.shape[0], 48), float("nan"), dtype=torch.float32, device=dev)
    warmstart.vertex_features(xt, bt, et, et.shape[0], pt_, pt_.shape[0], dhat, out=out)
    torch.cuda.synchronize()
    return out.cpu().numpy()


def test_features_pressed_cubes(dev, oracle):
    soup = scenes.two_cubes_on_ground(seed=5, n=6, gap=4e-4, jitter=2e-4)
    x = torch.from_numpy(soup.x).to(dev)
    S = ContactSolverBase(x, x.clone(), torch.from_numpy(soup.edges), torch.from_numpy(soup.triangles),
                          torch.from_numpy(soup.body), thickness=DHAT / 2, buffer_=DHAT / 2, kappa=1e8)
    S.detect_collision()
    assert S.n_contacts > 0 and S.n_contacts_pt > 0
    base = np.random.default_rng(0).normal(size=(soup.x.shape[0], 12)).astype(np.float32)
    bt = torch.from_numpy(base).to(dev)
    f = warmstart.vertex_features(x, bt, S.contacts_new.list, S.n_contacts, S.contacts_pt.list, S.n_contacts_pt, DHAT)
    ee = S.contacts_new.list[:S.n_contacts].cpu().numpy()
    pt = S.contacts_pt.list[:S.n_contacts_pt].cpu().numpy()
    ref = oracle.warmstart_features(soup.x, base, ee, pt, DHAT)
    got = f.cpu().numpy()
    assert np.array_equal(got, ref)
    # vertices with more records than slots exist in this scene
    counts = np.bincount(np.concatenate([ee.reshape(-1), pt.reshape(-1)]), minlength=soup.x.shape[0])
    assert counts.max() > 6


def test_features_crowded_random_records(dev, oracle):
    rng = np.random.default_rng(7)
    V = 300
    x = rng.random((V, 3)).astype(np.float32)
    def recs(n):
        r = np.stack([rng.choice(V, 4, replace=False) for _ in range(n)]).astype(np.int32)
        return r
    ee, pt = recs(5000), recs(5000)
    ref = oracle.warmstart_features(x, None, ee, pt, DHAT)
    got = _gpu(dev, x, None, ee, pt)
    assert np.array_equal(got, ref)
    assert (ref[:, 12 + 6 * 5 + 3] > 0).all()  # every vertex has more than 6 records


def test_features_ties_and_stack(dev, oracle):
    depth = np.array([7, 3, 9, 1, 5, 3, 8, 2, 10, 6], np.float32) * np.float32(1e-4)
    xs = [np.zeros(3, np.float32)]
    pts = []
    for d in depth:
        b0 = len(xs)
        xs += [np.array([-1, -d, -1], np.float32), np.array([1, -d, -1], np.float32), np.array([0, -d, 1], np.float32)]
        pts.append([0, b0 + 1, b0, b0 + 2])
    x = np.array(xs, np.float32)
    pt = np.array(pts, np.int32)
    # the same records also as EE records over the triangles' first edges
    ee = np.array([[p[1], p[2], p[2], p[3]] for p in pts], np.int32)
    ref = oracle.warmstart_features(x, None, ee, pt, DHAT)
    assert np.array_equal(_gpu(dev, x, None, ee, pt), ref)


def test_features_empty_and_invalid(dev, oracle):
    x = np.random.default_rng(1).normal(size=(40, 3)).astype(np.float32)
    base = np.random.default_rng(2).normal(size=(40, 12)).astype(np.float32)
    z = np.zeros((0, 4), np.int32)
    got = _gpu(dev, x, base, z, z)
    assert np.array_equal(got[:, :12], base) and (got[:, 12:] == 0).all()
    bad = np.array([[0, 1, 2, 40], [-1, 1, 2, 3], [4, 5, 6, 7]], np.int32)
    ref = oracle.warmstart_features(x, None, bad, bad, DHAT)
    got = _gpu(dev, x, None, bad, bad)
    assert np.array_equal(got, ref)
    assert (got[0:4, 12:] == 0).all() and (got[4, 12:] != 0).any()


def test_features_argument_checks(dev):
    x = torch.zeros((4, 3), dtype=torch.float32, device=dev)
    c = torch.zeros((1, 4), dtype=torch.int32, device=dev)
    with pytest.raises(Exception):
        warmstart.vertex_features(x, None, c, 1, c, 0, 0.0)
    with pytest.raises(ValueError):
        warmstart.vertex_features(x, None, c.long(), 1, c, 0, DHAT)

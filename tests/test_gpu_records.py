"""The reference's contact records (XConstraint, contact.py:61-69) across the
boundary: the library's DCD lists written as records the way append writes
them (contact.py:163-197), packed back to the int32 [n][4] lists bit-exactly,
and the Hessian of packed records equal to the Hessian of the library's own
lists; both the 36-byte (fp32 scalars) and 48-byte (fp64 scalars) layouts."""
import numpy as np
import pytest
import torch

from ipc_warp_amd import ops, scenes
from ipc_warp_amd._lib import check, lib
from ipc_warp_amd.contact import ContactSolverBase

pytestmark = pytest.mark.gpu

TH = 5e-4
LAYOUTS = {
    36: np.dtype([("a1a2b1b2", "<i4", 4), ("e0e1", "<i4", 2), ("l0", "<f4"), ("alpha", "<f4"), ("lam", "<f4")]),
    48: np.dtype([("a1a2b1b2", "<i4", 4), ("e0e1", "<i4", 2), ("l0", "<f8"), ("alpha", "<f8"), ("lam", "<f8")]),
}


@pytest.fixture(scope="module")
def lists(dev):
    soup = scenes.two_cubes_on_ground(seed=11, n=8, gap=4e-4, jitter=2e-4)
    x = torch.from_numpy(soup.x).to(dev)
    S = ContactSolverBase(x, x.clone(), torch.from_numpy(soup.edges), torch.from_numpy(soup.triangles),
                          torch.from_numpy(soup.body), thickness=TH, buffer_=TH, kappa=1e8)
    S.detect_collision()
    assert S.n_contacts > 100 and S.n_contacts_pt > 100
    return S


def _records(dev, n, nbytes, lam):
    rec = np.zeros(n, LAYOUTS[nbytes])
    rec["lam"] = lam
    return torch.from_numpy(rec.view(np.uint8).copy()).to(dev)


@pytest.mark.parametrize("nbytes", [36, 48])
@pytest.mark.parametrize("kind", ["ee", "pt"])
def test_unpack_then_pack_round_trip(dev, lists, nbytes, kind):
    S = lists
    C = S.contacts_new if kind == "ee" else S.contacts_pt
    n = S.n_contacts if kind == "ee" else S.n_contacts_pt
    l0 = float(np.float32(2 * TH))
    rec = _records(dev, n, nbytes, 7.0)
    check(lib().ipc_contact_records_unpack(C.list.data_ptr(), C.e0e1.data_ptr(), n, l0, None, rec.data_ptr(), nbytes,
                                           None), "unpack")
    torch.cuda.synchronize()
    r = rec.cpu().numpy().view(LAYOUTS[nbytes])
    assert np.array_equal(r["a1a2b1b2"], C.list[:n].cpu().numpy())
    assert np.array_equal(r["e0e1"], C.e0e1[:n].cpu().numpy())
    assert (r["l0"] == r["l0"].dtype.type(np.float32(l0))).all()
    assert (r["alpha"] == r["alpha"].dtype.type(1e-6)).all()
    assert (r["lam"] == 7.0).all()
    packed = torch.full((n, 4), -1, dtype=torch.int32, device=dev)
    bad = torch.zeros(1, dtype=torch.int32, device=dev)
    check(lib().ipc_contact_records_pack(rec.data_ptr(), nbytes, n, l0, packed.data_ptr(), bad.data_ptr(), None),
          "pack")
    torch.cuda.synchronize()
    assert torch.equal(packed, C.list[:n]) and int(bad.item()) == 0
    # one record with another l0 is counted
    r = rec.cpu().numpy().view(LAYOUTS[nbytes]).copy()
    r["l0"][n // 2] = 3e-3
    rec2 = torch.from_numpy(r.view(np.uint8).copy()).to(dev)
    check(lib().ipc_contact_records_pack(rec2.data_ptr(), nbytes, n, l0, packed.data_ptr(), bad.data_ptr(), None),
          "pack")
    torch.cuda.synchronize()
    assert int(bad.item()) == 1


def test_hessian_of_packed_records_equals_the_lists(dev, lists):
    S = lists
    n_ee, n_pt = S.n_contacts, S.n_contacts_pt
    l0 = float(np.float32(2 * TH))
    packed = {}
    for kind, C, n in (("ee", S.contacts_new, n_ee), ("pt", S.contacts_pt, n_pt)):
        rec = _records(dev, n, 36, 0.0)
        check(lib().ipc_contact_records_unpack(C.list.data_ptr(), C.e0e1.data_ptr(), n, l0, None, rec.data_ptr(), 36,
                                               None), "unpack")
        p = torch.empty((n, 4), dtype=torch.int32, device=dev)
        check(lib().ipc_contact_records_pack(rec.data_ptr(), 36, n, l0, p.data_ptr(), None, None), "pack")
        packed[kind] = p
    outs = []
    for ee, pt in ((S.contacts_new.list, S.contacts_pt.list), (packed["ee"], packed["pt"])):
        tot = n_ee + n_pt
        rows = torch.empty(16 * tot, dtype=torch.int32, device=dev)
        cols = torch.empty_like(rows)
        vals = torch.empty((16 * tot, 3, 3), dtype=torch.float32, device=dev)
        b = torch.zeros_like(S.x)
        ops.collision_hessian(S.x, S.xcs, ee, n_ee, pt, n_pt, l0, 1e8, rows, cols, vals, b)
        torch.cuda.synchronize()
        outs.append((rows.cpu(), cols.cpu(), vals.cpu(), b.cpu()))
    for a, c in zip(outs[0][:3], outs[1][:3]):
        assert torch.equal(a, c)
    # forces are float atomics: equal up to the order of the additions
    assert torch.allclose(outs[0][3], outs[1][3], rtol=1e-5, atol=1e-6 * float(outs[0][3].abs().max()))


def test_ccd_records_carry_alpha(dev, lists):
    S = lists
    n = S.n_contacts
    C = S.contacts_new
    alpha = torch.linspace(0.1, 0.9, n, dtype=torch.float64, device=dev)
    rec = _records(dev, n, 48, 0.5)
    check(lib().ipc_contact_records_unpack(C.list.data_ptr(), C.e0e1.data_ptr(), n, 1e-3, alpha.data_ptr(),
                                           rec.data_ptr(), 48, None), "unpack")
    torch.cuda.synchronize()
    r = rec.cpu().numpy().view(LAYOUTS[48])
    assert np.array_equal(r["alpha"], alpha.cpu().numpy()) and (r["lam"] == 0.5).all()

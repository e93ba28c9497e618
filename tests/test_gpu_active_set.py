"""The augmented-Lagrangian active set on the GPU (augmented_lagrangian.py:
105-197, 506-547, 705-717) against the oracle, bit for bit: cell_start,
every entry word (the zeroed tail included) and the vertex TOIs, with fp32
and fp64 scalars; random constraint sets over small and large tables (a
2-bucket table puts ~1500 entries in one bucket), merges with decayed and
re-detected entries, the CCD-recorded pairs of a pressed two-cube scene
through update_active_set, empty inputs and refused arguments."""
import numpy as np
import pytest
import torch

from ipc_warp_amd import scenes
from ipc_warp_amd._lib import IPC_EINVAL, lib
from ipc_warp_amd.active_set import (ALActiveSets, ActiveSet, ALConstraintView, accumulate_vertex_toi,
                                     xconstraint_records)
from ipc_warp_amd.contact import ContactSolverBase
from ipc_warp_amd.ops import stream_handle

pytestmark = pytest.mark.gpu

SCALARS = {torch.float32: np.float32, torch.float64: np.float64}
XW = {torch.float32: 9, torch.float64: 12}


def random_records(rng, n, nv, scalar, key_base=0):
    """XConstraint records as int32 words: vertex ids, distinct keys, alphas
    on a coarse grid (ties) with 1e-13 nudges (the 1e-12 filter)."""
    s = SCALARS[scalar]
    dt = np.dtype([("a", "<i4", 4), ("e", "<i4", 2), ("l0", s), ("alpha", s), ("lam", s)], align=True)
    rec = np.zeros(n, dt)
    rec["a"] = rng.integers(0, nv, (n, 4))
    i = rng.choice(1 << 16, n, replace=False) + key_base
    rec["e"][:, 0] = i
    rec["e"][:, 1] = i + 1 + rng.integers(0, 1000, n)
    rec["l0"] = 1e-3
    rec["alpha"] = (rng.integers(0, 32, n) / 32).astype(s) + s(1e-13) * rng.integers(0, 2, n)
    rec["lam"] = 3.0
    return rec.view(np.int32).reshape(n, -1)


def gpu_create(dev, words, nv, scalar, T, cap):
    r = torch.from_numpy(words).to(dev)
    vt = torch.ones(nv, dtype=scalar, device=dev)
    accumulate_vertex_toi(r, r.shape[0], vt)
    A = ActiveSet(cap, dev, scalar)
    A.hash.table_size = T
    A.hash.cell_start = torch.zeros(T + 1, dtype=torch.int32, device=dev)
    A.hash.cell_entries.fill_(-1)  # the tail must come back zeroed
    A.create_hash(r, r.shape[0], vt)
    return A, vt


def check_against_oracle(oracle, dev, words, nv, scalar, T, cap):
    A, vt = gpu_create(dev, words, nv, scalar, T, cap)
    vt_ref = oracle.al_vertex_toi(words, np.ones(nv))
    assert np.array_equal(vt.cpu().numpy(), vt_ref)
    cnt, cs, ent = oracle.al_create_hash(words, vt_ref, T, cap)
    torch.cuda.synchronize()
    assert A.count() == cnt
    assert np.array_equal(A.hash.cell_start.cpu().numpy(), cs)
    assert np.array_equal(A.hash.cell_entries.cpu().numpy(), ent)
    return A, cnt, cs, ent


@pytest.mark.parametrize("scalar", [torch.float32, torch.float64])
@pytest.mark.parametrize("T", [2, 61, 8192, 1 << 16])
def test_create_hash_matches_oracle(dev, oracle, scalar, T):
    rng = np.random.default_rng(T + 1)
    words = random_records(rng, 3000, 400, scalar)
    check_against_oracle(oracle, dev, words, 400, scalar, T, 4096)


@pytest.mark.parametrize("scalar", [torch.float32, torch.float64])
def test_merge_matches_oracle(dev, oracle, scalar):
    rng = np.random.default_rng(9)
    nv = 500
    old_words = random_records(rng, 2500, nv, scalar)
    A, n_o, cs_o, ent_o = check_against_oracle(oracle, dev, old_words, nv, scalar, 8192, 4096)
    # decay part of the old set and give every old entry a payload to carry
    view = ALConstraintView(A.hash.cell_entries[:n_o])
    gamma = torch.from_numpy(rng.choice([1.0, 0.5, 0.01, 0.009, 0.0], n_o)).to(dev)
    view.set_scalar("gamma", gamma)
    view.set_scalar("lam", torch.from_numpy(rng.normal(size=n_o)).to(dev))
    view.set_scalar("d0", torch.from_numpy(rng.normal(size=n_o)).to(dev))
    ent_o = A.hash.cell_entries.cpu().numpy()
    new_words = random_records(rng, 2500, nv, scalar, key_base=1 << 20)
    keys_old = ent_o[:n_o, 4:6]
    pick = rng.choice(n_o, n_o // 2, replace=False)
    new_words[:n_o // 2, 4:6] = keys_old[pick]
    B, n_n, _, ent_n = check_against_oracle(oracle, dev, new_words, nv, scalar, 8192, 4096)
    A.merge(B)
    cnt, cs, ent = oracle.al_merge(cs_o, ent_o[:n_o], 8192, n_o, ent_n[:n_n], n_n, 2 * (n_o + n_n), n_o + n_n)
    assert 0 < cnt < n_o + n_n
    assert A.count() == cnt and A.hash.table_size == 2 * (n_o + n_n)
    assert np.array_equal(A.hash.cell_start.cpu().numpy(), cs)
    assert np.array_equal(A.hash.cell_entries.cpu().numpy(), ent)


def test_update_active_set_on_ccd_pairs(dev, oracle):
    """update_active_set (augmented_lagrangian.py:705-717) twice on the CCD
    pairs of two cubes pressed together, against the oracle's composition
    of the same steps on the lists new_intersections recorded."""
    soup = scenes.two_cubes_on_ground(seed=15, n=8, gap=2e-3, jitter=1e-4)
    rng = np.random.default_rng(15)
    x = torch.from_numpy(soup.x).to(dev)
    S = ContactSolverBase(x, x.clone(), torch.from_numpy(soup.edges), torch.from_numpy(soup.triangles),
                          torch.from_numpy(soup.body), thickness=5e-4, buffer_=5e-4, kappa=1e8)
    sets = ALActiveSets(soup.x.shape[0], dev, initial_capacity=64)
    ref = {"ee": (np.zeros(129, np.int32), np.zeros((64, 26), np.int32), 128, 0),
           "pt": (np.zeros(129, np.int32), np.zeros((64, 26), np.int32), 128, 0)}
    for it in range(2):
        dx = np.zeros_like(soup.x)
        dx[soup.body == 1, 1] = 6e-3 * (it + 1)
        dx += (rng.normal(size=dx.shape) * 1e-3).astype(np.float32)
        x1 = torch.from_numpy((soup.x - dx).astype(np.float32)).to(dev)
        sets.update_active_set(S, x, x1)
        assert S.n_contacts > 0 and S.n_contacts_pt > 0
        recs = {}
        for kind, C, n in (("ee", S.contacts_new, S.n_contacts), ("pt", S.contacts_pt, S.n_contacts_pt)):
            recs[kind] = xconstraint_records(C.list, C.e0e1, C.alpha, n, S.l0).cpu().numpy()
        vt = np.ones(soup.x.shape[0], np.float32)
        for kind in ("ee", "pt"):
            vt = oracle.al_vertex_toi(recs[kind], vt)
        assert np.array_equal(sets.vertex_toi.cpu().numpy(), vt)
        for kind, det, act in (("ee", sets.detected_set_ee, sets.active_set_ee),
                               ("pt", sets.detected_set_pt, sets.active_set_pt)):
            n_det = det.count()
            d_cnt, d_cs, d_ent = oracle.al_create_hash(recs[kind], vt, det.hash.table_size, det.capacity)
            assert n_det == d_cnt
            assert np.array_equal(det.hash.cell_start.cpu().numpy(), d_cs)
            assert np.array_equal(det.hash.cell_entries.cpu().numpy(), d_ent)
            o_cs, o_ent, o_T, n_o = ref[kind]
            m = max(1, n_o + d_cnt)
            cnt, cs, ent = oracle.al_merge(o_cs, o_ent[:max(n_o, 1)], o_T, n_o, d_ent[:max(d_cnt, 1)], d_cnt,
                                           max(2, 2 * m), m)
            assert act.count() == cnt
            assert np.array_equal(act.hash.cell_start.cpu().numpy(), cs)
            assert np.array_equal(act.hash.cell_entries.cpu().numpy(), ent)
            ref[kind] = (cs, ent, max(2, 2 * m), cnt)
        x = x1


@pytest.mark.parametrize("scalar", [torch.float32, torch.float64])
def test_slack_and_force_updates_match_oracle(dev, oracle, scalar):
    """ActiveSet.update_slack / update_forces (augmented_lagrangian.py:
    326-345, 564-572) on a created set given random linearisations,
    multipliers and slacks (a third of them zero), bit for bit."""
    rng = np.random.default_rng(17)
    nv = 400
    words = random_records(rng, 3000, nv, scalar)
    A, n, _, _ = check_against_oracle(oracle, dev, words, nv, scalar, 8192, 4096)
    view = ALConstraintView(A.hash.cell_entries[:n])
    view.set_scalar("d0", torch.from_numpy(rng.normal(size=n)).to(dev))
    view.set_scalar("lam", torch.from_numpy(rng.normal(size=n)).to(dev))
    view.set_scalar("gamma", torch.from_numpy(rng.uniform(0, 1, n)).to(dev))
    view.set_scalar("slack", torch.from_numpy(np.where(rng.random(n) < 0.33, 0.0, rng.uniform(0, 1, n))).to(dev))
    g = torch.from_numpy(rng.normal(size=(n, 12))).to(scalar).contiguous().view(torch.int32)
    sw = 1 if scalar == torch.float32 else 2
    d0w = 13 if sw == 1 else 20
    A.hash.cell_entries[:n, d0w + sw:d0w + 13 * sw] = g.reshape(n, 12 * sw)
    x = torch.from_numpy(rng.normal(size=(nv, 3))).to(dev).to(scalar)
    before = A.hash.cell_entries[:n].cpu().numpy()
    for forces in (False, True):
        if forces:
            A.update_forces(x, 3.0)
        else:
            A.update_slack(x, 3.0)
        torch.cuda.synchronize()
        before = oracle.al_update(before, x.cpu().numpy(), 3.0, forces)
        assert np.array_equal(A.hash.cell_entries[:n].cpu().numpy(), before)


@pytest.mark.parametrize("scalar", [torch.float32, torch.float64])
@pytest.mark.parametrize("kind", ["pt", "ee"])
def test_linearize_matches_oracle(dev, oracle, scalar, kind):
    """ActiveSet.linearize (augmented_lagrangian.py:270-323, 549-562) on a
    synthetic pair stream (EE with near-parallel pairs, whose mollifier
    selects the point-edge form): mollifier and distance_type bit for bit,
    gradients within 2e-6 of each pair's largest (fp32 records; 1e-9 in
    fp64, where the kernel's closest-point form of grad(d^2) and the oracle's
    jets round differently), d0 within 2e-6 (1e-9) absolute."""
    s = SCALARS[scalar]
    x, c = scenes.pair_stream(kind, 4096, seed=23, dhat=1e-3, dtype=np.float64)
    quads = c.copy()
    if kind == "pt":
        quads[:, [1, 2]] = c[:, [2, 1]]
    dt = np.dtype([("a", "<i4", 4), ("e", "<i4", 2), ("l0", s), ("alpha", s), ("lam", s), ("gamma", s),
                   ("slack", s), ("mollifier", s), ("distance_type", "<i4"), ("d0", s), ("grad", s, 12)], align=True)
    rec = np.zeros(len(quads), dt)
    rec["a"] = quads
    rec["e"][:, 0] = np.arange(len(quads))
    rec["l0"] = 1e-3
    rec["gamma"] = 1.0
    words = rec.view(np.int32).reshape(len(quads), -1)
    A = ActiveSet(len(quads), dev, scalar)
    A.hash.cell_entries.copy_(torch.from_numpy(words))
    A.hash.cell_start[-1] = len(quads)
    xt = torch.from_numpy(x).to(dev).to(scalar)
    A.linearize(xt, xt, kind == "pt")
    torch.cuda.synchronize()
    got = A.hash.cell_entries.cpu().numpy().view(dt).reshape(-1)
    ref = oracle.al_linearize(words, x.astype(s), x.astype(s), kind == "pt").view(dt).reshape(-1)
    assert np.array_equal(got["distance_type"], ref["distance_type"])
    assert np.array_equal(got["mollifier"], ref["mollifier"])
    if kind == "ee":
        assert (ref["distance_type"] > 0).sum() > 50
    tol = 2e-6 if scalar == torch.float32 else 1e-9
    gg, rg = got["grad"].astype(np.float64), ref["grad"].astype(np.float64)
    scale = np.abs(rg).max(axis=1)
    assert (np.abs(gg - rg).max(axis=1) <= tol * scale).all()
    assert np.abs(got["d0"].astype(np.float64) - ref["d0"]).max() <= tol


def test_empty_inputs(dev, oracle):
    A = ActiveSet(8, dev)
    A.hash.cell_entries.fill_(5)
    A.hash.cell_start.fill_(5)
    A.create_hash(torch.zeros((0, 9), dtype=torch.int32, device=dev), 0, torch.ones(4, device=dev))
    assert A.count() == 0 and not A.hash.cell_entries.any() and not A.hash.cell_start.any()
    B = ActiveSet(8, dev)
    A.merge(B)
    assert A.count() == 0 and A.capacity == 1 and not A.hash.cell_entries.any()


def test_refused_arguments(dev):
    L = lib()
    s = stream_handle(dev)
    A = ActiveSet(4, dev)
    r = torch.zeros((8, 9), dtype=torch.int32, device=dev)
    vt = torch.ones(4, device=dev)
    h = A.hash
    # capacity below n, bad record size, table_size < 1
    assert L.ipc_al_create_hash(A._ws.handle, r.data_ptr(), 36, 8, vt.data_ptr(), 4, h.table_size,
                                h.cell_start.data_ptr(), h.cell_entries.data_ptr(), 4, s) == IPC_EINVAL
    assert L.ipc_al_create_hash(A._ws.handle, r.data_ptr(), 40, 2, vt.data_ptr(), 4, h.table_size,
                                h.cell_start.data_ptr(), h.cell_entries.data_ptr(), 4, s) == IPC_EINVAL
    assert L.ipc_al_create_hash(A._ws.handle, r.data_ptr(), 36, 2, vt.data_ptr(), 4, 0,
                                h.cell_start.data_ptr(), h.cell_entries.data_ptr(), 4, s) == IPC_EINVAL
    # the destination of a merge may not be one of its inputs
    assert L.ipc_al_merge(A._ws.handle, 104, h.cell_start.data_ptr(), h.cell_entries.data_ptr(), h.table_size, 0,
                          h.cell_entries.data_ptr(), 0, h.table_size, h.cell_start.data_ptr(),
                          h.cell_entries.data_ptr(), 4, s) == IPC_EINVAL
    # capacity below old + new
    assert L.ipc_al_merge(A._ws.handle, 104, h.cell_start.data_ptr(), h.cell_entries.data_ptr(), h.table_size, 3,
                          h.cell_entries.data_ptr(), 3, h.table_size, None, None, 4, s) == IPC_EINVAL
    assert L.ipc_al_vertex_toi(r.data_ptr(), 44, 8, vt.data_ptr(), 4, s) == IPC_EINVAL

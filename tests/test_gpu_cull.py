"""GPU candidate culling (contact.py:199-336,539-725; ccd/ccd.py:166-210,
360-427) against the oracle's exact pair enumeration.

Bar: the DCD contact sets (records and ids) and the CCD candidate sets are
bit-exact after sorting; the step bound equals the oracle's to 1e-12."""
import numpy as np
import pytest
import torch

from ipc_warp_amd import scenes
from ipc_warp_amd.contact import ContactSolverBase

pytestmark = pytest.mark.gpu

TH = 5e-4


def _solver(dev, soup, exclude=False, capacity=None):
    x = torch.from_numpy(soup.x).to(dev)
    return ContactSolverBase(x, x.clone(), torch.from_numpy(soup.edges), torch.from_numpy(soup.triangles),
                             torch.from_numpy(soup.body), thickness=TH, buffer_=TH, kappa=1e8,
                             disable_self_collision_=exclude, capacity=capacity)


def _sorted(rec, ids):
    rec = np.asarray(rec)
    ids = np.asarray(ids)
    order = np.lexsort((ids[:, 1], ids[:, 0]))
    return rec[order], ids[order]


def _check_dcd(S, soup, exclude):
    S.detect_collision()
    ee = S.contacts_new.list[:S.n_contacts].cpu().numpy()
    ee_ids = S.contacts_new.e0e1[:S.n_contacts].cpu().numpy()
    pt = S.contacts_pt.list[:S.n_contacts_pt].cpu().numpy()
    pt_ids = S.contacts_pt.e0e1[:S.n_contacts_pt].cpu().numpy()
    import oracle
    r_ee, r_ee_ids = oracle.dcd_ee(soup.x, soup.edges, soup.body, TH, TH, exclude)
    r_pt, r_pt_ids = oracle.dcd_pt(soup.x, soup.triangles, soup.body, TH, TH, exclude)
    g_ee, g_ee_ids = _sorted(ee, ee_ids)
    g_pt, g_pt_ids = _sorted(pt, pt_ids)
    assert np.array_equal(g_ee_ids, r_ee_ids) and np.array_equal(g_ee, r_ee)
    assert np.array_equal(g_pt_ids, r_pt_ids) and np.array_equal(g_pt, r_pt)
    return len(r_ee), len(r_pt)


def test_dcd_two_cubes_pressed(dev, oracle):
    soup = scenes.two_cubes_on_ground(seed=11, n=8, gap=4e-4, jitter=2e-4)
    n_ee, n_pt = _check_dcd(_solver(dev, soup), soup, False)
    assert n_ee > 100 and n_pt > 100


@pytest.mark.parametrize("offset", [0.0, 30.0, 37.0, 1000.0])
@pytest.mark.parametrize("buffer", [TH, 0.0, 1e-3 * TH])
def test_dcd_borderline_distances(dev, oracle, offset, buffer):
    """DCD boxes are inflated by thickness + min(buffer, slack) instead of the
    reference's thickness + buffer: on pairs whose fp32 distance sits within a
    few ulps of 2 thickness, at coordinates up to 1000 and with buffers
    smaller than the slack, the sets equal the oracle's, which applies the
    reference's own box test."""
    soup = scenes.borderline_pairs(offset, TH, seed=int(offset) + 3)
    x = torch.from_numpy(soup.x).to(dev)
    S = ContactSolverBase(x, x.clone(), torch.from_numpy(soup.edges), torch.from_numpy(soup.triangles),
                          torch.from_numpy(soup.body), thickness=TH, buffer_=buffer, kappa=1e8)
    S.detect_collision()
    g_ee = _sorted(S.contacts_new.list[:S.n_contacts].cpu().numpy(), S.contacts_new.e0e1[:S.n_contacts].cpu().numpy())
    g_pt = _sorted(S.contacts_pt.list[:S.n_contacts_pt].cpu().numpy(),
                   S.contacts_pt.e0e1[:S.n_contacts_pt].cpu().numpy())
    r_ee = oracle.dcd_ee(soup.x, soup.edges, soup.body, TH, buffer)
    r_pt = oracle.dcd_pt(soup.x, soup.triangles, soup.body, TH, buffer)
    assert 40 < len(r_ee[0]) < 256 and 40 < len(r_pt[0])
    assert np.array_equal(g_ee[1], r_ee[1]) and np.array_equal(g_ee[0], r_ee[0])
    assert np.array_equal(g_pt[1], r_pt[1]) and np.array_equal(g_pt[0], r_pt[0])


def test_dcd_nan_buffer_tests_every_pair(dev, oracle):
    """A NaN buffer makes every box NaN, and a NaN box overlaps everything
    (the reference's box test is a chain of > comparisons): the set is every
    pair the filters and the fp32 distance test accept, the oracle's set with
    a buffer wider than the scene."""
    soup = scenes.two_cubes_on_ground(seed=13, n=3, gap=4e-4, jitter=2e-4)
    x = torch.from_numpy(soup.x).to(dev)
    S = ContactSolverBase(x, x.clone(), torch.from_numpy(soup.edges), torch.from_numpy(soup.triangles),
                          torch.from_numpy(soup.body), thickness=TH, buffer_=float("nan"), kappa=1e8)
    S.detect_collision()
    g_ee = _sorted(S.contacts_new.list[:S.n_contacts].cpu().numpy(), S.contacts_new.e0e1[:S.n_contacts].cpu().numpy())
    g_pt = _sorted(S.contacts_pt.list[:S.n_contacts_pt].cpu().numpy(),
                   S.contacts_pt.e0e1[:S.n_contacts_pt].cpu().numpy())
    r_ee = oracle.dcd_ee(soup.x, soup.edges, soup.body, TH, 10.0)
    r_pt = oracle.dcd_pt(soup.x, soup.triangles, soup.body, TH, 10.0)
    assert len(r_ee[0]) > 10 and len(r_pt[0]) > 10
    assert np.array_equal(g_ee[1], r_ee[1]) and np.array_equal(g_ee[0], r_ee[0])
    assert np.array_equal(g_pt[1], r_pt[1]) and np.array_equal(g_pt[0], r_pt[0])


def test_dcd_overflow_grows_the_lists(dev, oracle):
    soup = scenes.two_cubes_on_ground(seed=12, n=6, gap=4e-4, jitter=2e-4)
    S = _solver(dev, soup, capacity=1)
    n_ee, n_pt = _check_dcd(S, soup, False)
    assert S.contacts_new.capacity >= n_ee and S.contacts_pt.capacity >= n_pt > 1


def test_dcd_icospheres_exclude_same_body(dev, oracle):
    soup, _ = scenes.icosphere_grid(seed=13, per_axis=2, spacing=1.0003, subdivisions=3)
    rng = np.random.default_rng(13)
    soup.x = (soup.x + rng.uniform(-2e-4, 2e-4, soup.x.shape)).astype(np.float32)
    n_ee, n_pt = _check_dcd(_solver(dev, soup, exclude=True), soup, True)
    assert n_ee + n_pt > 0


def _ccd_setup(seed):
    soup = scenes.two_cubes_on_ground(seed=seed, n=8, gap=2e-3, jitter=1e-4)
    rng = np.random.default_rng(seed)
    dx = np.zeros_like(soup.x)
    dx[soup.body == 1, 1] = 6e-3
    dx += (rng.normal(size=dx.shape) * 1e-3).astype(np.float32)
    return soup, dx.astype(np.float32)


@pytest.mark.parametrize("kind", ["ee", "pt"])
def test_ccd_candidate_sets(dev, oracle, kind):
    soup, dx = _ccd_setup(14)
    S = _solver(dev, soup)
    x0 = S.x
    x1 = x0 - torch.from_numpy(dx).to(dev)
    quads, ids = S.ccd_candidates(kind, x0, x1)
    q, i = _sorted(quads.cpu().numpy(), ids.cpu().numpy())
    fn = oracle.ccd_candidates_ee if kind == "ee" else oracle.ccd_candidates_pt
    topo = soup.edges if kind == "ee" else soup.triangles
    rq, ri = fn(soup.x, x1.cpu().numpy(), topo, soup.body)
    assert np.array_equal(i, ri) and np.array_equal(q, rq)
    assert len(rq) > 1000


def test_collision_free_step_matches_oracle(dev, oracle):
    soup, dx = _ccd_setup(15)
    S = _solver(dev, soup)
    dxt = torch.from_numpy(dx).to(dev)
    toi = S.raw_toi(dxt)
    x1 = soup.x - dx
    ref = oracle.ccd_toi(soup.x, x1, soup.edges, soup.triangles, soup.body, relative=True, threads=8)
    assert toi < 1.0
    assert abs(toi - ref) <= 1e-12
    assert S.collision_free_step(dxt) == pytest.approx(0.9 * toi)
    n_ee, n_pt = S.last_ccd_counts()
    r_ee, _ = oracle.ccd_candidates_ee(soup.x, x1, soup.edges, soup.body)
    r_pt, _ = oracle.ccd_candidates_pt(soup.x, x1, soup.triangles, soup.body)
    assert (n_ee, n_pt) == (len(r_ee), len(r_pt))
    # no motion, no impact
    assert S.collision_free_step(torch.zeros_like(dxt)) == 1.0


def test_dispatch_order_keeps_the_sets(dev, oracle):
    """The paired DCD and CCD query launches order their workgroups from the
    previous launch's step counts: repeated calls on moving positions keep
    the exact contact sets, CCD candidate counts and step bound."""
    soup, vel = scenes.icosphere_grid(seed=16, per_axis=2, spacing=1.0003, subdivisions=3)
    rng = np.random.default_rng(16)
    S = _solver(dev, soup, exclude=True)
    for _ in range(4):
        soup.x = (soup.x + rng.uniform(-2e-4, 2e-4, soup.x.shape)).astype(np.float32)
        S.x.copy_(torch.from_numpy(soup.x).to(dev))
        n_ee, n_pt = _check_dcd(S, soup, True)
        assert n_ee + n_pt > 0
        dx = (-1e-2 * vel[soup.body]).astype(np.float32)
        x1 = soup.x - dx
        toi = S.raw_toi(torch.from_numpy(dx).to(dev))
        ref = oracle.ccd_toi(soup.x, x1, soup.edges, soup.triangles, soup.body, 1e-7, True, relative=True, threads=8)
        assert abs(toi - ref) <= 1e-12
        r_ee, _ = oracle.ccd_candidates_ee(soup.x, x1, soup.edges, soup.body, 1e-7, True)
        r_pt, _ = oracle.ccd_candidates_pt(soup.x, x1, soup.triangles, soup.body, 1e-7, True)
        assert S.last_ccd_counts() == (len(r_ee), len(r_pt))
        assert len(r_ee) > 0 and len(r_pt) > 0


def test_leaf_order_refresh_keeps_the_sets(dev, oracle):
    """A tree's leaf order is re-sorted every 16 builds and reused, with
    only the boxes recomputed, in between. Forty DCD + CCD rounds (two
    builds of each tree per round) on drifting positions cross several
    re-sorts; every round keeps the exact contact sets and CCD candidate
    counts."""
    soup, vel = scenes.icosphere_grid(seed=17, per_axis=2, spacing=1.0003, subdivisions=2)
    rng = np.random.default_rng(17)
    S = _solver(dev, soup, exclude=True)
    for _ in range(40):
        soup.x = (soup.x + rng.uniform(-2e-3, 2e-3, soup.x.shape)).astype(np.float32)
        S.x.copy_(torch.from_numpy(soup.x).to(dev))
        _check_dcd(S, soup, True)
        dx = (-1e-2 * vel[soup.body]).astype(np.float32)
        x1 = soup.x - dx
        S.raw_toi(torch.from_numpy(dx).to(dev))
        r_ee, _ = oracle.ccd_candidates_ee(soup.x, x1, soup.edges, soup.body, 1e-7, True)
        r_pt, _ = oracle.ccd_candidates_pt(soup.x, x1, soup.triangles, soup.body, 1e-7, True)
        assert S.last_ccd_counts() == (len(r_ee), len(r_pt))


def test_non_finite_vertices_drop_out(dev, oracle):
    """A NaN vertex and a vertex at +inf (one coordinate), each of them in a
    contact of the finite scene: the contacts through them drop out and
    every other contact stays, as in the oracle, through a fresh build
    (NaN and infinite box centres in the leaf order) and a second build
    that reuses it. The CCD candidates with the infinite vertex moving (its
    boxes reach to +inf) are the oracle's too."""
    soup, vel = scenes.icosphere_grid(seed=18, per_axis=2, spacing=1.0003, subdivisions=2)
    r_ee, _ = oracle.dcd_ee(soup.x, soup.edges, soup.body, TH, TH, True)
    r_pt, _ = oracle.dcd_pt(soup.x, soup.triangles, soup.body, TH, TH, True)
    v_nan, v_inf = int(r_ee[0][0]), int(r_pt[-1][0])
    assert v_nan != v_inf
    x = soup.x.copy()
    x[v_inf, 1] = np.inf
    x_nan = x.copy()
    x_nan[v_nan] = np.nan
    S = _solver(dev, soup, exclude=True)
    for xs in (x_nan, x_nan, x):
        soup.x = xs
        S.x.copy_(torch.from_numpy(xs).to(dev))
        n_ee, n_pt = _check_dcd(S, soup, True)
        assert n_ee < len(r_ee) and n_pt < len(r_pt)
    dx = (-1e-2 * vel[soup.body]).astype(np.float32)
    x1 = x - dx
    for kind in ("ee", "pt"):
        quads, ids = S.ccd_candidates(kind, S.x, torch.from_numpy(x1).to(dev))
        q, i = _sorted(quads.cpu().numpy(), ids.cpu().numpy())
        fn = oracle.ccd_candidates_ee if kind == "ee" else oracle.ccd_candidates_pt
        rq, ri = fn(x, x1, soup.edges if kind == "ee" else soup.triangles, soup.body, 1e-7, True)
        assert len(rq) > 0
        assert np.array_equal(i, ri) and np.array_equal(q, rq)


def test_degenerate_soups(dev, oracle):
    # one edge and no triangles; two crossing edges
    x = np.array([[0, 0, 0], [1, 0, 0], [0.5, -0.5, 4e-4], [0.5, 0.5, 4e-4]], np.float32)
    for edges in (np.array([[0, 1]], np.int32), np.array([[0, 1], [2, 3]], np.int32)):
        soup = scenes.Soup(x=x, xcs=x.copy(), edges=edges, triangles=np.zeros((0, 3), np.int32),
                           body=np.array([0, 0, 1, 1], np.int32))
        S = _solver(dev, soup)
        S.detect_collision()
        assert S.n_contacts_pt == 0
        assert S.n_contacts == (1 if len(edges) == 2 else 0)
        assert S.collision_free_step(torch.zeros_like(S.x)) == 1.0


def test_kernel_timing_hooks(dev, oracle):
    from ipc_warp_amd._lib import check, lib

    soup, dx = _ccd_setup(16)
    S = _solver(dev, soup)
    dxt = torch.from_numpy(dx).to(dev)
    S.set_kernel_timing(True)
    S.kernel_times()
    # the combined calls launch each stage once for both halves (EE kinds)
    S.detect_collision()
    S.collision_free_step(dxt)
    t = S.kernel_times()
    for k in ("build_edge", "query_dcd_ee", "query_ccd_ee", "toi_ee"):
        ms, n = t[k]
        assert n >= 1 and ms > 0.0, k
    for k in ("build_tri", "query_dcd_pt", "query_ccd_pt", "toi_pt"):
        assert t[k][1] == 0, k
    # the single-kind calls record their own kinds
    S.new_intersections(S.x, (S.x - dxt).contiguous())
    c = S.contacts_pt
    check(lib().ipc_detect_collision_pt(S._ctx, S.x.data_ptr(), float(S.thickness), float(S.buffer), 1,
                                        c.list.data_ptr(), c.e0e1.data_ptr(), int(c.capacity), c.cnt.data_ptr(),
                                        S._stream()), "detect_collision_pt")
    t = S.kernel_times()
    S.set_kernel_timing(False)
    for k in ("build_edge", "build_tri", "query_dcd_pt", "query_ccd_ee", "query_ccd_pt", "toi_ee", "toi_pt"):
        ms, n = t[k]
        assert n >= 1 and ms > 0.0, k
    # reading resets the counters
    assert all(n == 0 for _, n in S.kernel_times().values())


def _intersections_ref(oracle, soup, x1, exclude):
    out = {}
    for kind, prims in (("ee", soup.edges), ("pt", soup.triangles)):
        fn = oracle.ccd_candidates_ee if kind == "ee" else oracle.ccd_candidates_pt
        q, ids = fn(soup.x, x1, prims, soup.body, 1e-7, exclude)
        t, _ = oracle.collision_time(kind, soup.x, x1, q, relative=True, threads=8)
        hit = t < 1.0
        out[kind] = (q[hit], ids[hit], t[hit])
    return out


@pytest.mark.parametrize("capacity", [None, 4])
def test_new_intersections_matches_oracle(dev, oracle, capacity):
    """new_intersections (contact.py:587-670): the TOI < 1 pairs with their
    alpha; the lists grow when they overflow."""
    soup, dx = _ccd_setup(15)
    S = _solver(dev, soup, capacity=capacity)
    x0 = torch.from_numpy(soup.x).to(dev)
    x1_np = (soup.x - dx).astype(np.float32)
    x1 = torch.from_numpy(x1_np).to(dev)
    ret = S.new_intersections(x0, x1)
    ref = _intersections_ref(oracle, soup, x1_np, False)
    assert len(ref["ee"][0]) > 0 and len(ref["pt"][0]) > 0
    for kind, C, n in (("ee", S.contacts_new, S.n_contacts), ("pt", S.contacts_pt, S.n_contacts_pt)):
        r_q, r_ids, r_t = ref[kind]
        assert n == len(r_q)
        g_q = C.list[:n].cpu().numpy()
        g_ids = C.e0e1[:n].cpu().numpy()
        g_t = C.alpha[:n].cpu().numpy()
        go = np.lexsort((g_ids[:, 1], g_ids[:, 0]))
        ro = np.lexsort((r_ids[:, 1], r_ids[:, 0]))
        assert np.array_equal(g_ids[go], r_ids[ro]) and np.array_equal(g_q[go], r_q[ro])
        assert np.abs(g_t[go] - r_t[ro]).max() <= 1e-12
    tmin = min(ref["ee"][2].min(), ref["pt"][2].min())
    assert ret == pytest.approx(0.9 * tmin, abs=1e-12)
    assert S.raw_toi(torch.from_numpy(dx).to(dev)) == pytest.approx(tmin, abs=1e-12)
    # no motion: nothing recorded
    assert S.new_intersections(x0, x0.clone()) == 1.0
    assert S.n_contacts == 0 and S.n_contacts_pt == 0


def test_new_intersections_exclude_same_body(dev, oracle):
    soup, vel = scenes.icosphere_grid(seed=3, per_axis=2, spacing=1.001)
    dx = (-0.05 * vel[soup.body]).astype(np.float32)
    S = _solver(dev, soup, exclude=True)
    x0 = torch.from_numpy(soup.x).to(dev)
    x1_np = (soup.x - dx).astype(np.float32)
    S.new_intersections(x0, torch.from_numpy(x1_np).to(dev))
    ref = _intersections_ref(oracle, soup, x1_np, True)
    assert S.n_contacts == len(ref["ee"][0]) and S.n_contacts_pt == len(ref["pt"][0])
    if S.n_contacts:
        body = soup.body
        q = S.contacts_new.list[:S.n_contacts].cpu().numpy()
        assert (body[q[:, 0]] != body[q[:, 2]]).all()


@pytest.mark.parametrize("shift", [0.0, -0.3])
def test_count_intersections_matches_oracle(dev, oracle, shift):
    """Edge-triangle crossing count (the Newton harness's penetration check)."""
    soup = scenes.two_cubes_on_ground(seed=3, n=4, gap=2e-3, jitter=1e-4)
    x = soup.x.copy()
    x[soup.body == 1, 1] += np.float32(shift)
    x[soup.body == 1, 0] += np.float32(0.137 if shift else 0.0)
    S = _solver(dev, soup)
    n, ids = S.count_intersections(torch.from_numpy(x).to(dev), capacity=1 << 14)
    r_n, r_ids = oracle.count_intersections(x, soup.edges, soup.triangles)
    assert n == r_n
    assert (n > 0) == (shift != 0.0)
    g = ids.cpu().numpy()
    g = g[np.lexsort((g[:, 1], g[:, 0]))]
    assert np.array_equal(g, r_ids)


def test_solver_hessian_and_energy_match_oracle(dev, oracle):
    """ContactSolverBase.compute_collision_hessian / collision_energy_raw (the
    contact part of compute_A and compute_collision_energy,
    dynamic_contacts.py:281-305) after its own DCD, against the oracle on the
    same contact lists; a new b tensor is checked afresh, a misshaped one
    refused."""
    soup = scenes.two_cubes_on_ground(seed=17, n=8, gap=4e-4, jitter=2e-4)
    S = _solver(dev, soup)
    S.detect_collision()
    ee = S.contacts_new.list[:S.n_contacts].cpu().numpy()
    pt = S.contacts_pt.list[:S.n_contacts_pt].cpu().numpy()
    n_ee, n_pt = len(ee), len(pt)
    assert n_ee > 50 and n_pt > 50
    l0 = np.float32(TH) * np.float32(2)
    r_ee = oracle.contact_hessian("ee", soup.x, ee, l0, 1e8, n_total=n_ee + n_pt, threads=8)
    r_pt = oracle.contact_hessian("pt", soup.x, pt, l0, 1e8, offset=n_ee, n_total=n_ee + n_pt, threads=8)
    r_b = r_ee[3] + r_pt[3]
    for _ in range(2):
        b = torch.zeros_like(S.x)
        trip = S.compute_collision_hessian(b)
        torch.cuda.synchronize()
        assert np.array_equal(trip.rows.cpu().numpy(), r_ee[0] + r_pt[0])
        assert np.array_equal(trip.cols.cpu().numpy(), r_ee[1] + r_pt[1])
        assert np.abs(b.cpu().numpy() - r_b).max() <= 1e-5 * np.abs(r_b).max()
    e_ref = oracle.contact_energy("ee", soup.x, ee, l0, 1e8, threads=8) + oracle.contact_energy("pt", soup.x, pt, l0,
                                                                                               1e8, threads=8)
    assert S.collision_energy_raw() == pytest.approx(e_ref, rel=1e-9)
    with pytest.raises(ValueError):
        S.compute_collision_hessian(torch.zeros((3, 3), dtype=torch.float32, device=dev))


def test_misaligned_lists_are_refused(dev, oracle):
    """Record lists are written as 16-byte vectors and id lists as 8-byte
    pairs: a pointer off that alignment is refused with IPC_EINVAL before any
    launch, and the count stays untouched."""
    from ipc_warp_amd._lib import IPC_EINVAL, lib

    soup = scenes.two_cubes_on_ground(seed=3, n=4)
    S = _solver(dev, soup)
    rec = torch.zeros(4 * 64 + 4, dtype=torch.int32, device=dev)
    ids = torch.zeros(2 * 64 + 2, dtype=torch.int32, device=dev)
    cnt = torch.full((1,), -7, dtype=torch.int32, device=dev)
    xp = S.x.data_ptr()
    for r_off, i_off in ((4, 0), (0, 4)):
        r, i = rec.data_ptr() + r_off, ids.data_ptr() + i_off
        assert lib().ipc_detect_collision_ee(S._ctx, xp, TH, TH, 0, r, i, 64, cnt.data_ptr(), None) == IPC_EINVAL
        assert lib().ipc_detect_collision_pt(S._ctx, xp, TH, TH, 0, r, i, 64, cnt.data_ptr(), None) == IPC_EINVAL
        assert lib().ipc_ccd_candidates_ee(S._ctx, xp, xp, 1e-7, 0, r, i, 64, cnt.data_ptr(), None) == IPC_EINVAL
        assert lib().ipc_ccd_candidates_pt(S._ctx, xp, xp, 1e-7, 0, r, i, 64, cnt.data_ptr(), None) == IPC_EINVAL
        assert lib().ipc_detect_collision(S._ctx, xp, TH, TH, 0, r, i, 64, cnt.data_ptr(), rec.data_ptr(),
                                          ids.data_ptr(), 64, cnt.data_ptr(), None) == IPC_EINVAL
        assert b"aligned" in lib().ipc_last_error()
    torch.cuda.synchronize()
    assert int(cnt.item()) == -7

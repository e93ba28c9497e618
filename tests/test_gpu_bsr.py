"""GPU block-sparse assembly (warp.sparse.bsr_from_triplets / bsr_axpy at
dynamic_contacts.py:290-297) against the oracle's restatement.

Bar: offsets, columns and values bit-exact (the kernels sum each block's
triplets in fp32 in ascending triplet order, as the oracle does), on random
triplets that reach every row-sort path (rows of <= 64, <= 1024 and more
entries), on the contact Hessian triplets of the pressed two-cube scene and
the screw-and-nut scene, through a pattern refill and through axpy."""
import ctypes

import numpy as np
import pytest
import torch

from ipc_warp_amd import scenes, sparse
from ipc_warp_amd._lib import IPC_EINVAL, lib
from ipc_warp_amd.contact import ContactSolverBase

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _host(M: sparse.BsrMatrix):
    torch.cuda.synchronize()
    nnz = M.nnz_sync()
    return M.offsets.cpu().numpy(), M.columns[:nnz].cpu().numpy(), M.values[:nnz].cpu().numpy()


def _same(a, b):
    assert np.array_equal(a[0], b[0]), "offsets differ"
    assert np.array_equal(a[1], b[1]), "columns differ"
    assert np.array_equal(a[2].view(np.uint32), b[2].view(np.uint32)), "values differ"


def _triplets(seed, nr, nc, n):
    rng = np.random.default_rng(seed)
    rows = rng.integers(0, nr, n).astype(np.int32)
    cols = rng.integers(0, nc, n).astype(np.int32)
    # one row far beyond a wave's LDS sort, one between 64 and 1024 entries
    rows[: n // 4] = 3
    rows[n // 4: n // 4 + 700] = 11
    rng.shuffle(rows)
    vals = (rng.normal(size=(n, 3, 3)) * 10.0 ** rng.integers(-3, 4, (n, 1, 1))).astype(np.float32)
    vals[rng.random(n) < 0.1] = 0.0
    vals[rng.random(n) < 0.01, 1, 1] = -0.0
    bad = rng.random(n) < 0.03
    rows[bad & (rng.random(n) < 0.5)] = -5
    cols[bad & (rng.random(n) < 0.5)] = nc + 2
    return rows, cols, vals


def _to_dev(dev, rows, cols, vals):
    return (torch.from_numpy(rows).to(dev), torch.from_numpy(cols).to(dev),
            torch.from_numpy(np.ascontiguousarray(vals, np.float32)).to(dev))


@pytest.mark.parametrize("seed", [0, 1])
def test_random_triplets_bit_exact(dev, oracle, seed):
    nr, nc, n = 500, 480, 20000
    rows, cols, vals = _triplets(seed, nr, nc, n)
    _say(f"bsr random triplets seed={seed} n={n}")
    M = sparse.bsr_from_triplets(nr, nc, *_to_dev(dev, rows, cols, vals))
    got = _host(M)
    ref = oracle.bsr_from_triplets(nr, nc, rows, cols, vals)
    _same(got, ref)
    # kept triplets per row: row 3 beyond a wave's LDS sort, row 11 within it
    kept = (rows >= 0) & (rows < nr) & (cols >= 0) & (cols < nc) & (vals.reshape(n, 9) != 0).any(1)
    per_row = np.bincount(rows[kept], minlength=nr)
    assert per_row[3] > 1024 and 64 < per_row[11] <= 1024 and (per_row <= 64).sum() > nr // 2


def test_nan_blocks_are_kept(dev, oracle):
    rows = np.array([0, 0, 1, 2], np.int32)
    cols = np.array([1, 1, 0, 2], np.int32)
    vals = np.zeros((4, 3, 3), np.float32)
    vals[0, 0, 0] = np.nan
    vals[1, 2, 2] = 1.5
    vals[3, 1, 1] = 2.0
    M = sparse.bsr_from_triplets(3, 3, *_to_dev(dev, rows, cols, vals))
    o, c, v = _host(M)
    ro, rc, rv = oracle.bsr_from_triplets(3, 3, rows, cols, vals)
    assert np.array_equal(o, ro) and np.array_equal(c, rc)
    assert np.array_equal(v, rv, equal_nan=True)


def _contact_triplets(dev, soup, th, kappa, exclude):
    x = torch.from_numpy(soup.x).to(dev)
    S = ContactSolverBase(x, x.clone(), torch.from_numpy(soup.edges), torch.from_numpy(soup.triangles),
                          torch.from_numpy(soup.body), thickness=th, buffer_=th, kappa=kappa,
                          disable_self_collision_=exclude)
    S.detect_collision()
    b = torch.zeros_like(x)
    trip = S.compute_collision_hessian(b)
    return S, trip


@pytest.fixture(scope="module")
def cubes():
    return scenes.two_cubes_on_ground(seed=11, n=8, gap=4e-4, jitter=2e-4), 5e-4, 1e8, False


@pytest.fixture(scope="module")
def screw():
    soup = scenes.screw_and_nut(seed=0)[0]
    return soup, 5e-4, 1e8, True


@pytest.mark.parametrize("name", ["cubes", "screw"])
def test_contact_hessian_triplets_bit_exact(dev, oracle, request, name):
    soup, th, kappa, exclude = request.getfixturevalue(name)
    S, trip = _contact_triplets(dev, soup, th, kappa, exclude)
    n = trip.rows.shape[0]
    _say(f"bsr [{name}] contacts ee={S.n_contacts} pt={S.n_contacts_pt} triplets={n}")
    assert n > 1000
    nv = soup.x.shape[0]
    M = sparse.bsr_from_triplets(nv, nv, trip.rows, trip.cols, trip.vals)
    got = _host(M)
    rows, cols, vals = trip.rows.cpu().numpy(), trip.cols.cpu().numpy(), trip.vals.cpu().numpy()
    ref = oracle.bsr_from_triplets(nv, nv, rows, cols, vals)
    _same(got, ref)
    # the contact Hessian is symmetric: block (j, i) is block (i, j)
    # transposed, up to the rounding of each pair's blocks
    o, c, v = got
    r = np.repeat(np.arange(nv), np.diff(o))
    key = r.astype(np.int64) * nv + c
    tkey = c.astype(np.int64) * nv + r
    order = np.argsort(tkey)
    assert np.array_equal(np.sort(key), tkey[order])
    a, bt = v[np.argsort(key)], v[order].transpose(0, 2, 1)
    scale = np.maximum(np.abs(a).reshape(-1, 9).max(1), 1e-30)
    assert (np.abs(a - bt).reshape(-1, 9).max(1) / scale).max() < 1e-4


def test_refill_equals_rebuild(dev, oracle, cubes):
    soup, th, kappa, exclude = cubes
    S, trip = _contact_triplets(dev, soup, th, kappa, exclude)
    nv = soup.x.shape[0]
    M = sparse.bsr_from_triplets(nv, nv, trip.rows, trip.cols, trip.vals)
    # new values on the same rows/columns; no kept block becomes zero
    g = torch.Generator(device=dev).manual_seed(3)
    vals2 = trip.vals * 1.5 + torch.randn(trip.vals.shape, device=dev, generator=g) * (trip.vals != 0)
    M.refill(vals2)
    got = _host(M)
    ref = oracle.bsr_from_triplets(nv, nv, trip.rows.cpu().numpy(), trip.cols.cpu().numpy(), vals2.cpu().numpy())
    _same(got, ref)
    # twice in a row gives the same bits
    M.refill(vals2)
    _same(_host(M), ref)


def test_axpy_bit_exact(dev, oracle, cubes):
    soup, th, kappa, exclude = cubes
    S, trip = _contact_triplets(dev, soup, th, kappa, exclude)
    nv = soup.x.shape[0]
    H = sparse.bsr_from_triplets(nv, nv, trip.rows, trip.cols, trip.vals)
    # an elastic-like matrix: the vertex adjacency of the surface triangles
    tri = soup.triangles
    rows = np.repeat(tri, 3, axis=1).reshape(-1).astype(np.int32)
    cols = np.tile(tri, (1, 3)).reshape(-1).astype(np.int32)
    vals = np.random.default_rng(4).normal(size=(rows.size, 3, 3)).astype(np.float32)
    K = sparse.bsr_from_triplets(nv, nv, *_to_dev(dev, rows, cols, vals))
    h2 = float(np.float32(1e-4))
    ref = oracle.bsr_axpy(_host(H), _host(K), h2, 1.0)
    Z = sparse.bsr_axpy(H, K, h2, 1.0)
    assert Z is K
    _same(_host(Z), ref)


def test_capacity_overflow_reports_the_block_count(dev, oracle):
    nr, nc, n = 64, 64, 4000
    rows, cols, vals = _triplets(5, nr, nc, n)
    rows = np.clip(rows, -1, nr - 1)
    ref = oracle.bsr_from_triplets(nr, nc, rows, cols, vals)
    nnz = len(ref[1])
    r, c, v = _to_dev(dev, rows, cols, vals)
    ws = sparse._Workspace()
    off = torch.empty(nr + 1, dtype=torch.int32, device=dev)
    cap = nnz - 1
    col = torch.full((cap,), -3, dtype=torch.int32, device=dev)
    val = torch.zeros((cap, 3, 3), dtype=torch.float32, device=dev)
    st = lib().ipc_bsr_from_triplets(ws.handle, nr, nc, r.data_ptr(), c.data_ptr(), v.data_ptr(), n, off.data_ptr(),
                                     col.data_ptr(), val.data_ptr(), cap, None)
    assert st == nnz
    assert (col == -3).all() and np.array_equal(off.cpu().numpy(), ref[0])
    # a refill without a kept pattern is refused
    assert lib().ipc_bsr_refill(ws.handle, v.data_ptr(), n, val.data_ptr(), None) == IPC_EINVAL
    col = torch.empty((nnz,), dtype=torch.int32, device=dev)
    val = torch.empty((nnz, 3, 3), dtype=torch.float32, device=dev)
    st = lib().ipc_bsr_from_triplets(ws.handle, nr, nc, r.data_ptr(), c.data_ptr(), v.data_ptr(), n, off.data_ptr(),
                                     col.data_ptr(), val.data_ptr(), nnz, None)
    assert st == 0
    torch.cuda.synchronize()
    _same((off.cpu().numpy(), col.cpu().numpy(), val.cpu().numpy()), ref)


def test_empty_inputs(dev, oracle):
    z = torch.zeros(0, dtype=torch.int32, device=dev)
    M = sparse.bsr_from_triplets(5, 5, z, z, torch.zeros((0, 3, 3), dtype=torch.float32, device=dev))
    o, c, v = _host(M)
    assert o.tolist() == [0] * 6 and c.size == 0
    M.refill(torch.zeros((0, 3, 3), dtype=torch.float32, device=dev))
    M0 = sparse.bsr_from_triplets(0, 0, z, z, torch.zeros((0, 3, 3), dtype=torch.float32, device=dev))
    assert _host(M0)[0].tolist() == [0]
    # all triplets dropped
    rows = torch.tensor([0, 1], dtype=torch.int32, device=dev)
    M = sparse.bsr_from_triplets(2, 2, rows, rows, torch.zeros((2, 3, 3), dtype=torch.float32, device=dev))
    assert _host(M)[0].tolist() == [0, 0, 0]
    Z = sparse.bsr_axpy(M, None, 2.0, 1.0)
    assert _host(Z)[0].tolist() == [0, 0, 0]


def test_few_long_rows(dev, oracle):
    """The affine-body pattern the reference assembles (16 body triplets per
    vertex triplet onto a few body rows): 4 rows of ~500k triplets each over
    32 columns, through the segmented radix sort of long rows."""
    import time

    rng = np.random.default_rng(9)
    n = 2_000_000
    rows = rng.integers(0, 4, n).astype(np.int32)
    cols = rng.integers(0, 32, n).astype(np.int32)
    vals = rng.normal(size=(n, 3, 3)).astype(np.float32)
    r, c, v = _to_dev(dev, rows, cols, vals)
    M = sparse.bsr_from_triplets(8, 32, r, c, v)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sparse.bsr_set_from_triplets(M, r, c, v)
    torch.cuda.synchronize()
    _say(f"bsr few long rows: {time.perf_counter() - t0:.4f} s")
    _same(_host(M), oracle.bsr_from_triplets(8, 32, rows, cols, vals))

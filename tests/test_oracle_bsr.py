"""Pins the oracle's BSR assembly (warp.sparse.bsr_from_triplets / bsr_axpy
as called at dynamic_contacts.py:290-297; Warp itself is absent, so its
published contract is restated) against scipy's COO -> CSR duplicate sum,
dense sums, and known answers for the summation order and the pruning rule."""
import numpy as np
import pytest
import scipy.sparse as sp


def _random_triplets(rng, nr, nc, n, zero_frac=0.1, oob_frac=0.05):
    rows = rng.integers(0, nr, n).astype(np.int32)
    cols = rng.integers(0, nc, n).astype(np.int32)
    vals = rng.normal(size=(n, 3, 3)).astype(np.float32)
    vals[rng.random(n) < zero_frac] = 0.0
    bad = rng.random(n) < oob_frac
    rows[bad & (rng.random(n) < 0.5)] = -1
    cols[bad & (rng.random(n) < 0.5)] = nc
    return rows, cols, vals


def _kept(rows, cols, vals, nr, nc):
    return (rows >= 0) & (rows < nr) & (cols >= 0) & (cols < nc) & (vals.reshape(-1, 9) != 0).any(1)


def _dense(nr, nc, offsets, columns, values):
    D = np.zeros((3 * nr, 3 * nc))
    for r in range(nr):
        for k in range(offsets[r], offsets[r + 1]):
            D[3 * r:3 * r + 3, 3 * columns[k]:3 * columns[k] + 3] += values[k]
    return D


def _scipy_sum(rows, cols, vals, nr, nc, keep):
    r, c, v = rows[keep], cols[keep], vals[keep].astype(np.float64)
    rr = (3 * r[:, None, None] + np.arange(3)[:, None]).repeat(3, 2)
    cc = (3 * c[:, None, None] + np.arange(3)[None, :]).repeat(3, 1)
    return sp.coo_matrix((v.ravel(), (rr.ravel(), cc.ravel())), shape=(3 * nr, 3 * nc)).toarray()


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_bsr_from_triplets_matches_scipy_coo_sum(oracle, seed):
    rng = np.random.default_rng(seed)
    nr, nc = 40, 37
    rows, cols, vals = _random_triplets(rng, nr, nc, 3000)
    offsets, columns, values = oracle.bsr_from_triplets(nr, nc, rows, cols, vals)
    keep = _kept(rows, cols, vals, nr, nc)
    # pattern: exactly the distinct (row, column) keys of the kept triplets,
    # row by row with sorted columns
    keys = np.unique(rows[keep].astype(np.int64) * nc + cols[keep])
    assert np.array_equal(np.repeat(np.arange(nr), np.diff(offsets)) * nc + columns, keys)
    assert offsets[0] == 0 and offsets[-1] == len(keys)
    ref = _scipy_sum(rows, cols, vals, nr, nc, keep)
    got = _dense(nr, nc, offsets, columns, values)
    # fp32 sums of up to ~10 unit normals against the fp64 sum
    assert np.abs(got - ref).max() <= 1e-5 * max(1.0, np.abs(ref).max())


def test_bsr_sums_in_triplet_order_from_plus_zero(oracle):
    # 1e8 + 1 - 1e8 in fp32 is 0 in this order and 1 in the order (1e8, -1e8, 1)
    vals = np.zeros((3, 3, 3), np.float32)
    vals[:, 0, 0] = [1e8, 1.0, -1e8]
    _, _, v = oracle.bsr_from_triplets(1, 1, [0, 0, 0], [0, 0, 0], vals)
    assert v[0, 0, 0] == 0.0
    vals[:, 0, 0] = [1e8, -1e8, 1.0]
    _, _, v = oracle.bsr_from_triplets(1, 1, [0, 0, 0], [0, 0, 0], vals)
    assert v[0, 0, 0] == 1.0
    # a single -0.0 entry in a kept block sums to +0.0
    vals = np.zeros((1, 3, 3), np.float32)
    vals[0, 0, 0] = -0.0
    vals[0, 1, 1] = 2.0
    _, _, v = oracle.bsr_from_triplets(1, 1, [0], [0], vals)
    assert v[0, 1, 1] == 2.0 and not np.signbit(v[0, 0, 0])


def test_bsr_pruning_rule(oracle):
    vals = np.zeros((5, 3, 3), np.float32)
    vals[1, 2, 2] = -0.0           # all zero (negative zero): dropped
    vals[2, 0, 1] = np.nan         # NaN is nonzero: kept
    vals[3, 1, 0] = 1.0            # kept, but its row is out of range
    vals[4, 1, 0] = 1.0            # kept
    rows = np.array([0, 1, 2, 5, 3], np.int32)
    cols = np.array([0, 1, 2, 0, 3], np.int32)
    offsets, columns, values = oracle.bsr_from_triplets(4, 4, rows, cols, vals)
    assert offsets.tolist() == [0, 0, 0, 1, 2]
    assert columns.tolist() == [2, 3]
    assert np.isnan(values[0, 0, 1])


def test_bsr_empty_and_single_row(oracle):
    o, c, v = oracle.bsr_from_triplets(3, 3, np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros((0, 3, 3)))
    assert o.tolist() == [0, 0, 0, 0] and c.size == 0 and v.shape == (0, 3, 3)
    rng = np.random.default_rng(5)
    rows = np.zeros(2000, np.int32)
    cols = rng.integers(0, 500, 2000).astype(np.int32)
    vals = rng.normal(size=(2000, 3, 3)).astype(np.float32)
    o, c, v = oracle.bsr_from_triplets(1, 500, rows, cols, vals)
    assert o.tolist() == [0, len(np.unique(cols))]
    assert np.all(np.diff(c) > 0)


@pytest.mark.parametrize("alpha,beta", [(1.0, 1.0), (1e-4, 1.0), (-2.0, 0.5)])
def test_bsr_axpy_is_the_dense_sum_over_the_union(oracle, alpha, beta):
    rng = np.random.default_rng(7)
    nr = 30
    x = oracle.bsr_from_triplets(nr, nr, *_random_triplets(rng, nr, nr, 500, oob_frac=0))
    y = oracle.bsr_from_triplets(nr, nr, *_random_triplets(rng, nr, nr, 700, oob_frac=0))
    z = oracle.bsr_axpy(x, y, alpha, beta)
    Dx, Dy, Dz = _dense(nr, nr, *x), _dense(nr, nr, *y), _dense(nr, nr, *z)
    ref = alpha * Dx + beta * Dy
    assert np.abs(Dz - ref).max() <= 1e-6 * np.abs(ref).max()
    kx = set(zip(np.repeat(np.arange(nr), np.diff(x[0])), x[1]))
    ky = set(zip(np.repeat(np.arange(nr), np.diff(y[0])), y[1]))
    kz = list(zip(np.repeat(np.arange(nr), np.diff(z[0])), z[1]))
    assert kz == sorted(kx | ky)
    # one-sided blocks are exactly alpha * x (or beta * y) in fp32
    for (r, c), blk in zip(kz, z[2]):
        if (r, c) in kx and (r, c) not in ky:
            k = x[0][r] + int(np.searchsorted(x[1][x[0][r]:x[0][r + 1]], c))
            assert np.array_equal(blk, np.float32(alpha) * x[2][k])

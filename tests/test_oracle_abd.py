"""Pins the oracle's affine-body contact reduction (affine_body_dynamics.py:
144-190) by the mathematics it restates: with J the Jacobian of the vertex
positions x_i = p_b + r_i.x a0_b + r_i.y a1_b + r_i.z a2_b in the body
degrees of freedom (fixed bodies' columns removed), the reduced Hessian is
scale * J^T H J and the reduced force J^T f; and against the reference's
literal 16-triplets-per-triplet expansion summed by scipy."""
import numpy as np
import pytest
import scipy.sparse as sp


def _scene(seed, nb=5, nv=40, m=600):
    rng = np.random.default_rng(seed)
    body = np.sort(rng.integers(0, nb, nv)).astype(np.int32)
    rest = rng.normal(size=(nv, 3)).astype(np.float32)
    fixed = np.zeros(nb, np.int32)
    fixed[3] = 1
    rows = rng.integers(0, nv, m).astype(np.int32)
    cols = rng.integers(0, nv, m).astype(np.int32)
    vals = rng.normal(size=(m, 3, 3)).astype(np.float32)
    vals[rng.random(m) < 0.1] = 0.0
    rows[:5] = -1  # out of range: dropped
    force = rng.normal(size=(nv, 3)).astype(np.float32)
    return nb, body, rest, fixed, rows, cols, vals, force


def _jacobian(nb, body, rest, fixed):
    nv = body.size
    J = np.zeros((3 * nv, 12 * nb))
    for i in range(nv):
        b = body[i]
        if fixed[b]:
            continue
        c = [1.0, *rest[i].astype(np.float64)]
        for a in range(4):
            J[3 * i:3 * i + 3, 12 * b + 3 * a:12 * b + 3 * a + 3] = c[a] * np.eye(3)
    return J


def _dense(n, offsets, columns, values):
    D = np.zeros((3 * n, 3 * n))
    for r in range(n):
        for k in range(offsets[r], offsets[r + 1]):
            D[3 * r:3 * r + 3, 3 * columns[k]:3 * columns[k] + 3] += values[k]
    return D


@pytest.mark.parametrize("seed", [0, 1])
def test_reduction_is_jt_h_j(oracle, seed):
    nb, body, rest, fixed, rows, cols, vals, force = _scene(seed)
    nv = body.size
    scale = 1e-4
    o, c, v, a, rf = oracle.abd_reduce(nb, body, rest, fixed, rows, cols, vals, force, scale)
    ok = rows >= 0
    rr = (3 * rows[ok, None, None] + np.arange(3)[:, None]).repeat(3, 2)
    cc = (3 * cols[ok, None, None] + np.arange(3)[None, :]).repeat(3, 1)
    H = sp.coo_matrix((vals[ok].astype(np.float64).ravel(), (rr.ravel(), cc.ravel())), shape=(3 * nv, 3 * nv))
    J = _jacobian(nb, body, rest, fixed)
    ref = scale * (J.T @ (H @ J))
    got = _dense(4 * nb, o, c, v)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.allclose(rf.reshape(-1), J.T @ force.astype(np.float64).reshape(-1), rtol=1e-12, atol=1e-12)
    # fixed body rows/columns carry nothing; the pattern is whole body pairs
    assert not got[12 * 3:12 * 4].any() and not got[:, 12 * 3:12 * 4].any()
    assert np.all(np.diff(o) % 4 == 0) and o[-1] == len(c)
    assert (a >= np.abs(v) - 1e-15).all()


def test_matches_the_literal_expansion(oracle):
    nb, body, rest, fixed, rows, cols, vals, force = _scene(7, m=300)
    nv = body.size
    scale = 2.5
    # the reference's 16 dst triplets per src triplet (affine_body_dynamics.py:172-190)
    R, C, V = [], [], []
    for t in range(rows.size):
        vi, vj = rows[t], cols[t]
        if vi < 0:
            continue
        bi, bj = body[vi], body[vj]
        for a in range(4):
            for c in range(4):
                R.append(4 * bi + a)
                C.append(4 * bj + c)
                if fixed[bi] or fixed[bj]:
                    V.append(np.zeros((3, 3)))
                    continue
                ca = 1.0 if a == 0 else float(rest[vi][a - 1])
                cc = 1.0 if c == 0 else float(rest[vj][c - 1])
                V.append(scale * ca * cc * vals[t].astype(np.float64))
    V = np.array(V)
    R, C = np.array(R), np.array(C)
    rr = (3 * R[:, None, None] + np.arange(3)[:, None]).repeat(3, 2)
    cc = (3 * C[:, None, None] + np.arange(3)[None, :]).repeat(3, 1)
    ref = sp.coo_matrix((V.ravel(), (rr.ravel(), cc.ravel())), shape=(12 * nb, 12 * nb)).toarray()
    o, c, v, _, _ = oracle.abd_reduce(nb, body, rest, fixed, rows, cols, vals, None, scale)
    assert np.abs(_dense(4 * nb, o, c, v) - ref).max() <= 1e-12 * np.abs(ref).max()

"""Pins of the oracle's augmented-Lagrangian active set
(augmented_lagrangian.py:68-197, 506-547; the reference ships no tests or
fixtures for it). The hash is pinned by hand-derived values, including
products that wrap and the INT_MIN XOR; create and merge by a set-based
restatement in numpy over records in numpy's C-aligned structured layouts
(the Warp struct layout, written independently of the oracle's word
offsets): which constraints enter, which bucket each lands in, the order
within a bucket (decreasing input index: the reference's atomics with its
threads run in index order), the fields of a fresh entry, old payloads
carried bit for bit, and zeroed tails."""
import numpy as np
import pytest

import oracle

XC = {
    np.float32: np.dtype([("a", "<i4", 4), ("e", "<i4", 2), ("l0", "<f4"), ("alpha", "<f4"), ("lam", "<f4")],
                         align=True),
    np.float64: np.dtype([("a", "<i4", 4), ("e", "<i4", 2), ("l0", "<f8"), ("alpha", "<f8"), ("lam", "<f8")],
                         align=True),
}
AL = {
    s: np.dtype([("a", "<i4", 4), ("e", "<i4", 2), ("l0", s), ("alpha", s), ("lam", s), ("gamma", s), ("slack", s),
                 ("mollifier", s), ("distance_type", "<i4"), ("d0", s), ("grad", s, 12)], align=True)
    for s in (np.float32, np.float64)
}


def test_struct_sizes_match_the_c_layouts():
    assert (XC[np.float32].itemsize, XC[np.float64].itemsize) == (36, 48)
    assert (AL[np.float32].itemsize, AL[np.float64].itemsize) == (104, 184)
    assert AL[np.float64].fields["d0"][1] == 80  # 4 bytes of padding after distance_type


@pytest.mark.parametrize("e0,e1,T,want", [
    (1, 0, 8191, 6037),             # 73856093 mod 8191
    (0, 1, 8191, 2521),             # 19349663 mod 8191
    (1, 1, 8191, 7888),             # 0x466F45D ^ 0x127409F = 88192194
    (30, 0, 8191, 7347),            # 30 * 73856093 wraps to -2079284506
    (30, 7, 1000, 681),             # XOR -1945999681, magnitude taken
    (12345, 678, 1 << 20, 589231),  # 1162411439 mod 2^20
    (-(1 << 31), 0, 1000, 648),     # XOR = INT_MIN: magnitude 2^31
    (-(1 << 31), 0, 7, 2),
])
def test_hash_known_answers(e0, e1, T, want):
    assert oracle.al_hash_pos(e0, e1, T) == want


def test_hash_matches_int32_numpy_arithmetic():
    rng = np.random.default_rng(3)
    e0 = rng.integers(0, 1 << 24, 4000).astype(np.int32)
    e1 = rng.integers(0, 1 << 24, 4000).astype(np.int32)
    for T in (2, 97, 8191, 1 << 17):
        with np.errstate(over="ignore"):
            v = (e0 * np.int32(73856093)) ^ (e1 * np.int32(19349663))
        want = np.abs(v.astype(np.int64)) % T
        got = np.array([oracle.al_hash_pos(a, b, T) for a, b in zip(e0, e1)])
        np.testing.assert_array_equal(got, want)


def random_constraints(rng, n, nv, scalar, key_base=0, quant=32):
    rec = np.zeros(n, XC[scalar])
    rec["a"] = rng.integers(0, nv, (n, 4))
    i = rng.choice(1 << 16, n, replace=False) + key_base
    rec["e"][:, 0] = i
    rec["e"][:, 1] = i + 1 + rng.integers(0, 1000, n)
    rec["l0"] = 1e-3
    rec["alpha"] = (rng.integers(0, quant, n) / quant).astype(scalar) + scalar(rng.integers(0, 2, n) * 1e-13)
    rec["lam"] = 7.0  # ignored: a fresh entry starts with lam 0
    return rec


def words(rec):
    return rec.view(np.int32).reshape(len(rec), -1)


def buckets_of(keys, T):
    return np.array([oracle.al_hash_pos(a, b, T) for a, b in keys], dtype=np.int64)


def check_table(cs, entries, count, T, members, bucket, capacity):
    """members: source indices accepted (any order); bucket: their buckets.
    Returns the source index of every entry, checking the bucket layout,
    the decreasing-index order and cell_start."""
    assert cs[T] == count == len(members)
    order = np.lexsort((-members, bucket))  # by bucket, then decreasing index
    want_src = members[order]
    want_cs = np.searchsorted(bucket[order], np.arange(T + 1))
    np.testing.assert_array_equal(cs, want_cs)
    assert not entries[count:capacity].any()
    return want_src


@pytest.mark.parametrize("scalar", [np.float32, np.float64])
def test_vertex_toi_is_the_per_vertex_minimum(scalar):
    rng = np.random.default_rng(0)
    rec = random_constraints(rng, 2000, 300, scalar)
    rec["a"][5, 2] = 10_000  # out of range: skipped
    vt = oracle.al_vertex_toi(words(rec), np.ones(300))
    want = np.ones(300, scalar)
    a = rec["a"].astype(np.int64)
    for k in range(4):
        ok = a[:, k] < 300
        np.minimum.at(want, a[ok, k], rec["alpha"][ok])
    np.testing.assert_array_equal(vt, want)


@pytest.mark.parametrize("scalar", [np.float32, np.float64])
@pytest.mark.parametrize("T", [2, 61, 8192])
def test_create_hash_keeps_each_vertex_earliest_pairs(scalar, T):
    rng = np.random.default_rng(T)
    n, nv, cap = 3000, 400, 4096
    rec = random_constraints(rng, n, nv, scalar)
    vt = oracle.al_vertex_toi(words(rec), np.ones(nv))
    count, cs, ent = oracle.al_create_hash(words(rec), vt, T, cap)
    eps = scalar(1e-12)
    accept = np.zeros(n, bool)
    for k in range(4):
        accept |= rec["alpha"] <= (vt[rec["a"][:, k]] + eps).astype(scalar)
    assert 0 < accept.sum() < n
    members = np.flatnonzero(accept)
    bucket = buckets_of(rec["e"][members], T)
    src = check_table(cs, ent, count, T, members, bucket, cap)
    out = ent[:count].copy().view(AL[scalar]).reshape(-1)
    np.testing.assert_array_equal(out["a"], rec["a"][src])
    np.testing.assert_array_equal(out["e"], rec["e"][src])
    np.testing.assert_array_equal(out["l0"], rec["l0"][src])
    np.testing.assert_array_equal(out["alpha"], rec["alpha"][src])
    for f, v in (("lam", 0), ("gamma", 1), ("slack", 0), ("mollifier", 1), ("distance_type", 0), ("d0", 0)):
        assert (out[f] == v).all(), f
    assert not out["grad"].any()


@pytest.mark.parametrize("scalar", [np.float32, np.float64])
def test_merge_keeps_live_old_entries_and_new_keys(scalar):
    rng = np.random.default_rng(5)
    nv = 500
    old_rec = random_constraints(rng, 2500, nv, scalar)
    n_o, cs_o, ent_o = oracle.al_create_hash(words(old_rec), oracle.al_vertex_toi(words(old_rec), np.ones(nv)),
                                             8192, 4096)
    old = ent_o[:n_o].copy().view(AL[scalar]).reshape(-1)
    old["gamma"] = rng.choice([1.0, 0.5, 0.01, 0.009, 0.0], n_o).astype(scalar)
    old["lam"] = rng.normal(size=n_o)
    old["grad"] = rng.normal(size=(n_o, 12))
    old["d0"] = rng.normal(size=n_o)
    old["distance_type"] = rng.integers(0, 5, n_o)
    ent_o[:n_o] = old.view(np.int32).reshape(n_o, -1)
    # new constraints: half re-detect old keys, half are fresh
    new_rec = random_constraints(rng, 2500, nv, scalar, key_base=1 << 20)
    n_reuse = n_o // 2
    reuse = rng.choice(n_o, n_reuse, replace=False)
    new_rec["e"][:n_reuse] = old["e"][reuse]
    n_n, _, ent_n = oracle.al_create_hash(words(new_rec), oracle.al_vertex_toi(words(new_rec), np.ones(nv)), 8192,
                                          4096)
    new = ent_n[:n_n].copy().view(AL[scalar]).reshape(-1)
    T, cap = 2 * (n_o + n_n), n_o + n_n
    count, cs, ent = oracle.al_merge(cs_o, ent_o[:n_o], 8192, n_o, ent_n[:n_n], n_n, T, cap)
    old_keys = {tuple(k) for k in old["e"]}
    keep_old = np.flatnonzero(old["gamma"] >= scalar(0.01))
    keep_new = np.array([i for i in range(n_n) if tuple(new["e"][i]) not in old_keys], dtype=np.int64)
    assert 0 < len(keep_old) < n_o and 0 < len(keep_new) < n_n
    members = np.concatenate([keep_old, n_o + keep_new])
    both = np.concatenate([old, new])
    bucket = buckets_of(both["e"][members], T)
    src = check_table(cs, ent, count, T, members, bucket, cap)
    np.testing.assert_array_equal(ent[:count], both.view(np.int32).reshape(len(both), -1)[src])


def test_merge_of_empty_tables():
    count, cs, ent = oracle.al_merge(np.zeros(3, np.int32), np.zeros((1, 26), np.int32), 2, 0,
                                     np.zeros((1, 26), np.int32), 0, 2, 1)
    assert count == 0 and not cs.any() and not ent.any()


def _entry(scalar, ids, d0, grad, lam, gamma, slack):
    e = np.zeros(1, AL[scalar])
    e["a"] = ids
    e["d0"], e["grad"], e["lam"], e["gamma"], e["slack"] = d0, grad, lam, gamma, slack
    return e.view(np.int32).reshape(1, -1)


@pytest.mark.parametrize("scalar", [np.float32, np.float64])
def test_slack_and_forces_by_hand(scalar):
    """c = d0 + grad . x, by hand on dyadic values (exact in both precisions):
    d0 = 0.5, the vertices' gradients (1, 0, 0), (0, 2, 0), (0, 0, 3),
    (1, 1, 1) at x = (1, 0, 0), (0, 1, 0), (0, 0, 1), (0.5, 0.25, 0.25):
    c = 0.5 + 1 + 2 + 3 + 1 = 7.5 (augmented_lagrangian.py:76-85)."""
    x = np.array([[1, 0, 0], [0, 1, 0], [0, 0, 1], [0.5, 0.25, 0.25]], dtype=scalar)
    grad = [1, 0, 0, 0, 2, 0, 0, 0, 3, 1, 1, 1]
    e = _entry(scalar, [0, 1, 2, 3], 0.5, grad, lam=2.0, gamma=0.5, slack=9.0)
    s = oracle.al_update(e, x, mu=4.0, forces=False).view(AL[scalar])[0]
    assert s["slack"] == 7.0  # max(0, 7.5 - 2 / 4)
    f = oracle.al_update(e, x, mu=4.0, forces=True).view(AL[scalar])[0]
    assert f["lam"] == 0.0 and f["gamma"] == scalar(0.5) * scalar(0.9)  # slack 9 != 0: released, gamma decays
    e0 = _entry(scalar, [0, 1, 2, 3], 0.5, grad, lam=2.0, gamma=0.5, slack=0.0)
    f0 = oracle.al_update(e0, x, mu=4.0, forces=True).view(AL[scalar])[0]
    assert f0["lam"] == 2.0 - 4.0 * 7.5 and f0["gamma"] == 1.0  # slack 0: lam -= mu c, gamma reset
    # a constraint the multiplier outweighs has zero slack
    e1 = _entry(scalar, [0, 1, 2, 3], 0.5, grad, lam=64.0, gamma=1.0, slack=1.0)
    assert oracle.al_update(e1, x, mu=4.0, forces=False).view(AL[scalar])[0]["slack"] == 0.0
    # an entry naming a vertex out of range stays as it was
    e2 = _entry(scalar, [0, 1, 2, 9], 0.5, grad, lam=2.0, gamma=0.5, slack=3.0)
    assert np.array_equal(oracle.al_update(e2, x, mu=4.0, forces=True), e2)


def _al_entries(scalar, quads, l0):
    e = np.zeros(len(quads), AL[scalar])
    e["a"] = quads
    e["l0"] = l0
    e["gamma"] = 1.0
    e["mollifier"] = 1.0
    return e.view(np.int32).reshape(len(quads), -1)


PE_OF_TYPE = {1: (0, 2, 3), 2: (1, 2, 3), 3: (2, 0, 1), 4: (3, 0, 1)}


def _line_distance(p, a, b):
    e = b - a
    t = np.dot(p - a, e) / max(np.dot(e, e), 1e-30)
    return np.linalg.norm(p - (a + t * e))


@pytest.mark.parametrize("scalar", [np.float32, np.float64])
@pytest.mark.parametrize("kind", ["pt", "ee"])
def test_linearize_is_the_distance_first_order_model(oracle, scalar, kind):
    """linearize_constraints (augmented_lagrangian.py:270-323): the
    constraint c(x) = d0 + grad . x is the first-order model of the
    linearised distance minus l0 -- the Ericson distance (separately pinned)
    for a regular feature, the point-line distance of the chosen endpoint
    for a mollified edge pair (distance_type 1..4) -- so c(x_safe) equals
    that distance - l0 (up to the fp32 casts the reference classifies and
    measures with), grad is its gradient (central differences, fp64 records)
    and the four vertices' gradients sum to zero."""
    from ipc_warp_amd import scenes
    x, c = scenes.pair_stream(kind, 64, seed=3, dhat=1e-3, dtype=np.float64)
    quads = c.copy()
    if kind == "pt":  # the AL records (p, t0, t1, t2); pair_stream writes (p, t1, t0, t2)
        quads[:, [1, 2]] = c[:, [2, 1]]
    e = _al_entries(scalar, quads, 1e-3)
    out = oracle.al_linearize(e, x, x, kind == "pt").view(AL[scalar]).reshape(-1)
    xs = x.astype(scalar).astype(np.float64)
    assert kind == "pt" or (out["distance_type"] > 0).any()

    def dist(xx, q, t):
        if t > 0:
            i, j, k = PE_OF_TYPE[t]
            return _line_distance(xx[q[i]], xx[q[j]], xx[q[k]])
        if kind == "pt":
            return np.sqrt(oracle.triangle_distance_sq(xx[q[0]], xx[q[1]], xx[q[2]], xx[q[3]]))
        return np.sqrt(oracle.segment_distance_sq(xx[q[0]], xx[q[1]], xx[q[2]], xx[q[3]]))

    for i, q in enumerate(quads):
        t = int(out["distance_type"][i])
        g = out["grad"][i].astype(np.float64).reshape(4, 3)
        d = dist(xs, q, t)
        assert np.abs(g.sum(axis=0)).max() <= (1e-4 if scalar == np.float32 else 1e-9) * max(1.0, np.abs(g).max())
        model = float(out["d0"][i]) + float((g * xs[q]).sum())
        assert abs(model - (d - 1e-3)) <= (2e-6 if scalar == np.float32 else 2e-7), (i, t, model, d)
        if scalar == np.float32 or d < 5e-5:
            continue
        for k in range(4):
            for a in range(3):
                h = 1e-3 * d  # well inside the distance's curvature radius
                xp, xm = xs.copy(), xs.copy()
                xp[q[k], a] += h
                xm[q[k], a] -= h
                fd = (dist(xp, q, t) - dist(xm, q, t)) / (xp[q[k], a] - xm[q[k], a])
                # grad = grad(d^2) / (2 d_stored), d_stored the fp32 distance
                # the reference divides by (model + l0 in a fp64 record)
                scaled = g[k, a] * (model + 1e-3) / d
                assert abs(fd - scaled) <= 1e-6 * max(1.0, abs(fd)), (i, k, a, fd, scaled)


def test_linearize_parallel_edges_choose_an_overlapping_endpoint(oracle):
    """Two parallel edges 1e-4 apart (|ea x eb|^2 = 0 < eps_x, so the
    mollifier is 0): the point-edge representation takes an endpoint whose
    projection lies on the other edge, the first by the rule's order
    (augmented_lagrangian.py:208-267) -- here x2, projecting into edge a --
    and the distance is the 1e-4 gap."""
    x = np.array([[0.0, 0, 0], [1.0, 0, 0], [0.5, 1e-4, 0], [1.5, 1e-4, 0]])
    e = _al_entries(np.float64, np.array([[0, 1, 2, 3]]), 5e-5)
    out = oracle.al_linearize(e, x, x, False).view(AL[np.float64]).reshape(-1)[0]
    # x0 projects before x2..x3 (alpha < 0); x1 projects inside it; x2 inside x0..x1
    assert out["mollifier"] == 0.0
    assert out["distance_type"] == 2  # x1 onto edge b: the first in-range candidate of equal distance
    g = out["grad"].reshape(4, 3)
    assert np.allclose(g[1], [0, -1, 0]) and np.allclose(g[2] + g[3], [0, 1, 0])
    assert abs(out["d0"] + (g * x).sum() - (1e-4 - 5e-5)) < 1e-15

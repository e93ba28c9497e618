"""Pins of the oracle's fp64-scalar contact drivers (scalar_types.py with
USE_FLOAT32 = False; dynamic_contacts.py:151-246,374-398 at double
positions): each contact's blocks and force equal the single-pair
evaluation contact_pair(mode="fp64"), itself pinned by finite differences
and closed forms; the driver's energy is the sum of the pairs' energies; and
the driver's force is minus the central difference of its energy in double
positions, which an fp32 pipeline cannot resolve."""
import numpy as np
import pytest

from ipc_warp_amd import scenes

L0 = 2.0 ** -10  # exact in float and double, so contact_pair's float l0 agrees
KAPPA = 1e8


@pytest.mark.parametrize("kind", ["pt", "ee"])
def test_driver_equals_the_pinned_pair_evaluation(oracle, kind):
    x, c = scenes.pair_stream(kind, 200, seed=4, dhat=L0, dtype=np.float64)
    rows, cols, vals, b, act = oracle.contact_hessian_f64(kind, x, c, L0, KAPPA)
    assert act.all()
    for i in range(0, 200, 7):
        x12 = x[c[i]].reshape(12)
        e, f, H, _ = oracle.contact_pair(kind, x12, L0, KAPPA, mode="fp64")
        blocks = vals[16 * i:16 * i + 16].reshape(4, 4, 3, 3).transpose(0, 2, 1, 3).reshape(12, 12)
        np.testing.assert_array_equal(blocks, H)
        np.testing.assert_array_equal(b[c[i]].reshape(12), f)
        assert (rows[16 * i:16 * i + 16] == np.repeat(c[i], 4)).all()
        assert (cols[16 * i:16 * i + 16] == np.tile(c[i], 4)).all()
    E = oracle.contact_energy_f64(kind, x, c, L0, KAPPA)
    pairs = sum(oracle.contact_pair(kind, x[c[i]].reshape(12), L0, KAPPA, mode="fp64")[0] for i in range(200))
    assert E == pytest.approx(pairs, rel=1e-13)


@pytest.mark.parametrize("kind", ["pt", "ee"])
def test_force_is_minus_the_energy_gradient_in_double(oracle, kind):
    x, c = scenes.pair_stream(kind, 24, seed=9, dhat=L0, dtype=np.float64)
    _, _, _, b, act = oracle.contact_hessian_f64(kind, x, c, L0, KAPPA)
    assert act.all()
    rng = np.random.default_rng(1)
    for v in rng.choice(x.shape[0], 12, replace=False):
        for k in range(3):
            h = 1e-9 * max(1.0, abs(x[v, k]))
            xp, xm = x.copy(), x.copy()
            xp[v, k] += h
            xm[v, k] -= h
            g = (oracle.contact_energy_f64(kind, xp, c, L0, KAPPA, xcs=x) -
                 oracle.contact_energy_f64(kind, xm, c, L0, KAPPA, xcs=x)) / (xp[v, k] - xm[v, k])
            assert -g == pytest.approx(b[v, k], rel=2e-4, abs=1e-6 * np.abs(b).max())


def test_inactive_contacts_write_nothing(oracle):
    x, c = scenes.pair_stream("pt", 8, seed=2, dhat=L0, dtype=np.float64)
    x[c[3, 0]] += 1.0
    rows, cols, vals, b, act = oracle.contact_hessian_f64("pt", x, c, L0, KAPPA)
    assert not act[3] and act[[0, 1, 2, 4, 5, 6, 7]].all()
    assert not rows[48:64].any() and not vals[48:64].any() and not b[c[3]].any()


@pytest.mark.parametrize("kind", ["pt", "ee"])
def test_toi_f64_equals_the_double_driver_on_fp32_inputs(oracle, kind):
    """collision_time_f64 evaluates ccd/ccd.py:304-357 in double exactly like
    the fp32-position driver's double mode: on positions that fp32 holds
    exactly the two agree bit for bit."""
    x0, x1, q, _ = scenes.ccd_stream(kind, 4096, seed=12)
    r32, _ = oracle.collision_time(kind, x0, x1, q, relative=True, threads=4)
    r64 = oracle.collision_time_f64(kind, x0.astype(np.float64), x1.astype(np.float64), q, threads=4)
    assert np.array_equal(r64.view(np.uint64), r32.view(np.uint64))


@pytest.mark.parametrize("kind", ["pt", "ee"])
def test_toi_f64_resolves_the_unrounded_construction(oracle, kind):
    """On the unrounded double construction (coordinates ~10, impacts at a
    known time t) every impact comes back within 2e-10 of t; the same
    stream rounded to fp32 positions misses t by ~1e-6 at the median, so a
    driver that dropped the double positions would fail."""
    x0, x1, q, t = scenes.ccd_stream(kind, 4096, seed=13, dtype=np.float64)
    hits = t < 1
    r64 = oracle.collision_time_f64(kind, x0, x1, q, threads=4)
    assert np.abs(r64[hits] - t[hits]).max() <= 2e-10
    r32, _ = oracle.collision_time(kind, x0.astype(np.float32), x1.astype(np.float32), q, relative=True, threads=4)
    assert np.median(np.abs(r32[hits] - t[hits])) > 1e-7

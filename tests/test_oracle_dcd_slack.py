"""The DCD box slack against the reference's own fp32 distance routines.

The library inflates a DCD box by thickness + min(buffer, sigma_i), sigma_i
= thickness 2^-10 + 2^-16 M_i with M_i the primitive's largest coordinate
magnitude, where the reference (contact.py:199-336) inflates by thickness +
buffer. With one uniform slack sigma_min (M the smallest such magnitude in
the scene) every box is at most as wide as the library's, so if the oracle's
sets at buffer = sigma_min equal its sets at the reference's buffer, the
library's sets lie between the two and equal them too. Checked on pairs at
distances within a few ulps of 2 thickness, at coordinate scales up to the
largest where sigma stays below the buffer."""
import numpy as np
import pytest

from ipc_warp_amd import scenes

TH = 5e-4


def _sigma_min(x):
    m = float(np.abs(x).max(axis=1).min())
    return float(np.float32(TH * 2.0 ** -10 + m * 2.0 ** -16))


@pytest.mark.parametrize("offset", [0.0, 1.0, 8.0, 30.0])
@pytest.mark.parametrize("seed", [1, 2])
def test_slack_keeps_every_pair_the_distance_test_accepts(oracle, offset, seed):
    soup = scenes.borderline_pairs(offset, TH, seed=seed, n_case=512)
    s = _sigma_min(soup.x)
    assert s < TH  # the library takes the slack, not the buffer, at these scales
    for fn, topo in ((oracle.dcd_ee, soup.edges), (oracle.dcd_pt, soup.triangles)):
        ref = fn(soup.x, topo, soup.body, TH, TH)
        tight = fn(soup.x, topo, soup.body, TH, s)
        assert 100 < len(ref[0]) < 512
        assert np.array_equal(ref[0], tight[0]) and np.array_equal(ref[1], tight[1])


def test_an_inflation_short_by_ulps_drops_pairs(oracle):
    """The check has teeth: boxes inflated by 64 ulps of the coordinates less
    than thickness lose the axis-aligned pairs whose fp32 distance is within
    a few dozen ulps below 2 thickness."""
    soup = scenes.borderline_pairs(30.0, TH, seed=1, n_case=512)
    short = -64 * float(np.spacing(np.float32(30.0)))
    for fn, topo in ((oracle.dcd_ee, soup.edges), (oracle.dcd_pt, soup.triangles)):
        assert len(fn(soup.x, topo, soup.body, TH, short)[0]) < len(fn(soup.x, topo, soup.body, TH, TH)[0])

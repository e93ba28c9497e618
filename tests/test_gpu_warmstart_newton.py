"""The learned Newton warm start on a held-out cube drop (seed 100, 20
steps): every step converges to the 1e-6 relative energy tolerance, no
edge-triangle pair crosses at the end of any warm-started step, and the warm
start takes fewer Newton iterations in total than the plain start (the
previous state, as the reference does, time_integrator_base.py:375-401).
The saving is an average over steps (impact steps dominate it), so the run
spans 20 steps."""
import os

import pytest
import torch

from ipc_warp_amd import warmstart

pytestmark = pytest.mark.gpu

MODEL = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "ipc_warp_amd", "data",
                     "warmstart_cube_drop.npz")


def test_warm_start_saves_newton_iterations(dev):
    if not os.path.exists(MODEL):
        pytest.skip("no trained warm-start model committed")
    ws = warmstart.WarmStart.load(MODEL, dev)
    r = warmstart.compare_newton(100, 20, ws, dev, modes=("plain", "warm_start"))
    assert all(r["plain"]["converged"]) and all(r["warm_start"]["converged"])
    assert max(r["warm_start"]["intersections"]) == 0
    assert sum(r["warm_start"]["iterations"]) < sum(r["plain"]["iterations"])


def test_warm_start_candidates_are_finite(dev):
    if not os.path.exists(MODEL):
        pytest.skip("no trained warm-start model committed")
    ws = warmstart.WarmStart.load(MODEL, dev)
    sim = warmstart.drop_simulation(7, dev, n=4)
    cands = ws(sim)
    assert isinstance(cands, list) and len(cands) == 2
    for d in cands:
        assert d.shape == (sim.V, 3) and torch.isfinite(d).all()
        assert (d[sim.fixed] == 0).all()

"""The affine-body Newton harness (affine_body_dynamics.py:378-690) on the
screw turning in its fixed nut and on the 64-icosphere grid: the potential
decreases within each step, no edge-triangle pair crosses, the fixed nut
does not move, and the screw turns with its motor. Screw steps converge to
a 1e-6 relative decrease of the incremental potential; a grid step may
instead stall when the remaining decrease is below the energy's own
resolution (barrier terms of gaps ~1e-3 evaluated at fp32 positions), which
ends its line search."""
import math

import numpy as np
import pytest
import torch

from ipc_warp_amd import scenes
from ipc_warp_amd.abd import AbdIntegrator, screw_and_nut_integrator

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def test_screw_and_nut_steps(dev):
    I = screw_and_nut_integrator(0, dev)
    q_nut = I.q[1].clone()
    for s in range(3):
        st = I.step()
        _say(f"screw step {s}: {st.iterations} iterations, reference criterion {st.reference_criterion_iterations}")
        assert st.converged
        assert all(b <= a for a, b in zip(st.energies, st.energies[1:]))
        assert I.contact.count_intersections()[0] == 0
    assert torch.equal(I.q[1], q_nut)
    # the screw's frame turned about y by roughly the motor's angle
    a0 = I.q[0, 1]
    ang = math.atan2(-float(a0[2]), float(a0[0]))
    assert -3 * math.pi * 0.01 * 1.5 < ang < 0.0
    assert I.contact.n_contacts > 1000


def test_icosphere_grid_steps(dev):
    soup, vel = scenes.icosphere_grid(seed=0)
    I = AbdIntegrator(soup, dev, h=0.01, kappa=1e8)
    I.qdot[:, 0] = torch.from_numpy(np.asarray(vel, np.float64)).to(dev)
    for s in range(3):
        st = I.step()
        _say(f"grid step {s}: {st.iterations} iterations, converged {st.converged}, "
             f"energy {st.energies[0]:.9e} -> {st.energies[-1]:.9e}")
        assert st.converged or st.line_search_failures == 1
        assert all(b <= a for a, b in zip(st.energies, st.energies[1:]))
        assert st.iterations < I.max_iter
        assert I.contact.count_intersections()[0] == 0

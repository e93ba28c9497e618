"""GPU MLP kernels (forward, backward, loss, Adam) against the float64
oracle. Bar: fp32 forward within 1e-4 relative of the output scale,
gradients within 1e-3 relative of each tensor's scale (fp32 accumulation
over the batch), Adam update within 1e-6."""
import ctypes

import numpy as np
import pytest
import torch

from ipc_warp_amd import warmstart
from oracle import mlp

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 37, 3000])
def test_forward_backward_match_oracle(dev, n):
    rng = np.random.default_rng(n)
    p = mlp.init(3)
    x = rng.standard_normal((n, mlp.IN)).astype(np.float32)
    t = rng.standard_normal((n, mlp.OUT)).astype(np.float32)
    w = rng.random(n).astype(np.float32) + 0.5
    net = warmstart.MLP(dev, params=torch.from_numpy(p.astype(np.float32)).to(dev))
    xt = torch.from_numpy(x).to(dev)
    y, work = net.forward(xt, keep=True)
    loss, dy = net.mse(y, torch.from_numpy(t).to(dev), torch.from_numpy(w).to(dev))
    grad = torch.zeros_like(net.params)
    net.backward(xt, work, dy, grad)
    torch.cuda.synchronize()
    yr, cache = mlp.forward(p.astype(np.float32).astype(np.float64), x.astype(np.float64))
    lr_, dyr = mlp.mse(yr, t.astype(np.float64), w.astype(np.float64))
    gr = mlp.backward(p.astype(np.float32).astype(np.float64), cache, dyr)
    yg = y.cpu().numpy()
    assert np.abs(yg - yr).max() <= 1e-4 * max(1.0, np.abs(yr).max())
    assert loss == pytest.approx(lr_, rel=1e-4)
    gg = grad.cpu().numpy().astype(np.float64)
    o = 0
    for name, shape in mlp.SHAPES:
        k = int(np.prod(shape))
        a, b = gg[o:o + k], gr[o:o + k]
        assert np.abs(a - b).max() <= 1e-3 * max(1e-6, np.abs(b).max()), name
        o += k


def test_adam_step_matches_oracle(dev):
    rng = np.random.default_rng(5)
    n = 10007
    p, g = rng.standard_normal(n), rng.standard_normal(n)
    m, v = rng.standard_normal(n) * 0.1, rng.random(n) * 0.1
    pt, gt, mt, vt = (torch.from_numpy(a.astype(np.float32)).to(dev) for a in (p, g, m, v))
    warmstart.adam_step(pt, gt, mt, vt, lr=1e-3, step=7)
    torch.cuda.synchronize()
    pr, mr, vr = mlp.adam(p.astype(np.float32).astype(np.float64), g.astype(np.float32).astype(np.float64),
                          m.astype(np.float32).astype(np.float64), v.astype(np.float32).astype(np.float64),
                          1e-3, 0.9, 0.999, 1e-8, 7)
    assert np.abs(pt.cpu().numpy() - pr).max() <= 1e-6
    assert np.abs(mt.cpu().numpy() - mr).max() <= 1e-6


def test_training_reduces_loss(dev):
    rng = np.random.default_rng(6)
    n = 4096
    x = rng.standard_normal((n, mlp.IN)).astype(np.float32)
    A = rng.standard_normal((mlp.IN, mlp.OUT)) * 0.3
    t = (np.tanh(x @ A)).astype(np.float32)
    net = warmstart.MLP(dev, seed=0)
    xt, tt = torch.from_numpy(x).to(dev), torch.from_numpy(t).to(dev)
    first = net.train_epochs(xt, tt, None, epochs=1, batch=1024, lr=1e-3)
    last = net.train_epochs(xt, tt, None, epochs=30, batch=1024, lr=1e-3)
    assert last < 0.2 * first

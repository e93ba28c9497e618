"""Learned Newton warm start: the MLP and its training loop over the HIP kernels.

The network (48 -> 256 -> 256 -> 256 -> 3, GELU, fp32) maps per-vertex
features to the vertex's converged Newton displacement over a time step.
Every layer product, activation, loss and optimiser step runs in
libipc_warp_amd (``ipc_mlp_*``); PyTorch only provides the buffers.

There is no counterpart in the reference: its Newton loop starts from the
previous state (time_integrator_base.py:375-401).
"""
from __future__ import annotations

import math

import numpy as np
import torch

from ._lib import check, lib
from .ops import stream_handle

IN, HID, OUT = 48, 256, 3
N_PARAMS = HID * IN + HID + HID * HID + HID + HID * HID + HID + OUT * HID + OUT


def _p(t: torch.Tensor) -> int:
    if t.device.type != "cuda" or not t.is_contiguous():
        raise ValueError("expected a contiguous GPU tensor")
    return t.data_ptr()


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights (fan-in), zero biases, last layer scaled by 0.1."""
    rng = np.random.default_rng(seed)
    parts = []
    for shape, scale in (((HID, IN), 1.0), ((HID,), 0.0), ((HID, HID), 1.0), ((HID,), 0.0), ((HID, HID), 1.0),
                         ((HID,), 0.0), ((OUT, HID), 0.1), ((OUT,), 0.0)):
        if len(shape) == 2:
            parts.append(rng.standard_normal(shape).reshape(-1) * math.sqrt(2.0 / shape[1]) * scale)
        else:
            parts.append(np.zeros(shape))
    return np.concatenate(parts).astype(np.float32)


def adam_step(p, g, m, v, lr=1e-3, step=1, beta1=0.9, beta2=0.999, eps=1e-8):
    check(lib().ipc_mlp_adam(_p(p), _p(g), _p(m), _p(v), p.numel(), float(lr), float(beta1), float(beta2), float(eps),
                             int(step), stream_handle(p.device)), "mlp_adam")


class MLP:
    """Parameters, Adam state and HIP forward/backward of the warm-start network."""

    def __init__(self, device, params: torch.Tensor | None = None, seed: int = 0):
        self.device = device
        if params is None:
            params = torch.from_numpy(init_params(seed)).to(device)
        if params.numel() != N_PARAMS or params.dtype != torch.float32:
            raise ValueError(f"params must be {N_PARAMS} float32 values")
        self.params = params.contiguous()
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.grad = torch.zeros_like(self.params)
        self.step = 0
        # feature normalisation (set by fit_normalisation)
        self.x_mean = torch.zeros(IN, device=device)
        self.x_scale = torch.ones(IN, device=device)
        self.y_scale = torch.ones(OUT, device=device)

    # ------------------------------------------------------------------ kernels ----
    def forward(self, x: torch.Tensor, keep: bool = False):
        n = x.shape[0]
        if x.shape[1] != IN or x.dtype != torch.float32:
            raise ValueError(f"x must be float32 [n, {IN}]")
        y = torch.empty((n, OUT), dtype=torch.float32, device=self.device)
        if keep:
            work = torch.empty(6 * HID * n, dtype=torch.float32, device=self.device)
            check(lib().ipc_mlp_forward(_p(self.params), n, _p(x), _p(y), _p(work), None, stream_handle(self.device)),
                  "mlp_forward")
            return y, work
        scratch = torch.empty(3 * HID * max(n, 1), dtype=torch.float32, device=self.device)
        check(lib().ipc_mlp_forward(_p(self.params), n, _p(x), _p(y), None, _p(scratch), stream_handle(self.device)),
              "mlp_forward")
        return y, None

    def backward(self, x, work, dy, grad):
        n = x.shape[0]
        scratch = torch.empty(2 * HID * max(n, 1), dtype=torch.float32, device=self.device)
        check(lib().ipc_mlp_backward(_p(self.params), n, _p(x), _p(work), _p(dy), _p(grad), _p(scratch),
                                     stream_handle(self.device)), "mlp_backward")

    def mse(self, y, t, w=None):
        n = y.shape[0]
        wsum = float(n) if w is None else float(w.sum().item())
        dy = torch.empty_like(y)
        loss = torch.zeros(1, dtype=torch.float64, device=self.device)
        check(lib().ipc_mlp_mse(_p(y), _p(t), None if w is None else _p(w), n, wsum, _p(dy), _p(loss),
                                stream_handle(self.device)), "mlp_mse")
        return float(loss.item()), dy

    # ----------------------------------------------------------------- training ----
    def fit_normalisation(self, x: torch.Tensor, y: torch.Tensor):
        """Per-feature standardisation of inputs and per-axis scale of targets."""
        self.x_mean = x.mean(0)
        self.x_scale = x.std(0).clamp_min(1e-12)
        self.y_scale = y.abs().mean(0).clamp_min(1e-12)

    def normalise_x(self, x):
        return ((x - self.x_mean) / self.x_scale).contiguous()

    def train_epochs(self, x, t, w=None, epochs=1, batch=4096, lr=1e-3, seed=0) -> float:
        """Adam over shuffled mini-batches of already-normalised (x, t); returns the last epoch's mean loss."""
        n = x.shape[0]
        g = torch.Generator(device="cpu").manual_seed(seed + self.step)
        last = 0.0
        for _ in range(epochs):
            perm = torch.randperm(n, generator=g).to(self.device)
            tot, cnt = 0.0, 0
            for s0 in range(0, n, batch):
                idx = perm[s0:s0 + batch]
                xb, tb = x[idx].contiguous(), t[idx].contiguous()
                wb = None if w is None else w[idx].contiguous()
                y, work = self.forward(xb, keep=True)
                loss, dy = self.mse(y, tb, wb)
                check(lib().ipc_zero_f32(_p(self.grad), self.grad.numel(), stream_handle(self.device)), "zero_f32")
                self.backward(xb, work, dy, self.grad)
                self.step += 1
                adam_step(self.params, self.grad, self.m, self.v, lr=lr, step=self.step)
                tot += loss * xb.shape[0]
                cnt += xb.shape[0]
            last = tot / max(cnt, 1)
        return last

    def predict(self, x_raw: torch.Tensor) -> torch.Tensor:
        """Displacements for raw (unnormalised) features."""
        y, _ = self.forward(self.normalise_x(x_raw))
        return y * self.y_scale

    def state_dict(self):
        return {"params": self.params.cpu(), "x_mean": self.x_mean.cpu(), "x_scale": self.x_scale.cpu(),
                "y_scale": self.y_scale.cpu(), "step": self.step}

    def load_state_dict(self, d):
        self.params.copy_(d["params"].to(self.device))
        self.x_mean = d["x_mean"].to(self.device)
        self.x_scale = d["x_scale"].to(self.device)
        self.y_scale = d["y_scale"].to(self.device)
        self.step = int(d["step"])

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


BASE, SLOTS, SLOT_DIM = 12, 6, 6


def vertex_features(x: torch.Tensor, base: torch.Tensor | None, ee: torch.Tensor, n_ee: int, pt: torch.Tensor,
                    n_pt: int, dhat: float, out: torch.Tensor | None = None,
                    slots: torch.Tensor | None = None) -> torch.Tensor:
    """float32 [V, 48] network inputs: `base` [V, 12] followed by the vertex's 6 nearest
    contact records of the lists ee / pt (ipc_warmstart_features)."""
    V = x.shape[0]
    if x.dtype != torch.float32 or x.dim() != 2 or x.shape[1] != 3:
        raise ValueError("x must be float32 [V, 3]")
    if base is not None and (base.shape != (V, BASE) or base.dtype != torch.float32):
        raise ValueError(f"base must be float32 [V, {BASE}]")
    for name, c, n in (("ee", ee, n_ee), ("pt", pt, n_pt)):
        if n > 0 and (c.dtype != torch.int32 or c.dim() != 2 or c.shape[1] != 4 or c.shape[0] < n):
            raise ValueError(f"{name} must be int32 [>= {n}, 4]")
    if out is None:
        out = torch.empty((V, IN), dtype=torch.float32, device=x.device)
    elif out.shape != (V, IN) or out.dtype != torch.float32:
        raise ValueError(f"out must be float32 [V, {IN}]")
    if slots is None or slots.numel() < SLOTS * V:
        slots = torch.empty(SLOTS * max(V, 1), dtype=torch.int64, device=x.device)
    check(lib().ipc_warmstart_features(_p(x), V, None if base is None else _p(base), _p(ee) if n_ee else None, n_ee,
                                       _p(pt) if n_pt else None, n_pt, float(dhat), _p(slots), _p(out),
                                       stream_handle(x.device)), "warmstart_features")
    return out


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
        self.x_scale = x.std(0).clamp_min(1e-8)
        self.y_scale = y.abs().mean(0).clamp_min(1e-12)

    def normalise_x(self, x):
        # clipped so that a feature that was (nearly) constant in training cannot blow up
        return ((x - self.x_mean) / self.x_scale).clamp(-20.0, 20.0).contiguous()

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


# ------------------------------------------------------------ the warm start ----
FORCE_SCALE = 1.0e-4  # length scale of the asinh compression of the force features


def base_features(sim) -> torch.Tensor:
    """float32 [V, 12] at the start of a step of a ContactIntegrator `sim`:
    h v, c(h^2/m f_elastic), c(h^2/m f_contact), h (v - v_prev), with
    c(u) = s asinh(u / s), s = FORCE_SCALE; zero on fixed vertices. Runs
    the DCD at the previous state (the lists then feed the contact slots)."""
    h = sim.h
    sim._set(sim.x0.clone())
    grad_el, _ = sim.elastic.gradient_hessian(sim.x0)
    sim.contact.detect_collision()
    force = torch.zeros_like(sim.x)
    sim.contact.compute_collision_hessian(force)
    minv = (h * h / sim.m)[:, None]

    def c(u):
        return FORCE_SCALE * torch.asinh(u / FORCE_SCALE)

    base = torch.cat([h * sim.v, c(-minv * grad_el), c(minv * force.double()), h * (sim.v - sim.v_prev)], dim=1)
    base[sim.fixed] = 0.0
    return base.float().contiguous()


def step_features(sim) -> torch.Tensor:
    """float32 [V, 48]: base_features and the 6 nearest contact records."""
    base = base_features(sim)
    c = sim.contact
    return vertex_features(sim.x, base, c.contacts_new.list, c.n_contacts, c.contacts_pt.list, c.n_contacts_pt,
                           float(c.l0))


def extrapolation(sim) -> torch.Tensor:
    """Constant-acceleration extrapolation of the next displacement,
    h v + h (v - v_prev) (float64 [V, 3], zero on fixed vertices): the
    baseline the network corrects."""
    d = sim.h * (2.0 * sim.v - sim.v_prev)
    d[sim.fixed] = 0.0
    return d


def truncate_locally(sim, delta: torch.Tensor, rounds: int = 3, factor: float = 0.9) -> torch.Tensor:
    """Scales each vertex's displacement by `factor` times the smallest time of
    impact of the pairs it takes part in over the move x0 -> x0 + delta
    (ContactSolverBase.new_intersections records every pair with TOI < 1),
    so that a colliding corner does not cut the whole prediction short.
    Repeated while pairs still collide, at most `rounds` times; whatever
    remains is left to the integrator's CCD-bounded line search."""
    c = sim.contact
    x0 = sim.x0.float().contiguous()
    for _ in range(rounds):
        x1 = (sim.x0 + delta).float().contiguous()
        c.new_intersections(x0, x1)
        if c.n_contacts + c.n_contacts_pt == 0:
            break
        f = torch.ones(sim.V, dtype=torch.float64, device=delta.device)
        for C, n in ((c.contacts_new, c.n_contacts), (c.contacts_pt, c.n_contacts_pt)):
            if n:
                f.scatter_reduce_(0, C.list[:n].long().reshape(-1), (factor * C.alpha[:n]).repeat_interleave(4),
                                  reduce="amin")
        delta = delta * f[:, None]
    return delta


def neighbour_average(sim, u: torch.Tensor, rounds: int) -> torch.Tensor:
    """`rounds` passes of averaging u over each vertex and its tet neighbours
    (removes vertex-to-vertex noise from a per-vertex prediction)."""
    if rounds <= 0:
        return u
    A = getattr(sim, "_vertex_average", None)
    if A is None:
        T = sim.tets.long()
        i = T[:, :, None].expand(-1, 4, 4).reshape(-1)
        j = T[:, None, :].expand(-1, 4, 4).reshape(-1)
        n = sim.V
        ones = torch.ones(i.numel(), dtype=torch.float64, device=u.device)
        A = torch.sparse_coo_tensor(torch.stack([i, j]), ones, (n, n)).coalesce()
        idx = A.indices()
        val = torch.ones_like(A.values())  # each neighbour once
        deg = torch.zeros(n, dtype=torch.float64, device=u.device).index_add_(0, idx[0], val)
        A = torch.sparse_coo_tensor(idx, val / deg[idx[0]], (n, n)).coalesce()
        sim._vertex_average = A
    for _ in range(rounds):
        u = torch.sparse.mm(A, u) if u.dim() == 2 else u
    return u


class WarmStart:
    """Callable for ContactIntegrator.step(warm_start=...): the displacement
    extrapolation(sim) + MLP(features) (float64 [V, 3], zero on fixed
    vertices), optionally truncated per vertex (truncate_locally with
    `local_factor`) before the integrator's CCD-bounded line search. With
    mlp = None it is the extrapolation alone. With `gate`, the extrapolation
    alone is offered as a second candidate and the integrator keeps whichever
    reaches the lower incremental potential; `smooth` averages the network's
    correction over tet neighbourhoods that many times."""

    def __init__(self, mlp: MLP | None, local_factor: float | None = None, gate: bool = True, smooth: int = 2,
                 scale: float = 1.0):
        self.mlp = mlp
        self.local_factor = local_factor
        self.gate = gate
        self.smooth = smooth
        self.scale = scale

    def __call__(self, sim):
        base = extrapolation(sim)
        delta = base
        if self.mlp is not None:
            y = self.mlp.predict(step_features(sim)).double()
            y = neighbour_average(sim, y, self.smooth) * self.scale
            delta = base + y
            delta[sim.fixed] = 0.0
        if self.local_factor is not None:
            delta = truncate_locally(sim, delta, factor=self.local_factor)
        if self.gate and self.mlp is not None:
            return [delta, base]
        return delta

    def save(self, path: str):
        d = self.mlp.state_dict()
        np.savez(path, params=d["params"].numpy(), x_mean=d["x_mean"].numpy(), x_scale=d["x_scale"].numpy(),
                 y_scale=d["y_scale"].numpy(), step=np.int64(d["step"]), force_scale=np.float64(FORCE_SCALE))

    @classmethod
    def load(cls, path: str, device) -> "WarmStart":
        z = np.load(path)  # plain arrays, no pickled objects
        if float(z["force_scale"]) != FORCE_SCALE:
            raise ValueError("model was trained with another feature scaling")
        mlp = MLP(device, params=torch.from_numpy(z["params"].astype(np.float32)).to(device))
        mlp.x_mean = torch.from_numpy(z["x_mean"]).float().to(device)
        mlp.x_scale = torch.from_numpy(z["x_scale"]).float().to(device)
        mlp.y_scale = torch.from_numpy(z["y_scale"]).float().to(device)
        mlp.step = int(z["step"])
        return cls(mlp)


def record_trajectory(sim, steps: int, log=None):
    """Plain-Newton trajectory: per step, the features at the previous state and
    the target (converged displacement minus the extrapolation) of every free
    vertex."""
    X, T = [], []
    free = ~sim.fixed
    for s in range(steps):
        f = step_features(sim)
        x0 = sim.x0.clone()
        base = extrapolation(sim)
        st = sim.step()
        X.append(f[free])
        T.append((sim.x0 - x0 - base)[free].float())
        if log:
            log(s, st)
    return torch.cat(X), torch.cat(T)


def train(X: torch.Tensor, T: torch.Tensor, epochs: int = 60, batch: int = 4096, lr: float = 1e-3,
          seed: int = 0, log=None) -> MLP:
    """Adam with a cosine learning-rate decay over standardised features and
    per-axis scaled targets."""
    mlp = MLP(X.device, seed=seed)
    mlp.fit_normalisation(X, T)
    Xn = mlp.normalise_x(X)
    # a few impact steps carry residuals tens of scales wide: clip them
    Tn = (T / mlp.y_scale).clamp(-20.0, 20.0).contiguous()
    for e in range(epochs):
        lr_e = 1e-5 + 0.5 * (lr - 1e-5) * (1.0 + math.cos(math.pi * e / max(epochs, 1)))
        loss = mlp.train_epochs(Xn, Tn, epochs=1, batch=batch, lr=lr_e, seed=seed + 7919 * e)
        if log:
            log(e, loss)
    return mlp


def drop_simulation(seed: int, device, n: int = 8):
    """ContactIntegrator over scenes.cube_drop(seed) with its initial velocities."""
    from . import scenes
    from .integrator import ContactIntegrator

    soup, v, fixed = scenes.cube_drop(seed, n=n)
    sim = ContactIntegrator(soup, device, fixed=fixed)
    sim.v = torch.from_numpy(v).to(device)
    sim.v_prev = sim.v.clone()  # moving steadily before the first step
    return sim


def compare_newton(seed: int, steps: int, warm: WarmStart, device, n: int = 8, log=None,
                   modes=("plain", "extrapolation", "warm_start")) -> dict:
    """Newton iterations per step on the same seeded drop from the previous
    state (plain, the reference's start), from the extrapolation alone and
    from the learned warm start. Every step's final state is checked for
    penetration with the exact edge-triangle crossing count."""
    starts = {"plain": None, "extrapolation": WarmStart(None), "warm_start": warm}
    out = {}
    for mode in modes:
        sim = drop_simulation(seed, device, n=n)
        rec = {"iterations": [], "converged": [], "accepted": [], "choice": [], "intersections": [],
               "max_dx_over_h": []}
        for s in range(steps):
            st = sim.step(warm_start=starts[mode])
            n_cross, _ = sim.contact.count_intersections(sim.x0.float().contiguous())
            rec["iterations"].append(st.iterations)
            rec["converged"].append(bool(st.converged))
            rec["accepted"].append(bool(st.warm_start_accepted))
            rec["choice"].append(int(st.warm_start_choice))
            rec["intersections"].append(n_cross)
            rec["max_dx_over_h"].append(st.max_dx_over_h)
            if log:
                log(mode, s, st, n_cross)
        rec["mean_iterations"] = float(np.mean(rec["iterations"]))
        rec["crossing_rejections"] = int(sim.crossing_rejections)
        out[mode] = rec
    return out

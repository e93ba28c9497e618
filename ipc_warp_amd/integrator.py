"""Newton harness: implicit Euler with Neo-Hookean tets and the IPC barrier.

This reproduces the solve the contact loop serves — RodBCBase.step and
line_search (time_integrator_base.py:375-401,514-546) with RodComplexBC's
contact terms (dynamic_contacts.py:275-308) — so that Newton iteration
counts with and without the learned warm start can be compared:

* incremental potential E(x) = 1/2 |x - x_tilde|_M^2 + h^2 Psi(x) + h^2 B(x),
  x_tilde = x0 + h v + h^2 g, lumped masses, kinematic (fixed) vertices;
* every Newton iteration: DCD + barrier gradient/PSD Hessian (HIP library),
  elastic gradient/PSD Hessian, a Jacobi-preconditioned CG solve, a CCD
  upper bound (HIP library, 0.9 toi) and backtracking until the energy
  decreases and the end state has no edge-triangle crossing (at most 64
  halvings, alpha = 0 ends the solve);
* convergence: the relative decrease of E between iterations falls below
  `tol` (1e-6), or the line search finds no decrease while the Newton
  decrement is already within it (the full step's predicted decrease
  g.dx / 2 <= tol |E|: the search then fails at the rounding floor of the
  fp32 contact terms, not short of the minimum). The reference's own test
  max|dx| <= tol_v h is reported too.

The elasticity and the linear solve are harness code on PyTorch tensors;
they are not part of the contact path.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import torch

from .contact import ContactSolverBase


def lame(E: float, nu: float):
    return E / (2.0 * (1.0 + nu)), E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu))


@dataclass
class StepStats:
    iterations: int = 0
    line_search_failures: int = 0
    energies: list = field(default_factory=list)
    converged: bool = False
    max_dx_over_h: float = 0.0
    warm_start_accepted: bool = False
    warm_start_alpha: float = 0.0
    warm_start_choice: int = -1
    reference_criterion_iterations: int = 0


class NeoHookeanTets:
    """Compressible Neo-Hookean energy Psi = mu/2 (|F|^2 - 3) - mu ln J + lam/2 ln^2 J per tet,
    with the 9x9 dP/dF projected to PSD per tet (float64)."""

    def __init__(self, rest: torch.Tensor, tets: torch.Tensor, mu: float, lam: float):
        self.T = tets.long()
        X = rest.double()[self.T]  # [t, 4, 3]
        Dm = torch.stack([X[:, 1] - X[:, 0], X[:, 2] - X[:, 0], X[:, 3] - X[:, 0]], dim=2)  # columns
        self.W = Dm.det().abs() / 6.0
        self.Bm = torch.linalg.inv(Dm)  # [t, 3, 3]
        self.mu, self.lam = mu, lam
        # G: d vec(F) / d x (9 x 12) with vec row-major F[i][j] -> 3 i + j
        t = self.T.shape[0]
        G = torch.zeros((t, 9, 12), dtype=torch.float64, device=rest.device)
        for i in range(3):
            for j in range(3):
                for a in range(3):  # F = Ds Bm, Ds[:, a] = x_{a+1} - x_0
                    coef = self.Bm[:, a, j]
                    G[:, 3 * i + j, 3 * (a + 1) + i] += coef
                    G[:, 3 * i + j, i] -= coef
        self.G = G

    def deformation(self, x):
        Xc = x.double()[self.T]
        Ds = torch.stack([Xc[:, 1] - Xc[:, 0], Xc[:, 2] - Xc[:, 0], Xc[:, 3] - Xc[:, 0]], dim=2)
        return Ds @ self.Bm

    def energy(self, x) -> float:
        F = self.deformation(x)
        J = F.det()
        if bool((J <= 0).any()):
            return math.inf
        lnJ = torch.log(J)
        psi = 0.5 * self.mu * ((F * F).sum((1, 2)) - 3.0) - self.mu * lnJ + 0.5 * self.lam * lnJ * lnJ
        return float((self.W * psi).sum())

    def gradient_hessian(self, x):
        """(gradient [V,3], per-tet 12x12 PSD Hessians [t,12,12]); inverted tets contribute nothing."""
        F = self.deformation(x)
        J = F.det()
        ok = J > 0
        Fs = torch.where(ok[:, None, None], F, torch.eye(3, dtype=F.dtype, device=F.device).expand_as(F))
        Jp = torch.where(ok, J, torch.ones_like(J))
        lnJ = torch.log(Jp)
        FinvT = torch.linalg.inv(Fs).transpose(1, 2)
        P = self.mu * Fs + (self.lam * lnJ - self.mu)[:, None, None] * FinvT
        Finv = FinvT.transpose(1, 2)
        # dP_ij / dF_kl = mu d_ik d_jl + (mu - lam lnJ) Finv_li Finv_jk + lam FinvT_ij FinvT_kl
        t = F.shape[0]
        eye9 = torch.eye(9, dtype=F.dtype, device=F.device)
        A = self.mu * eye9.expand(t, 9, 9).clone()
        c2 = (self.mu - self.lam * lnJ)[:, None, None, None, None]
        T2 = c2 * torch.einsum("bli,bjk->bijkl", Finv, Finv)
        A = A + T2.reshape(t, 9, 9)
        v = FinvT.reshape(t, 9)
        A = A + self.lam * v[:, :, None] * v[:, None, :]
        A = 0.5 * (A + A.transpose(1, 2))
        ev, Q = torch.linalg.eigh(A)
        A = (Q * ev.clamp_min(0.0)[:, None, :]) @ Q.transpose(1, 2)
        w = (self.W * ok)[:, None, None]
        H = w * (self.G.transpose(1, 2) @ A @ self.G)
        gl = (self.W * ok)[:, None] * (self.G.transpose(1, 2) @ P.reshape(t, 9, 1)).squeeze(2)  # [t, 12]
        grad = torch.zeros((x.shape[0], 3), dtype=torch.float64, device=x.device)
        grad.index_add_(0, self.T.reshape(-1), gl.reshape(-1, 3))
        return grad, H


class ContactIntegrator:
    """Implicit Euler over a tet soup with fixed vertices, IPC contact and CCD line search."""

    def __init__(self, soup, device, h=1e-2, E=1e5, nu=0.4, rho=1000.0, gravity=(0.0, -9.8, 0.0),
                 thickness=5e-4, kappa=1e9, fixed=None, tol=1e-6, max_iter=100, cg_rtol=1e-6, cg_maxiter=2000):
        self.device = device
        self.h = h
        self.tol = tol
        self.max_iter = max_iter
        self.cg_rtol, self.cg_maxiter = cg_rtol, cg_maxiter
        self.V = soup.x.shape[0]
        self.x = torch.from_numpy(soup.x).to(device)  # float32, shared with the contact solver
        self.xd = self.x.double()                    # float64 master copy of the iterate
        self.x0 = self.xd.clone()
        self.v = torch.zeros_like(self.xd)
        self.v_prev = torch.zeros_like(self.xd)
        rest = torch.from_numpy(soup.xcs).to(device)
        self.tets = torch.from_numpy(soup.tets).to(device)
        mu, lam = lame(E, nu)
        self.elastic = NeoHookeanTets(rest, self.tets, mu, lam)
        m = torch.zeros(self.V, dtype=torch.float64, device=device)
        m.index_add_(0, self.tets.long().reshape(-1), (rho * self.elastic.W / 4.0).repeat_interleave(4))
        fixed = np.zeros(self.V, bool) if fixed is None else np.asarray(fixed, bool)
        self.fixed = torch.from_numpy(fixed).to(device)
        m[m == 0] = 1.0  # vertices without tets (the kinematic ground) carry a unit mass
        self.m = m
        self.free = (~self.fixed).double()[:, None].expand(-1, 3).reshape(-1)
        self.g = torch.tensor(gravity, dtype=torch.float64, device=device)
        self.contact = ContactSolverBase(self.x, rest, torch.from_numpy(soup.edges), torch.from_numpy(soup.triangles),
                                         torch.from_numpy(soup.body), thickness=thickness, buffer_=thickness,
                                         kappa=kappa, h=h)
        # COO pattern of the elastic Hessian blocks
        T = self.tets.long()
        vi = T[:, :, None].expand(-1, 4, 4)
        vj = T[:, None, :].expand(-1, 4, 4)
        ri = (3 * vi[..., None, None] + torch.arange(3, device=device)[:, None]).expand(-1, 4, 4, 3, 3)
        ci = (3 * vj[..., None, None] + torch.arange(3, device=device)[None, :]).expand(-1, 4, 4, 3, 3)
        self._el_rows = ri.reshape(-1)
        self._el_cols = ci.reshape(-1)
        self.frame = 0
        self.check_crossings = True
        self.crossing_rejections = 0

    # ------------------------------------------------------------------ energy ----
    def x_tilde(self):
        xt = self.x0 + self.h * self.v + self.h * self.h * self.g
        return torch.where(self.fixed[:, None], self.x0, xt)

    def _set(self, xd):
        self.xd = xd
        self.x.copy_(xd.float())

    def energy(self, xd, xt) -> float:
        self._set(xd)
        d = xd - xt
        e = 0.5 * float((self.m[:, None] * d * d * (~self.fixed)[:, None]).sum())
        psi = self.elastic.energy(xd)
        if not math.isfinite(psi):
            return math.inf
        e += self.h * self.h * psi
        e += self.contact.compute_collision_energy()
        return e

    # ----------------------------------------------------------------- newton ----
    def _system(self, xt):
        h2 = self.h * self.h
        self._set(self.xd)
        grad_el, H_el = self.elastic.gradient_hessian(self.xd)
        self.contact.detect_collision()
        force = torch.zeros_like(self.x)
        trip = self.contact.compute_collision_hessian(force)
        g = self.m[:, None] * (self.xd - xt) + h2 * grad_el - h2 * force.double()
        rows = [self._el_rows]
        cols = [self._el_cols]
        vals = [h2 * H_el.reshape(-1, 4, 3, 4, 3).permute(0, 1, 3, 2, 4).reshape(-1)]
        if trip.rows.numel():
            r = trip.rows.long()
            c = trip.cols.long()
            rr = (3 * r[:, None, None] + torch.arange(3, device=self.device)[:, None]).expand(-1, 3, 3)
            cc = (3 * c[:, None, None] + torch.arange(3, device=self.device)[None, :]).expand(-1, 3, 3)
            rows.append(rr.reshape(-1))
            cols.append(cc.reshape(-1))
            vals.append(h2 * trip.vals.double().reshape(-1))
        idx = torch.arange(3 * self.V, device=self.device)
        rows.append(idx)
        cols.append(idx)
        vals.append(self.m.repeat_interleave(3))
        A = torch.sparse_coo_tensor(torch.stack([torch.cat(rows), torch.cat(cols)]), torch.cat(vals),
                                    (3 * self.V, 3 * self.V)).coalesce().to_sparse_csr()
        return A, g.reshape(-1) * self.free

    def _cg(self, A, b):
        diag = torch.ones_like(b)
        crow, col, val = A.crow_indices(), A.col_indices(), A.values()
        rows = torch.repeat_interleave(torch.arange(b.numel(), device=b.device), crow[1:] - crow[:-1])
        on = rows == col
        diag.index_copy_(0, rows[on], val[on])
        Minv = self.free / diag
        x = torch.zeros_like(b)
        r = b.clone()
        z = Minv * r
        p = z.clone()
        rz = torch.dot(r, z)
        bn = torch.linalg.norm(b)
        if float(bn) == 0.0:
            return x
        for _ in range(self.cg_maxiter):
            Ap = (A @ p.unsqueeze(1)).squeeze(1) * self.free
            alpha = rz / torch.dot(p, Ap)
            x += alpha * p
            r -= alpha * Ap
            if float(torch.linalg.norm(r)) <= self.cg_rtol * float(bn):
                break
            z = Minv * r
            rz_new = torch.dot(r, z)
            p = z + (rz_new / rz) * p
            rz = rz_new
        return x

    def line_search(self, dx, xt, E0, max_halvings=64):
        """time_integrator_base.py:514-546 with the CCD bound of contact.py:539-585:
        x - alpha dx, alpha from the CCD bound, halved until the energy decreases;
        on failure the iterate is restored and alpha = 0."""
        x_start = self.xd
        upper = self.contact.collision_free_step(dx.float().contiguous())
        alpha = upper
        for _ in range(max_halvings):
            E1 = self.energy(x_start - alpha * dx, xt)
            if math.isfinite(E1) and E1 < E0:
                # The zero-thickness CCD certifies the straight path, but a
                # pair whose gap has collapsed below the fp32 resolution of
                # the positions can still end up crossed once the iterate is
                # rounded; such a move is rejected (exact crossing count).
                if self.check_crossings and self.contact.count_intersections()[0] > 0:
                    self.crossing_rejections += 1
                else:
                    return alpha, E1
            alpha *= 0.5
            if alpha <= 1.1920928955078125e-07 * max(1.0, upper):
                break
        self._set(x_start)
        return 0.0, E0

    def newton(self, xt, stats: StepStats):
        E0 = self.energy(self.xd, xt)
        stats.energies.append(E0)
        for it in range(self.max_iter):
            A, g = self._system(xt)
            dx = self._cg(A, g).reshape(-1, 3)
            alpha, E1 = self.line_search(dx, xt, E0)
            stats.iterations = it + 1
            stats.max_dx_over_h = float(dx.abs().max()) / self.h
            if alpha == 0.0:
                stats.line_search_failures += 1
                stats.converged = 0.5 * float(torch.dot(g, dx.reshape(-1))) <= self.tol * abs(E0)
                break
            stats.energies.append(E1)
            rel = (E0 - E1) / max(abs(E1), 1e-300)
            E0 = E1
            if rel <= self.tol:
                stats.converged = True
                break
        return stats

    def step(self, initial=None, warm_start=None) -> StepStats:
        """One time step. The first Newton iterate is the previous state, unless
        `initial` (float64 [V,3]) lowers the incremental potential, or
        `warm_start(self)` returns a displacement (float64 [V,3]), or a list of
        candidate displacements, along which a CCD-bounded backtracking search
        from the previous state finds a lower potential (the lowest over the
        candidates is kept). Neither counts as a Newton iteration (no linear
        solve)."""
        xt = self.x_tilde()
        stats = StepStats()
        self._set(self.x0.clone())
        if initial is not None:
            E_plain = self.energy(self.x0, xt)
            E_ws = self.energy(initial, xt)
            if math.isfinite(E_ws) and E_ws < E_plain:
                self._set(initial)
                stats.warm_start_accepted = True
            else:
                self._set(self.x0.clone())
        elif warm_start is not None:
            cands = warm_start(self)
            if isinstance(cands, torch.Tensor):
                cands = [cands]
            E_plain = self.energy(self.x0, xt)
            best_E, best_x, best_alpha, best_k = E_plain, None, 0.0, -1
            for k, delta in enumerate(cands):
                self._set(self.x0.clone())
                alpha, E1 = self.line_search(-delta, xt, E_plain, max_halvings=8)
                if alpha > 0.0 and E1 < best_E:
                    best_E, best_x, best_alpha, best_k = E1, self.xd.clone(), alpha, k
            self._set(self.x0.clone() if best_x is None else best_x)
            stats.warm_start_accepted = best_x is not None
            stats.warm_start_alpha = best_alpha
            stats.warm_start_choice = best_k
        self.newton(xt, stats)
        self.v_prev = self.v
        self.v = (self.xd - self.x0) / self.h
        self.v[self.fixed] = 0.0
        self.x0 = self.xd.clone()
        self.frame += 1
        return stats

    def safe_prediction(self, delta):
        """x0 + alpha delta with alpha = the CCD bound of the move (0.9 toi)."""
        dx = (-delta).float().contiguous()
        self._set(self.x0.clone())
        alpha = self.contact.collision_free_step(dx)
        return self.x0 + alpha * delta

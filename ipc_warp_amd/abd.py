"""Affine-body contact reduction (affine_body_dynamics.py:144-190) on the
GPU: the vertex contact Hessian triplets and forces carried onto each body's
affine degrees of freedom q_b = [p, a0, a1, a2] by ``ipc_abd_reduce_contacts``.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from ._lib import check, lib
from .ops import _need, stream_handle
from .sparse import BsrMatrix


class AbdReducer:
    """An ``ipc_abd`` workspace plus the capacity of its last output."""

    def __init__(self):
        h = ctypes.c_void_p()
        check(lib().ipc_abd_create(ctypes.byref(h)), "abd_create")
        self.handle = h
        self.capacity = 0

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                lib().ipc_abd_destroy(h)
            except Exception:  # interpreter shutdown: the process frees it
                pass
            self.handle = None

    def reduce(self, n_bodies: int, body: torch.Tensor, rest: torch.Tensor, fixed: torch.Tensor, rows: torch.Tensor,
               cols: torch.Tensor, vals: torch.Tensor, force: torch.Tensor | None = None, scale: float = 1.0,
               reduced_force: torch.Tensor | None = None):
        """_reduce_contact_hessian (scaled by `scale`) as a BSR over
        4 n_bodies block rows, and, when `force` is given,
        _reduce_contact_force's sums added into `reduced_force` (float32
        [4 n_bodies, 3], zeros when not given). Returns (BsrMatrix, force)."""
        dev = vals.device
        nb = int(n_bodies)
        n = int(rows.shape[0])
        if cols.shape[0] != n or vals.shape[0] != n:
            raise ValueError("rows, cols and vals must hold the same number of triplets")
        if force is not None and reduced_force is None:
            reduced_force = torch.zeros((4 * nb, 3), dtype=torch.float32, device=dev)
        offsets = torch.empty(4 * nb + 1, dtype=torch.int32, device=dev)
        while True:
            cap = max(self.capacity, 16)
            columns = torch.empty(cap, dtype=torch.int32, device=dev)
            values = torch.empty((cap, 3, 3), dtype=torch.float32, device=dev)
            st = check(lib().ipc_abd_reduce_contacts(
                self.handle, nb, _need(body, torch.int32, "body"), _need(rest, torch.float32, "rest", (3,)),
                _need(fixed, torch.int32, "fixed"), int(body.shape[0]), _need(rows, torch.int32, "rows"),
                _need(cols, torch.int32, "cols"), _need(vals, torch.float32, "vals", (3, 3)), n,
                None if force is None else _need(force, torch.float32, "force", (3,)), float(scale),
                offsets.data_ptr(), columns.data_ptr(), values.data_ptr(), cap,
                None if force is None else _need(reduced_force, torch.float32, "reduced_force", (3,)),
                stream_handle(dev)), "abd_reduce_contacts")
            if st == 0:
                break
            self.capacity = 1 << (int(st) - 1).bit_length()
        M = BsrMatrix(4 * nb, 4 * nb, offsets, columns, values)
        return M, reduced_force


# --------------------------------------------------------------------------
# Affine-body Newton harness (affine_body_dynamics.py:240-300,378-505,600-690)
# --------------------------------------------------------------------------

def mass_properties(world: np.ndarray, faces: np.ndarray, body: np.ndarray):
    """affine_body_dynamics.py:317-342: per body the volume centroid and
    volume of its closed surface (signed tets from the origin) and per vertex
    a third of its incident triangles' area."""
    nb = int(body.max()) + 1
    centers = np.zeros((nb, 3))
    volumes = np.zeros(nb)
    weights = np.zeros(len(world))
    for b in range(nb):
        fb = faces[body[faces[:, 0]] == b]
        if len(fb):
            a, c, d = world[fb[:, 0]], world[fb[:, 1]], world[fb[:, 2]]
            signed = np.einsum("ij,ij->i", a, np.cross(c, d)) / 6.0
            total = signed.sum()
            if abs(total) > 1.0e-12:
                centers[b] = np.sum(signed[:, None] * (a + c + d) / 4.0, axis=0) / total
                volumes[b] = abs(total)
            else:
                centers[b] = world[body == b].mean(axis=0)
                volumes[b] = max(float(np.prod(np.ptp(world[body == b], axis=0))), 1.0e-12)
            area = 0.5 * np.linalg.norm(np.cross(c - a, d - a), axis=1)
            np.add.at(weights, fb.reshape(-1), np.repeat(area / 3.0, 3))
        else:
            centers[b] = world[body == b].mean(axis=0)
            volumes[b] = 1.0
            weights[body == b] = 1.0
    return centers, volumes, weights


def mass_moments(local: np.ndarray, body: np.ndarray, weights: np.ndarray, volumes: np.ndarray, density: float):
    """affine_body_dynamics.py:344-358: the 4x4 moment matrix of [1, r] under
    the body's mass spread over its vertices by area weight."""
    nb = volumes.shape[0]
    M = np.zeros((nb, 4, 4))
    for b in range(nb):
        ids = np.flatnonzero(body == b)
        w = weights[ids]
        total = density * float(volumes[b])
        w = np.full(len(ids), total / len(ids)) if w.sum() == 0 else w / w.sum() * total
        coeff = np.column_stack((np.ones(len(ids)), local[ids]))
        M[b] = coeff.T @ (w[:, None] * coeff)
    return M


class AbdIntegrator:
    """Implicit Euler over affine bodies (q_b = [p, a0, a1, a2]) with the
    orthogonality potential, IPC contact through the HIP kernels and the
    contact reduction above, and a CCD-bounded backtracking line search. The
    body system is small (12 unknowns per body) and solved densely in fp64."""

    def __init__(self, soup, device, h=1e-2, density=1000.0, affine_stiffness=1.0e8, fixed_bodies=(), motors=None,
                 gravity=(0.0, 0.0, 0.0), thickness=5e-4, kappa=1e8, tol=1e-6, max_iter=100,
                 velocity_tolerance=5e-2):
        from .contact import ContactSolverBase

        self.device, self.h, self.tol, self.max_iter = device, h, tol, max_iter
        self.velocity_tolerance = velocity_tolerance
        world = soup.x.astype(np.float64)
        body = soup.body.astype(np.int64)
        centers, volumes, weights = mass_properties(world, soup.triangles.astype(np.int64), body)
        local = world - centers[body]
        self.nb = nb = centers.shape[0]
        self.V = world.shape[0]
        self.body = torch.from_numpy(body).to(device)
        self.body32 = self.body.int().contiguous()
        self.rest = torch.from_numpy(local).to(device)
        self.rest32 = self.rest.float().contiguous()
        M = mass_moments(local, body, weights, volumes, density)
        self.M = torch.from_numpy(M).to(device)                              # [nb, 4, 4]
        self.kv = torch.from_numpy(affine_stiffness * volumes).to(device)   # stiffness * volume per body
        fixed = np.zeros(nb, np.int32)
        fixed[list(fixed_bodies)] = 1
        self.fixed = torch.from_numpy(fixed).to(device)
        self.free = (self.fixed == 0)
        q = np.zeros((nb, 4, 3))
        q[:, 0] = centers
        q[:, 1:] = np.eye(3)[None]
        self.q = torch.from_numpy(q).to(device)
        self.q0 = self.q.clone()
        self.qdot = torch.zeros_like(self.q)
        self.g = torch.tensor(gravity, dtype=torch.float64, device=device)
        self.motor = torch.zeros((nb, 3), dtype=torch.float64, device=device)
        self.motor_on = torch.zeros(nb, dtype=torch.bool, device=device)
        for b, w in (motors or {}).items():
            self.motor[int(b)] = torch.tensor(w, dtype=torch.float64)
            self.motor_on[int(b)] = True
        self.x = self.vertices(self.q).float().contiguous()
        self.contact = ContactSolverBase(self.x, self.rest32, torch.from_numpy(soup.edges),
                                         torch.from_numpy(soup.triangles), torch.from_numpy(soup.body),
                                         thickness=thickness, buffer_=thickness, kappa=kappa, h=h,
                                         disable_self_collision_=True)
        self.reducer = AbdReducer()
        self.frame = 0
        self.check_crossings = True
        self.crossing_rejections = 0
        self.project_body_hessian = True

    # ------------------------------------------------------------ kinematics ----
    def vertices(self, q):
        """x_i = p + r.x a0 + r.y a1 + r.z a2 (affine_body_dynamics.py:41-47)."""
        qb = q[self.body]
        return qb[:, 0] + (self.rest[:, :, None] * qb[:, 1:]).sum(1)

    def vertex_direction(self, dq):
        """The vertex move of a body move dq (the map is linear in q,
        affine_body_dynamics.py:50-55)."""
        return self.vertices(dq)

    def q_tilde(self):
        """affine_body_dynamics.py:63-75 (_prepare_tilde)."""
        h = self.h
        qt = self.q0 + h * self.qdot
        qt[:, 0] = self.q0[:, 0] + h * self.qdot[:, 0] + h * h * self.g
        rot = torch.cross(self.motor[:, None, :].expand(-1, 3, -1), self.q0[:, 1:], dim=2)
        qt[:, 1:] = torch.where(self.motor_on[:, None, None], self.q0[:, 1:] + h * rot, qt[:, 1:])
        return qt

    def _set(self, q):
        self.q = q
        self.x.copy_(self.vertices(q).float())

    # --------------------------------------------------------------- energies ----
    def _body_terms(self, q, qt, need_hessian):
        """Inertia 0.5 (q - qt)^T M (q - qt) and the orthogonality potential
        kappa V h^2 (sum (a_i.a_i - 1)^2 + 2 sum_{i<j} (a_i.a_j)^2), with
        gradient and Hessian (affine_body_dynamics.py:86-141,200-238)."""
        h2 = self.h * self.h
        d = q - qt
        Md = torch.einsum("bij,bjk->bik", self.M, d)                        # [nb, 4, 3]
        inertia = 0.5 * (d * Md).sum((1, 2))
        A = q[:, 1:]                                                        # rows = a_i
        G = torch.einsum("bik,bjk->bij", A, A)                              # a_i . a_j
        eye3 = torch.eye(3, dtype=q.dtype, device=q.device)
        off = G - torch.diag_embed(torch.diagonal(G, dim1=1, dim2=2))
        ortho = (torch.diagonal(G, dim1=1, dim2=2) - 1.0).pow(2).sum(1) + off.pow(2).sum((1, 2))
        s = self.kv * h2
        grad = Md.clone()
        # 4 s [(a_i.a_i - 1) a_i + sum_{k != i} (a_i.a_k) a_k]
        og = torch.einsum("bik,bkj->bij", G - eye3, A)
        grad[:, 1:] += 4.0 * s[:, None, None] * og
        e = inertia + s * ortho
        e = torch.where(self.free, e, torch.zeros_like(e))
        grad[~self.free] = 0.0
        if not need_hessian:
            return float(e.sum()), grad, None
        nb = self.nb
        H = torch.zeros((nb, 4, 3, 4, 3), dtype=q.dtype, device=q.device)
        H += self.M[:, :, None, :, None] * eye3[None, None, :, None, :]
        sc = (4.0 * s)[:, None, None]
        for i in range(3):
            ai = A[:, i]
            blk = (G[:, i, i] - 1.0)[:, None, None] * eye3 + 2.0 * ai[:, :, None] * ai[:, None, :]
            for k in range(3):
                if k != i:
                    ak = A[:, k]
                    blk = blk + ak[:, :, None] * ak[:, None, :]
            H[:, i + 1, :, i + 1, :] += sc * blk
            for j in range(3):
                if j != i:
                    aj = A[:, j]
                    H[:, i + 1, :, j + 1, :] += sc * (G[:, i, j][:, None, None] * eye3 + aj[:, :, None] * ai[:, None, :])
        H = H.reshape(nb, 12, 12)
        if self.project_body_hessian:
            # the orthogonality Hessian is indefinite away from orthonormal
            # frames; its positive part keeps the Newton direction a descent
            # direction (the contact part is PSD already)
            ev, Q = torch.linalg.eigh(0.5 * (H + H.transpose(1, 2)))
            H = (Q * ev.clamp_min(0.0)[:, None, :]) @ Q.transpose(1, 2)
        H[~self.free] = torch.eye(12, dtype=q.dtype, device=q.device)
        return float(e.sum()), grad, H

    def energy(self, q, qt) -> float:
        self._set(q)
        e, _, _ = self._body_terms(q, qt, False)
        c = self.contact.compute_collision_energy()
        return e + c if math.isfinite(c) else math.inf

    # ----------------------------------------------------------------- newton ----
    def _system(self, qt):
        nb, h2 = self.nb, self.h * self.h
        self._set(self.q)
        _, grad, Hb = self._body_terms(self.q, qt, True)
        S = self.contact
        S.detect_collision()
        force = torch.zeros_like(self.x)
        trip = S.compute_collision_hessian(force)
        Mc, rf = self.reducer.reduce(nb, self.body32, self.rest32, self.fixed, trip.rows, trip.cols, trip.vals,
                                     force, h2)
        A = torch.zeros((12 * nb, 12 * nb), dtype=torch.float64, device=self.device)
        idx = torch.arange(nb, device=self.device)
        rr = (12 * idx[:, None, None] + torch.arange(12, device=self.device)[None, :, None]).expand(-1, 12, 12)
        cc = (12 * idx[:, None, None] + torch.arange(12, device=self.device)[None, None, :]).expand(-1, 12, 12)
        A[rr.reshape(-1), cc.reshape(-1)] = Hb.reshape(-1)
        nnz = int(Mc.offsets[-1].item())
        if nnz:
            rows = torch.repeat_interleave(torch.arange(4 * nb, device=self.device),
                                           (Mc.offsets[1:] - Mc.offsets[:-1]).long())
            r3 = (3 * rows[:, None, None] + torch.arange(3, device=self.device)[:, None]).expand(-1, 3, 3)
            c3 = (3 * Mc.columns[:nnz].long()[:, None, None] + torch.arange(3, device=self.device)[None, :]).expand(
                -1, 3, 3)
            A.index_put_((r3.reshape(-1), c3.reshape(-1)), Mc.values[:nnz].double().reshape(-1), accumulate=True)
        g = grad.reshape(nb, 4, 3) - h2 * rf.double().reshape(nb, 4, 3)
        g[~self.free] = 0.0
        return A, g.reshape(-1)

    def line_search(self, dq, qt, E0, max_halvings=64):
        """affine_body_dynamics.py:600-646: q - alpha dq, alpha from the CCD
        bound of the vertex move, halved until the energy decreases."""
        q_start = self.q
        self._set(q_start)
        upper = self.contact.collision_free_step(self.vertex_direction(dq).float().contiguous())
        alpha = upper
        for _ in range(max_halvings):
            E1 = self.energy(q_start - alpha * dq, qt)
            if math.isfinite(E1) and E1 < E0:
                if self.check_crossings and self.contact.count_intersections()[0] > 0:
                    self.crossing_rejections += 1
                else:
                    return alpha, E1
            alpha *= 0.5
            if alpha <= 1.1920928955078125e-07 * max(1.0, upper):
                break
        self._set(q_start)
        return 0.0, E0

    def newton(self, qt, stats):
        E0 = self.energy(self.q, qt)
        stats.energies.append(E0)
        for it in range(self.max_iter):
            A, g = self._system(qt)
            dq = torch.linalg.solve(A, g).reshape(self.nb, 4, 3)
            alpha, E1 = self.line_search(dq, qt, E0)
            stats.iterations = it + 1
            stats.max_dx_over_h = float(dq.abs().max()) / self.h
            if stats.reference_criterion_iterations == 0 and alpha > 0.0 and \
                    stats.max_dx_over_h <= self.velocity_tolerance:
                # time_integrator_base.py:392-396: max|dx| <= velocity_tolerance * h
                stats.reference_criterion_iterations = it + 1
            if alpha == 0.0:
                stats.line_search_failures += 1
                # no decrease left: converged when the predicted decrease is
                # within tolerance or the vertex move is below the fp32
                # resolution of the positions the energy is evaluated at
                move = float(self.vertex_direction(dq).abs().max())
                stats.converged = (0.5 * float(torch.dot(g, dq.reshape(-1))) <= self.tol * abs(E0)
                                   or move <= 4.0 * 2.0 ** -24 * float(self.x.abs().max()))
                break
            stats.energies.append(E1)
            rel = (E0 - E1) / max(abs(E1), 1e-300)
            E0 = E1
            if rel <= self.tol:
                stats.converged = True
                break
        return stats

    def step(self, initial=None):
        """One time step from the previous state (the reference's start,
        time_integrator_base.py:375-401), or from `initial` (float64
        [nb, 4, 3]) when it lowers the incremental potential and the move
        there is collision-free."""
        from .integrator import StepStats

        qt = self.q_tilde()
        stats = StepStats()
        self._set(self.q0.clone())
        if initial is not None:
            E_plain = self.energy(self.q0, qt)
            self._set(self.q0.clone())
            alpha, E1 = self.line_search(self.q0 - initial, qt, E_plain, max_halvings=8)
            stats.warm_start_accepted = alpha > 0.0
            stats.warm_start_alpha = alpha
        self.newton(qt, stats)
        self.qdot = (self.q - self.q0) / self.h
        self.qdot[~self.free] = 0.0
        self.q0 = self.q.clone()
        self.frame += 1
        return stats

    def extrapolation(self):
        """The inertial prediction q0 + h qdot (motor bodies turned), the
        initial iterate a plain extrapolating warm start would take."""
        qt = self.q_tilde()
        qt[~self.free] = self.q0[~self.free]
        return qt


def screw_and_nut_integrator(seed: int, device, kappa: float = 1.0e9):
    """ScrewAndNut (affine_body_dynamics.py:758-770) on the procedural screw
    and nut: h = 1e-2, nut fixed, no gravity, affine stiffness 1e8, the
    screw driven by a motor of -pi rad/s about its axis (y)."""
    from . import scenes

    soup = scenes.screw_and_nut(seed=seed)[0]
    return AbdIntegrator(soup, device, h=0.01, fixed_bodies=[1], motors={0: (0.0, -math.pi, 0.0)},
                         affine_stiffness=1.0e8, kappa=kappa)


def compare_newton(seed: int, steps: int, device, modes=("plain", "extrapolation")):
    """Newton iterations per step of the screw and nut started from the
    previous state (the reference's start) and from the inertial
    extrapolation, each on its own copy of the scene."""
    out = {}
    for mode in modes:
        I = screw_and_nut_integrator(seed, device)
        its, ref_its, conv, cross = [], [], [], []
        for _ in range(steps):
            st = I.step(initial=I.extrapolation() if mode == "extrapolation" else None)
            its.append(st.iterations)
            ref_its.append(st.reference_criterion_iterations or st.iterations)
            conv.append(bool(st.converged))
            cross.append(int(I.contact.count_intersections()[0]))
        out[mode] = {"iterations": its, "reference_criterion_iterations": ref_its, "converged": conv,
                     "intersections": cross}
    return out

"""Host-side mirror of the reference's contact solver over the HIP library.

``ContactSolverBase`` keeps the reference's method names, argument meaning
and error behaviour (contact.py:444-750, dynamic_contacts.py:249-308) so a
time integrator calls it the way RodBCBase / RodComplexBC do:

* ``detect_collision()`` fills ``contacts_new`` / ``contacts_pt`` and sets
  ``n_contacts`` / ``n_contacts_pt``, growing the lists to the next power of
  two on overflow (contact.py:523-537,687-725);
* ``collision_free_step(dx)`` returns ``0.9 * toi`` if ``toi < 1`` else 1
  (contact.py:539-585);
* ``compute_collision_hessian(b)`` launches contact_hessian_ee then
  contact_hessian_pt at offset n_ee into freshly zeroed triplets and adds
  the force into ``b`` (dynamic_contacts.py:280-288);
* ``compute_collision_energy()`` re-runs DCD and returns h^2 times the
  barrier energy (dynamic_contacts.py:299-308).

Every device buffer is a torch tensor on the GPU; every computation runs in
libipc_warp_amd. Module globals mirror contact.py:14-22.
"""
from __future__ import annotations

import ctypes
import weakref
from dataclasses import dataclass

import torch

from . import ops
from ._lib import IPCError, check, lib

_thickness = 0.01
contact_volume = 10000
buffer = _thickness
disable_self_collision = False
verbose = False
CCD_PADDING = 1e-7


@dataclass
class Contacts:
    """A contact list (contact.py:75-80) as SoA device tensors.

    ``list`` holds XConstraint.a1a2b1b2 (int32 [capacity, 4]) and ``e0e1``
    XConstraint.e0e1 (int32 [capacity, 2]); ``cnt`` is the device counter.
    Every contact's l0 is 2 * thickness (contact.py:180).
    """

    list: torch.Tensor
    e0e1: torch.Tensor
    cnt: torch.Tensor
    capacity: int
    alpha: torch.Tensor | None = None  # XConstraint.alpha (float64 [capacity]), set by new_intersections


@dataclass
class Triplets:
    """COO blocks (fem/fem.py:15-19): rows, cols int32 [n]; vals float32 [n, 3, 3]."""

    rows: torch.Tensor
    cols: torch.Tensor
    vals: torch.Tensor


def _new_contacts(capacity: int, device, cnt: torch.Tensor | None = None) -> Contacts:
    return Contacts(
        list=torch.zeros((capacity, 4), dtype=torch.int32, device=device),
        e0e1=torch.zeros((capacity, 2), dtype=torch.int32, device=device),
        cnt=torch.zeros(1, dtype=torch.int32, device=device) if cnt is None else cnt,
        capacity=capacity,
        alpha=torch.zeros(capacity, dtype=torch.float64, device=device),
    )


class ContactSolverBase:
    """Contact detection, barrier assembly and CCD over one mesh soup.

    x: float32 [V,3] device tensor of current positions (the reference's
    ``soup.x_transformed``, aliased to ``states.x``); xcs: rest positions;
    edges int32 [E,2]; triangles int32 [F,3]; body int32 [V].
    """

    def __init__(self, x, xcs, edges, triangles, body, thickness=None, buffer_=None, kappa=1e9, h=1.0,
                 disable_self_collision_=None, capacity=None):
        dev = x.device
        self.device = dev
        self.x = x
        self.xcs = xcs
        self._nv = int(x.shape[0])  # the vertex count the library context is built for
        self.edges = edges.to(dev, torch.int32).contiguous().reshape(-1, 2)
        self.triangles = triangles.to(dev, torch.int32).contiguous().reshape(-1, 3)
        self.body = body.to(dev, torch.int32).contiguous()
        self.thickness = float(_thickness if thickness is None else thickness)
        self.buffer = float(buffer if buffer_ is None else buffer_)
        self.contact_stiffness = float(kappa)
        self.h = float(h)
        self.disable_self_collision = bool(disable_self_collision if disable_self_collision_ is None
                                           else disable_self_collision_)
        self.has_triangles = self.triangles.shape[0] > 0
        cap = contact_volume if capacity is None else int(capacity)
        # both lists' device counters live in one tensor: one readback per DCD
        self._counts = torch.zeros(2, dtype=torch.int32, device=dev)
        self.contacts_new = _new_contacts(cap, dev, self._counts[0:1])
        self.contacts_pt = _new_contacts(cap, dev, self._counts[1:2])
        self.n_contacts = 0
        self.n_contacts_pt = 0
        self.collision_triplets = None
        handle = ctypes.c_void_p()
        check(lib().ipc_ctx_create(ctypes.byref(handle), int(x.shape[0]), self.edges.data_ptr(),
                                   int(self.edges.shape[0]), self.triangles.data_ptr(), int(self.triangles.shape[0]),
                                   self.body.data_ptr()), "ctx_create")
        self._ctx = handle

    def __del__(self):
        ctx = getattr(self, "_ctx", None)
        if ctx is not None and ctx.value:
            try:
                lib().ipc_ctx_destroy(ctx)
            except Exception:
                pass
            self._ctx = None

    # ------------------------------------------------------------- helpers ----
    @property
    def l0(self) -> float:
        """2 * thickness in fp32 (the reference's l0), cached per thickness."""
        cached = getattr(self, "_l0_cache", None)
        if cached is None or cached[0] != self.thickness:
            import numpy as np

            cached = (self.thickness, float(np.float32(self.thickness) * np.float32(2.0)))
            self._l0_cache = cached
        return cached[1]

    def _stream(self):
        return ops.stream_handle(self.device)

    def _ptr(self, name: str, t: torch.Tensor, dtype: torch.dtype, like: torch.Tensor | None = None) -> int:
        """t.data_ptr() after the ops argument checks (float32 [V, 3] with the
        context's V, or a tensor of dtype; shaped like `like` when given),
        which run again only when the tensor under this name is a different
        object or storage. The cache holds weak references: a caller's
        tensor is not kept alive by it."""
        cache = self.__dict__.setdefault("_checked", {})
        hit = cache.get(name)
        if hit is not None and hit[0]() is t:
            ptr = t.data_ptr()
            if hit[1] == ptr:
                return ptr
        ptr = ops._need(t, dtype, name, (3,) if dtype == torch.float32 else None)
        if dtype == torch.float32 and t.shape[0] != self._nv:
            raise ValueError(f"{name} must have {self._nv} rows (got {t.shape[0]})")
        if like is not None and t.shape != like.shape:
            raise ValueError(f"{name} must have the shape of x")
        cache[name] = (weakref.ref(t), ptr)
        return ptr

    def _grow_contact_list(self, contacts: Contacts, required: int) -> Contacts:
        """contact.py:523-537: regrow to the next power of two >= required."""
        capacity = 1 << (int(required) - 1).bit_length()
        grown = _new_contacts(capacity, self.device, contacts.cnt)
        if verbose:
            print(f"contact buffer grown to {capacity}")
        return grown

    # ---------------------------------------------------------- reference API ----
    def compute_V(self, ret=True):
        # x_transformed is tied to the solver's x (contact.py:683-685).
        return None

    def detect_collision(self):
        """contact.py:687-725: EE and PT lists in one library call (the two halves
        run concurrently) and one readback of both counts; a list that overflowed
        grows to the next power of two and the call runs again."""
        self.compute_V(ret=False)
        xp = self._ptr("x", self.x, torch.float32)
        while True:
            ee, pt = self.contacts_new, self.contacts_pt
            check(lib().ipc_detect_collision(self._ctx, xp, float(self.thickness), float(self.buffer),
                                             int(self.disable_self_collision), ee.list.data_ptr(), ee.e0e1.data_ptr(),
                                             int(ee.capacity), ee.cnt.data_ptr(), pt.list.data_ptr(),
                                             pt.e0e1.data_ptr(), int(pt.capacity) if self.has_triangles else 0,
                                             pt.cnt.data_ptr(), self._stream()), "detect_collision")
            n_ee, n_pt = self._counts.tolist()
            grown = False
            if n_ee > ee.capacity:
                self.contacts_new = self._grow_contact_list(ee, n_ee)
                grown = True
            if n_pt > pt.capacity:
                self.contacts_pt = self._grow_contact_list(pt, n_pt)
                grown = True
            if not grown:
                break
        self.n_contacts, self.n_contacts_pt = n_ee, n_pt
        if verbose:
            print(f"n ee contacts = {self.n_contacts}, n pt contacts = {self.n_contacts_pt}")

    def collision_free_step(self, dx) -> float:
        """Zero-thickness CCD upper bound for x -= alpha * dx (contact.py:539-585)."""
        toi = self.raw_toi(dx)
        return 0.9 * toi if toi < 1.0 else 1.0

    def raw_toi(self, dx, x=None) -> float:
        x = self.x if x is None else x
        xp = self._ptr("x" if x is self.x else "x_toi", x, torch.float32)
        dxp = self._ptr("dx", dx, torch.float32, x)
        out = ctypes.c_double(1.0)
        check(lib().ipc_collision_free_step(self._ctx, xp, dxp, float(CCD_PADDING),
                                            int(self.disable_self_collision), ctypes.byref(out), self._stream()),
              "collision_free_step")
        return float(out.value)

    def new_intersections(self, x0, x1) -> float:
        """contact.py:587-670 (the augmented-Lagrangian active set): every EE pair
        (edge_edge_intersections) and PT pair (point_triangle_intersections,
        contact.py:370-442) whose time of impact between x0 and x1 is < 1
        replaces contacts_new / contacts_pt, with alpha = that time
        (append_ccd, contact.py:185-197); the lists grow to the next power of
        two on overflow. Returns 0.9 * toi if toi < 1 else 1."""
        for t in (x0, x1):
            if t.shape != self.x.shape or t.dtype != torch.float32 or not t.is_contiguous():
                raise ValueError("x0 and x1 must be contiguous float32 tensors shaped like x")
        toi = 1.0
        kinds = [("ee", lib().ipc_ccd_intersections_ee)]
        if self.has_triangles:
            kinds.append(("pt", lib().ipc_ccd_intersections_pt))
        else:
            self.n_contacts_pt = 0
        for kind, fn in kinds:
            contacts = self.contacts_new if kind == "ee" else self.contacts_pt
            while True:
                t = ctypes.c_double(1.0)
                check(fn(self._ctx, x0.data_ptr(), x1.data_ptr(), float(CCD_PADDING), int(self.disable_self_collision),
                         contacts.list.data_ptr(), contacts.e0e1.data_ptr(), contacts.alpha.data_ptr(),
                         int(contacts.capacity), contacts.cnt.data_ptr(), ctypes.byref(t), self._stream()),
                      f"ccd_intersections_{kind}")
                n = int(contacts.cnt.item())
                if n <= contacts.capacity:
                    break
                contacts = self._grow_contact_list(contacts, n)
            toi = min(toi, t.value)
            if kind == "ee":
                self.contacts_new, self.n_contacts = contacts, n
            else:
                self.contacts_pt, self.n_contacts_pt = contacts, n
        return 0.9 * toi if toi < 1.0 else 1.0

    def count_intersections(self, x=None, capacity: int = 0):
        """Edge-triangle crossings of the soup at x (default: the solver's x);
        returns (count, int32 [min(count, capacity), 2] (edge, triangle) ids)."""
        x = self.x if x is None else x
        if x.shape != self.x.shape or x.dtype != torch.float32 or not x.is_contiguous():
            raise ValueError("x must be a contiguous float32 tensor shaped like the solver's x")
        ids = torch.zeros((max(capacity, 1), 2), dtype=torch.int32, device=self.device)
        cnt = torch.zeros(1, dtype=torch.int32, device=self.device)
        check(lib().ipc_count_intersections(self._ctx, x.data_ptr(), ids.data_ptr() if capacity else None,
                                            int(capacity), cnt.data_ptr(), self._stream()), "count_intersections")
        n = int(cnt.item())
        return n, ids[:min(n, capacity)]

    def ccd_candidates(self, kind: str, x0, x1, padding=CCD_PADDING, capacity=1 << 12):
        """Candidate quadruples and primitive ids of edge_edge_toi (kind 'ee') or
        point_triangle_toi ('pt') between x0 and x1 (ccd/ccd.py:360-427)."""
        fn = lib().ipc_ccd_candidates_ee if kind == "ee" else lib().ipc_ccd_candidates_pt
        for t, nm in ((x0, "x0"), (x1, "x1")):
            if t.shape != self.x.shape or t.dtype != torch.float32 or not t.is_contiguous():
                raise ValueError(f"{nm} must be a contiguous float32 tensor shaped like x")
        while True:
            quads = torch.zeros((capacity, 4), dtype=torch.int32, device=self.device)
            ids = torch.zeros((capacity, 2), dtype=torch.int32, device=self.device)
            cnt = torch.zeros(1, dtype=torch.int32, device=self.device)
            check(fn(self._ctx, x0.data_ptr(), x1.data_ptr(), float(padding), int(self.disable_self_collision),
                     quads.data_ptr(), ids.data_ptr(), int(capacity), cnt.data_ptr(), self._stream()),
                  f"ccd_candidates_{kind}")
            n = int(cnt.item())
            if n <= capacity:
                return quads[:n], ids[:n]
            capacity = 1 << (n - 1).bit_length()

    def last_ccd_counts(self):
        a, b = ctypes.c_int(), ctypes.c_int()
        check(lib().ipc_ctx_last_ccd_counts(self._ctx, ctypes.byref(a), ctypes.byref(b)), "ctx_last_ccd_counts")
        return a.value, b.value

    def set_kernel_timing(self, enable: bool = True):
        """Bracket the context's builds, queries and TOI launches with HIP events."""
        check(lib().ipc_ctx_set_timing(self._ctx, int(enable)), "ctx_set_timing")

    def kernel_times(self) -> dict:
        """{kind: (total_ms, launches)} since the last call (waits for the events)."""
        from ._lib import TIMER_KINDS

        ms = (ctypes.c_double * len(TIMER_KINDS))()
        n = (ctypes.c_longlong * len(TIMER_KINDS))()
        check(lib().ipc_ctx_kernel_times(self._ctx, ms, n), "ctx_kernel_times")
        return {k: (ms[i], n[i]) for i, k in enumerate(TIMER_KINDS)}

    def line_search_upper_bound(self, dx) -> float:
        return self.collision_free_step(dx)

    def new_triplets(self) -> Triplets:
        """Triplets for (n_ee + n_pt) * 16 blocks (dynamic_contacts.py:282-286).

        The kernels write every block they own (inactive contacts get zero
        blocks), so the buffers need no zeroing.
        """
        nnz = (self.n_contacts + self.n_contacts_pt) * 16
        return Triplets(
            rows=torch.empty(nnz, dtype=torch.int32, device=self.device),
            cols=torch.empty(nnz, dtype=torch.int32, device=self.device),
            vals=torch.empty((nnz, 3, 3), dtype=torch.float32, device=self.device),
        )

    def contact_hessian_ee(self, trip: Triplets, b: torch.Tensor):
        """wp.launch(contact_hessian_ee, dim=n_ee, ...) (dynamic_contacts.py:287)."""
        ops.contact_hessian_ee(self.x, self.xcs, self.contacts_new.list, self.n_contacts, self.l0,
                               self.contact_stiffness, trip.rows, trip.cols, trip.vals, b, 0)

    def contact_hessian_pt(self, trip: Triplets, b: torch.Tensor):
        """wp.launch(contact_hessian_pt, dim=n_pt, ..., offset=n_ee) (dynamic_contacts.py:288)."""
        ops.contact_hessian_pt(self.x, self.contacts_pt.list, self.n_contacts_pt, self.l0, self.contact_stiffness,
                               trip.rows, trip.cols, trip.vals, b, self.n_contacts)

    def compute_collision_hessian(self, b: torch.Tensor) -> Triplets:
        """Contact part of RodComplexBC.compute_A (dynamic_contacts.py:280-288):
        triplets for (n_ee + n_pt) * 16 blocks, the EE kernel, then the PT
        kernel at offset n_ee; forces are added into b."""
        trip = self.new_triplets()
        # The contact lists and triplets are the solver's own (sized by
        # detect_collision and new_triplets); x, xcs and b are checked once
        # per tensor (the library checks pointers and sizes on every call).
        x = self.x
        xp = self._ptr("x", x, torch.float32)
        check(lib().ipc_collision_hessian(xp, self._ptr("xcs", self.xcs, torch.float32, x), x.shape[0],
                                          self.contacts_new.list.data_ptr(), self.n_contacts,
                                          self.contacts_pt.list.data_ptr(), self.n_contacts_pt, self.l0,
                                          self.contact_stiffness, trip.rows.data_ptr(), trip.cols.data_ptr(),
                                          trip.vals.data_ptr(), self._ptr("b", b, torch.float32, x),
                                          self._stream()), "collision_hessian")
        self.collision_triplets = trip
        return trip

    def collision_energy_raw(self) -> float:
        """Barrier energy of the current contact lists (no DCD, no h^2)."""
        e = torch.empty(1, dtype=torch.float64, device=self.device)
        self.collision_energy_into(e)
        return float(e.item())

    def collision_energy_into(self, e: torch.Tensor):
        """The barrier energy of the current contact lists into the device
        double e[0], without a host transfer (dynamic_contacts.py:300-305)."""
        x = self.x
        xp = self._ptr("x", x, torch.float32)
        check(lib().ipc_collision_energy(xp, self._ptr("xcs", self.xcs, torch.float32, x), x.shape[0],
                                         self.contacts_new.list.data_ptr(), self.n_contacts,
                                         self.contacts_pt.list.data_ptr(), self.n_contacts_pt, self.l0,
                                         self.contact_stiffness, self._ptr("e", e, torch.float64), self._stream()),
              "collision_energy")

    def compute_collision_energy(self) -> float:
        """dynamic_contacts.py:299-308: DCD, then h^2 * sum of barrier energies."""
        self.detect_collision()
        return self.collision_energy_raw() * self.h * self.h


__all__ = ["ContactSolverBase", "Contacts", "Triplets", "IPCError"]


class FixedBodyContactSolver(ContactSolverBase):
    """Contact solver whose broad phase uses per-body BVHs built once in the
    bodies' rest frames (FixedBodyContactSolver, fixed_bvh_contacts.py:503-827).

    ``xcs`` holds body-local rest coordinates and ``affine_q`` is a float32
    device tensor of four vec3 blocks per body, ``[p, A[:,0], A[:,1], A[:,2]]``
    ([n_bodies * 4, 3]); ``x`` must equal A xcs + p per body.
    """

    def initialize_fixed_body_bvhs(self, affine_q, n_bodies: int | None = None, fixed_bodies=None):
        """fixed_bvh_contacts.py:511-597."""
        n_bodies = int(self.body.max().item()) + 1 if n_bodies is None else int(n_bodies)
        self._check_q(affine_q, n_bodies)
        check(lib().ipc_ctx_init_fixed_bodies(self._ctx, n_bodies, self.xcs.data_ptr()), "ctx_init_fixed_bodies")
        self.affine_q = affine_q
        self.n_bodies = n_bodies
        if fixed_bodies is None:
            self.abd_fixed = None
        else:
            self.abd_fixed = torch.as_tensor(fixed_bodies, dtype=torch.int32).to(self.device).contiguous()
            if self.abd_fixed.shape != (n_bodies,):
                raise ValueError("fixed_bodies must hold one flag per body")

    def _check_q(self, q, n_bodies):
        if q.dtype != torch.float32 or not q.is_contiguous() or q.numel() != 12 * n_bodies or q.device != self.device:
            raise ValueError(f"affine q must be a contiguous float32 device tensor of {n_bodies} x 4 vec3")

    def detect_collision(self):
        """fixed_bvh_contacts.py:758-827: both lists at once, regrown and rerun on overflow."""
        while True:
            c_ee, c_pt = self.contacts_new, self.contacts_pt
            check(lib().ipc_detect_collision_fixed(self._ctx, self.x.data_ptr(), self.affine_q.data_ptr(),
                                                   float(self.thickness), c_ee.list.data_ptr(), c_ee.e0e1.data_ptr(),
                                                   int(c_ee.capacity), c_ee.cnt.data_ptr(), c_pt.list.data_ptr(),
                                                   c_pt.e0e1.data_ptr(), int(c_pt.capacity), c_pt.cnt.data_ptr(),
                                                   self._stream()), "detect_collision_fixed")
            n_ee, n_pt = self._counts.tolist()
            grown = False
            if n_ee > c_ee.capacity:
                self.contacts_new = self._grow_contact_list(c_ee, n_ee)
                grown = True
            if n_pt > c_pt.capacity:
                self.contacts_pt = self._grow_contact_list(c_pt, n_pt)
                grown = True
            if not grown:
                self.n_contacts, self.n_contacts_pt = n_ee, n_pt
                return

    def raw_toi_fixed(self, dx, affine_dx):
        """Minimum TOI of the fixed-BVH CCD and whether it fell back to the world path."""
        if dx.shape != self.x.shape or dx.dtype != torch.float32 or not dx.is_contiguous():
            raise ValueError("dx must be a contiguous float32 tensor shaped like x")
        self._check_q(affine_dx, self.n_bodies)
        out = ctypes.c_double(1.0)
        fb = ctypes.c_int(0)
        check(lib().ipc_collision_free_step_fixed(self._ctx, self.x.data_ptr(), dx.data_ptr(), self.affine_q.data_ptr(),
                                                  affine_dx.data_ptr(),
                                                  None if self.abd_fixed is None else self.abd_fixed.data_ptr(),
                                                  float(CCD_PADDING), ctypes.byref(out), ctypes.byref(fb),
                                                  self._stream()), "collision_free_step_fixed")
        return float(out.value), bool(fb.value)

    def collision_free_step(self, dx, affine_dx=None) -> float:
        """fixed_bvh_contacts.py:710-756; without the affine step (no ABD state)
        it is the inherited world-space CCD."""
        if affine_dx is None:
            return ContactSolverBase.collision_free_step(self, dx)
        toi, _ = self.raw_toi_fixed(dx, affine_dx)
        return 0.9 * toi if toi < 1.0 else 1.0

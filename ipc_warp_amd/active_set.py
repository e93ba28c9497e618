"""The augmented-Lagrangian active set (augmented_lagrangian.py:42-197,
506-547, 705-717): hash tables of ALConstraint records built and merged by
the HIP kernels behind ``ipc_al_*`` in the C ABI.

A table is the reference's ``Hash``: ``table_size`` buckets, ``cell_start``
int32 [table_size + 1] (bucket b holds entries cell_start[b] ..
cell_start[b+1], cell_start[-1] is the count) and ``cell_entries``, here an
int32 [capacity, words] tensor holding ALConstraint records in the
reference's byte layout (``ALConstraintView`` names the fields).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import torch

from ._lib import check, lib
from .ops import _need, stream_handle

XCONSTRAINT_WORDS = {torch.float32: 9, torch.float64: 12}
ALCONSTRAINT_WORDS = {torch.float32: 26, torch.float64: 46}

DECAY_FACTOR = 0.9
ACTIVE_SET_DECAY_THRESHOLD = 0.01
TOI_FILTER_EPS = 1.0e-12
EE_MOLLIFIER_POINT_EDGE_THRESHOLD = 1.0


def _scalar_of(words: int, table: dict) -> torch.dtype:
    for dtype, w in table.items():
        if w == words:
            return dtype
    raise ValueError(f"records of {words} int32 words match no scalar type ({table})")


class ALConstraintView:
    """Field views of ALConstraint records (int32 [n, 26] with fp32 scalars,
    [n, 46] with fp64): a1a2b1b2, e0e1, l0, alpha, lam, gamma, slack,
    mollifier, distance_type, d0, grad (augmented_lagrangian.py:42-56)."""

    def __init__(self, records: torch.Tensor):
        self.records = records
        self.scalar = _scalar_of(records.shape[1], ALCONSTRAINT_WORDS)
        sw = 1 if self.scalar == torch.float32 else 2
        self._sw = sw
        self._d0 = 13 if sw == 1 else 20

    def _s(self, word: int, count: int = 1) -> torch.Tensor:
        w = self.records[:, word:word + count * self._sw]
        out = w.contiguous().view(self.scalar)
        return out[:, 0] if count == 1 else out

    @property
    def a1a2b1b2(self):
        return self.records[:, 0:4]

    @property
    def e0e1(self):
        return self.records[:, 4:6]

    @property
    def l0(self):
        return self._s(6)

    @property
    def alpha(self):
        return self._s(6 + self._sw)

    @property
    def lam(self):
        return self._s(6 + 2 * self._sw)

    @property
    def gamma(self):
        return self._s(6 + 3 * self._sw)

    @property
    def slack(self):
        return self._s(6 + 4 * self._sw)

    @property
    def mollifier(self):
        return self._s(6 + 5 * self._sw)

    @property
    def distance_type(self):
        return self.records[:, 6 + 6 * self._sw]

    @property
    def d0(self):
        return self._s(self._d0)

    @property
    def grad(self):
        return self._s(self._d0 + self._sw, 12)

    def set_scalar(self, name: str, values: torch.Tensor) -> None:
        """Write one scalar field (gamma, lam, ...) of every record."""
        word = {"l0": 6, "alpha": 6 + self._sw, "lam": 6 + 2 * self._sw, "gamma": 6 + 3 * self._sw,
                "slack": 6 + 4 * self._sw, "mollifier": 6 + 5 * self._sw, "d0": self._d0}[name]
        v = values.to(self.scalar).contiguous().view(torch.int32).reshape(self.records.shape[0], self._sw)
        self.records[:, word:word + self._sw] = v


def xconstraint_records(contacts: torch.Tensor, e0e1: torch.Tensor, alpha: torch.Tensor, n: int, l0: float,
                        scalar: torch.dtype = torch.float32) -> torch.Tensor:
    """XConstraint records (contact.py:61-67) of n contacts: int32 [n, 9]
    with fp32 scalars, [n, 12] with fp64; l0 for all, alpha float64 [n]
    (append_ccd, contact.py:185-197), lam 0."""
    words = XCONSTRAINT_WORDS[scalar]
    out = torch.zeros((max(int(n), 1), words), dtype=torch.int32, device=contacts.device)
    if n:
        check(lib().ipc_contact_records_unpack(_need(contacts, torch.int32, "contacts", (4,)),
                                               _need(e0e1, torch.int32, "e0e1", (2,)), int(n), float(l0),
                                               _need(alpha, torch.float64, "alpha"), out.data_ptr(), 4 * words,
                                               stream_handle(contacts.device)), "contact_records_unpack")
    return out[:n]


class _Workspace:
    """An ``ipc_active_set`` handle (device workspace of create and merge)."""

    def __init__(self):
        h = ctypes.c_void_p()
        check(lib().ipc_active_set_create(ctypes.byref(h)), "active_set_create")
        self.handle = h

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                lib().ipc_active_set_destroy(h)
            except Exception:  # interpreter shutdown: the process frees it
                pass
            self.handle = None


@dataclass
class Hash:
    """augmented_lagrangian.py:59-65."""

    table_size: int
    cell_start: torch.Tensor
    cell_entries: torch.Tensor


def accumulate_vertex_toi(constraints: torch.Tensor, n: int, vertex_toi: torch.Tensor) -> None:
    """augmented_lagrangian.py:105-113: vertex_toi[v] = min(vertex_toi[v],
    alpha) over the constraints on v. constraints are XConstraint records;
    vertex_toi is `scalar` [V]."""
    if n == 0:
        return
    scalar = _scalar_of(constraints.shape[1], XCONSTRAINT_WORDS)
    check(lib().ipc_al_vertex_toi(_need(constraints, torch.int32, "constraints"), 4 * constraints.shape[1], int(n),
                                  _need(vertex_toi, scalar, "vertex_toi"), int(vertex_toi.shape[0]),
                                  stream_handle(constraints.device)), "al_vertex_toi")


class ActiveSet:
    """augmented_lagrangian.py:506-572 (the hash operations): tables grow
    as the reference's do (create_hash: to the next power of two of the
    constraint count; merge: a fresh table sized for old + new)."""

    def __init__(self, contact_volume: int, device, scalar: torch.dtype = torch.float32):
        self.device = torch.device(device)
        self.scalar = scalar
        self._ws = _Workspace()
        self.hash = self.reserve(max(1, int(contact_volume)))

    def reserve(self, contact_volume: int) -> Hash:
        table_size = max(2, 2 * contact_volume)
        return Hash(table_size, torch.zeros(table_size + 1, dtype=torch.int32, device=self.device),
                    torch.zeros((contact_volume, ALCONSTRAINT_WORDS[self.scalar]), dtype=torch.int32,
                                device=self.device))

    @property
    def capacity(self) -> int:
        return int(self.hash.cell_entries.shape[0])

    def count(self) -> int:
        return int(self.hash.cell_start[-1].item())

    def entries(self, n: int | None = None) -> ALConstraintView:
        n = self.count() if n is None else n
        return ALConstraintView(self.hash.cell_entries[:n])

    def create_hash(self, constraints: torch.Tensor, n_constraints: int, vertex_toi: torch.Tensor) -> None:
        if constraints.shape[1] != XCONSTRAINT_WORDS[self.scalar]:
            raise ValueError(f"constraints must be XConstraint records of {XCONSTRAINT_WORDS[self.scalar]} words")
        required = max(1, int(n_constraints))
        if required > self.capacity:
            self.hash = self.reserve(1 << (required - 1).bit_length())
        h = self.hash
        check(lib().ipc_al_create_hash(self._ws.handle, constraints.data_ptr() if n_constraints else None,
                                       4 * XCONSTRAINT_WORDS[self.scalar], int(n_constraints),
                                       _need(vertex_toi, self.scalar, "vertex_toi") if n_constraints else None,
                                       int(vertex_toi.shape[0]), int(h.table_size), h.cell_start.data_ptr(),
                                       h.cell_entries.data_ptr(), self.capacity, stream_handle(self.device)),
              "al_create_hash")

    def merge(self, new: "ActiveSet") -> None:
        n_old = self.count()
        n_new = new.count()
        destination = self.reserve(max(1, n_old + n_new))
        words = ALCONSTRAINT_WORDS[self.scalar]
        check(lib().ipc_al_merge(self._ws.handle, 4 * words, self.hash.cell_start.data_ptr(),
                                 self.hash.cell_entries.data_ptr(), int(self.hash.table_size), n_old,
                                 new.hash.cell_entries.data_ptr(), n_new, int(destination.table_size),
                                 destination.cell_start.data_ptr(), destination.cell_entries.data_ptr(),
                                 int(destination.cell_entries.shape[0]), stream_handle(self.device)), "al_merge")
        self.hash = destination

    def linearize(self, x_safe: torch.Tensor, x_rest: torch.Tensor, is_point_triangle: bool,
                  point_edge_threshold: float = EE_MOLLIFIER_POINT_EDGE_THRESHOLD) -> None:
        """augmented_lagrangian.py:549-562 (linearize_constraints, :270-323)."""
        n = self.count()
        if n == 0:
            return
        check(lib().ipc_al_linearize(self.hash.cell_entries.data_ptr(), 4 * ALCONSTRAINT_WORDS[self.scalar], n,
                                     _need(x_safe, self.scalar, "x_safe", (3,)),
                                     _need(x_rest, self.scalar, "x_rest", (3,)), int(x_safe.shape[0]),
                                     int(bool(is_point_triangle)), float(point_edge_threshold),
                                     stream_handle(x_safe.device)), "al_linearize")

    def update_forces(self, x_hat: torch.Tensor, mu: float) -> None:
        """augmented_lagrangian.py:564-567 (update_active_set_forces)."""
        _update("forces", self.hash, self.count(), self.scalar, x_hat, mu)

    def update_slack(self, x_hat: torch.Tensor, mu: float) -> None:
        """augmented_lagrangian.py:569-572 (update_slack_variables)."""
        _update("slack", self.hash, self.count(), self.scalar, x_hat, mu)


def _update(kind: str, hash_: Hash, n: int, scalar: torch.dtype, x_hat: torch.Tensor, mu: float) -> None:
    if n == 0:
        return
    fn = lib().ipc_al_update_forces if kind == "forces" else lib().ipc_al_update_slack
    check(fn(hash_.cell_entries.data_ptr(), 4 * ALCONSTRAINT_WORDS[scalar], int(n),
             _need(x_hat, scalar, "x_hat", (3,)), int(x_hat.shape[0]), float(mu), stream_handle(x_hat.device)),
          f"al_update_{kind}")


class ALActiveSets:
    """The four tables and the per-vertex TOI of RodComplexAL
    (augmented_lagrangian.py:595-600) and update_active_set (:705-717)."""

    def __init__(self, n_vertices: int, device, initial_capacity: int = 1 << 12,
                 scalar: torch.dtype = torch.float32):
        self.scalar = scalar
        self.vertex_toi = torch.ones(n_vertices, dtype=scalar, device=device)
        self.active_set_ee = ActiveSet(initial_capacity, device, scalar)
        self.active_set_pt = ActiveSet(initial_capacity, device, scalar)
        self.detected_set_ee = ActiveSet(initial_capacity, device, scalar)
        self.detected_set_pt = ActiveSet(initial_capacity, device, scalar)

    def update_active_set(self, solver, x_safe: torch.Tensor, x: torch.Tensor) -> float:
        """Algorithm 3: every pair that meets between x_safe and x, kept when
        it is the earliest at one of its vertices, joins the active sets;
        entries whose gamma decayed below 0.01 leave. Returns
        new_intersections' step bound."""
        alpha = solver.new_intersections(x_safe, x)
        recs = []
        for contacts, n in ((solver.contacts_new, solver.n_contacts), (solver.contacts_pt, solver.n_contacts_pt)):
            recs.append(xconstraint_records(contacts.list, contacts.e0e1, contacts.alpha, n, solver.l0,
                                            self.scalar) if n else None)
        self.vertex_toi.fill_(1.0)
        for r in recs:
            if r is not None:
                accumulate_vertex_toi(r, r.shape[0], self.vertex_toi)
        empty = torch.zeros((0, XCONSTRAINT_WORDS[self.scalar]), dtype=torch.int32, device=x.device)
        self.detected_set_ee.create_hash(recs[0] if recs[0] is not None else empty, solver.n_contacts,
                                         self.vertex_toi)
        self.detected_set_pt.create_hash(recs[1] if recs[1] is not None else empty, solver.n_contacts_pt,
                                         self.vertex_toi)
        self.active_set_ee.merge(self.detected_set_ee)
        self.active_set_pt.merge(self.detected_set_pt)
        return alpha

"""Affine-body contact reduction (affine_body_dynamics.py:144-190) on the
GPU: the vertex contact Hessian triplets and forces carried onto each body's
affine degrees of freedom q_b = [p, a0, a1, a2] by ``ipc_abd_reduce_contacts``.
"""
from __future__ import annotations

import ctypes

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

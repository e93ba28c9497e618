"""Block-sparse matrices of 3x3 fp32 blocks: the ``warp.sparse`` calls the
Newton system assembly makes after the contact Hessian
(``bsr_from_triplets`` / ``bsr_axpy`` at dynamic_contacts.py:290-297,
affine_body_dynamics.py:500-503, augmented_lagrangian.py:631-632), run by the
HIP kernels behind ``ipc_bsr_*`` in the C ABI.
"""
from __future__ import annotations

import ctypes

import torch

from ._lib import check, lib
from .ops import _need, stream_handle


class _Workspace:
    """An ``ipc_bsr`` handle: device workspace plus the last built pattern."""

    def __init__(self):
        h = ctypes.c_void_p()
        check(lib().ipc_bsr_create(ctypes.byref(h)), "bsr_create")
        self.handle = h

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                lib().ipc_bsr_destroy(h)
            except Exception:  # interpreter shutdown: the process frees it
                pass
            self.handle = None


class BsrMatrix:
    """warp.sparse.BsrMatrix's fields for 3x3 fp32 blocks: ``offsets`` int32
    [nrow + 1], ``columns`` int32 [>= nnz] (sorted and distinct per row),
    ``values`` float32 [>= nnz, 3, 3]. ``nnz`` is known on the host after
    ``nnz_sync()``, as in Warp."""

    def __init__(self, nrow, ncol, offsets, columns, values, workspace=None, n_triplets=None):
        self.nrow, self.ncol = int(nrow), int(ncol)
        self.offsets, self.columns, self.values = offsets, columns, values
        self._ws = workspace
        self._n_triplets = n_triplets
        self.nnz = None

    @property
    def shape(self):
        return (self.nrow, self.ncol)

    @property
    def capacity(self) -> int:
        return int(self.columns.shape[0])

    def nnz_sync(self) -> int:
        self.nnz = int(self.offsets[self.nrow].item())
        return self.nnz

    def refill(self, values: torch.Tensor) -> "BsrMatrix":
        """New values for the triplets this matrix was built from (same rows
        and columns): each block sums its triplets in the build's order over
        the build's pattern (the pattern reuse of dynamic_contacts.py:314-333)."""
        if self._ws is None or self._n_triplets is None:
            raise ValueError("refill needs a matrix built by bsr_from_triplets whose pattern is unchanged")
        n = int(self._n_triplets)
        if values.shape[0] != n:
            raise ValueError(f"refill expects {n} triplet values (got {values.shape[0]})")
        check(lib().ipc_bsr_refill(self._ws.handle, _need(values, torch.float32, "values", (3, 3)), n,
                                   _need(self.values, torch.float32, "bsr values", (3, 3)),
                                   stream_handle(values.device)), "bsr_refill")
        return self


def bsr_from_triplets(rows_of_blocks: int, cols_of_blocks: int, rows: torch.Tensor, columns: torch.Tensor,
                      values: torch.Tensor) -> BsrMatrix:
    """warp.sparse.bsr_from_triplets (dynamic_contacts.py:290-296): triplets
    out of range or with nine zero values are dropped, duplicates summed in
    fp32 in triplet order. rows/columns int32 [n], values float32 [n, 3, 3]."""
    dest = BsrMatrix(rows_of_blocks, cols_of_blocks, None, None, None)
    return bsr_set_from_triplets(dest, rows, columns, values)


def bsr_set_from_triplets(dest: BsrMatrix, rows: torch.Tensor, columns: torch.Tensor,
                          values: torch.Tensor) -> BsrMatrix:
    """warp.sparse.bsr_set_from_triplets: bsr_from_triplets into an existing
    matrix, reusing its device workspace and arrays when they are large
    enough (no allocation on a repeated call of the same or smaller size)."""
    n = int(rows.shape[0])
    if columns.shape[0] != n or values.shape[0] != n:
        raise ValueError("rows, columns and values must hold the same number of triplets")
    dev = values.device
    if dest._ws is None:
        dest._ws = _Workspace()
    if dest.offsets is None or dest.offsets.shape[0] != dest.nrow + 1 or dest.offsets.device != dev:
        dest.offsets = torch.empty(dest.nrow + 1, dtype=torch.int32, device=dev)
    # capacity >= n: the pattern always fits, so the call never synchronises
    if dest.columns is None or dest.columns.shape[0] < max(n, 1) or dest.columns.device != dev:
        dest.columns = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
        dest.values = torch.empty((max(n, 1), 3, 3), dtype=torch.float32, device=dev)
    dest._n_triplets = None
    check(lib().ipc_bsr_from_triplets(dest._ws.handle, dest.nrow, dest.ncol, _need(rows, torch.int32, "rows"),
                                      _need(columns, torch.int32, "columns"),
                                      _need(values, torch.float32, "values", (3, 3)), n, dest.offsets.data_ptr(),
                                      dest.columns.data_ptr(), dest.values.data_ptr(), dest.capacity,
                                      stream_handle(dev)), "bsr_from_triplets")
    dest._n_triplets = n
    dest.nnz = None
    return dest


def bsr_zeros(rows_of_blocks: int, cols_of_blocks: int, device) -> BsrMatrix:
    offsets = torch.zeros(int(rows_of_blocks) + 1, dtype=torch.int32, device=device)
    return BsrMatrix(rows_of_blocks, cols_of_blocks, offsets, torch.empty(1, dtype=torch.int32, device=device),
                     torch.empty((1, 3, 3), dtype=torch.float32, device=device))


_axpy_ws: dict = {}


def bsr_axpy(x: BsrMatrix, y: BsrMatrix | None = None, alpha: float = 1.0, beta: float = 1.0) -> BsrMatrix:
    """warp.sparse.bsr_axpy (dynamic_contacts.py:297): y = alpha x + beta y
    over the union of the patterns; y's arrays are replaced and y returned."""
    if y is None:
        y = bsr_zeros(x.nrow, x.ncol, x.values.device)
    if x.shape != y.shape:
        raise ValueError(f"bsr_axpy: shapes differ ({x.shape} vs {y.shape})")
    dev = x.values.device
    ws = _axpy_ws.get(dev)
    if ws is None:
        ws = _axpy_ws[dev] = _Workspace()
    # the capacities bound the block counts, so the union always fits
    cap = x.capacity + y.capacity
    zo = torch.empty(x.nrow + 1, dtype=torch.int32, device=dev)
    zc = torch.empty(cap, dtype=torch.int32, device=dev)
    zv = torch.empty((cap, 3, 3), dtype=torch.float32, device=dev)
    check(lib().ipc_bsr_axpy(ws.handle, x.nrow, _need(x.offsets, torch.int32, "x offsets"),
                             _need(x.columns, torch.int32, "x columns"),
                             _need(x.values, torch.float32, "x values", (3, 3)), x.capacity, float(alpha),
                             _need(y.offsets, torch.int32, "y offsets"), _need(y.columns, torch.int32, "y columns"),
                             _need(y.values, torch.float32, "y values", (3, 3)), y.capacity, float(beta),
                             zo.data_ptr(), zc.data_ptr(), zv.data_ptr(), cap, stream_handle(dev)), "bsr_axpy")
    y.offsets, y.columns, y.values = zo, zc, zv
    y.nnz = None
    # y's pattern changed: a refill over its old pattern would be wrong
    y._ws, y._n_triplets = None, None
    return y


def bsr_mv(A: BsrMatrix, x: torch.Tensor, y: torch.Tensor | None = None, alpha: float = 1.0,
           beta: float = 0.0) -> torch.Tensor:
    """warp.sparse.bsr_mv: y = alpha A x + beta y (x, y float32 [n, 3])."""
    if y is None:
        y = torch.zeros((A.nrow, 3), dtype=torch.float32, device=x.device)
        beta = 0.0
    if x.shape[0] != A.ncol or y.shape[0] != A.nrow:
        raise ValueError("bsr_mv: x / y do not match the matrix shape")
    check(lib().ipc_bsr_mv(A.nrow, _need(A.offsets, torch.int32, "offsets"), A.columns.data_ptr(),
                           A.values.data_ptr(), _need(x, torch.float32, "x", (3,)), _need(y, torch.float32, "y", (3,)),
                           float(alpha), float(beta), stream_handle(x.device)), "bsr_mv")
    return y


class _CgWorkspace:
    def __init__(self):
        h = ctypes.c_void_p()
        check(lib().ipc_cg_create(ctypes.byref(h)), "cg_create")
        self.handle = h

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and h.value:
            try:
                lib().ipc_cg_destroy(h)
            except Exception:  # interpreter shutdown: the process frees it
                pass
            self.handle = None


_cg_ws: dict = {}


def cg(A: BsrMatrix, b: torch.Tensor, x: torch.Tensor, tol: float = 1e-4, maxiter: int = 0, check_every: int = 10,
       jacobi: bool = False):
    """warp.optim.linear.cg(A, b, x, tol, maxiter) (time_integrator_base.py:
    500-507): x (the initial guess) is overwritten with the solution; returns
    (iterations, |r|, tol |b|) at the last convergence check."""
    if b.shape != x.shape or x.shape[0] != A.nrow:
        raise ValueError("cg: b and x must be float32 [n_rows, 3]")
    dev = x.device
    ws = _cg_ws.get(dev)
    if ws is None:
        ws = _cg_ws[dev] = _CgWorkspace()
    it = ctypes.c_int(0)
    res = ctypes.c_double(0.0)
    check(lib().ipc_bsr_cg(ws.handle, A.nrow, _need(A.offsets, torch.int32, "offsets"), A.columns.data_ptr(),
                           A.values.data_ptr(), _need(b, torch.float32, "b", (3,)), _need(x, torch.float32, "x", (3,)),
                           float(tol), int(maxiter), int(check_every), int(bool(jacobi)), ctypes.byref(it),
                           ctypes.byref(res), stream_handle(dev)), "bsr_cg")
    return it.value, res.value, float(tol) * float(torch.linalg.norm(b.double()))

"""Tensor-level launches of the libipc_warp_amd kernels.

PyTorch-ROCm tensors are only buffers here: each function checks device,
dtype and contiguity, then hands raw device pointers and torch's current HIP
stream to the C ABI. Names and argument meaning follow the reference's Warp
kernels (file:line given per function).
"""
from __future__ import annotations

import struct

import torch

from ._lib import check, lib


def _need(t: torch.Tensor, dtype: torch.dtype, name: str, shape_tail=None) -> int:
    # One cheap combined test on the hot path (these wrappers run between a
    # step's device synchronisations); the specific error only on failure.
    if (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == dtype and t.is_contiguous()
            and (shape_tail is None or t.shape[1:] == shape_tail)):
        return t.data_ptr()
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if not t.is_cuda:
        raise ValueError(f"{name} must live on the GPU (got {t.device})")
    if t.dtype != dtype:
        raise TypeError(f"{name} must be {dtype} (got {t.dtype})")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    raise ValueError(f"{name} must have shape [n, {', '.join(map(str, shape_tail))}] (got {tuple(t.shape)})")


def _opt(t, dtype, name, shape_tail=None):
    return None if t is None else _need(t, dtype, name, shape_tail)


def _rows(t: torch.Tensor, n: int, name: str) -> int:
    """Pointer to an int32 [>=n, 4] record stream."""
    p = _need(t, torch.int32, name, (4,))
    if int(n) < 0 or int(n) > t.shape[0]:
        raise ValueError(f"{name} holds {t.shape[0]} records, fewer than n = {n}")
    return p


def _vec_len(t, n, name):
    if t is not None and t.shape[0] < int(n):
        raise ValueError(f"{name} holds {t.shape[0]} entries, fewer than n = {n}")


_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)


def stream_handle(device=None) -> int:
    """Torch's current HIP stream on `device` (default: the current device)."""
    if _raw_stream is not None:
        index = device.index if isinstance(device, torch.device) else device
        return _raw_stream(torch.cuda.current_device() if index is None else int(index))
    return torch.cuda.current_stream(device).cuda_stream


def contact_hessian_ee(x, xcs, contacts, n, l0, kappa, rows, cols, vals, b, offset=0):
    """contact_hessian_ee (dynamic_contacts.py:151-207), dim = n.

    x, xcs: float32 [V,3]; contacts: int32 [>=n,4]; rows/cols int32 [16*total];
    vals float32 [16*total,3,3]; b float32 [V,3] (force accumulator).
    """
    _check_triplets(rows, cols, vals, offset, n, x, b)
    if xcs.shape != x.shape:
        raise ValueError("xcs must have the shape of x")
    check(lib().ipc_contact_hessian_ee(_need(x, torch.float32, "x", (3,)), _need(xcs, torch.float32, "xcs", (3,)),
                                       x.shape[0], _rows(contacts, n, "contacts"), int(n), float(l0), float(kappa),
                                       int(offset), _need(rows, torch.int32, "rows"), _need(cols, torch.int32, "cols"),
                                       _need(vals, torch.float32, "vals", (3, 3)), _need(b, torch.float32, "b", (3,)),
                                       stream_handle(x.device)), "contact_hessian_ee")


def contact_hessian_pt(x, contacts, n, l0, kappa, rows, cols, vals, b, offset=0):
    """contact_hessian_pt (dynamic_contacts.py:209-246), dim = n, triplets at offset."""
    _check_triplets(rows, cols, vals, offset, n, x, b)
    check(lib().ipc_contact_hessian_pt(_need(x, torch.float32, "x", (3,)), x.shape[0],
                                       _rows(contacts, n, "contacts"), int(n), float(l0), float(kappa),
                                       int(offset), _need(rows, torch.int32, "rows"), _need(cols, torch.int32, "cols"),
                                       _need(vals, torch.float32, "vals", (3, 3)), _need(b, torch.float32, "b", (3,)),
                                       stream_handle(x.device)), "contact_hessian_pt")


def _check_triplets(rows, cols, vals, offset, n, x=None, b=None):
    if x is not None and b is not None and b.shape != x.shape:
        raise ValueError("b must have the shape of x")
    need = 16 * (int(offset) + int(n))
    if int(offset) < 0:
        raise ValueError("offset must be >= 0")
    if rows.numel() < need or cols.numel() < need or vals.shape[0] < need:
        raise ValueError(f"triplet buffers hold fewer than 16 * (offset + n) = {need} blocks")


def contact_energy_ee(x, xcs, contacts, n, l0, kappa, e):
    """contact_energy_ee (dynamic_contacts.py:374-388): adds the sum into e[0] (float64)."""
    if xcs.shape != x.shape:
        raise ValueError("xcs must have the shape of x")
    check(lib().ipc_contact_energy_ee(_need(x, torch.float32, "x", (3,)), _need(xcs, torch.float32, "xcs", (3,)),
                                      x.shape[0], _rows(contacts, n, "contacts"), int(n), float(l0), float(kappa),
                                      _need(e, torch.float64, "e"), stream_handle(x.device)), "contact_energy_ee")


def contact_energy_pt(x, contacts, n, l0, kappa, e):
    """contact_energy_pt (dynamic_contacts.py:390-398): adds the sum into e[0] (float64)."""
    check(lib().ipc_contact_energy_pt(_need(x, torch.float32, "x", (3,)), x.shape[0],
                                      _rows(contacts, n, "contacts"), int(n), float(l0), float(kappa),
                                      _need(e, torch.float64, "e"), stream_handle(x.device)), "contact_energy_pt")


def contact_hessian_f64(kind, x, contacts, n, l0, kappa, rows, cols, vals, b, offset=0, xcs=None):
    """contact_hessian_ee / _pt with fp64 scalars (scalar_types.py with
    USE_FLOAT32 = False): x, xcs, b float64 [V,3], vals float64 [16*total,3,3],
    l0 in double. kind is 'ee' (xcs required) or 'pt'."""
    _check_triplets(rows, cols, vals, offset, n, x, b)
    args = (_rows(contacts, n, "contacts"), int(n), float(l0), float(kappa), int(offset),
            _need(rows, torch.int32, "rows"), _need(cols, torch.int32, "cols"),
            _need(vals, torch.float64, "vals", (3, 3)), _need(b, torch.float64, "b", (3,)), stream_handle(x.device))
    if kind == "ee":
        if xcs is None or xcs.shape != x.shape:
            raise ValueError("xcs (float64, shaped like x) is required for edge-edge contacts")
        check(lib().ipc_contact_hessian_ee_f64(_need(x, torch.float64, "x", (3,)), _need(xcs, torch.float64, "xcs", (3,)),
                                               x.shape[0], *args), "contact_hessian_ee_f64")
    else:
        check(lib().ipc_contact_hessian_pt_f64(_need(x, torch.float64, "x", (3,)), x.shape[0], *args),
              "contact_hessian_pt_f64")


def contact_energy_f64(kind, x, contacts, n, l0, kappa, e, xcs=None):
    """contact_energy_ee / _pt with fp64 scalars: adds the sum into e[0]."""
    args = (_rows(contacts, n, "contacts"), int(n), float(l0), float(kappa), _need(e, torch.float64, "e"),
            stream_handle(x.device))
    if kind == "ee":
        if xcs is None or xcs.shape != x.shape:
            raise ValueError("xcs (float64, shaped like x) is required for edge-edge contacts")
        check(lib().ipc_contact_energy_ee_f64(_need(x, torch.float64, "x", (3,)), _need(xcs, torch.float64, "xcs", (3,)),
                                              x.shape[0], *args), "contact_energy_ee_f64")
    else:
        check(lib().ipc_contact_energy_pt_f64(_need(x, torch.float64, "x", (3,)), x.shape[0], *args),
              "contact_energy_pt_f64")


def collision_hessian(x, xcs, ee, n_ee, pt, n_pt, l0, kappa, rows, cols, vals, b):
    """The contact part of RodComplexBC.compute_A (dynamic_contacts.py:281-288):
    contact_hessian_ee at triplet offset 0 and contact_hessian_pt at offset
    n_ee, in one launch; rows/cols/vals hold 16 * (n_ee + n_pt) blocks."""
    _check_triplets(rows, cols, vals, n_ee, n_pt, x, b)
    if xcs.shape != x.shape:
        raise ValueError("xcs must have the shape of x")
    check(lib().ipc_collision_hessian(_need(x, torch.float32, "x", (3,)), _need(xcs, torch.float32, "xcs", (3,)),
                                      x.shape[0], _rows(ee, n_ee, "ee"), int(n_ee), _rows(pt, n_pt, "pt"), int(n_pt),
                                      float(l0), float(kappa), _need(rows, torch.int32, "rows"),
                                      _need(cols, torch.int32, "cols"), _need(vals, torch.float32, "vals", (3, 3)),
                                      _need(b, torch.float32, "b", (3,)), stream_handle(x.device)), "collision_hessian")


def collision_energy(x, xcs, ee, n_ee, pt, n_pt, l0, kappa, e):
    """The kernels of compute_collision_energy (dynamic_contacts.py:300-305):
    e[0] (float64) is set to the EE + PT barrier energy, without a host
    transfer."""
    if xcs.shape != x.shape:
        raise ValueError("xcs must have the shape of x")
    check(lib().ipc_collision_energy(_need(x, torch.float32, "x", (3,)), _need(xcs, torch.float32, "xcs", (3,)),
                                     x.shape[0], _rows(ee, n_ee, "ee"), int(n_ee), _rows(pt, n_pt, "pt"), int(n_pt),
                                     float(l0), float(kappa), _need(e, torch.float64, "e"), stream_handle(x.device)),
          "collision_energy")


def contact_points_ee(x, contacts, n, points=None, dists=None, features=None, xcs=None):
    """get_contact_points (contact.py:148-153) over fetch_dist_v0v1 (contact.py:88-110)."""
    for t, nm in ((points, "points"), (dists, "dists"), (features, "features")):
        _vec_len(t, n, nm)
    if xcs is not None and xcs.shape != x.shape:
        raise ValueError("xcs must have the shape of x")
    check(lib().ipc_contact_points_ee(_need(x, torch.float32, "x", (3,)), _opt(xcs, torch.float32, "xcs", (3,)),
                                      x.shape[0], _rows(contacts, n, "contacts"), int(n),
                                      _opt(points, torch.float32, "points", (3,)), _opt(dists, torch.float32, "dists"),
                                      _opt(features, torch.int32, "features"), stream_handle(x.device)),
          "contact_points_ee")


def contact_points_pt(x, contacts, n, points=None, dists=None, features=None):
    """get_contact_points_pt (contact.py:155-161) over fetch_dist_v0v1_pt (contact.py:116-145)."""
    for t, nm in ((points, "points"), (dists, "dists"), (features, "features")):
        _vec_len(t, n, nm)
    check(lib().ipc_contact_points_pt(_need(x, torch.float32, "x", (3,)), x.shape[0],
                                      _rows(contacts, n, "contacts"), int(n),
                                      _opt(points, torch.float32, "points", (3,)), _opt(dists, torch.float32, "dists"),
                                      _opt(features, torch.int32, "features"), stream_handle(x.device)),
          "contact_points_pt")


ONE_BITS = struct.unpack("<q", struct.pack("<d", 1.0))[0]


def new_toi_min(device) -> torch.Tensor:
    """A device int64 holding the bit pattern of 1.0 (the reference's ccd_toi.fill_(1.0))."""
    return torch.full((1,), ONE_BITS, dtype=torch.int64, device=device)


def toi_value(toi_min: torch.Tensor) -> float:
    return struct.unpack("<d", struct.pack("<q", int(toi_min.item())))[0]


def collision_time(kind, x0, x1, quads, n, toi=None, toi_min=None):
    """pt_collision_time / ee_collision_time (ccd/ccd.py:304-357) over candidate quadruples."""
    fn = lib().ipc_collision_time_ee if kind == "ee" else lib().ipc_collision_time_pt
    if x1.shape != x0.shape:
        raise ValueError("x1 must have the shape of x0")
    _vec_len(toi, n, "toi")
    check(fn(_need(x0, torch.float32, "x0", (3,)), _need(x1, torch.float32, "x1", (3,)), x0.shape[0],
             _rows(quads, n, "quads"), int(n), _opt(toi, torch.float64, "toi"),
             _opt(toi_min, torch.int64, "toi_min"), stream_handle(x0.device)), f"collision_time_{kind}")


def collision_time_f64(kind, x0, x1, quads, n, toi=None, toi_min=None):
    """pt_collision_time / ee_collision_time (ccd/ccd.py:304-357) with fp64
    scalars: x0, x1 float64 [V,3]; otherwise as collision_time."""
    fn = lib().ipc_collision_time_ee_f64 if kind == "ee" else lib().ipc_collision_time_pt_f64
    if x1.shape != x0.shape:
        raise ValueError("x1 must have the shape of x0")
    _vec_len(toi, n, "toi")
    check(fn(_need(x0, torch.float64, "x0", (3,)), _need(x1, torch.float64, "x1", (3,)), x0.shape[0],
             _rows(quads, n, "quads"), int(n), _opt(toi, torch.float64, "toi"), _opt(toi_min, torch.int64, "toi_min"),
             stream_handle(x0.device)), f"collision_time_{kind}_f64")


def conservative_time(kind: str, x0, x1, quads, n, toi=None, toi_min=None):
    """conservative_ee_toi / conservative_pt_toi (ccd/ccd.py:86-163) per
    candidate: toi float64 [>=n] (optional), toi_min from new_toi_min()
    (optional, folded in as the minimum)."""
    fn = lib().ipc_conservative_time_ee if kind == "ee" else lib().ipc_conservative_time_pt
    if x1.shape != x0.shape:
        raise ValueError("x1 must have the shape of x0")
    _vec_len(toi, n, "toi")
    check(fn(_need(x0, torch.float32, "x0", (3,)), _need(x1, torch.float32, "x1", (3,)), x0.shape[0],
             _rows(quads, n, "quads"), int(n), _opt(toi, torch.float64, "toi"), _opt(toi_min, torch.int64, "toi_min"),
             stream_handle(x0.device)), "conservative_time")


def tet_rest(xcs, tets):
    """compute_Dm (fem/fem.py:21-34): (Bm float32 [n,3,3], W float32 [n])."""
    n = int(tets.shape[0])
    Bm = torch.empty((n, 3, 3), dtype=torch.float32, device=xcs.device)
    W = torch.empty(n, dtype=torch.float32, device=xcs.device)
    check(lib().ipc_tet_rest(_need(xcs, torch.float32, "xcs", (3,)), xcs.shape[0], _need(tets, torch.int32, "tets", (4,)),
                             n, Bm.data_ptr(), W.data_ptr(), stream_handle(xcs.device)), "tet_rest")
    return Bm, W


def tet_hessian(x, tets, Bm, W, mu, lam, rows, cols, vals, b=None, offset=0):
    """tet_kernel_sparse (fem/fem.py:138-203) into triplets at `offset`."""
    n = int(tets.shape[0])
    need = 16 * (int(offset) + n)
    if rows.numel() < need or cols.numel() < need or vals.shape[0] < need:
        raise ValueError(f"triplet arrays hold fewer than {need} blocks")
    check(lib().ipc_tet_hessian(_need(x, torch.float32, "x", (3,)), x.shape[0], _need(tets, torch.int32, "tets", (4,)), n,
                                _need(Bm, torch.float32, "Bm", (3, 3)), _need(W, torch.float32, "W"), float(mu),
                                float(lam), int(offset), _need(rows, torch.int32, "rows"), _need(cols, torch.int32, "cols"),
                                _need(vals, torch.float32, "vals", (3, 3)), _opt(b, torch.float32, "b", (3,)),
                                stream_handle(x.device)), "tet_hessian")


def tet_energy(x, tets, Bm, W, mu, lam, psi=None, energy=None):
    """compute_Psi (time_integrator_base.py:124-137): per-tet W psi(F) into
    psi and their sum added into energy (float64 [1])."""
    n = int(tets.shape[0])
    check(lib().ipc_tet_energy(_need(x, torch.float32, "x", (3,)), x.shape[0], _need(tets, torch.int32, "tets", (4,)), n,
                               _need(Bm, torch.float32, "Bm", (3, 3)), _need(W, torch.float32, "W"), float(mu),
                               float(lam), _opt(psi, torch.float32, "psi"), _opt(energy, torch.float64, "energy"),
                               stream_handle(x.device)), "tet_energy")

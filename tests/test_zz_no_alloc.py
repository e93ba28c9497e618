"""ipc_mlp_forward_noalloc (inference that allocates nothing) against
ipc_mlp_forward with a scratch argument, through the C ABI: bit-identical
outputs on two calls in a row (the second with fewer rows, so stale rows of
the shared device array must not leak), and sizes beyond the maximum
rejected before any launch."""
import ctypes

import pytest
import torch

import ipc_warp_amd
from ipc_warp_amd import warmstart

pytestmark = pytest.mark.gpu

IN, HID, OUT = 48, 256, 3
MAX_ROWS = 4096


def _forward_scratch(L, params, x):
    n = x.shape[0]
    y = torch.full((n, OUT), float("nan"), dtype=torch.float32, device=x.device)
    tmp = torch.empty(3 * HID * n, dtype=torch.float32, device=x.device)
    assert L.ipc_mlp_forward(ctypes.c_void_p(params.data_ptr()), n, ctypes.c_void_p(x.data_ptr()),
                             ctypes.c_void_p(y.data_ptr()), None, ctypes.c_void_p(tmp.data_ptr()), None) == 0
    return y


def _forward_noalloc(L, params, x):
    n = x.shape[0]
    y = torch.full((n, OUT), float("nan"), dtype=torch.float32, device=x.device)
    st = L.ipc_mlp_forward_noalloc(ctypes.c_void_p(params.data_ptr()), n, ctypes.c_void_p(x.data_ptr()),
                                   ctypes.c_void_p(y.data_ptr()), None)
    assert st == 0, L.ipc_last_error()
    return y


def test_noalloc_forward_equals_scratch_forward(dev):
    L = ipc_warp_amd.lib()
    params = torch.from_numpy(warmstart.init_params(seed=3)).to(dev)
    g = torch.Generator(device="cpu").manual_seed(7)
    for n in (MAX_ROWS, 1000):
        print(f"noalloc forward n={n}", flush=True)
        x = torch.randn((n, IN), generator=g).to(dev)
        ref = _forward_scratch(L, params, x)
        got = _forward_noalloc(L, params, x)
        torch.cuda.synchronize()
        assert torch.isfinite(got).all()
        assert torch.equal(got, ref)


def test_noalloc_rejects_oversized_batches(dev):
    L = ipc_warp_amd.lib()
    st = L.ipc_mlp_forward_noalloc(None, MAX_ROWS + 1, None, None, None)
    assert st == -100000
    assert b"IPC_MLP_NOALLOC_MAX_ROWS" in L.ipc_last_error()
    assert L.ipc_mlp_forward_noalloc(None, 0, None, None, None) == 0

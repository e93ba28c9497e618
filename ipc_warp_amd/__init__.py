"""ipc_warp_amd — MI355X-native IPC contact inner loop.

Hand-written HIP kernels for gfx950 (``csrc/``) behind the C ABI declared in
``include/ipc_warp_amd.h``, reached from Python through ctypes with
PyTorch-ROCm tensors as buffers. The Python layer keeps the call signatures of
the reference's contact solver (Luke-Skycrawler/ipc_warp) so a time
integrator can use it as a drop-in.
"""
from . import ops  # noqa: F401
from ._lib import EXPORTED, IPCError, lib  # noqa: F401

__all__ = ["ops", "lib", "IPCError", "EXPORTED"]

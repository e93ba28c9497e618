"""The C-ABI library loads and exports exactly the entry points include/*.h
declares; argument validation rejects bad calls before any launch."""
import ctypes
import glob
import os
import re

import pytest

import ipc_warp_amd
from ipc_warp_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        names |= set(re.findall(r"\b(ipc_[a-z0-9_]+)\s*\(", src))
    return names


def test_library_exports_every_declared_symbol():
    L = ipc_warp_amd.lib()
    names = declared()
    assert names, "no declarations found"
    for n in sorted(names):
        assert hasattr(L, n), f"{n} declared in include/ but not exported"
    assert names == set(_lib.EXPORTED), "ctypes signature table out of sync with include/"


def test_abi_version_and_validation():
    L = ipc_warp_amd.lib()
    assert L.ipc_abi_version() == 2
    # negative sizes are rejected on the host before any HIP call
    st = L.ipc_contact_hessian_pt(None, 4, None, -1, ctypes.c_float(1e-3), 1e8, 0, None, None, None, None, None)
    assert st == -100000
    assert b"negative" in L.ipc_last_error()
    st = L.ipc_collision_time_ee(None, None, 4, None, -5, None, None, None)
    assert st == -100000
    # zero-size calls are no-ops
    assert L.ipc_contact_energy_pt(None, 4, None, 0, ctypes.c_float(1e-3), 1e8, None, None) == 0
    with pytest.raises(_lib.IPCError):
        _lib.check(-100000, "probe")


def test_misaligned_pointers_are_refused_before_any_launch():
    """The Hessian and energy entry points check 16-byte alignment of the
    record and triplet arrays on the host; the stand-in addresses below are
    never dereferenced."""
    L = ipc_warp_amd.lib()
    ok, off = 1 << 20, (1 << 20) + 4
    st = L.ipc_contact_hessian_ee(ok, ok, 4, off, 1, ctypes.c_float(1e-3), 1e8, 0, ok, ok, ok, ok, None)
    assert st == _lib.IPC_EINVAL and b"aligned" in L.ipc_last_error()
    st = L.ipc_contact_hessian_pt(ok, 4, ok, 1, ctypes.c_float(1e-3), 1e8, 0, ok, ok, off, ok, None)
    assert st == _lib.IPC_EINVAL and b"aligned" in L.ipc_last_error()
    st = L.ipc_contact_energy_pt(ok, 4, off, 1, ctypes.c_float(1e-3), 1e8, ok, None)
    assert st == _lib.IPC_EINVAL and b"aligned" in L.ipc_last_error()


def test_status_constant_matches_header():
    src = open(os.path.join(ROOT, "include", "ipc_warp_amd.h")).read()
    assert int(re.search(r"#define IPC_EINVAL \((-?\d+)\)", src).group(1)) == _lib.IPC_EINVAL


def test_contact_record_layout_is_checked():
    L = ipc_warp_amd.lib()
    assert L.ipc_contact_records_pack(1 << 20, 40, 4, ctypes.c_float(1e-3), 1 << 20, None, None) == _lib.IPC_EINVAL
    assert b"36" in L.ipc_last_error()
    assert L.ipc_contact_records_unpack(1 << 20, 1 << 20, 4, ctypes.c_float(1e-3), None, 1 << 20, 12, None) == \
        _lib.IPC_EINVAL
    # misaligned output list
    assert L.ipc_contact_records_pack(1 << 20, 36, 4, ctypes.c_float(1e-3), (1 << 20) + 4, None, None) == \
        _lib.IPC_EINVAL
    assert L.ipc_contact_records_pack(None, 36, 0, ctypes.c_float(1e-3), None, None, None) == 0

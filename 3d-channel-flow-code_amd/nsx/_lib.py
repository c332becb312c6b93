"""Loading of the in-tree shared libraries (ctypes).

libnsx.so is the HIP drop-in (csrc/, C-ABI of include/nsx.h); libnsx_host.so
is the host set-up (host/).  Both are built in-tree by build.py; loading fails
loudly when a library is missing: there is no fallback path.
"""
from __future__ import annotations

import ctypes
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(PKG_DIR, "_build")

_host = None
_dev = None

c_int = ctypes.c_int
c_dbl = ctypes.c_double
c_vp = ctypes.c_void_p
c_cp = ctypes.c_char_p


def _load(name: str) -> ctypes.CDLL:
    path = os.path.join(LIB_DIR, name)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
    return ctypes.CDLL(path)


def host() -> ctypes.CDLL:
    global _host
    if _host is None:
        lib = _load("libnsx_host.so")
        lib.nsx_case_create.restype = c_vp
        lib.nsx_case_create.argtypes = [c_int, c_int, c_int, c_dbl, c_dbl, c_dbl, c_dbl, c_dbl, c_int,
                                        c_dbl, c_dbl, ctypes.c_uint64]
        lib.nsx_case_create_walls.restype = c_vp
        lib.nsx_case_create_walls.argtypes = [c_int, c_int, c_int, c_dbl, c_dbl, c_dbl, c_dbl, c_dbl, c_int,
                                              c_dbl, c_dbl, ctypes.c_uint64, c_int, c_int, c_int, c_int]
        lib.nsx_case_save.argtypes = [c_vp, c_cp]
        lib.nsx_case_load.restype = c_vp
        lib.nsx_case_load.argtypes = [c_cp]
        lib.nsx_case_free.argtypes = [c_vp]
        lib.nsx_case_status.argtypes = [c_vp, c_cp, c_int]
        lib.nsx_case_info.argtypes = [c_vp, ctypes.POINTER(c_int)]
        lib.nsx_case_total_volume.argtypes = [c_vp]
        lib.nsx_case_total_volume.restype = c_dbl
        lib.nsx_case_num_arrays.argtypes = [c_vp]
        lib.nsx_case_array_name.argtypes = [c_vp, c_int]
        lib.nsx_case_array_name.restype = c_cp
        lib.nsx_case_get.argtypes = [c_vp, c_cp, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(c_int)]
        lib.nsx_case_get.restype = c_vp
        lib.nsx_case_from_mesh.restype = c_vp
        lib.nsx_case_from_mesh.argtypes = [c_int, c_int, c_int] + [c_vp] * 7 + [c_dbl] * 8 + [ctypes.c_uint64]
        lib.nsx_read_mesh.argtypes = [c_cp, c_int, c_int] + [c_vp] * 7
        lib.nsx_write_mesh.argtypes = [c_cp, c_int, c_int] + [c_vp] * 7
        lib.nsx_rng_u01.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint64]
        lib.nsx_rng_u01.restype = c_dbl
        _host = lib
    return _host


def device() -> ctypes.CDLL:
    """The HIP library.  Importing it does not touch the GPU."""
    global _dev
    if _dev is None:
        _dev = _load("libnsx.so")
        from . import abi
        abi.declare(_dev)
    return _dev

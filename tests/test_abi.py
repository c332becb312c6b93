"""The C-ABI libraries load on the CPU and export every entry point of
include/nsx.h (libnsx.so) and include/nsx_setup.h (libnsx_host.so)."""
import ctypes
import os
import re

from conftest import REPO
from nsx import abi

HEADER = os.path.join(REPO, "include", "nsx.h")
LIB = os.path.join(REPO, "3d-channel-flow-code_amd", "nsx", "_build", "libnsx.so")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:int|const char \*)\s*(nsx_\w+)\(", txt, re.M)))


def test_header_and_binding_agree():
    assert declared() == sorted(abi.EXPORTS)


def test_library_exports_every_symbol():
    assert os.path.exists(LIB), "build first: python -c 'import __graft_entry__ as g; g.build()'"
    lib = ctypes.CDLL(LIB)
    for name in declared():
        assert hasattr(lib, name), name


def test_create_rejects_bad_sizes_without_touching_the_gpu():
    lib = ctypes.CDLL(LIB)
    abi.declare(lib)
    h = ctypes.c_void_p()
    assert lib.nsx_create(0, 0, 0, 0, ctypes.byref(h)) == -1
    assert lib.nsx_create(10, 5, 4, 6, ctypes.byref(h)) == -1


def test_fortran_interfaces_cover_the_header():
    """Every entry point of include/nsx.h has a bind(C) interface in the
    Fortran binding, under its C name."""
    f90 = open(os.path.join(REPO, "3d-channel-flow-code_amd", "fortran", "nsx_iface.F90")).read()
    bound = set(re.findall(r'bind\(C,\s*name="(nsx_\w+)"\)', f90))
    assert set(declared()) == bound, sorted(set(declared()) ^ bound)


def test_setup_library_exports_every_symbol():
    txt = open(os.path.join(REPO, "include", "nsx_setup.h")).read()
    names = sorted(set(re.findall(r"^(?:int|void|double|const char \*|nsx_case \*|void \*)\s*\*?(nsx_\w+)\(",
                                  txt, re.M)))
    assert len(names) == 15, names
    lib = ctypes.CDLL(os.path.join(REPO, "3d-channel-flow-code_amd", "nsx", "_build", "libnsx_host.so"))
    for name in names:
        assert hasattr(lib, name), name


def test_format_real_rejects_out_of_range_widths():
    """nsx_format_real builds an Ew.d field in fixed buffers: digit counts
    outside 1..30 and widths outside 0..63 are refused, not written."""
    lib = ctypes.CDLL(LIB)
    fr = lib.nsx_format_real
    fr.argtypes = [ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_int]
    fr.restype = ctypes.c_int
    buf = ctypes.create_string_buffer(128)
    assert fr(1.5, 0, 12, 5, buf, 128) == 0 and buf.value == b" 0.15000E+01"
    assert fr(0.0, 0, 12, 5, buf, 128) == 0 and buf.value == b" 0.00000E+00"
    for w, d in ((12, 0), (12, -3), (12, 31), (12, 40), (-1, 5), (64, 5), (200, 5)):
        assert fr(1.5, 0, w, d, buf, 128) == -1, (w, d)
        assert fr(0.0, 0, w, d, buf, 128) == -1, (w, d)
    assert fr(1.5, 2, 12, 5, buf, 128) == -1
    assert fr(1.5, 1, 0, 0, buf, 128) == 0 and buf.value == b"1.5"

"""The C-ABI library loads and exports every function include/*.h declares
(no device calls)."""
import ctypes
import os
import re

from gpis_extend_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    inc = os.path.join(ROOT, "include")
    txt = "".join(open(os.path.join(inc, f)).read() for f in sorted(os.listdir(inc)) if f.endswith(".h"))
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gpis_[a-z0-9_]+)\s*\(", txt)))


def test_header_matches_binding_list():
    assert declared_functions() == sorted(_capi.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_capi.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name


def test_struct_layouts_match_header_sizes():
    # field counts of the C structs mirrored in _capi
    assert ctypes.sizeof(_capi.RaySoA) == 17 * ctypes.sizeof(ctypes.c_void_p)
    assert ctypes.sizeof(_capi.ReplySoA) == 7 * ctypes.sizeof(ctypes.c_void_p)
    assert ctypes.sizeof(_capi.PathReply) == 10 * ctypes.sizeof(ctypes.c_void_p)

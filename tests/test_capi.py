"""The C-ABI library: loads without a GPU and exports every symbol that
include/cbx.h declares (no compute calls here)."""
import ctypes
import os
import re

from cbclib_v2_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, 'include', 'cbx.h')).read()
    return sorted(set(re.findall(r'CBX_API\s+(?:const\s+)?\w+\s*\*?\s*(cbx_\w+)\s*\(', text)))


def test_header_lists_the_abi():
    names = declared()
    assert set(names) == set(_native.EXPORTS), names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_native.LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), name


def test_host_only_entry_points():
    L = _native.lib()
    assert L.cbx_version().startswith(b'cbx')
    for i, k in enumerate(['biweight', 'gaussian', 'parabolic', 'rectangular', 'triangular']):
        assert L.cbx_kernel_id(k.encode()) == i == _native.KERNELS[k]
    assert L.cbx_kernel_id(b'cosine') == -1

"""CPU checks of the C-ABI boundary: the library loads, exports every symbol
include/tractor_hip.h declares, and its struct layouts match the binding.
No GPU call is made."""
import ctypes
import os
import re

import pytest

from tractor_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, 'include', 'tractor_hip.h')).read()
    txt = re.sub(r'/\*.*?\*/', '', txt, flags=re.S)
    return sorted(set(re.findall(r'\b(trh_[a-z0-9_]+)\s*\(', txt)))


def test_library_built():
    assert os.path.exists(_lib.LIB_PATH), 'run python -m tractor_amd.build'


def test_exports_match_header():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    names = header_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    # the binding declares exactly the header's functions
    assert sorted(_lib.SIGNATURES) == names


def test_struct_layouts():
    lib = _lib.load()      # checks sizes against the numpy dtypes
    fs, ms, As = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    assert lib.trh_struct_sizes(ctypes.byref(fs), ctypes.byref(ms), ctypes.byref(As)) == 0
    assert fs.value == _lib.FFT_JOB.itemsize == 328
    assert ms.value == _lib.MOG_JOB.itemsize == 64
    assert As.value == _lib.MOG_APPROX_JOB.itemsize == 80


def test_error_path_without_gpu():
    lib = _lib.load()
    # argument validation happens before any device call
    assert lib.trh_render_fft_batch(None, 64, None, 1, None, None, None, None, None, None) == -1
    assert b'null' in lib.trh_last_error()

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


def test_every_entry_point_rejects_bad_arguments_without_a_device():
    """Argument validation runs on the host before any device call: a null
    handle (or empty/invalid sizes) gives TRH_E_ARG and a message, never a
    device access."""
    lib = _lib.load()
    N = None
    calls = {
        'trh_psf_spectrum_batch': (N, N, 1, 63, 63, 64, N, N, N, N),
        'trh_render_fft_batch': (N, 64, N, 1, N, N, N, N, N, N),
        'trh_render_mog_batch': (N, N, 1, 1) + (N,) * 11,
        'trh_render_mog_approx3_batch': (N, N, 1, 10) + (N,) * 8,
        'trh_lanczos_shift_batch': (N, N, N, 1, 4, 4, N, N, N),
        'trh_render_point_batch': (N, N, 1, N, 3, 3, N, N, N, N, N),
        'trh_normal_eq_dense': (N, N, 10, 3, N, N),
        'trh_solve_normal_f64': (N, N, 3, 0., 0, N, N, N),
        'trh_set_fft_config': (N, 64, 0, 0),
        'trh_psf_combine': (N, N, 9, 100, N, N, 4, N, 1, N),
        'trh_psfex_terms': (N, 1, N, N, 2, 0., 0., 1., 1., N, N, N),
        'trh_fp_keys': (N, 1, N, 10, 10, 4, N, N),
        'trh_fp_plan': (N, 1, N, N, N, 10, 10, 3, 3, N, N, N, N, N, N, N, N),
        'trh_fp_jobs_bucketed': (N, 1, N, N, N, N, N, N, N),
        'trh_fp_overlaps': (N, 1, N, N, 1, 1, N, N, 0, N, N, N, N, N, N),
        'trh_fp_pair_dots_csr': (N, 1, N, N, N, N, N, N, N, N, N),
        'trh_fp_pair_dots': (N, N, N, 1, N, N, N, N, 10, N, N),
        'trh_fp_rhs': (N, 1, N, N, N, N, N, N, 10, N, N, N),
        'trh_fp_block_inv': (N, N, N, N, N, N, 1, 8, N, N, N),
        'trh_fp_precond': (N, N, N, 1, 8, N, N, N, N),
        'trh_fp_spmv': (N, 4, N, N, N, N, N, N),
        'trh_fp_pcg': (N, 4, N, N, N, N, N, 1, 8, N, N, 1e-12, 0, 1, N, N, N, N),
        'trh_fp_fitstats': (N, 1, N, N, N, N, N, N, N, 10, 1., 0., N, N),
        'trh_fp_imstats': (N, 1, N, N, N, 0., N, N),
        'trh_fp_add_patches': (N, 1, N, N, N, N, 10, N, N),
        'trh_em_fit_psf': (N, 1, N, 9, 9, -4, -4, 2, 0., 10, N, N, N, N, N),
        'trh_cnn_conv_fwd': (N, N, N, N, N, 1, 32, 32, 1, N),
        'trh_cnn_conv_bwd_data': (N, N, N, N, N, 1, 32, 32, N),
        'trh_cnn_conv_bwd_weight': (N, N, N, N, N, 1, 32, 32, 1, N),
        'trh_cnn_head': (N, N, N, 1, N, N, N, 1.0, N, N, N, N, N, N),
        'trh_cnn_adam': (N, N, N, N, N, 10, 1e-3, 0.9, 0.999, 1e-8, 0., N),
        'trh_cnn_pack': (N, N, 1, N, N, N),
        'trh_cnn_cutouts': (N, N, N, N, 1, N, N, N),
        'trh_fitter_iterate': (N, 0., 0., 100, N),
        'trh_fitter_iterate_graph': (N, 0., 0., 100, 1, N),
        'trh_fitter_flags': (N, N, N),
        'trh_fitter_set_branching': (N, 100),
        'trh_fitter_count_active': (N, N),
        'trh_fitter_count_read': (N, 1, N),
        'trh_jacobian_normal_eq_batch': (N, 1) + (N,) * 11 + (0.01, N, N),
        'trh_solve_damped_cholesky_f64_batch': (N, 1, N, N, 0., N, N, N, N, N, N),
        'trh_line_search_chi2_batch': (N, N, N, N, N),
    }
    assert set(calls) <= set(_lib.SIGNATURES)
    for name, args in calls.items():
        rc = getattr(lib, name)(*args)
        assert rc == -1, (name, rc)
        assert lib.trh_last_error(), name

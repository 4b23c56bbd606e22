"""ctypes binding of libtractor_hip.so (the C ABI declared in include/tractor_hip.h).

The library is required: every compute entry point of tractor_amd goes
through it, and loading fails loudly if it has not been built
(``python -m tractor_amd.build``).  There is no CPU fallback.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libtractor_hip.so')

TRH_STORE, TRH_ADD, TRH_CHI2 = 0, 1, 2
TRH_MAXK = 16
# per-source flags of the batched fitter (trh_fitter_flags)
TRH_FIT_F_JAC_SIZE, TRH_FIT_F_LS_SIZE, TRH_FIT_F_MAXITER = 1, 2, 4
TRH_FIT_F_LIMITED, TRH_FIT_F_MODEL_SIZE, TRH_FIT_F_HYBRID = 8, 16, 32
TRH_NOALLOC_SCRATCH_FLOATS = 8 * 1024 * 1024

# struct trh_fft_job (include/tractor_hip.h)
FFT_JOB = np.dtype([
    ('out_offset', '<i8'), ('out_stride', '<i4'), ('mw', '<i4'), ('mh', '<i4'),
    ('sx', '<i4'), ('sy', '<i4'), ('psf', '<i4'), ('K', '<i4'), ('mode', '<i4'),
    ('slot', '<i4'), ('pad0', '<i4'), ('mux', '<f8'), ('muy', '<f8'),
    ('scale', '<f4'), ('pad1', '<f4'), ('amp', '<f4', (TRH_MAXK,)),
    ('var', '<f4', (TRH_MAXK, 3)),
], align=True)

# struct trh_mog_job
MOG_JOB = np.dtype([
    ('out_offset', '<i8'), ('mask_offset', '<i8'), ('out_stride', '<i4'),
    ('x0', '<i4'), ('y0', '<i4'), ('w', '<i4'), ('h', '<i4'), ('comp0', '<i4'),
    ('K', '<i4'), ('cutoff_mode', '<i4'), ('mode', '<i4'), ('want_derivs', '<i4'),
    ('scale', '<f4'), ('pad', '<f4'),
], align=True)

# struct trh_mog_approx_job
MOG_APPROX_JOB = np.dtype([
    ('out_offset', '<i8'), ('out_stride', '<i4'), ('x0', '<i4'), ('y0', '<i4'),
    ('w', '<i4'), ('h', '<i4'), ('comp0', '<i4'), ('K', '<i4'), ('xc', '<i4'), ('yc', '<i4'),
    ('minradius', '<i4'), ('want_derivs', '<i4'), ('pad', '<i4'), ('fx', '<f8'), ('fy', '<f8'),
    ('minval', '<f8'),
], align=True)
TRH_APPROX_MAXK = 64
TRH_APPROX_MAX_WH = 8192

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_PP = ctypes.POINTER(ctypes.c_void_p)

# name -> (restype, argtypes); the exact export set of include/tractor_hip.h
SIGNATURES = {
    'trh_last_error': (ctypes.c_char_p, []),
    'trh_version': (_I, []),
    'trh_struct_sizes': (_I, [_P, _P, _P]),
    'trh_copy_to_host': (_I, [_P, _P, ctypes.c_uint64, _P]),
    'trh_create': (_I, [_PP, _I]),
    'trh_destroy': (_I, [_P]),
    'trh_psf_spectrum_batch': (_I, [_P, _P, _I, _I, _I, _I, _P, _P, _P, _P]),
    'trh_render_fft_batch': (_I, [_P, _I, _P, _I, _P, _P, _P, _P, _P, _P]),
    'trh_set_fft_config': (_I, [_P, _I, _I, _I]),
    'trh_normal_eq_dense': (_I, [_P, _P, _I, _I, _P, _P]),
    'trh_solve_normal_f64': (_I, [_P, _P, _I, _D, _I, _P, _P, _P]),
    'trh_render_mog_batch': (_I, [_P, _P, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    'trh_render_mog_approx3_batch': (_I, [_P, _P, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P]),
    'trh_lanczos_shift_batch': (_I, [_P, _P, _P, _I, _I, _I, _P, _P, _P]),
    'trh_lanczos_shift_batch_noalloc': (_I, [_P, _P, _P, _I, _I, _I, _P, _P, _P]),
    'trh_render_point_batch': (_I, [_P, _P, _I, _P, _I, _I, _P, _P, _P, _P, _P]),
    'trh_psf_combine': (_I, [_P, _P, _I, ctypes.c_int64, _P, _P, _I, _P, _I, _P]),
    'trh_psfex_terms': (_I, [_P, _I, _P, _P, _I, _D, _D, _D, _D, _P, _P, _P]),
    'trh_fp_keys': (_I, [_P, _I, _P, _I, _I, _I, _P, _P]),
    'trh_fp_plan': (_I, [_P, _I, _P, _P, _P, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _P]),
    'trh_fp_jobs_bucketed': (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _P]),
    'trh_fp_overlaps': (_I, [_P, _I, _P, _P, _I, _I, _P, _P, _I, _P, _P, _P, _P, _P, _P]),
    'trh_fp_pair_dots_csr': (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P]),
    'trh_fp_pair_dots': (_I, [_P, _P, _P, _I, _P, _P, _P, _P, _I, _P, _P]),
    'trh_fp_rhs': (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _I, _P, _P, _P]),
    'trh_fp_block_inv': (_I, [_P, _P, _P, _P, _P, _P, _I, _I, _P, _P, _P]),
    'trh_fp_precond': (_I, [_P, _P, _P, _I, _I, _P, _P, _P, _P]),
    'trh_fp_spmv': (_I, [_P, _I, _P, _P, _P, _P, _P, _P]),
    'trh_fp_pcg': (_I, [_P, _I, _P, _P, _P, _P, _P, _I, _I, _P, _P, _D, _I, _I, _P, _P, _P, _P]),
    'trh_fp_fitstats': (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _P, _I, _D, _D, _P, _P]),
    'trh_fp_imstats': (_I, [_P, ctypes.c_int64, _P, _P, _P, _D, _P, _P]),
    'trh_fp_add_patches': (_I, [_P, _I, _P, _P, _P, _P, _I, _P, _P]),
    'trh_em_fit_psf': (_I, [_P, _I, _P, _I, _I, _I, _I, _I, _D, _I, _P, _P, _P, _P, _P]),
    'trh_cnn_conv_fwd': (_I, [_P, _P, _P, _P, _P, _I, _I, _I, _I, _P]),
    'trh_cnn_conv_bwd_data': (_I, [_P, _P, _P, _P, _P, _I, _I, _I, _P]),
    'trh_cnn_conv_bwd_weight': (_I, [_P, _P, _P, _P, _P, _I, _I, _I, _I, _P]),
    'trh_cnn_head': (_I, [_P, _P, _P, _I, _P, _P, _P, ctypes.c_float, _P, _P, _P, _P, _P, _P]),
    'trh_cnn_adam': (_I, [_P, _P, _P, _P, _P, _I, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                          ctypes.c_float, ctypes.c_float, _P]),
    'trh_cnn_pack': (_I, [_P, _P, _I, _P, _P, _P]),
    'trh_cnn_cutouts': (_I, [_P, _P, _P, _P, _I, _P, _P, _P]),
    'trh_fitter_create': (_I, [_P, _PP, _I, _P, _P, _P, _P, _P, _P, _P, _I, _I, _P]),
    'trh_fitter_destroy': (_I, [_P]),
    'trh_fitter_set_params': (_I, [_P, _P, _P]),
    'trh_fitter_get_params': (_I, [_P, _P, _P]),
    'trh_fitter_set_sersic': (_I, [_P, _P, _I]),
    'trh_fitter_set_hybrid': (_I, [_P, _I, _P, _P, _P, _D, _D]),
    'trh_fitter_set_chain': (_I, [_P, _I, _P, _P, _P, _D, _P]),
    'trh_fitter_chain_state': (_I, [_P, _P, _P]),
    'trh_fitter_iterate': (_I, [_P, _D, _D, _I, _P]),
    'trh_fitter_iterate_graph': (_I, [_P, _D, _D, _I, _I, _P]),
    'trh_fitter_set_branching': (_I, [_P, _I]),
    'trh_fitter_count_active': (_I, [_P, _P]),
    'trh_fitter_count_read': (_I, [_P, _I, _P]),
    'trh_fitter_render_models': (_I, [_P, _P, _I, _P]),
    'trh_fitter_status': (_I, [_P, _P, _P, _P, _P, _P, _P]),
    'trh_fitter_flags': (_I, [_P, _P, _P]),
    'trh_jacobian_normal_eq_batch': (_I, [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P,
                                          _D, _P, _P]),
    'trh_solve_damped_cholesky_f64_batch': (_I, [_P, _I, _P, _P, _D, _P, _P, _P, _P, _P, _P]),
    'trh_line_search_chi2_batch': (_I, [_P, _P, _P, _P, _P]),
    'trh_fitter_debug': (_I, [_P, _PP, _PP, _PP, _PP]),
    'trh_fitter_timing': (_I, [_P, _P]),
    'trh_fitter_profile': (_I, [_P, _I]),
    'trh_fitter_profile_read': (_I, [_P, _P]),
}

_lib = None


class TractorHipError(RuntimeError):
    pass


def load():
    """Load and declare the library (raises if it is missing)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise TractorHipError(
            'libtractor_hip.so is not built (%s); run `python -m tractor_amd.build`' % LIB_PATH)
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    fs, ms, As = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_int(0)
    lib.trh_struct_sizes(ctypes.byref(fs), ctypes.byref(ms), ctypes.byref(As))
    if (fs.value != FFT_JOB.itemsize or ms.value != MOG_JOB.itemsize
            or As.value != MOG_APPROX_JOB.itemsize):
        raise TractorHipError(
            'struct layout mismatch: fft_job %d vs %d, mog_job %d vs %d, mog_approx_job %d vs %d'
            % (fs.value, FFT_JOB.itemsize, ms.value, MOG_JOB.itemsize, As.value,
               MOG_APPROX_JOB.itemsize))
    _lib = lib
    return lib


def check(rc, what=''):
    if rc != 0:
        msg = load().trh_last_error()
        raise TractorHipError('%s failed (%d): %s' % (what, rc, msg.decode() if msg else ''))


def ptr(x):
    """Device or host address of a torch tensor / numpy array (or None)."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if isinstance(x, np.ndarray):
        return x.ctypes.data
    return x.data_ptr()


class Handle(object):
    """Library handle bound to one HIP device."""
    def __init__(self, device=0):
        lib = load()
        h = ctypes.c_void_p()
        check(lib.trh_create(ctypes.byref(h), device), 'trh_create')
        self._h = h
        self.device = device

    @property
    def value(self):
        return self._h

    def close(self):
        if self._h:
            load().trh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_handles = {}


def handle(device=0):
    """Process-wide handle per device."""
    h = _handles.get(device)
    if h is None:
        h = Handle(device)
        _handles[device] = h
    return h

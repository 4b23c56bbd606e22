"""ctypes binding of the oracle's C++ FP64 fit (oracle/fit_fp64.cpp, built by
oracle/Makefile into oracle/liborcfit.so).

TEST INFRASTRUCTURE ONLY -- see ``oracle/__init__.py``.  It fits the same
SourceProblems as ``tractor_oracle.fit_source`` (kinds 0-2 on P x P windows)
on many host threads; the batched-fit benchmark times it as its CPU baseline
and checks the GPU fit's chi^2 against it on a large sample.
"""
import ctypes
import os
import platform
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'liborcfit.so')
_lib = None


def build():
    """make -C oracle (g++; no GPU code)."""
    subprocess.check_call(['make', '-s', '-C', _HERE])


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError('oracle/liborcfit.so is not built: make -C oracle')
        lib = ctypes.CDLL(LIB_PATH)
        p = ctypes.c_void_p
        lib.orc_fit_sources.restype = ctypes.c_int
        lib.orc_fit_sources.argtypes = [ctypes.c_int, p, p, p, p, p, p, p, ctypes.c_int,
                                        ctypes.c_int, p, ctypes.c_double, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_double, p, p, p, p]
        _lib = lib
    return _lib


def cpu_model():
    """The host CPU's model name (lscpu 'Model name', else /proc/cpuinfo)."""
    try:
        out = subprocess.run(['lscpu'], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.strip().startswith('Model name'):
                return line.split(':', 1)[1].strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        with open('/proc/cpuinfo') as fh:
            for line in fh:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or 'unknown'


def fit_sources(kind, P, pix_off, data, inverr, start, psf, cdi=np.eye(2), dchisq=0.,
                max_iter=100, nthreads=1, budget_s=0.):
    """Fit the sources in order (with budget_s > 0, no thread starts another
    source beyond its first after budget_s seconds) -> (n_done, params (n, 6),
    iterations (n,), chisq (n,)) with the first n_done rows filled.
    iterations -1: a galaxy whose centre left its window (the reference
    raises there without a hybrid PSF)."""
    lib = _load()
    kind = np.ascontiguousarray(kind, np.int32)
    n = len(kind)
    P = np.ascontiguousarray(P, np.int32)
    pix_off = np.ascontiguousarray(pix_off, np.int64)
    data = np.ascontiguousarray(data, np.float32)
    inverr = np.ascontiguousarray(inverr, np.float32)
    st = np.zeros((n, 6))
    start = np.asarray(start, np.float64)
    st[:, :start.shape[1]] = start[:, :6]
    psf = np.ascontiguousarray(psf, np.float32)
    cdi = np.ascontiguousarray(cdi, np.float64).ravel()
    params = np.zeros((n, 6))
    iters = np.zeros(n, np.int32)
    chisq = np.zeros(n)
    done = np.zeros(1, np.int32)
    a = lambda x: x.ctypes.data_as(ctypes.c_void_p)
    rc = lib.orc_fit_sources(n, a(kind), a(P), a(pix_off), a(data), a(inverr), a(st), a(psf),
                             psf.shape[0], psf.shape[1], a(cdi), float(dchisq), int(max_iter),
                             int(nthreads), float(budget_s), a(params), a(iters), a(chisq),
                             a(done))
    if rc != 0:
        raise ValueError('orc_fit_sources: unsupported source kind (0, 1, 2 only)')
    return int(done[0]), params, iters, chisq

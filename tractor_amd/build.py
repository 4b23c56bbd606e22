"""Build libtractor_hip.so (gfx950) in-tree with hipcc.

Usage: python -m tractor_amd.build   (or tractor_amd.build.build()).
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'libtractor_hip.so')
SOURCES = ['tractor_hip.hip', 'fitter.hip', 'mog_approx.hip', 'cnn.hip', 'forced.hip',
           'psf_fit.hip']
HEADERS = ['fft_common.h', 'render_fft.h', 'internal.h']

HIPCC_FLAGS = [
    '--offload-arch=gfx950', '-O3', '-fPIC', '-shared', '-std=c++17',
    # packed FP32 (v_pk_*_f32) off for every kernel
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wno-unused-result',
]


def hipcc():
    for c in (os.environ.get('HIPCC'), shutil.which('hipcc'), '/opt/rocm/bin/hipcc'):
        if c and os.path.exists(c):
            return c
    raise RuntimeError('hipcc not found')


def build(verbose=False):
    """Compile every HIP source into tractor_amd/libtractor_hip.so (always
    recompiles)."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB + '.tmp'
    cmd = [hipcc()] + HIPCC_FLAGS + ['-I', os.path.join(ROOT, 'include'), '-o', tmp] + srcs
    if verbose:
        print(' '.join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError('hipcc failed (exit %d)' % r.returncode)
    if verbose and r.stderr.strip():
        sys.stderr.write(r.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    print(build(verbose=True))

"""Build librbx.so in-tree with hipcc for gfx950.

    python -m radiobear_amd.build

Every HIP source is compiled with packed FP32 disabled
(-target-feature -packed-fp32-ops) and linked into one shared library that
exports the C ABI declared in include/rbx.h.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'librbx.so')
SOURCES = ['catalog.hip', 'alpha.hip', 'raypath.hip', 'integrate.hip', 'atmosphere.hip']


def hipcc():
    for cand in (os.environ.get('HIPCC'), '/opt/rocm/bin/hipcc', 'hipcc'):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    raise RuntimeError('hipcc not found')


def command(out=LIB, extra=()):
    return [hipcc(), '--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC', '-shared',
            '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
            '-Wno-unused-command-line-argument',
            '-I', os.path.join(REPO, 'include'), '-I', CSRC,
            *extra, '-o', out] + [os.path.join(CSRC, s) for s in SOURCES]


def build(verbose=False):
    tmp = LIB + '.tmp'
    cmd = command(tmp)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError('hipcc failed building librbx.so')
    if verbose and res.stderr.strip():
        sys.stderr.write(res.stderr)
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    print(build(verbose=True))

"""Compile the HIP sources of ssqueezepy_amd/csrc into libssqk.so (gfx950).

Used by ``__graft_entry__.build()`` and runnable directly:
    python -m ssqueezepy_amd._build
Objects go to ssqueezepy_amd/_build/, the library next to this file.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
INCLUDE = os.path.join(os.path.dirname(HERE), 'include')
OBJDIR = os.path.join(HERE, '_build')
LIB = os.path.join(HERE, 'libssqk.so')
SOURCES = ['api.hip', 'cwt.hip', 'stft.hip', 'inverse.hip', 'ridge.hip']
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
FLAGS = ['-O3', '-std=c++17', '-fPIC', '--offload-arch=gfx950',
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-Wno-unused-result', '-I' + INCLUDE]


def _compile(src):
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + '.o')
    cmd = [HIPCC] + FLAGS + ['-c', os.path.join(CSRC, src), '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed for %s:\n%s\n%s" % (src, ' '.join(cmd),
                                                               r.stderr[-8000:]))
    return obj


def build(verbose=True):
    os.makedirs(OBJDIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(len(SOURCES), 4)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + '.tmp%d' % os.getpid()
    cmd = [HIPCC, '-shared', '-fPIC', '--offload-arch=gfx950', '-o', tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s" % r.stderr[-4000:])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == '__main__':
    build()
    sys.exit(0)

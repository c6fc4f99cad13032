"""Build libfestung_hip.so in-tree for gfx950.

Host C++ (mesh, tables, C-ABI) is compiled with g++, the HIP sources with
hipcc --offload-arch=gfx950, and everything is linked into
festung_amd/libfestung_hip.so next to this file.

    python -m festung_amd.build [--jobs N]
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
BUILD = os.path.join(HERE, '_build')
INCLUDE = os.path.join(os.path.dirname(HERE), 'include')
LIB = os.path.join(HERE, 'libfestung_hip.so')
ROCM = os.environ.get('ROCM_PATH', '/opt/rocm')
ARCH = 'gfx950'

HOST_SRCS = ['tables.cpp', 'grid.cpp', 'checkpoint.cpp', 'capi.cpp']
HIP_SRCS = ['kernels.hip', 'problem.hip']

CXXFLAGS = ['-O2', '-std=c++17', '-fPIC', '-Wall', '-Wno-unused-function', '-I' + INCLUDE, '-I' + CSRC]
# -amdgpu-mfma-vgpr-form: the matrix instructions read and write ordinary
# VGPRs instead of accumulation registers.  The p = 4 stage kernel then has no
# v_accvgpr_read / write moves between its MFMAs and the pointwise products
# (56 per point block; a gfx950 SIMD issues no VALU instruction while an FP64
# MFMA runs, so every such move adds to the launch).
HIPFLAGS = ['-O3', '-std=c++17', '-fPIC', '--offload-arch=' + ARCH, '-I' + INCLUDE, '-I' + CSRC,
            '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
            '-munsafe-fp-atomics', '-Wno-unused-command-line-argument',
            '-mllvm', '-amdgpu-mfma-vgpr-form=1']


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('command failed: %s\n%s%s' % (' '.join(cmd), r.stdout, r.stderr))
    return r


def build(jobs=None, verbose=False):
    """Compile every source and link the shared library; returns its path."""
    os.makedirs(BUILD, exist_ok=True)
    hipcc = shutil.which('hipcc') or os.path.join(ROCM, 'bin', 'hipcc')
    cxx = shutil.which('g++') or 'g++'
    cmds = []
    objs = []
    for s in HOST_SRCS:
        o = os.path.join(BUILD, s + '.o')
        cmds.append([cxx] + CXXFLAGS + ['-I' + os.path.join(ROCM, 'include'), '-c', os.path.join(CSRC, s), '-o', o])
        objs.append(o)
    for s in HIP_SRCS:
        o = os.path.join(BUILD, s + '.o')
        cmds.append([hipcc] + HIPFLAGS + ['-c', os.path.join(CSRC, s), '-o', o])
        objs.append(o)
    with cf.ThreadPoolExecutor(max_workers=jobs or min(8, len(cmds))) as ex:
        for r in ex.map(_run, cmds):
            if verbose and (r.stdout or r.stderr):
                sys.stderr.write(r.stdout + r.stderr)
    tmp = LIB + '.tmp'
    _run([hipcc, '-shared', '-fPIC', '--offload-arch=' + ARCH, '-o', tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--jobs', type=int, default=None)
    ap.add_argument('-v', '--verbose', action='store_true')
    a = ap.parse_args()
    print(build(a.jobs, a.verbose))


if __name__ == '__main__':
    main()

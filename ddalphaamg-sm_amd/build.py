"""Builds libddamg.so (HIP for gfx950) in-tree: ddalphaamg-sm_amd/lib/libddamg.so.

    python ddalphaamg-sm_amd/build.py

Each source compiles to its own object in parallel; the objects link into one shared
library with the extern "C" boundary of include/ddamg.h.  Every HIP line carries
-Xclang -target-feature -Xclang -packed-fp32-ops (no packed FP32 instructions).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
BUILD = os.path.join(HERE, '_build')
LIB = os.path.join(HERE, 'lib', 'libddamg.so')
SOURCES = (['kernels.hip', 'coarsest.hip', 'sap_f32.hip', 'sap_f64.hip', 'solver.hip', 'methods.hip', 'capi.hip'] +
           [f'sap_{p}_d{d}.hip' for p in ('f32', 'f64') for d in (8, 16, 24, 32)])
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
FLAGS = ['--offload-arch=gfx950', '-O3', '-std=c++17', '-fPIC', '-Xclang', '-target-feature', '-Xclang',
         '-packed-fp32-ops', '-Wno-unused-command-line-argument']


def _compile(src: str) -> str:
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + '.o')
    cmd = [HIPCC] + FLAGS + ['-c', os.path.join(CSRC, src), '-o', obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f'compile failed: {" ".join(cmd)}\n{p.stdout}\n{p.stderr}')
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(os.cpu_count() or 8, 16, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + '.tmp'
    cmd = [HIPCC] + FLAGS + ['-shared', '-Wl,--no-undefined', '-o', tmp] + objs
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f'link failed: {" ".join(cmd)}\n{p.stdout}\n{p.stderr}')
    os.replace(tmp, LIB)
    if verbose:
        print('built', LIB)
    return LIB


if __name__ == '__main__':
    build(verbose=True)
    sys.exit(0)

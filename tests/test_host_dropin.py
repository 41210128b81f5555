"""CPU checks of the C++ host API as a drop-in for the reference's.

- The reference's own caller, src/main.cpp and src/methods.cpp, compiles unchanged
  against ddalphaamg-sm_amd/host/include, with MPI's header (MPICH here) and without it
  (the single-process definitions of ddamg_mpi.h), and links with libddamg_host and
  libddamg.so.  Compile and link only: nothing of the reference is copied, and the test
  skips where the reference is absent.
- The generic FGMRES and SAP_C algorithms that run a caller's own hooks on the host (a
  host Wilson-Dirac operator in tests/host_hooks_check.cpp) agree with the oracle's
  fgmres and sap (iteration counts equal, solutions to 1e-12), which the reference's
  golden fixtures pin.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import ddamg_oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, 'ddalphaamg-sm_amd', 'host')
LIB = os.path.join(ROOT, 'ddalphaamg-sm_amd', 'lib')
HOSTLIB = os.path.join(LIB, 'libddamg_host_64x64.a')
REF_SRC = '/root/reference/src'
MPI_INC = '/opt/conda/include'

needs_gxx = pytest.mark.skipif(shutil.which('g++') is None, reason='no g++')
needs_host = pytest.mark.skipif(not (os.path.exists(HOSTLIB) and os.path.exists(os.path.join(LIB, 'libddamg.so'))),
                                reason='host API not built (make -C ddalphaamg-sm_amd/host)')


def _gxx(args, cwd):
    p = subprocess.run(['g++', '-std=c++17', '-O1', '-w'] + args, capture_output=True, text=True, cwd=cwd)
    assert p.returncode == 0, p.stderr[-4000:]


@needs_gxx
@needs_host
@pytest.mark.skipif(not os.path.isdir(REF_SRC), reason='reference sources absent')
@pytest.mark.parametrize('with_mpi', [False, True])
def test_reference_caller_compiles_unchanged(tmp_path, with_mpi):
    inc = ['-I', os.path.join(HOST, 'include'), '-DNS=64', '-DNT=64']
    if with_mpi:
        if not os.path.exists(os.path.join(MPI_INC, 'mpi.h')):
            pytest.skip('no MPI header')
        inc = ['-DMPICH_SKIP_MPICXX', '-I', MPI_INC] + inc
    for f in ('main', 'methods'):
        _gxx(inc + ['-c', os.path.join(REF_SRC, f + '.cpp'), '-o', str(tmp_path / (f + '.o'))], tmp_path)
    if not with_mpi:   # link the single-process build with the library
        _gxx([str(tmp_path / 'main.o'), str(tmp_path / 'methods.o'), HOSTLIB, '-L', LIB, '-lddamg',
              '-Wl,-rpath,' + LIB, '-o', str(tmp_path / 'dropin')], tmp_path)
        assert os.path.exists(tmp_path / 'dropin')


@needs_gxx
@needs_host
def test_generic_fgmres_and_sap_hooks_match_oracle(tmp_path):
    exe = str(tmp_path / 'hooks')
    _gxx(['-I', os.path.join(HOST, 'include'), '-DNS=64', '-DNT=64', os.path.join(ROOT, 'tests', 'host_hooks_check.cpp'),
          HOSTLIB, '-L', LIB, '-lddamg', '-Wl,-rpath,' + LIB, '-o', exe], tmp_path)
    rng = np.random.default_rng(12)
    Lx, Lt, m0, nbx, nbt, nu = 8, 12, -0.3, 2, 2, 3
    U = np.exp(1j * rng.uniform(-np.pi, np.pi, (Lx, Lt, 2)) * 0.6)
    b = rng.standard_normal((Lx, Lt, 2)) + 1j * rng.standard_normal((Lx, Lt, 2))
    blob = (np.array([Lx, Lt], np.int32).tobytes() + np.array([m0]).tobytes() +
            np.array([nbx, nbt, nu], np.int32).tobytes() + U.astype(np.complex128).tobytes() + b.tobytes())
    p = subprocess.run([exe], input=blob, capture_output=True, timeout=120)
    assert p.returncode == 0, p.stderr.decode()[-2000:]
    out = p.stdout
    n = Lx * Lt * 2 * 16
    its = int(np.frombuffer(out[:4], np.int32)[0])
    x = np.frombuffer(out[4:4 + n], np.complex128).reshape(Lx, Lt, 2)
    sits = int(np.frombuffer(out[4 + n:8 + n], np.int32)[0])
    xs = np.frombuffer(out[8 + n:8 + 2 * n], np.complex128).reshape(Lx, Lt, 2)
    # oracle: FGMRES(10) x 20 to 1e-10 with no preconditioner
    ref = orc.fgmres(lambda v: np.stack([orc.D_phi(U, w, m0) for w in v]), None, b[None], np.zeros_like(b)[None],
                     10, 20, 1e-10)
    assert its == int(ref.iterations[0])
    assert np.abs(x - ref.x[0]).max() < 1e-12 * np.abs(ref.x[0]).max()
    # oracle: SAP on (Lx/nbx) x (Lt/nbt)-site blocks, nu iterations to 1e-12
    xs_ref, sits_ref = orc.sap(orc.dirac_stencil(U, m0), orc.SapBlocks(Lx, Lt, Lx // nbx, Lt // nbt), b,
                               np.zeros_like(b), nu, 1e-12)
    assert sits == sits_ref
    assert np.abs(xs - xs_ref).max() < 1e-12 * np.abs(xs_ref).max()

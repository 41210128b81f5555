"""Runs the compiled reference driver (oracle/_ref/ddamg_ref_<N>) under mpirun.

TEST INFRASTRUCTURE ONLY: used by tests/golden/make_golden.py to write fixtures and
by bench.py's cpu_baseline leg to time the reference's own CPU solve.

The reference reads its per-level blocking from a parameters file whose rows are
`level block_x block_t ntest sap_block_x sap_block_t` (include/params.h:7-57):
block_x/t are GLOBAL counts of aggregate blocks, sap_block_x/t are SAP blocks PER
RANK.  `write_params` converts the site-based LevelSpec of the oracle into that form.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import tempfile

import numpy as np

from . import ddamg_oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, '_ref')
MPIRUN = '/opt/conda/bin/mpirun'


def ref_binary(N: int) -> str:
    return os.path.join(REF_DIR, f'ddamg_ref_{N}')


def available(N: int) -> bool:
    return os.path.exists(ref_binary(N)) and os.path.exists(MPIRUN)


def write_params(path: str, N: int, levels: list[orc.LevelSpec], ranks: tuple[int, int]) -> None:
    rows = []
    nx = nt = N
    for l, spec in enumerate(levels):
        bx, bt = nx // spec.agg[0], nt // spec.agg[1]
        sx = (nx // ranks[0]) // spec.sap_block[0]
        st = (nt // ranks[1]) // spec.sap_block[1]
        rows.append(f'{l} {bx} {bt} {spec.nv} {sx} {st}')
        nx, nt = bx, bt
    with open(path, 'w') as f:
        f.write('\n'.join(rows) + '\n')


def run(mode: str, U: np.ndarray, rhs: np.ndarray, levels: list[orc.LevelSpec], m0: float,
        ranks: tuple[int, int] = (2, 2), workdir: str | None = None, timeout: float = 3600,
        **opts) -> tuple[str, str]:
    """Run one driver mode; returns (output directory, stdout)."""
    N = U.shape[0]
    assert U.shape[1] == N, 'the reference build is square (NS = NT)'
    wd = workdir or tempfile.mkdtemp(prefix='ddamg_ref_')
    os.makedirs(wd, exist_ok=True)
    conf, rhsf, par = (os.path.join(wd, n) for n in ('conf.bin', 'rhs.bin', 'parameters'))
    orc.write_records(conf, U)
    orc.write_records(rhsf, rhs)
    write_params(par, N, levels, ranks)
    args = [MPIRUN, '-n', str(ranks[0] * ranks[1]), ref_binary(N), f'mode={mode}', f'out={wd}',
            f'ranks_x={ranks[0]}', f'ranks_t={ranks[1]}', f'levels={len(levels) + 1}',
            f'm0={m0!r}', f'conf={conf}', f'rhs={rhsf}', f'params={par}']
    args += [f'{k}={v}' for k, v in opts.items()]
    env = dict(os.environ)
    env.setdefault('HOSTNAME', 'localhost')
    p = subprocess.run(args, cwd=wd, env=env, capture_output=True, text=True, timeout=timeout)
    if p.returncode != 0:
        raise RuntimeError(f'reference run failed ({p.returncode}):\n{p.stdout}\n{p.stderr}')
    return wd, p.stdout


def load(wd: str, name: str, ranks: tuple[int, int]) -> np.ndarray:
    return orc.read_rank_dump(os.path.join(wd, name), ranks[0] * ranks[1])


def result(wd: str) -> dict:
    out = {}
    with open(os.path.join(wd, 'result.txt')) as f:
        for line in f:
            k, v = line.split()
            out[k] = float(v) if '.' in v or 'e' in v else int(v)
    return out


def cleanup(wd: str) -> None:
    shutil.rmtree(wd, ignore_errors=True)

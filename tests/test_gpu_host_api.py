"""The C++ host API (reference class names) end to end, and the multigrid cycle itself
against the oracle."""
import os
import subprocess

import numpy as np
import pytest

from oracle import ddamg_oracle as orc

pytestmark = pytest.mark.gpu
ddamg = pytest.importorskip('ddamg')
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MAIN = os.path.join(ROOT, 'ddalphaamg-sm_amd', 'lib', 'ddamg_main_64x64')


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.skipif(not os.path.exists(MAIN), reason='host example not built (make -C ddalphaamg-sm_amd/host)')
def test_reference_style_driver(golden, tmp_path):
    meta, a = golden('g64_solve')
    U = np.exp(1j * a['theta'])
    orc.write_records(str(tmp_path / 'conf.bin'), U)
    orc.write_records(str(tmp_path / 'rhs.bin'), a['rhs'])
    (tmp_path / 'parameters').write_text('0 16 16 8 8 8\n')   # 4x4 aggregates, 4x4-site SAP blocks per 2x2 rank
    stdin = f"2\n2\n2\n{meta['m0']}\n{tmp_path / 'conf.bin'}\n{tmp_path / 'rhs.bin'}\n{tmp_path / 'parameters'}\n"
    p = subprocess.run([MAIN], input=stdin, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    assert p.stdout.count('Solution verified') == 5
    its = [int(line.split('(')[1].split()[0]) for line in p.stdout.splitlines() if 'time elapsed FGMRES AMG' in line]
    count = lambda key: [int(line.split('(')[1].split()[0]) for line in p.stdout.splitlines()
                         if line.startswith(f'time elapsed {key}:')][0]
    # the comparison solvers through the host API: same counts as through the binding,
    # FGMRES+SAP on 4x4 blocks per rank of the 2x2 grid (8x8 sites at 64^2)
    with ddamg.Context(64, 64) as ctx:
        ctx.set_gauge(U)
        kw = dict(tol=1e-10, max_iter=10000, m=20, restarts=50, sap_block=(8, 8))
        assert count('CG') == int(ctx.solve_method('cg', a['rhs'], meta['m0'], **kw)[1][0])
        assert count('Bi-CGstab') == int(ctx.solve_method('bicgstab', a['rhs'], meta['m0'], **kw)[1][0])
        assert count('FGMRES_SAP') == int(ctx.solve_method('fgmres_sap', a['rhs'], meta['m0'], **kw)[1][0])
    # same hierarchy and knobs through the Python binding (reference default nu2 = 2)
    prm = ddamg.params(2, agg=[(4, 4)], sap_block=[(4, 4)], nv=[8], m0=meta['m0'], nu2=2, fgmres_m=20,
                       rng_ranks=(2, 2))
    with ddamg.Context(64, 64) as ctx:
        ctx.set_gauge(U)
        expect = []
        for cyc in (0, 1):
            prm.cycle = cyc
            ctx.setup(prm)
            expect.append(int(ctx.solve(a['rhs'])[1][0]))
    assert its == expect


@pytest.mark.parametrize('kcycle', [True, False])
def test_cycle_matches_oracle(golden, kcycle):
    """One preconditioner application (K- or V-cycle from level 0, zero initial guess)
    equals the oracle's AlgebraicMG::k_cycle / v_cycle on the same field."""
    meta, a = golden('g16_setup3')
    U = np.exp(1j * a['theta'])
    lv = meta['levels']
    levels = [orc.LevelSpec(tuple(l['agg']), tuple(l['sap_block']), l['nv']) for l in lv]
    prm_o = orc.AmgParams(levels=levels, m0=meta['m0'], rng_ranks=tuple(meta['ranks']))
    amg = orc.AMG(U, prm_o)
    amg.setup()
    ref = (amg.k_cycle if kcycle else amg.v_cycle)(0, a['rhs'], np.zeros_like(a['rhs']))
    prm = ddamg.params(3, agg=[tuple(l['agg']) for l in lv], sap_block=[tuple(l['sap_block']) for l in lv],
                       nv=[l['nv'] for l in lv], m0=meta['m0'], rng_ranks=tuple(meta['ranks']),
                       ortho_coarse=0)
    with ddamg.Context(16, 16) as ctx:
        ctx.set_gauge(U)
        ctx.setup(prm)
        out = ctx.cycle(0, a['rhs'], np.zeros_like(a['rhs']), kcycle=kcycle)
    assert rel(out, ref) < 1e-9

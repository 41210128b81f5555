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
# the reference's own src/main.cpp + src/methods.cpp, compiled unchanged against the host
# API and linked with this library (make -C oracle dropin, where the reference is present)
DROPIN = os.path.join(ROOT, 'oracle', '_ref', 'dropin_main_64')


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


LEVEL_STEPS = os.path.join(ROOT, 'ddalphaamg-sm_amd', 'lib', 'level_steps_16x16')


def test_setup_steps_through_the_abi(golden):
    """dd_setup_storage + dd_set_test_vectors + dd_orthonormalize + dd_make_coarse_links
    rebuild the reference's level-1 operator from its own (orthonormal) level-0 test
    vectors (fixture g16_setup2), and dd_get_links at level 0 gives the Wilson-Dirac link
    matrices B = (1 - sigma) U / 2, C = (1 + sigma) U*(x - mu) / 2."""
    meta, a = golden('g16_setup2')
    U = np.exp(1j * a['theta'])
    lv = meta['levels']
    prm = ddamg.params(2, agg=[tuple(l['agg']) for l in lv], sap_block=[tuple(l['sap_block']) for l in lv],
                       nv=[l['nv'] for l in lv], m0=meta['m0'], rng_ranks=tuple(meta['ranks']))
    with ddamg.Context(16, 16) as ctx:
        ctx.set_gauge(U)
        ctx.setup_storage(prm)
        ctx.set_test_vectors(0, a['tvec_l0'])
        ctx.orthonormalize(0)
        assert rel(ctx.test_vectors(0), a['tvec_l0']) < 1e-13
        ctx.make_coarse_links(0)
        st = orc.ref_links_to_stencil(a['G1_l1'], a['G2_l1'], a['G3_l1'], meta['m0'])
        A, B, Cm = ctx.links(1)
        assert rel(A, st.A) < 1e-12 and rel(B, st.B) < 1e-12 and rel(Cm, st.C) < 1e-12
        assert rel(ctx.level_apply(1, ctx.restrict(0, a['rhs'])), a['Dop1_Pdagg_rhs']) < 1e-11
        A0, B0, C0 = ctx.links(0)
        d = orc.dirac_stencil(U, meta['m0'])
        assert np.abs(A0).max() == 0
        assert rel(B0, d.B) < 1e-15 and rel(C0, d.C) < 1e-15


@pytest.mark.skipif(not os.path.exists(LEVEL_STEPS), reason='host example not built (make -C ddalphaamg-sm_amd/host NS=16 NT=16)')
def test_level_objects_step_by_step(golden, tmp_path):
    """The reference's Level objects built by hand (Level(l, U), tvec filled by the
    caller, orthonormalize, makeCoarseLinks) reproduce the reference's hierarchy: the
    level-1 links in the reference's G1/G2/G3 layout, P^dagger b, D_1 P^dagger b and
    P P^dagger b of fixture g16_setup2."""
    meta, a = golden('g16_setup2')
    orc.write_records(str(tmp_path / 'conf.bin'), np.exp(1j * a['theta']))
    orc.write_records(str(tmp_path / 'rhs.bin'), a['rhs'])
    for cc in range(a['tvec_l0'].shape[0]):
        orc.write_records(str(tmp_path / f'tv{cc}.bin'), a['tvec_l0'][cc])
    (tmp_path / 'parameters').write_text('0 4 4 4 2 2\n')   # 4x4 aggregates, 4x4-site SAP blocks per 2x2 rank
    stdin = (f"2 2 {meta['m0']} {tmp_path / 'conf.bin'} {tmp_path / 'rhs.bin'} {tmp_path / 'parameters'} "
             f"{tmp_path / 'tv'} {tmp_path / 'out_'}\n")
    p = subprocess.run([LEVEL_STEPS], input=stdin, capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    assert 'are orthonormalized' in p.stdout and 'level steps done' in p.stdout
    load = lambda name, shape: np.fromfile(str(tmp_path / f'out_{name}.bin'), np.complex128).reshape(shape)
    nv = meta['levels'][0]['nv']
    for g in ('G1', 'G2', 'G3'):
        ref = a[f'{g}_l1']
        assert rel(load(g, ref.shape), ref) < 1e-11, g
    assert rel(load('Pdagg', a['Pdagg_rhs'].shape), a['Pdagg_rhs']) < 1e-11
    assert rel(load('D1', a['Dop1_Pdagg_rhs'].shape), a['Dop1_Pdagg_rhs']) < 1e-11
    assert rel(load('P', a['P_Pdagg_rhs'].shape), a['P_Pdagg_rhs']) < 1e-11
    assert rel(load('tvec', (nv, 16, 16, 2)), a['tvec_l0']) < 1e-13


@pytest.mark.skipif(not os.path.exists(DROPIN), reason='drop-in build absent (make -C oracle dropin)')
def test_reference_main_on_the_library(golden, tmp_path):
    """The reference's main() (CG on the normal equations, then FGMRES with the V- and the
    K-cycle and its own check_solution) runs on the GPU through the host API, and gives
    the reference's iteration counts for its default knobs (fixture g64_defaults, the
    compiled reference on the same field, right-hand side and 2x2 rank grid)."""
    meta, a = golden('g64_solve')
    with open(os.path.join(ROOT, 'tests', 'golden', 'g64_defaults.json')) as f:
        ref = __import__('json').load(f)
    orc.write_records(str(tmp_path / 'conf.bin'), np.exp(1j * a['theta']))
    orc.write_records(str(tmp_path / 'rhs.bin'), a['rhs'])
    (tmp_path / 'parameters').write_text('0 16 16 8 8 8\n')   # 4x4 aggregates, 4x4-site SAP blocks per 2x2 rank
    stdin = f"2\n2\n2\n{meta['m0']}\n{tmp_path / 'conf.bin'}\n{tmp_path / 'rhs.bin'}\n{tmp_path / 'parameters'}\n"
    p = subprocess.run([DROPIN], input=stdin, capture_output=True, text=True, timeout=120, cwd=tmp_path)
    print(p.stdout[-3000:])
    assert p.returncode == 0, p.stderr
    assert p.stdout.count('Solution verified') == 2
    assert 'CG for D^+D converged' in p.stdout
    its = [int(line.split()[1]) for line in p.stdout.splitlines() if line.startswith('With ')]
    assert its == [ref['V_iterations'], ref['K_iterations']], its   # main() runs V, then K
    assert (tmp_path / 'metadata_64x64_Levels2_Kcycle.dat').exists()


@pytest.mark.parametrize('nu1', [0, 1])
@pytest.mark.parametrize('kcycle', [True, False])
def test_cycle_matches_oracle(golden, kcycle, nu1):
    """One preconditioner application (K- or V-cycle from level 0, zero initial guess,
    with and without pre-smoothing) equals the oracle's AlgebraicMG::k_cycle / v_cycle on
    the same field."""
    meta, a = golden('g16_setup3')
    U = np.exp(1j * a['theta'])
    lv = meta['levels']
    levels = [orc.LevelSpec(tuple(l['agg']), tuple(l['sap_block']), l['nv']) for l in lv]
    prm_o = orc.AmgParams(levels=levels, m0=meta['m0'], rng_ranks=tuple(meta['ranks']), nu1=nu1)
    amg = orc.AMG(U, prm_o)
    amg.setup()
    ref = (amg.k_cycle if kcycle else amg.v_cycle)(0, a['rhs'], np.zeros_like(a['rhs']))
    prm = ddamg.params(3, agg=[tuple(l['agg']) for l in lv], sap_block=[tuple(l['sap_block']) for l in lv],
                       nv=[l['nv'] for l in lv], m0=meta['m0'], rng_ranks=tuple(meta['ranks']),
                       ortho_coarse=0, nu1=nu1)
    with ddamg.Context(16, 16) as ctx:
        ctx.set_gauge(U)
        ctx.setup(prm)
        out = ctx.cycle(0, a['rhs'], np.zeros_like(a['rhs']), kcycle=kcycle)
    assert rel(out, ref) < 1e-9


@pytest.mark.parametrize('cycle', ['K', 'V'])
def test_presmoothing_solve_matches_oracle(golden, cycle):
    """With one pre-smoothing SAP iteration (nu1 = 1) the solve, whose cycles start from a
    zero initial guess inside the FGMRES preconditioner (so the library forms the zero
    guess, the pre-smoothing and the residual itself), gives the oracle's iteration count
    and solution in FP64, and the mixed-precision solve reaches the same FP64 residual
    bound (src/fgmres.cpp:77-90)."""
    meta, a = golden('g16_setup3')
    U = np.exp(1j * a['theta'])
    lv = meta['levels']
    levels = [orc.LevelSpec(tuple(l['agg']), tuple(l['sap_block']), l['nv']) for l in lv]
    amg = orc.AMG(U, orc.AmgParams(levels=levels, m0=meta['m0'], rng_ranks=tuple(meta['ranks']), nu1=1))
    amg.setup()
    res = amg.solve(a['rhs'], cycle)
    prm = ddamg.params(3, agg=[tuple(l['agg']) for l in lv], sap_block=[tuple(l['sap_block']) for l in lv],
                       nv=[l['nv'] for l in lv], m0=meta['m0'], rng_ranks=tuple(meta['ranks']),
                       ortho_coarse=0, nu1=1, cycle=1 if cycle == 'K' else 0)
    with ddamg.Context(16, 16) as ctx:
        ctx.set_gauge(U)
        ctx.setup(prm)
        x, its, info = ctx.solve(a['rhs'])
        prm.precision = 1
        ctx.setup(prm)
        xm, itm, infom = ctx.solve(a['rhs'])
    rr = lambda y: np.linalg.norm(a['rhs'] - orc.D_phi(U, y, meta['m0'])) / np.linalg.norm(a['rhs'])
    print(f'nu1 = 1, {cycle}-cycle: {its[0]} iterations (oracle {res.iterations[0]}), relres {rr(x):.2e}, '
          f'vs oracle x {rel(x, res.x[0]):.2e}; mixed {itm[0]} iterations, relres {rr(xm):.2e}')
    assert int(its[0]) == int(res.iterations[0])
    assert rel(x, res.x[0]) < 1e-9 and rr(x) < 1e-10
    assert rr(xm) < 1e-10 and infom.relres < 1e-10

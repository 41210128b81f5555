"""Writes the golden fixtures under tests/golden/ by running the compiled reference.

Run in the development container (where the reference sources compile):

    make -C oracle ref N="16 32 64"
    python tests/golden/make_golden.py [fixture ...]      # all fixtures by default

Each fixture is an .npz of inputs and the reference's outputs (complex128), plus a
JSON sidecar with the parameters and scalar results.  The reference itself never
travels; only these vectors do.

Fixtures:
  g16_dirac   D_phi, D_dagger_phi, Level::D_operator (level 0) on a 16^2 Metropolis field
  g16_setup2  2-level hierarchy (Nv=4, 4x4 aggregates, 4x4 SAP blocks): test vectors,
              level-1 links, P^dagger b, D_1 P^dagger b, P P^dagger b, SAP(nu=1) b
  g16_setup3  3-level hierarchy (2x2 aggregates, Nv=4): test vectors of levels 0 and 1,
              level-2 links, transfer ops, D_1, SAP at levels 0 and 1
  g64_solve   cfg1 shape (64^2, m0=-0.1, 2 levels, Nv=8, nu2=1, FGMRES(20)): K- and
              V-cycle solutions, iteration counts and residuals
  g256_solve  cfg2 shape (256^2, beta=2, m0=-0.18, 3 levels, 4x4 aggregates, Nv=12,
              4x4-site SAP blocks, nu2=2, FGMRES(30), 2x4 ranks) on the field that the
              library's device Metropolis generator makes from seed 256: K- and V-cycle
              iteration counts, residuals and the solution at 512 sampled sites
  g512_solve  cfg3 shape (512^2, as g256_solve, seed 512): K-cycle iteration count,
              residual and the solution at 512 sampled sites
              (both need the field as .npy: python tests/golden/make_golden.py
               g256_solve=theta_256.npy g512_solve=theta_512.npy)
  g64_defaults g64_solve's inputs with the reference's default knobs (nu2=2, FGMRES(20)):
              K- and V-cycle iteration counts, as the reference's main() runs them
  g32_methods the comparison solvers of src/methods.cpp on a 32^2 field (m0=-0.1, tol
              1e-10, 2x2 ranks so that FGMRES_SAP's 4x4 blocks per rank are 4x4 sites):
              CG on the normal equations, BiCGStab, GMRES(20) and FGMRES(20)+SAP --
              iteration counts, final residuals and solutions
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle import ddamg_oracle as orc  # noqa: E402
from oracle import run_ref  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
RANKS = (2, 2)


def save(name, meta, **arrays):
    np.savez(os.path.join(OUT, name + '.npz'), **arrays)
    with open(os.path.join(OUT, name + '.json'), 'w') as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    print('wrote', name, {k: v.shape for k, v in arrays.items()})


def field(N, seed, sweeps=2000):
    theta = orc.gauge_metropolis(N, N, 2.0, sweeps, seed=seed)
    return theta, np.exp(1j * theta)


def gaussian(N, seed, dof=2):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((N, N, dof)) + 1j * rng.standard_normal((N, N, dof))


def level_meta(levels):
    return [dict(agg=list(s.agg), sap_block=list(s.sap_block), nv=s.nv) for s in levels]


def g32_methods():
    N, m0 = 32, -0.1
    theta, U = field(N, seed=3201, sweeps=500)
    b = gaussian(N, 3202)
    levels = [orc.LevelSpec((4, 4), (4, 4), 4)]
    wd, _ = run_ref.run('methods', U, b, levels, m0, ranks=RANKS, gmres_m=20, gmres_restarts=200, max_iter=20000)
    res = run_ref.result(wd)
    meta = dict(N=N, m0=m0, beta=2.0, sweeps=500, seed=3201, ranks=list(RANKS), tol=1e-10, gmres_m=20,
                gmres_restarts=200, max_iter=20000, sap_block=[4, 4], **res)
    save('g32_methods', meta, theta=theta, rhs=b, x_cg=run_ref.load(wd, 'x_cg', RANKS),
         x_bicgstab=run_ref.load(wd, 'x_bicgstab', RANKS), x_gmres=run_ref.load(wd, 'x_gmres', RANKS),
         x_fgmres_sap=run_ref.load(wd, 'x_fgmres_sap', RANKS))
    run_ref.cleanup(wd)


def sample_sites(N, count=512, seed=7):
    """Site indices (x*N + t) at which the large-lattice fixtures keep the solution."""
    return np.sort(np.random.default_rng(seed).choice(N * N, size=count, replace=False))


def big_solve(name, N, theta_path, cycles):
    """Reference solves at the cfg2/cfg3 shape on a field from the library's device
    generator (seed = N).  The fixture keeps the field's SHA-256, so a test can tell
    that it regenerated the same field, the iteration counts and residuals, and the
    solution at sampled sites."""
    import hashlib
    seed, m0, ranks = N, -0.18, (2, 4)
    theta = np.load(theta_path)
    assert theta.shape == (N, N, 2) and theta.dtype == np.float64
    U = np.exp(1j * theta)
    b = gaussian(N, seed + 1)
    levels = [orc.LevelSpec((4, 4), (4, 4), 12), orc.LevelSpec((4, 4), (4, 4), 12)]
    sites = sample_sites(N)
    meta = dict(N=N, m0=m0, beta=2.0, sweeps=2000, seed=seed, step=0.5, rhs_seed=seed + 1,
                theta_sha256=hashlib.sha256(np.ascontiguousarray(theta).tobytes()).hexdigest(),
                levels=level_meta(levels), ranks=list(ranks), nu1=0, nu2=2, fgmres_m=30, fgmres_restarts=50,
                tol=1e-10, plaquette=float(np.mean(np.cos(orc.plaquette_angle(theta)))))
    arr = dict(sites=sites)
    for cyc in cycles:
        wd, _ = run_ref.run('solve', U, b, levels, m0, ranks=ranks, nu2=2, fgmres_m=30, cycle=cyc, timeout=3600)
        res = run_ref.result(wd)
        meta[f'{cyc}_iterations'] = res['iterations']
        meta[f'{cyc}_relres'] = res['relres']
        meta[f'{cyc}_setup_s'] = res['setup_s']
        meta[f'{cyc}_solve_s'] = res['solve_s']
        x = run_ref.load(wd, 'x', ranks)
        arr[f'x_{cyc}'] = x.reshape(N * N, 2)[sites]
        run_ref.cleanup(wd)
    save(name, meta, **arr)


def g64_defaults():
    """The g64_solve field and right-hand side solved with the reference's default knobs
    (nu2 = 2, FGMRES(20), src/variables.cpp:90-134), as its own main() runs them: K- and
    V-cycle iteration counts and residuals (JSON only; the inputs are g64_solve's)."""
    meta0 = json.load(open(os.path.join(OUT, 'g64_solve.json')))
    a = dict(np.load(os.path.join(OUT, 'g64_solve.npz')))
    U = np.exp(1j * a['theta'])
    levels = [orc.LevelSpec(tuple(l['agg']), tuple(l['sap_block']), l['nv']) for l in meta0['levels']]
    meta = dict(N=meta0['N'], m0=meta0['m0'], ranks=meta0['ranks'], levels=meta0['levels'], nu1=0, nu2=2,
                fgmres_m=20, fgmres_restarts=50, tol=1e-10, inputs='g64_solve')
    for cyc in ('K', 'V'):
        wd, _ = run_ref.run('solve', U, a['rhs'], levels, meta0['m0'], ranks=tuple(meta0['ranks']), nu2=2,
                            fgmres_m=20, cycle=cyc, dump_x=0)
        res = run_ref.result(wd)
        meta[f'{cyc}_iterations'] = res['iterations']
        meta[f'{cyc}_relres'] = res['relres']
        run_ref.cleanup(wd)
    with open(os.path.join(OUT, 'g64_defaults.json'), 'w') as f:
        json.dump(meta, f, indent=1, sort_keys=True)
    print('wrote g64_defaults', meta)


def main(names=()):
    if names:
        for nm in names:
            if nm.startswith('g256_solve='):
                big_solve('g256_solve', 256, nm.split('=', 1)[1], ('K', 'V'))
            elif nm.startswith('g512_solve='):
                big_solve('g512_solve', 512, nm.split('=', 1)[1], ('K',))
            else:
                {'g32_methods': g32_methods, 'g64_defaults': g64_defaults}[nm]()
        return
    g32_methods()
    # ---- 16^2 operator fixtures -------------------------------------------------------
    N, m0 = 16, -0.1
    theta, U = field(N, seed=1601)
    b = gaussian(N, 1602)
    levels = [orc.LevelSpec((4, 4), (4, 4), 4)]
    wd, _ = run_ref.run('dirac', U, b, levels, m0, ranks=RANKS)
    save('g16_dirac', dict(N=N, m0=m0, beta=2.0, sweeps=2000, seed=1601, ranks=list(RANKS)),
         theta=theta, rhs=b, Dphi=run_ref.load(wd, 'Dphi', RANKS),
         Ddagphi=run_ref.load(wd, 'Ddagphi', RANKS), Dop0=run_ref.load(wd, 'Dop0', RANKS))
    run_ref.cleanup(wd)

    wd, _ = run_ref.run('setup', U, b, levels, m0, ranks=RANKS, nu2=2, sap_nu=1)
    arr = dict(theta=theta, rhs=b,
               tvec_l0=np.stack([run_ref.load(wd, f'tvec_l0_{c}', RANKS) for c in range(4)]),
               G1_l1=run_ref.load(wd, 'G1_l1', RANKS), G2_l1=run_ref.load(wd, 'G2_l1', RANKS),
               G3_l1=run_ref.load(wd, 'G3_l1', RANKS), Pdagg_rhs=run_ref.load(wd, 'Pdagg_rhs', RANKS),
               Dop1_Pdagg_rhs=run_ref.load(wd, 'Dop1_Pdagg_rhs', RANKS),
               P_Pdagg_rhs=run_ref.load(wd, 'P_Pdagg_rhs', RANKS), SAP0_rhs=run_ref.load(wd, 'SAP0_rhs', RANKS))
    save('g16_setup2', dict(N=N, m0=m0, levels=level_meta(levels), ranks=list(RANKS), sap_nu=1,
                            setup_sap_iters=4), **arr)
    run_ref.cleanup(wd)

    levels3 = [orc.LevelSpec((2, 2), (4, 4), 4), orc.LevelSpec((2, 2), (2, 2), 4)]
    wd, _ = run_ref.run('setup', U, b, levels3, m0, ranks=RANKS, nu2=2, sap_nu=1)
    arr = dict(theta=theta, rhs=b,
               tvec_l0=np.stack([run_ref.load(wd, f'tvec_l0_{c}', RANKS) for c in range(4)]),
               tvec_l1=np.stack([run_ref.load(wd, f'tvec_l1_{c}', RANKS) for c in range(4)]),
               G1_l2=run_ref.load(wd, 'G1_l2', RANKS), G2_l2=run_ref.load(wd, 'G2_l2', RANKS),
               G3_l2=run_ref.load(wd, 'G3_l2', RANKS), Pdagg_rhs=run_ref.load(wd, 'Pdagg_rhs', RANKS),
               Dop1_Pdagg_rhs=run_ref.load(wd, 'Dop1_Pdagg_rhs', RANKS),
               P_Pdagg_rhs=run_ref.load(wd, 'P_Pdagg_rhs', RANKS), SAP0_rhs=run_ref.load(wd, 'SAP0_rhs', RANKS),
               SAP1_Pdagg_rhs=run_ref.load(wd, 'SAP1_Pdagg_rhs', RANKS))
    save('g16_setup3', dict(N=N, m0=m0, levels=level_meta(levels3), ranks=list(RANKS), sap_nu=1,
                            setup_sap_iters=4), **arr)
    run_ref.cleanup(wd)

    # ---- 64^2 full solve (cfg1 shape) ---------------------------------------------------
    N, m0 = 64, -0.1
    theta, U = field(N, seed=6401)
    b = gaussian(N, 6402)
    levels = [orc.LevelSpec((4, 4), (4, 4), 8)]
    meta = dict(N=N, m0=m0, beta=2.0, sweeps=2000, seed=6401, levels=level_meta(levels), ranks=list(RANKS),
                nu1=0, nu2=1, fgmres_m=20, fgmres_restarts=50, tol=1e-10,
                plaquette=float(np.mean(np.cos(orc.plaquette_angle(theta)))))
    xs = {}
    for cyc in ('K', 'V'):
        wd, _ = run_ref.run('solve', U, b, levels, m0, ranks=RANKS, nu2=1, fgmres_m=20, cycle=cyc)
        res = run_ref.result(wd)
        meta[f'{cyc}_iterations'] = res['iterations']
        meta[f'{cyc}_relres'] = res['relres']
        xs[cyc] = run_ref.load(wd, 'x', RANKS)
        run_ref.cleanup(wd)
    save('g64_solve', meta, theta=theta, rhs=b, x_K=xs['K'])


if __name__ == '__main__':
    main(sys.argv[1:])

"""Benchmark of the DD-alphaAMG solve on one MI355X.

Metric: ms per right-hand side to reach 1e-10 relative residual (true FP64 D_W
residual) at 512^2, beta=2, m0=-0.18, 3 levels, 4x4 aggregates, Nv=12, 4x4 SAP
blocks, outer FGMRES(30).  The measured form is the mixed-precision batched solve (FP32
hierarchy inside the FP64 outer FGMRES, a batch of RHS per launch) with the V-cycle and
the coarsest GMRES's degree-12 polynomial preconditioner; the same batches are timed
with the K-cycle and without the polynomial beside it, and the plain FP64 single-RHS
solve of the same system with the reference's K-cycle (and the V-cycle).

A step is one batched solve of --nrhs Gaussian right-hand sides already resident in
HBM; every timed step solves its own batch (seeded), so the mean covers the spread of
outer iteration counts over batches.  Setup (test vectors + Galerkin operators) is done once before the timed region
and reported separately.  The CPU baseline is the reference's own CPU solve
(oracle/_ref, built from the reference sources) under mpirun on the host cores, on the
same gauge field and right-hand side 0.

    python bench.py [--steps K] [--warmup W] [--N 512] [--nrhs 32]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, 'ddalphaamg-sm_amd'))


def log(*a):
    print(*a, file=sys.stderr, flush=True)


COARSE_BASIS = None   # None: the library default; 0 FP32, 1 binary16 (--coarse-basis)
SAP_LINKS = None      # None: the library default; 0 FP32, 1 binary16 (--coarse-sap-links)
BOOTSTRAP = 0         # bootstrap passes after the setup (--bootstrap; 0 = the reference's setup)
SAP_FUSED = None      # None: the library default; 0 separate launches per colour, 1 one launch per iteration
GRAPH = 0             # 1: the cycles replayed from captured HIP graphs (--graph)
CYCLE = 0             # cycle of the mixed rows (--cycle): 0 V-cycle (FGMRES_AMG_v_cycle), 1 K-cycle
                      # (FGMRES_AMG_k_cycle); the plain FP64 rows use the K-cycle unless asked
COARSE_POLY = 12      # degree of the coarsest GMRES's polynomial preconditioner in the FP32 (mixed)
                      # hierarchies (--coarse-poly; 0: the reference's plain GMRES); FP64 rows keep plain GMRES


def hierarchy_params(ddamg, N, precision, nrhs, m0=-0.18, coarse_poly=None, cycle=None):
    if cycle is None:
        cycle = CYCLE if precision == 1 else 1
    p = ddamg.params(3, agg=[(4, 4), (4, 4)], sap_block=[(4, 4), (4, 4)], nv=[12, 12], m0=m0, nu2=2,
                     fgmres_m=30, cycle=cycle, precision=precision, max_rhs=nrhs, rng_ranks=(4, 4))
    if COARSE_BASIS is not None:
        p.coarse_basis = COARSE_BASIS
    if SAP_LINKS is not None:
        p.coarse_sap_links = SAP_LINKS
    p.bootstrap = BOOTSTRAP
    if SAP_FUSED is not None:
        p.sap_fused = SAP_FUSED
    p.graph = GRAPH
    p.coarse_poly = (COARSE_POLY if precision == 1 else 0) if coarse_poly is None else coarse_poly
    return p


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=4)
    ap.add_argument('--warmup', type=int, default=1)
    ap.add_argument('--N', type=int, default=512)
    ap.add_argument('--nrhs', type=int, default=32)
    ap.add_argument('--sweeps', type=int, default=2000)
    ap.add_argument('--seed', type=int, default=2024)
    ap.add_argument('--cpu-baseline', choices=['reference', 'none'], default='reference')
    ap.add_argument('--cpu-ranks', type=str, default='4x4')
    ap.add_argument('--skip-plain', action='store_true')
    ap.add_argument('--skip-mixed', action='store_true')
    ap.add_argument('--no-secondary', action='store_true',
                    help='skip the secondary rows (256^2 plain FP64, 256^2 32-RHS and the ragged batch)')
    ap.add_argument('--calibrate-copy', action='store_true',
                    help='only run three device copies of 256 MiB (scale of the HBM byte counters)')
    ap.add_argument('--probe', choices=['wilson', 'sap0'], default=None,
                    help='only run one kernel family on 32 FP64 systems at --N (counter runs): the Wilson-Dirac '
                         'stencil 5 times, or one level-0 SAP iteration (4x4-site blocks)')
    ap.add_argument('--coarse-basis', type=int, choices=[0, 1], default=None,
                    help='storage of the coarsest GMRES basis in the FP32 hierarchy (0 FP32, 1 binary16)')
    ap.add_argument('--coarse-sap-links', type=int, choices=[0, 1], default=None,
                    help='links read by the matrix-core coarse SAP of the FP32 hierarchy (0 FP32, 1 binary16 copy)')
    ap.add_argument('--skip-fp32-basis', action='store_true',
                    help='skip the row that repeats the batch with an FP32 coarsest Krylov basis')
    ap.add_argument('--bootstrap', type=int, default=0,
                    help='bootstrap passes after the setup (0: the reference setup only)')
    ap.add_argument('--sap-fused', type=int, choices=[0, 1], default=None,
                    help='coarse SAP iterations in one launch (1) or per colour with separate residuals (0)')
    ap.add_argument('--graph', type=int, choices=[0, 1], default=0,
                    help='replay the cycles from captured HIP graphs (K-cycle FGMRES device-controlled)')
    ap.add_argument('--coarse-poly', type=int, default=12,
                    help='degree of the GMRES polynomial right preconditioner of the coarsest solve in the mixed '
                         'hierarchies (0: the plain GMRES of the reference, which the FP64 rows always use)')
    ap.add_argument('--cycle', type=int, choices=[0, 1], default=0,
                    help='cycle of the mixed rows: 0 V-cycle (default), 1 K-cycle (the plain FP64 rows run the '
                         'K-cycle and, beside it, the V-cycle)')
    ap.add_argument('--skip-no-poly', action='store_true',
                    help='skip the row that repeats the batches without the coarsest polynomial preconditioner')
    ap.add_argument('--setup-only', action='store_true',
                    help='only generate the field and build the mixed and the FP64 hierarchies (setup profiles)')
    args = ap.parse_args()
    global COARSE_BASIS, SAP_LINKS, BOOTSTRAP, SAP_FUSED, GRAPH, COARSE_POLY, CYCLE
    BOOTSTRAP = args.bootstrap
    CYCLE = args.cycle
    COARSE_POLY = args.coarse_poly
    GRAPH = args.graph
    SAP_FUSED = args.sap_fused
    COARSE_BASIS = args.coarse_basis
    SAP_LINKS = args.coarse_sap_links
    if args.gpus != 1:
        print('bench.py: this build runs on one GPU only (the multi-GPU domain decomposition is '
              'design-only); run with --gpus 1', file=sys.stderr)
        sys.exit(2)
    import torch
    import ddamg

    if args.calibrate_copy:
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device='cuda')
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log('calibration copy: 3 x (268435456 bytes read + 268435456 bytes written)')
        return

    if args.probe:
        probe(ddamg, torch, args.probe, args.N, args.nrhs, args.seed)
        return

    N, nrhs = args.N, args.nrhs
    m0 = -0.18
    dev = torch.device('cuda:0')
    # gauge field: seeded quenched Metropolis on the device (beta = 2, cold start)
    ctx = ddamg.Context(N, N)
    t0 = time.perf_counter()
    theta = ctx.generate_gauge(2.0, args.sweeps, args.seed)
    log(f'gauge {N}^2 {args.sweeps} sweeps: {time.perf_counter() - t0:.2f}s, '
        f'plaquette {np.mean(np.cos(theta[..., 0] + np.roll(theta[..., 1], -1, 1) - np.roll(theta[..., 0], -1, 0) - theta[..., 1])):.4f}')
    if args.setup_only:
        # setup phase alone (test vectors, smoothing, orthonormalisation, Galerkin products),
        # twice per precision: the second build runs on loaded code
        out = {}
        for prec, name in ((1, 'mixed'), (0, 'fp64')):
            ts = [ctx.setup(hierarchy_params(ddamg, N, prec, nrhs if prec else 1, m0)) for _ in range(2)]
            out[name] = ts
            log(f'setup {name}: ' + ', '.join(f'{t:.3f}s' for t in ts))
        ctx.close()
        print(json.dumps({'setup_only': True, 'N': N, 'setup_s': out}))
        return
    # one batch of right-hand sides per timed step (batch k from seed + 1 + k), all resident
    # in HBM before the timed region: the mean over distinct batches, not one batch's draw
    # of outer iterations, is what the steps measure
    nb = max(1, args.steps)
    Bs = []
    for k in range(nb):
        rng = np.random.default_rng(args.seed + 1 + k)
        Bs.append(rng.standard_normal((nrhs, N, N, 2)) + 1j * rng.standard_normal((nrhs, N, N, 2)))
    B = Bs[0]
    b_devs = [torch.from_numpy(Bk.view(np.float64).copy()).to(dev) for Bk in Bs]
    x_devs = [torch.zeros_like(v) for v in b_devs]
    b_dev, x_dev = b_devs[0], x_devs[0]
    torch.cuda.synchronize()

    # ---- mixed precision, batched -----------------------------------------------------
    if args.skip_mixed:
        nrhs = 1
    setup_mixed = ctx.setup(hierarchy_params(ddamg, N, 0 if args.skip_mixed else 1, nrhs, m0))
    tune = ctx.tuning(1)
    log(f'setup mixed: {setup_mixed:.3f}s, autotuned coarse SAP variant {tune[0]} ({tune[1]:.3f} ms/half-sweep)')
    # warm-up solves with every kernel family timed: per-family breakdown
    fams = {ddamg.K_WILSON: 'wilson', ddamg.K_COARSE: 'coarse_stencil', ddamg.K_SAP0: 'sap_level0',
            ddamg.K_SAPC: 'sap_coarse', ddamg.K_MULTIDOT: 'gs_multidot', ddamg.K_MULTIAXPY: 'gs_multiaxpy',
            ddamg.K_RESTRICT: 'restrict', ddamg.K_PROLONG: 'prolong', ddamg.K_COARSEST: 'coarsest_gmres',
            ddamg.K_DOT: 'dot'}
    breakdown = {}
    if args.warmup > 0:
        for k in fams:
            ddamg.timer_enable(k, True)
        ddamg.timer_reset()
        for _ in range(args.warmup):
            its, info = ctx.solve_dev(b_dev.data_ptr(), x_dev.data_ptr(), nrhs)
        for k, name in fams.items():
            ms, n, by = ddamg.timer_read(k)
            if n:
                breakdown[name] = {'ms_per_solve': ms / args.warmup, 'launches': n // args.warmup,
                                   'GB_per_s': by / (ms * 1e-3) / 1e9 if ms > 0 else None}
            ddamg.timer_enable(k, False)
        log('kernel breakdown (warm-up, per solve): ' + json.dumps(breakdown))
    dom = max(breakdown, key=lambda k: breakdown[k]['ms_per_solve']) if breakdown else 'sap_coarse'
    dom_id = {v: k for k, v in fams.items()}[dom]
    torch.cuda.synchronize()
    ddamg.timer_enable(dom_id, True)
    ddamg.timer_reset()
    step_its = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        its, info = ctx.solve_dev(b_devs[k].data_ptr(), x_devs[k].data_ptr(), nrhs)
        step_its.append(its)
    torch.cuda.synchronize()
    t_mixed = (time.perf_counter() - t0) / args.steps
    dom_ms, dom_n, dom_bytes = ddamg.timer_read(dom_id)
    ddamg.timer_enable(dom_id, False)
    dom_flops = ddamg.timer_flops(dom_id)
    hbm = dom_bytes / (dom_ms * 1e-3) / 1e9 if dom_ms > 0 else 0.0
    if dom_flops > 0:
        # a matrix-core kernel that counts its own products (the coarse SAP, FP32): against
        # the FP32 matrix peak; its HBM rate alongside
        peak = 157.3
        achieved = dom_flops / (dom_ms * 1e-3) / 1e12 if dom_ms > 0 else 0.0
        roofline = {'bound': 'mfma', 'kernel': dom, 'achieved': achieved, 'peak': peak, 'unit': 'TFLOP/s',
                    'frac': achieved / peak, 'traffic': measured_traffic(dom, N, nrhs),
                    'avg_launch_us': dom_ms * 1e3 / max(dom_n, 1), 'launches': dom_n,
                    'algorithmic_flops_per_launch': dom_flops / max(dom_n, 1),
                    'algorithmic_bytes_per_launch': dom_bytes / max(dom_n, 1), 'hbm_GB_per_s': hbm}
        log(f'roofline {dom}: {achieved:.2f} TFLOP/s (FP32 matrix cores) over {dom_n} launches, {hbm:.1f} GB/s')
    else:
        peak = 8000.0
        achieved = hbm
        roofline = {'bound': 'hbm', 'kernel': dom, 'achieved': achieved, 'peak': peak, 'unit': 'GB/s',
                    'frac': achieved / peak, 'traffic': measured_traffic(dom, N, nrhs),
                    'avg_launch_us': dom_ms * 1e3 / max(dom_n, 1), 'launches': dom_n,
                    'algorithmic_bytes_per_launch': dom_bytes / max(dom_n, 1)}
        log(f'roofline {dom}: {achieved:.1f} GB/s over {dom_n} launches')
    x_mixed = x_dev.cpu().numpy().view(np.complex128).reshape(nrhs, N, N, 2)
    # true FP64 residual of every system of every timed batch
    r_dev = torch.zeros_like(b_dev)
    relres = []
    for k in range(args.steps):
        ctx.apply_D_dev(x_devs[k].data_ptr(), r_dev.data_ptr(), nrhs, m0)
        torch.cuda.synchronize()
        Dx = r_dev.cpu().numpy().view(np.complex128).reshape(nrhs, N, N, 2)
        relres.append(np.linalg.norm((Bs[k] - Dx).reshape(nrhs, -1), axis=1) /
                      np.linalg.norm(Bs[k].reshape(nrhs, -1), axis=1))
    relres = np.concatenate(relres) if relres else np.zeros(1)
    its = np.concatenate(step_its) if step_its else its
    outer_per_step = [int(v.max()) for v in step_its]
    log(f'mixed batched: {t_mixed * 1e3:.1f} ms per {nrhs}-RHS solve (mean of {args.steps} batches), '
        f'its {its.min()}..{its.max()} (per batch max {outer_per_step}), max relres {relres.max():.3e}')
    cst = ctx.coarsest_stats()
    nsolve = args.warmup + args.steps
    coarsest_mixed = {k: (v / nsolve if k not in ('persistent',) else v) for k, v in cst.items()}
    log('coarsest solver per outer solve: ' + json.dumps(coarsest_mixed))
    # the same batch with the coarsest Krylov basis stored in FP32 instead of binary16
    fp32_basis = None
    if not args.skip_mixed and not args.skip_fp32_basis and (COARSE_BASIS is None or COARSE_BASIS == 1):
        p32 = hierarchy_params(ddamg, N, 1, nrhs, m0)
        p32.coarse_basis = 0
        ctx.setup(p32)
        x32 = torch.zeros_like(b_dev)
        ctx.solve_dev(b_dev.data_ptr(), x32.data_ptr(), nrhs)   # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        its32, info32 = ctx.solve_dev(b_dev.data_ptr(), x32.data_ptr(), nrhs)
        torch.cuda.synchronize()
        t32 = time.perf_counter() - t0
        fp32_basis = {'ms_per_rhs': t32 * 1e3 / nrhs, 'outer_iterations': [int(its32.min()), int(its32.max())],
                      'max_relres': info32.relres,
                      'ms_per_outer_iteration': t32 * 1e3 / max(int(its32.max()), 1)}
        log(f'mixed, FP32 coarsest basis: {t32 * 1e3:.1f} ms per solve, its {its32.min()}..{its32.max()}')
    def repeat(p, what):
        # the same batches (warm-up on batch 0, then every batch timed) with hierarchy p
        ctx.setup(p)
        xs = [torch.zeros_like(v) for v in b_devs]
        ctx.solve_dev(b_devs[0].data_ptr(), xs[0].data_ptr(), nrhs)   # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = [ctx.solve_dev(b_devs[k].data_ptr(), xs[k].data_ptr(), nrhs) for k in range(args.steps)]
        torch.cuda.synchronize()
        tn = (time.perf_counter() - t0) / max(args.steps, 1)
        per = [int(r[0].max()) for r in res]
        log(f'mixed, {what}: {tn * 1e3:.1f} ms per solve, per batch {per}')
        return {'ms_per_rhs': tn * 1e3 / nrhs, 'outer_iterations_per_batch': per,
                'ms_per_outer_iteration': tn * 1e3 * args.steps / max(sum(per), 1),
                'max_relres': max(r[1].relres for r in res)}
    # the same batches without the coarsest polynomial preconditioner (the reference's plain
    # coarsest GMRES inside the same mixed hierarchy), and with the other cycle
    no_poly = other_cycle = None
    if not args.skip_mixed and not args.skip_no_poly and COARSE_POLY > 0:
        no_poly = repeat(hierarchy_params(ddamg, N, 1, nrhs, m0, coarse_poly=0),
                         'without the coarsest polynomial')
    if not args.skip_mixed and not args.skip_no_poly:
        other_cycle = repeat(hierarchy_params(ddamg, N, 1, nrhs, m0, cycle=1 - CYCLE),
                             'K-cycle' if CYCLE == 0 else 'V-cycle')
    ctx.close()

    # ---- plain FP64, one RHS -------------------------------------------------------------
    t_plain, its_plain, err_vs_fp64, setup_plain, coarsest_plain, plain_v = None, None, None, None, None, None
    if not args.skip_plain:
        ctx = ddamg.Context(N, N)
        ctx.set_gauge_theta(theta)
        setup_plain = ctx.setup(hierarchy_params(ddamg, N, 0, 1, m0))
        xp = torch.zeros_like(b_dev[:1])
        if not args.skip_mixed:
            ctx.solve_dev(b_dev.data_ptr(), xp.data_ptr(), 1)
        torch.cuda.synchronize()
        plain_times = []
        for _ in range(3):
            t0 = time.perf_counter()
            its_plain, info_p = ctx.solve_dev(b_dev.data_ptr(), xp.data_ptr(), 1)
            torch.cuda.synchronize()
            plain_times.append(time.perf_counter() - t0)
        t_plain = float(np.median(plain_times))
        x64 = xp.cpu().numpy().view(np.complex128).reshape(N, N, 2)
        err_vs_fp64 = float(np.linalg.norm(x_mixed[0] - x64) / np.linalg.norm(x64))
        log(f'plain FP64: setup {setup_plain:.3f}s, solve {t_plain * 1e3:.1f} ms, its {its_plain[0]}, '
            f'relres {info_p.relres:.3e}, mixed-vs-FP64 rel. error {err_vs_fp64:.3e}')
        cst = ctx.coarsest_stats()
        nsolve = 3 if args.skip_mixed else 4
        coarsest_plain = {k: (v / nsolve if k not in ('persistent',) else v) for k, v in cst.items()}
        log('coarsest solver per plain solve: ' + json.dumps(coarsest_plain))
        # the plain FP64 solve with the V-cycle as well (median of 3 after a warm-up)
        ctx.setup(hierarchy_params(ddamg, N, 0, 1, m0, cycle=0))
        ctx.solve_dev(b_dev.data_ptr(), xp.data_ptr(), 1)
        torch.cuda.synchronize()
        vt = []
        for _ in range(3):
            t0 = time.perf_counter()
            its_pv, info_pv = ctx.solve_dev(b_dev.data_ptr(), xp.data_ptr(), 1)
            torch.cuda.synchronize()
            vt.append(time.perf_counter() - t0)
        plain_v = {'ms_per_rhs': float(np.median(vt)) * 1e3, 'iterations': int(its_pv[0]), 'relres': info_pv.relres}
        log(f'plain FP64, V-cycle: {plain_v["ms_per_rhs"]:.1f} ms, its {its_pv[0]}, relres {info_pv.relres:.3e}')
        ctx.close()

    secondary, field256, ragged = ((None, None, None) if args.no_secondary
                                   else secondary_rows(ddamg, torch, args.sweeps, args.seed))

    # ---- CPU baseline: the reference's own solve on the host cores --------------------------
    cpu = None
    if args.cpu_baseline == 'reference':
        cpu = cpu_reference(theta, B[0], N, m0, args.cpu_ranks)
        if field256 is not None:   # the 256^2 plain FP64 configuration on the same host cores
            cpu['cfg_256'] = cpu_reference(field256['theta'], field256['b0'], 256, -0.18, args.cpu_ranks)
        if ragged is not None:   # the ragged batch's fields, right-hand side 0 of each
            cpu['ragged'] = cpu_ragged(ragged)

    ms_rhs = t_mixed * 1e3 / nrhs
    line = {
        'metric': 'ms/RHS to 1e-10 rel. residual at 512^2, mixed+batched vs plain FP64 on GPU',
        'value': ms_rhs, 'unit': 'ms/RHS', 'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': t_mixed * 1e3, 'higher_is_better': False, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'c64 hierarchy (FP32 SAP, coarse operators, test vectors) with a binary16-stored coarsest Krylov '
                 'basis (products in FP32), c128 outer FGMRES and residual', 'data': 'synthetic (seeded Metropolis gauge field, Gaussian RHS)',
        'config': {'workload': f'DD-alphaAMG {"K" if CYCLE else "V"}-cycle solve, {N}^2 Schwinger Wilson-Dirac, beta=2, m0={m0}, '
                               f'3 levels, 4x4 aggregates, Nv=12, 4x4 SAP blocks, FGMRES(30), tol 1e-10, '
                               f'{nrhs} RHS per solve; coarsest GMRES(400) to 0.1 with a degree-{COARSE_POLY} '
                               f'GMRES polynomial right preconditioner (0: none)',
                   'lattice': [N, N], 'nrhs': nrhs, 'coarse_poly': COARSE_POLY},
        'outer_iterations': [int(its.min()), int(its.max())],
        'outer_iterations_per_batch': outer_per_step,
        'ms_per_outer_iteration': t_mixed * 1e3 * args.steps / max(sum(outer_per_step), 1),
        'mixed_fp32_coarsest_basis': fp32_basis,
        'mixed_no_coarse_poly': no_poly,
        ('mixed_kcycle' if CYCLE == 0 else 'mixed_vcycle'): other_cycle,
        'max_relres': float(relres.max()),
        'setup_s_mixed': setup_mixed,
        'autotuned_coarse_sap_variant': tune[0],
        'plain_fp64_ms_per_rhs': None if t_plain is None else t_plain * 1e3,
        'plain_fp64_ms_samples': None if t_plain is None else [t * 1e3 for t in plain_times],
        'plain_fp64_iterations': None if its_plain is None else int(its_plain[0]),
        'plain_fp64_vcycle': plain_v,
        'setup_s_fp64': setup_plain,
        'speedup_vs_plain_fp64': None if t_plain is None else t_plain * 1e3 / ms_rhs,
        'speedup_vs_plain_fp64_no_coarse_poly': None if (t_plain is None or no_poly is None)
        else t_plain * 1e3 / no_poly['ms_per_rhs'],
        'speedup_vs_plain_fp64_vcycle': None if plain_v is None else plain_v['ms_per_rhs'] / ms_rhs,
        'mixed_vs_fp64_rel_error': err_vs_fp64,
        'cpu_baseline': cpu,
        'roofline': roofline,
        'kernel_breakdown_ms_per_solve': breakdown,
        'coarsest_solver_mixed': coarsest_mixed,
        'coarsest_solver_plain': coarsest_plain,
        'secondary': secondary,
    }
    print(json.dumps(line))


def probe(ddamg, torch, which, N, nrhs, seed):
    """One kernel family alone, for counter runs: k_wilson (5 launches of D on nrhs FP64
    vectors) or one level-0 SAP iteration (k_sap0 on the red, then the black blocks, with
    the Wilson residuals between).  Prints the launches' algorithmic bytes as the library
    counts them."""
    rng = np.random.default_rng(seed)
    ctx = ddamg.Context(N, N)
    ctx.generate_gauge(2.0, 200, seed)
    v = torch.from_numpy((rng.standard_normal((nrhs, N, N, 2)) + 1j * rng.standard_normal((nrhs, N, N, 2)))
                         .view(np.float64).copy()).cuda()
    out = torch.zeros_like(v)
    kid = ddamg.K_WILSON if which == 'wilson' else ddamg.K_SAP0
    ddamg.timer_enable(kid, True)
    ddamg.timer_reset()
    torch.cuda.synchronize()
    if which == 'wilson':
        for _ in range(5):
            ctx.apply_D_dev(v.data_ptr(), out.data_ptr(), nrhs, -0.18)
    else:
        vb = v.cpu().numpy().view(np.complex128).reshape(nrhs, N, N, 2)
        ctx.sap_fine(vb, np.zeros_like(vb), 1, 1e-12, (4, 4), -0.18)
    torch.cuda.synchronize()
    ms, n, by = ddamg.timer_read(kid)
    ddamg.timer_enable(kid, False)
    log(f'probe {which}: {n} launches, {ms / max(n, 1) * 1e3:.1f} us each, '
        f'{by / max(n, 1):.6e} algorithmic bytes per launch, {by / (ms * 1e-3) / 1e9:.1f} GB/s')
    print(json.dumps({'probe': which, 'N': N, 'nrhs': nrhs, 'launches': n, 'us_per_launch': ms / max(n, 1) * 1e3,
                      'algorithmic_bytes_per_launch': by / max(n, 1)}))
    ctx.close()


def secondary_rows(ddamg, torch, sweeps, seed):
    """Further workloads: the 256^2 plain FP64 single-RHS solve, 32 RHS on one 256^2 field
    (mixed, batched; K-cycle eager and from captured graphs, V-cycle eager and from captured
    graphs), and a ragged
    batch of 8 fields of sizes {64,96,128,128,192,256,256,256}^2 with m0 drawn from
    [-0.18, -0.10] and 4 RHS each, all solved at the same time (dd_solve_many_dev).
    Gauge fields: seeded quenched Metropolis, beta = 2."""
    out = {}
    rng = np.random.default_rng(seed + 7)

    def field_ctx(N, s):
        ctx = ddamg.Context(N, N)
        ctx.theta = ctx.generate_gauge(2.0, sweeps, s)
        return ctx

    def rhs(nr, N):
        B = rng.standard_normal((nr, N, N, 2)) + 1j * rng.standard_normal((nr, N, N, 2))
        return torch.from_numpy(B.view(np.float64).copy()).cuda(), B

    # 256^2 plain FP64, 1 RHS, and mixed 32 RHS on the same field
    ctx = field_ctx(256, seed + 11)
    b, B = rhs(32, 256)
    field256 = {'theta': ctx.theta, 'b0': B[0]}
    x = torch.zeros_like(b)
    for prec, nr in ((0, 1), (1, 32)):
        ctx.setup(hierarchy_params(ddamg, 256, prec, nr, -0.18, cycle=1))
        ctx.solve_dev(b.data_ptr(), x.data_ptr(), nr)   # warm-up
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        its, info = ctx.solve_dev(b.data_ptr(), x.data_ptr(), nr)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        key = 'plain_fp64_256_1rhs' if prec == 0 else 'mixed_256_32rhs'
        out[key] = {'ms_per_rhs': t * 1e3 / nr, 'iterations': [int(its.min()), int(its.max())],
                    'max_relres': info.relres}
        log(f'{key}: {t * 1e3 / nr:.2f} ms/RHS, its {its.min()}..{its.max()}')
    # the same 32 RHS with the K-cycle replayed from captured HIP graphs (its FGMRES solves
    # device-controlled), then with V-cycles, eager and replayed from captured graphs
    for cyc, graph in ((1, 1), (0, 0), (0, 1)):
        p = hierarchy_params(ddamg, 256, 1, 32, -0.18)
        p.cycle = cyc
        p.graph = graph
        ctx.setup(p)
        ctx.solve_dev(b.data_ptr(), x.data_ptr(), 32)   # warm-up (captures the graphs)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        its, info = ctx.solve_dev(b.data_ptr(), x.data_ptr(), 32)
        torch.cuda.synchronize()
        t = time.perf_counter() - t0
        key = ('mixed_256_32rhs_kcycle_graph' if cyc else
               'mixed_256_32rhs_vcycle_graph' if graph else 'mixed_256_32rhs_vcycle')
        out[key] = {'ms_per_rhs': t * 1e3 / 32, 'iterations': [int(its.min()), int(its.max())],
                    'max_relres': info.relres, 'graphs': ctx.graph_count()}
        log(f'{key}: {t * 1e3 / 32:.2f} ms/RHS, its {its.min()}..{its.max()}, graphs {ctx.graph_count()}')
    ctx.close()

    # ragged batch: 8 fields, 4 RHS each, mixed precision
    sizes = [64, 96, 128, 128, 192, 256, 256, 256]
    m0s = rng.uniform(-0.18, -0.10, len(sizes))
    ctxs, bs, xs, ragged = [], [], [], []
    t0 = time.perf_counter()
    for i, (N, m0) in enumerate(zip(sizes, m0s)):
        c = field_ctx(N, seed + 100 + i)
        c.setup(hierarchy_params(ddamg, N, 1, 4, float(m0)))
        bb, Bh = rhs(4, N)
        ctxs.append(c)
        bs.append(bb)
        xs.append(torch.zeros_like(bb))
        ragged.append({'N': N, 'm0': float(m0), 'theta': c.theta, 'b0': Bh[0]})
    t_setup = time.perf_counter() - t0
    args = ([4] * len(sizes), [v.data_ptr() for v in bs], [v.data_ptr() for v in xs])
    ddamg.solve_many_dev(ctxs, *args)   # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    infos = ddamg.solve_many_dev(ctxs, *args)
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    out['ragged_8_fields_4rhs'] = {'sizes': sizes, 'm0': [float(v) for v in m0s], 'ms_total': t * 1e3,
                                   'ms_per_rhs': t * 1e3 / (4 * len(sizes)), 'setup_s_all': t_setup,
                                   'iterations': [i.iterations for i in infos],
                                   'max_relres': max(i.relres for i in infos)}
    log(f'ragged batch: {t * 1e3:.1f} ms for 8 fields x 4 RHS ({t * 1e3 / 32:.2f} ms/RHS), '
        f'its {[i.iterations for i in infos]}')
    for c in ctxs:
        c.close()
    return out, field256, ragged


def sources_sha256(files):
    """SHA-256 over the contents of the listed source files (paths relative to the
    repository root), in the listed order."""
    import hashlib
    h = hashlib.sha256()
    for rel in files:
        with open(os.path.join(ROOT, rel), 'rb') as f:
            h.update(f.read())
    return h.hexdigest()


def measured_traffic(kernel, N, nrhs):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    measurement (profiles/), when it covers this workload and was taken on the kernel's
    current sources (the file records the SHA-256 of the sources it measured); None
    otherwise."""
    path = os.path.join(ROOT, 'profiles', f'{kernel}_{N}_mixed{nrhs}_pmc.json')
    if not os.path.exists(path):
        return None
    with open(path) as f:
        rec = json.load(f)
    srcs = rec.get('sources')
    if not srcs or sources_sha256(srcs) != rec.get('sources_sha256'):
        log(f'traffic of {kernel}: {os.path.basename(path)} measured other sources; reported as null')
        return None
    return rec['traffic_bytes_per_dispatch']


def cpu_reference(theta, b, N, m0, ranks_s):
    """The reference's FGMRES + K-cycle solve (oracle/_ref, compiled from the reference
    sources) under mpirun, same field, RHS 0 and hierarchy parameters."""
    try:
        from oracle import ddamg_oracle as orc
        from oracle import run_ref
    except Exception as e:  # noqa: BLE001
        return {'error': f'oracle import failed: {e}'}
    if not run_ref.available(N):
        return {'error': f'oracle/_ref/ddamg_ref_{N} not built'}
    rx, rt = (int(v) for v in ranks_s.split('x'))
    levels = [orc.LevelSpec((4, 4), (4, 4), 12), orc.LevelSpec((4, 4), (4, 4), 12)]
    t0 = time.perf_counter()
    try:
        wd, _ = run_ref.run('solve', np.exp(1j * theta), b, levels, m0, ranks=(rx, rt), nu2=2, fgmres_m=30,
                            cycle='K', dump_x=0, timeout=900)
        res = run_ref.result(wd)
        run_ref.cleanup(wd)
    except Exception as e:  # noqa: BLE001
        return {'error': str(e)[-400:]}
    log(f'cpu reference: {res} ({time.perf_counter() - t0:.1f}s wall)')
    try:
        affinity = len(os.sched_getaffinity(0))
    except AttributeError:
        affinity = None
    return {'value': res['solve_s'] * 1e3, 'unit': 'ms/RHS', 'cores': rx * rt, 'kind': 'reference',
            'sample': f'one full reference solve (FGMRES(30)+K-cycle, {N}^2, same field and RHS 0) under '
                      f'mpirun -n {rx * rt} ({rx}x{rt} ranks); setup {res["setup_s"]:.2f}s excluded',
            'iterations': res['iterations'], 'relres': res['relres'], 'setup_s': res['setup_s'],
            'host_cpus_visible': os.cpu_count(), 'host_cpus_affinity': affinity,
            'host_cpu_share': os.environ.get('OMP_NUM_THREADS')}


# MPI rank grids for the reference on the ragged sizes: the reference colours its SAP
# blocks per rank, so every rank's level-0 and level-1 sub-lattices must hold an even
# number of 4x4 blocks in each direction
RAGGED_RANKS = {64: '2x2', 96: '3x3', 128: '4x4', 192: '3x3', 256: '4x4'}


def cpu_ragged(fields):
    """The reference's solve of right-hand side 0 of each ragged field, one after the other
    (FP64, the same field, m0 and hierarchy parameters), each on its rank grid."""
    rows = [cpu_reference(f['theta'], f['b0'], f['N'], f['m0'], RAGGED_RANKS[f['N']]) for f in fields]
    if any('error' in r for r in rows):
        return {'error': [r.get('error') for r in rows]}
    tot = sum(r['value'] for r in rows)
    return {'value': tot / len(rows), 'unit': 'ms/RHS', 'kind': 'reference',
            'cores': max(r['cores'] for r in rows),
            'sample': 'right-hand side 0 of each of the 8 ragged fields, solved one after the other by the '
                      'reference under mpirun on the rank grids ' +
                      ', '.join(f'{N}^2: {RAGGED_RANKS[N]}' for N in sorted({f['N'] for f in fields})) +
                      '; the value is the mean solve time per field (setup excluded)',
            'ms_total': tot, 'per_field_ms': [r['value'] for r in rows],
            'iterations': [r['iterations'] for r in rows]}


if __name__ == '__main__':
    main()

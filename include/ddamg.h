/* C ABI of the MI355X DD-alphaAMG library (libddamg.so).
 *
 * The boundary sits under the reference's C++ host API: each entry point below
 * replaces the reference function or class method cited next to it (paths relative to
 * the reference root).  Plain pointers and sizes only; no exceptions cross it.
 *
 * Data format at the boundary: complex128 (interleaved re, im) vectors in the
 * reference's interior site order without halos, x outer, t inner, then the degree of
 * freedom (dof = colour*2 + spin):  v[((x*Nt + t)*dof + d)*2 + {0:re, 1:im}].
 * A batch of nrhs vectors is stored back to back.  The gauge field is U_mu(x) with
 * mu = 0 the t direction: U[((x*Nt + t)*2 + mu)*2 + {re, im}].
 * Pointers are host pointers unless the entry point's name ends in _dev (device
 * pointers on the context's device, e.g. torch tensors).
 *
 * Every function returns a status code (DD_OK = 0).
 */
#ifndef DDAMG_H
#define DDAMG_H

#ifdef __cplusplus
extern "C" {
#endif

#define DD_OK 0
#define DD_ERR_NOT_CONVERGED 1 /* solve ended without reaching the tolerance */
#define DD_ERR_BAD_SHAPE 2     /* sizes that do not divide / unsupported shapes */
#define DD_ERR_HIP 3           /* HIP runtime error (details on stderr) */
#define DD_ERR_STATE 4         /* call out of order (e.g. solve before setup) */
#define DD_ERR_ARG 5           /* null pointer or invalid argument */

#define DD_MAX_LEVELS 8

typedef struct dd_ctx dd_ctx;

/* Solver parameters.  Defaults (dd_default_params) are the reference's hard-coded
 * knobs, src/variables.cpp:90-134, and its per-level parameters file,
 * include/params.h:7-57 (here in SITES per aggregate / SAP block, not block counts). */
typedef struct dd_params {
    int size;                      /* sizeof(dd_params) in the caller's build; set by
                                      dd_default_params.  Every entry point that takes
                                      a dd_params rejects (DD_ERR_ARG) a size other than
                                      the library's, so a caller compiled against
                                      another version of this header fails loudly
                                      instead of reading or writing past its struct. */
    int levels;                    /* number of levels L (>= 2)                     */
    int agg_x[DD_MAX_LEVELS];      /* aggregate size in sites at level l < L-1      */
    int agg_t[DD_MAX_LEVELS];
    int sap_x[DD_MAX_LEVELS];      /* SAP block size in sites at level l < L-1      */
    int sap_t[DD_MAX_LEVELS];
    int nv[DD_MAX_LEVELS];         /* test vectors building level l+1               */
    double m0;                     /* bare mass                                     */
    int nu1, nu2;                  /* pre/post smoothing SAP iterations (0, 2)      */
    int setup_sap_iters;           /* SAP iterations smoothing test vectors (4)     */
    int sap_m, sap_restarts;       /* block GMRES restart length / restarts (5, 5)  */
    double sap_tol;                /* block GMRES tolerance (1e-3)                  */
    int coarse_m, coarse_restarts; /* coarsest GMRES (400, 10)                      */
    double coarse_tol;             /* (0.1)                                         */
    int kcycle_m, kcycle_restarts; /* K-cycle FGMRES (5, 2)                         */
    double kcycle_tol;             /* (0.1)                                         */
    int fgmres_m, fgmres_restarts; /* outer FGMRES (20, 50)                         */
    double tol;                    /* outer tolerance (1e-10)                       */
    int cycle;                     /* 0 = V-cycle, 1 = K-cycle                      */
    int precision;                 /* 0 = FP64 throughout, 1 = FP32 hierarchy inside
                                      the FP64 outer FGMRES (mixed)                 */
    int rng_ranks_x, rng_ranks_t;  /* virtual MPI grid whose per-rank std::mt19937
                                      streams draw the level-0 test vectors (2, 2)  */
    int ortho_outer;               /* 0 = modified Gram-Schmidt (reference), 1 = CGS2 */
    int ortho_inner;               /* same, for the K-cycle FGMRES                  */
    int ortho_coarse;              /* coarsest GMRES: 0, 1 as above (launch per step),
                                      2 = CGS with delayed re-orthogonalisation in one
                                      persistent kernel, 3 = the faster of 1 and 2,
                                      timed at setup (default; 2 without autotune)  */
    int max_rhs;                   /* largest batch solve() will be asked for (1)   */
    int autotune;                  /* 1: time the coarse-SAP kernel variants on the
                                      built hierarchy at setup and keep the fastest  */
    int sap_variant;               /* coarse-SAP kernel variant for every coarse level
                                      (2*systems_per_workgroup + batched_link_loads;
                                      32 = matrix cores, FP32; 33 = matrix cores on
                                      the binary16 link copy); -1: autotuned/default */
    int coarse_basis;              /* storage of the persistent coarsest GMRES's Krylov
                                      basis in the FP32 hierarchy: 0 = FP32, 1 = IEEE
                                      binary16 (compressed basis, products in FP32;
                                      default).  Ignored by the FP64 hierarchy.     */
    int coarse_sap_links;          /* links read by the matrix-core coarse SAP of the
                                      FP32 hierarchy: 0 = the FP32 links (default),
                                      1 = a binary16 copy made at setup (each element
                                      within 2^-11 relative; the level operator keeps
                                      FP32).  Ignored by the FP64 hierarchy.         */
    int bootstrap;                 /* bootstrap passes after the setup (0, the
                                      reference's behaviour): each replaces every
                                      level's test vectors w by one cycle's solution of
                                      D x = w - D w, re-orthonormalises them and
                                      rebuilds the coarse operators.  Needs max_rhs
                                      systems of room per cycle (batches of max_rhs). */
    int graph;                     /* 1: with cycle = 0, solve() captures the V-cycle
                                      (SAP, transfers, persistent coarsest GMRES) as a
                                      HIP graph per batch mask and replays it; 0 (the
                                      default) runs it as separate launches.        */
    int sap_fused;                 /* 1 (default): on coarse levels smoothed on the matrix
                                      cores, each SAP iteration (red blocks, black blocks,
                                      residual norm, stopping test) is one launch whose
                                      blocks form their right-hand sides v - D x from the
                                      iterate; 0: separate launches with full-lattice
                                      residuals between the colours.                */
    int coarse_poly;               /* degree d of a GMRES polynomial right
                                      preconditioner of the coarsest GMRES: its roots
                                      are the harmonic Ritz values of d Arnoldi steps
                                      on the coarsest operator (a fixed start vector),
                                      computed when that operator is built; the solve
                                      then runs on D_c p(D_c), which needs far fewer
                                      Arnoldi steps to the same tolerance.  0 (the
                                      default): the reference's plain GMRES.        */
} dd_params;

typedef struct dd_solve_info {
    int iterations;     /* reference count convention k*m + maxIt (max over the batch) */
    double relres;      /* max over the batch of ||b - D x|| / ||b|| (FP64 D)          */
    double seconds;     /* wall time of the solve                                      */
} dd_solve_info;

/* Fills p with the reference defaults for a `levels`-level hierarchy with 4x4
 * aggregates, 4x4 SAP blocks and Nv = 8.  `size` is sizeof(dd_params) as the caller
 * compiled it: on a mismatch nothing is written and DD_ERR_ARG is returned.
 * dd_default_params() passes it. */
int dd_default_params_sized(dd_params* p, int size, int levels);
static inline int dd_default_params(dd_params* p, int levels) {
    return dd_default_params_sized(p, (int)sizeof(dd_params), levels);
}
/* sizeof(dd_params) of the library (bindings check their mirror of the struct). */
int dd_params_size(void);

/* Releases the library's process-wide HIP objects (the kernel timers' events).  Call
 * after the last dd_destroy and before the HIP runtime shuts down (the Python binding
 * registers it with atexit); later calls re-create what they need. */
int dd_finalize(void);

/* Context for an Nx x Nt lattice on HIP device `device`.  Replaces the global set-up
 * of src/main.cpp:57-72 (allocate_lattice_arrays, initializeMPI, boundaries). */
int dd_create(dd_ctx** ctx, int Nx, int Nt, int device);
int dd_destroy(dd_ctx* ctx);

/* Stream ordering of the _dev entry points.  Each _dev call first makes the context's
 * own stream wait for all work queued so far on `stream` (a hipStream_t; NULL, the
 * default, is the null stream, which is where torch queues work unless the caller
 * switched streams), so device inputs written there are complete; it returns after its
 * own work has finished.  No reference counterpart (the reference is host-only). */
int dd_set_stream(dd_ctx* ctx, void* stream);

/* Gauge field (complex128, layout above).  Replaces read_binary of the configuration
 * src/io.cpp:4-55 followed by the halo set-up of Level::makeDirac src/level.cpp:286-331. */
int dd_set_gauge(dd_ctx* ctx, const double* U);
/* Link phases theta_mu(x) (float64 [Nx][Nt][2]); U = exp(i theta). */
int dd_set_gauge_theta(dd_ctx* ctx, const double* theta);
/* Seeded quenched Metropolis field at coupling beta (cold start, `sweeps` sweeps,
 * proposal step `step` rad), generated on the device and installed as the gauge
 * field; theta_out (nullable) receives the phases.  No reference counterpart: the
 * reference reads external configurations (src/tests.cpp:556-566). */
int dd_generate_gauge(dd_ctx* ctx, double beta, int sweeps, unsigned long long seed, double step,
                      double* theta_out);
int dd_get_gauge(dd_ctx* ctx, double* U);

/* out = D in (dagger = 0) or D^dagger in (dagger = 1), FP64 (precision 0) or FP32
 * arithmetic (precision 1), for nrhs level-0 vectors.
 * Replaces D_phi / D_dagger_phi, src/dirac_operator.cpp:24-85. */
int dd_apply_D(dd_ctx* ctx, const double* in, double* out, int nrhs, double m0, int dagger, int precision);
int dd_apply_D_dev(dd_ctx* ctx, const double* in, double* out, int nrhs, double m0, int dagger, int precision);

/* Builds the hierarchy: AlgebraicMG(U, m0, nu1, nu2) + AlgebraicMG::setUpPhase,
 * include/amg.h:96-114, src/amg.cpp:4-62 (and the FGMRES_AMG constructor that runs it,
 * include/amg.h:150-165). */
int dd_setup(dd_ctx* ctx, const dd_params* p);
double dd_setup_seconds(dd_ctx* ctx);

/* The setup phase step by step, for callers that build the hierarchy themselves as the
 * reference's Level objects allow (include/level.h:147-233, 334-467):
 * dd_setup_storage allocates the hierarchy of p (test vectors and coarse operators
 * zero, no setup phase run); dd_set_test_vectors installs level l's test vectors
 * (complex128 [nv][V][dof], the layout of dd_get_test_vectors); dd_orthonormalize
 * orthonormalises them per aggregate and spin (Level::orthonormalize,
 * src/level.cpp:193-284); dd_make_coarse_links builds level l+1's operator as the
 * Galerkin product of level l's operator with them (Level::makeCoarseLinks,
 * src/level.cpp:431-542; level 0's operator is the Wilson-Dirac operator of the gauge
 * field, Level::makeDirac, src/level.cpp:286-331).  precision selects the FP64 (0) or
 * FP32 (1) hierarchy, as p->precision made it. */
int dd_setup_storage(dd_ctx* ctx, const dd_params* p);
int dd_set_test_vectors(dd_ctx* ctx, int level, const double* tv, int precision);
int dd_orthonormalize(dd_ctx* ctx, int level, int precision);
int dd_make_coarse_links(dd_ctx* ctx, int level, int precision);

/* Solves D x = b for nrhs right-hand sides (x0 = 0) with FGMRES preconditioned by the
 * K- or V-cycle.  Replaces FGMRES_AMG_{k,v}_cycle::fgmres, include/amg.h:240-270 and
 * src/fgmres.cpp:5-97.  iters (nullable, nrhs ints) receives each system's count.
 * Returns DD_ERR_NOT_CONVERGED if any system missed the tolerance. */
int dd_solve(dd_ctx* ctx, int nrhs, const double* b, double* x, int* iters, dd_solve_info* info);
int dd_solve_dev(dd_ctx* ctx, int nrhs, const double* b, double* x, int* iters, dd_solve_info* info);

/* Ragged batch: solves the systems of n contexts (each its own lattice, gauge field,
 * hierarchy and nrhs[i] right-hand sides; device pointers b[i], x[i]) at the same time,
 * each context's solve on its own stream from its own host thread.  The reference has no
 * batched or multi-lattice solve.  info (nullable) receives one entry per context.  The
 * kernel timers are not thread-safe: leave them off during this call. */
int dd_solve_many_dev(int n, dd_ctx* const* ctx, const int* nrhs, const double* const* b, double* const* x,
                      dd_solve_info* info);

/* Same with the initial guess taken from x (x0 = x on entry; the reference's
 * fgmres(phi, x0, x) signature, src/fgmres.cpp:5-10). */
int dd_solve_x0(dd_ctx* ctx, int nrhs, const double* b, double* x, int* iters, dd_solve_info* info);

/* Comparison solvers of the reference on the FP64 Wilson-Dirac operator D(m0) of the
 * context's gauge field (no dd_setup needed), batched over nrhs <= 64 independent
 * systems (each follows the reference's recurrence and stopping test).  info->relres is
 * the final relative residual of the system solved.  Returns DD_ERR_NOT_CONVERGED if any
 * system missed its tolerance.
 *
 * Conjugate gradient on the normal equations D^dagger D x = phi from x0 = phi, stopping
 * when ||phi - D^dagger D x|| < tol ||phi||.  Replaces conjugate_gradient
 * (src/conjugate_gradient.cpp:4-84); Methods::CG (src/methods.cpp:43-59) passes
 * phi = D^dagger b (dd_apply_D with dagger = 1).  A system that does not converge
 * reports max_iter (the reference returns 0). */
int dd_solve_cg(dd_ctx* ctx, int nrhs, const double* phi, double* x, double m0, int max_iter, double tol,
                int* iters, dd_solve_info* info);
/* BiCGStab on D x = b from x0 (nullable: zero).  Replaces bi_cgstab
 * (src/conjugate_gradient.cpp:87-176) as Methods::BiCG calls it (src/methods.cpp:3-19);
 * a system that does not converge reports max_iter. */
int dd_solve_bicgstab(dd_ctx* ctx, int nrhs, const double* b, const double* x0, double* x, double m0, int max_iter,
                      double tol, int* iters, dd_solve_info* info);
/* GMRES(m) with `restarts` cycles on D x = b from x0 (nullable: zero), count k*m + maxIt.
 * sap_bx = 0: no preconditioner, replaces FGMRES_fine_level (include/fgmres.h:153-234,
 * Methods::GMRES src/methods.cpp:21-41).
 * sap_bx > 0: flexible GMRES preconditioned by one red-black SAP iteration from a zero
 * guess over blocks of sap_bx x sap_bt sites (a power of two, at most 64; blocks solved by
 * GMRES(5) x 5 to 1e-3), replaces FGMRES_SAP (include/sap.h:236-333, Methods::FGMRES_sap
 * src/methods.cpp:83-104).  The reference's FGMRES_SAP uses 4 x 4 blocks per MPI rank
 * (SAPV::sap_block_x/t): sap_bx = Nx / (4 ranks_x), sap_bt = Nt / (4 ranks_t). */
int dd_solve_gmres(dd_ctx* ctx, int nrhs, const double* b, const double* x0, double* x, double m0, int m,
                   int restarts, double tol, int sap_bx, int sap_bt, int* iters, dd_solve_info* info);

/* Red-black SAP as a stand-alone solver on the FP64 D(m0) of the context's gauge field
 * (no dd_setup needed): x (in: initial guess, out: result) after at most nu iterations,
 * each system stopping once ||v - D x|| < tol ||v||; blocks of sap_bx x sap_bt sites (a
 * power of two, at most 64, an even number of blocks), each solved by GMRES(5) x 5 to
 * 1e-3.  iters[r] is SAP_C::SAP's return value (the 0-based iteration at which the test
 * held, or nu); relres (nullable) the final ||v - D x|| / ||v||.  Replaces
 * SAP_fine_level::SAP, include/sap.h:192-234 and src/sap.cpp:81-165, as Methods::SAP
 * calls it (src/methods.cpp:61-78; there SAPV-style blocks of Nx/(ranks_x*xblocks) sites).
 * Returns DD_ERR_NOT_CONVERGED if a system did not reach tol. */
int dd_sap_fine(dd_ctx* ctx, int nrhs, const double* v, double* x, int nu, double tol, int sap_bx, int sap_bt,
                double m0, int* iters, double* relres);

/* Adjoint solve D^dagger y = eta on the same hierarchy through sigma3-hermiticity:
 * y = sigma3 D^-1 sigma3 eta.  The reference has D_dagger_phi only
 * (src/dirac_operator.cpp:56-85); no adjoint solver. */
int dd_solve_adjoint(dd_ctx* ctx, int nrhs, const double* eta, double* y, int* iters, dd_solve_info* info);
int dd_solve_adjoint_dev(dd_ctx* ctx, int nrhs, const double* eta, double* y, int* iters, dd_solve_info* info);

/* out = sigma3 in (the spin-1 components negated) for nrhs level-0 vectors: the
 * gamma5 of the reference's D^dagger = sigma3 D sigma3 (src/dirac_operator.cpp:56-85). */
int dd_apply_sigma3(dd_ctx* ctx, const double* in, double* out, int nrhs);

/* out[r] = sum_i a_r[i] conj(b_r[i]) (complex128, out as re, im pairs) for nrhs pairs of
 * vectors of n complex entries each, stored back to back; FP64, summed in a fixed order
 * (deterministic).  Replaces dot(A, B) of include/utils.h:68-84 (and the FGMRES dot
 * hooks, include/fgmres.h:97-98).  dd_dot_dev: device pointers, out on the device. */
int dd_dot(dd_ctx* ctx, long long n, int nrhs, const double* a, const double* b, double* out);
int dd_dot_dev(dd_ctx* ctx, long long n, int nrhs, const double* a, const double* b, double* out);
/* dd_dot_dev for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops); no counterpart in the reference.  Same result as
 * dd_dot_dev, bit for bit.  It allocates nothing: its partial sums live in a fixed
 * device array of the library (256 KiB per GPU, kept until the process ends), so nrhs
 * must be at most 64 (DD_ERR_ARG otherwise).  The kernels go to the stream set with
 * dd_set_stream (the null stream by default) and the call returns without waiting:
 * the caller synchronises before reading `out`.  Calls that overlap in time on one GPU
 * share that array and race. */
int dd_dot_dev_noalloc(dd_ctx* ctx, long long n, int nrhs, const double* a, const double* b, double* out);

/* grad[x][t][mu] += scale * sum_r Re[ y_r^H (dD/dtheta_mu(x)) x_r ] for U = e^{i theta};
 * with scale = -1 this is dL/dtheta of L = Re<eta, D^-1 b> for y = D^-dagger eta,
 * x = D^-1 b.  No reference counterpart. */
int dd_dD_dtheta(dd_ctx* ctx, int nrhs, const double* y, const double* x, double* grad, double scale);

/* ---- per-component entry points (tests, benchmarks) ---------------------------------
 * Level-l vectors have dof = 2 (l = 0) or 2*nv[l-1]. */
/* One K-cycle (cycle = 1) or V-cycle (cycle = 0) from `level` on psi (in/out):
 * AlgebraicMG::k_cycle / v_cycle, src/amg.cpp:64-174. */
int dd_cycle(dd_ctx* ctx, int level, int cycle, const double* eta, double* psi, int nrhs, int precision);
/* Level::D_operator, src/level.cpp:394-426 (level 0: the Wilson stencil). */
int dd_level_apply(dd_ctx* ctx, int level, const double* in, double* out, int nrhs, int precision);
/* Level::Pdagg_v / Level::P_vc, src/level.cpp:145-189 / :92-141 (level -> level+1). */
int dd_restrict(dd_ctx* ctx, int level, const double* in, double* out, int nrhs, int precision);
int dd_prolong(dd_ctx* ctx, int level, const double* in, double* out, int nrhs, int precision);
/* SAP_C::SAP(v, x, nu, tol), src/sap.cpp:81-165, on level `level` (x in/out). */
int dd_sap(dd_ctx* ctx, int level, const double* v, double* x, int nu, double tol, int nrhs, int precision);
/* Test vectors of level l (after setup): out[nv][V][dof]. */
int dd_get_test_vectors(dd_ctx* ctx, int level, double* out, int precision);
/* Links of level l as [V][5][dof][dof] row-major (A, B_0, B_1, C_0, C_1; at level 0 the
   Wilson-Dirac ones, A = 0, B_mu = (1 - sigma_mu) U_mu / 2, C_mu = (1 + sigma_mu) U*_mu / 2);
   precision 0 FP64, 1 FP32, 2 the FP32 hierarchy's binary16 copy read by the
   matrix-core coarse SAP (coarse_sap_links = 1). */
int dd_get_links(dd_ctx* ctx, int level, double* out, int precision);
/* Autotuner result for level l >= 1: chosen coarse-SAP kernel variant
 * (2*systems_per_workgroup + batched_link_loads) and its measured half-sweep time. */
int dd_get_tuning(dd_ctx* ctx, int level, int* variant, double* ms);
/* Coarsest-solver statistics since setup (no reference counterpart; diagnostics):
 * out[0] coarsest solves, out[1] Arnoldi steps (max over the batch per step), out[2..10]
 * milliseconds spent by the persistent kernel's first work-group (wave 0) in: reduction
 * pass, publication of its partial sums (its other waves, its stores), per-system scalar
 * phase, wait for the owners' coefficient flags, basis update, publication of the update,
 * cycle end, wait for the neighbours' flags and stencil, gathering of the partial sums
 * (with the waits for them); out[11] = 1 when the persistent kernel is in use. */
int dd_coarsest_stats(dd_ctx* ctx, double* out);
/* Number of V-cycle graphs the hierarchy has captured (dd_params.graph = 1); no
 * reference counterpart (diagnostics). */
int dd_graph_count(dd_ctx* ctx, int* count);
/* Roots of the coarsest GMRES's polynomial preconditioner (dd_params.coarse_poly > 0),
 * in the order they are applied: roots[2i], roots[2i+1] = re, im of root i; *degree
 * receives their number (0 without the preconditioner).  No reference counterpart. */
int dd_get_coarse_poly(dd_ctx* ctx, int* degree, double* roots);
/* Host computation behind it, for any Arnoldi relation: from H ((d+1) x d complex,
 * row-major, H[2(i*d+j)] = re h_ij) the harmonic Ritz values in Leja order, as
 * roots[2i], roots[2i+1].  Needs no device.  No reference counterpart. */
int dd_poly_roots(const double* H, int d, double* roots);
/* Shape of level l: Nx, Nt, dof. */
int dd_level_shape(dd_ctx* ctx, int level, int* Nx, int* Nt, int* dof);
/* ---- live kernel timing -----------------------------------------------------------------
 * HIP events recorded on the library's stream around every launch of one kernel family,
 * for benchmarks (no reference counterpart; the reference times whole methods with
 * MPI_Wtime, src/methods.cpp:10-163).  Kernel ids: */
#define DD_K_WILSON 0    /* level-0 Wilson-Dirac stencil                 */
#define DD_K_COARSE 1    /* coarse-level stencil                         */
#define DD_K_SAP0 2      /* level-0 SAP block solves (one colour)         */
#define DD_K_SAPC 3      /* coarse-level SAP block solves (one colour)    */
#define DD_K_MULTIDOT 4  /* Gram-Schmidt multi-dot                        */
#define DD_K_MULTIAXPY 5 /* Gram-Schmidt multi-axpy                       */
#define DD_K_RESTRICT 6
#define DD_K_PROLONG 7
#define DD_K_COARSEST 8  /* coarsest-level GMRES (persistent kernel)      */
#define DD_K_DOT 9
int dd_timer_enable(int kernel, int on);
int dd_timer_reset(void);
/* total_ms: summed launch durations; bytes: summed algorithmic HBM bytes of the timed
 * launches (every operand counted once: links, test vectors and the vectors read or
 * written; re-reads from cache are not counted). */
int dd_timer_read(int kernel, double* total_ms, long long* launches, double* bytes);
/* Algorithmic floating-point operations of the timed launches of `kernel`, counted on the
 * device by the kernels themselves: the matrix-core products of the coarse SAP (DD_K_SAPC:
 * 128 dof^2 per site, hop and tile of 16 systems, for the block operators, the fused
 * right-hand sides and the norm stencil); 0 for the other families.  Synchronises. */
int dd_timer_flops(int kernel, double* flops);

/* Last error message (thread-unsafe, per process). */
const char* dd_last_error(void);

#ifdef __cplusplus
}
#endif
#endif

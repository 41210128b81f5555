// Host-side orchestration of the DD-alphaAMG solve on one GCD: device buffers, the
// Krylov engine, the multigrid hierarchy with its setup and cycles, and the top-level
// solver object behind the C ABI.
#pragma once
#include <map>
#include <tuple>
#include <complex>
#include <functional>
#include <memory>
#include <random>
#include <vector>

#include "../../include/ddamg.h"
#include "coarsest.h"
#include "devmem.h"
#include "kernels.h"

namespace dd {

inline uint64_t full_mask(int n) { return n >= 64 ? ~0ull : ((1ull << n) - 1ull); }
inline int mask_count(uint64_t m) { return m ? 64 - __builtin_clzll(m) : 0; }   // highest set bit + 1

// ---- Krylov engine -------------------------------------------------------------------
// Restates FGMRES::fgmres (src/fgmres.cpp:5-97) for a batch of independent systems:
// each system follows the reference's control flow (break on |g_{j+1}| < tol ||phi||,
// maxIt kept across restarts, true residual after each cycle, k*m + maxIt count),
// with Z_j persisting across restarts of one call.  Vectors of system r live at
// base + r*n.  Orthogonalisation: modified Gram-Schmidt as the reference, or
// classical Gram-Schmidt with one re-orthogonalisation (CGS2, two reductions per step).
template <typename R>
class Krylov {
   public:
    using Func = std::function<void(const cplx<R>* in, cplx<R>* out, uint64_t mask)>;
    using Prec = std::function<void(const cplx<R>* in, cplx<R>* inout, uint64_t mask)>;

    Krylov(hipStream_t s, size_t n, int m, int restarts, double tol, int cap, bool flexible, int ortho);
    // Returns a mask of the systems that converged; iters[r] (cap entries) per system.
    uint64_t solve(const cplx<R>* phi, const cplx<R>* x0, cplx<R>* x, uint64_t mask, const Func& A,
                   const Prec* M, int* iters, double* relres);
    // The same solve with x0 = phi and every control decision on the device, for a
    // captured graph: no host read-back.  The systems are those of mask that the calling
    // scope's device mask runs (DevMaskScope).  The Givens rotations, the break on
    // |g_{j+1}| < tol ||phi||, maxIt kept across restarts and the true-residual test of
    // each cycle are made per system by one-wave control kernels.  Every launch is masked
    // on the device by the cycle's live or the step's run systems, so all m * restarts
    // steps are issued and the systems that have stopped do nothing in them.
    void solve_dc(const cplx<R>* phi, cplx<R>* x, uint64_t mask, const Func& A, const Prec* M);
    // allocates the device control state of solve_dc (outside any graph capture)
    void alloc_ctl();
    size_t n() const { return n_; }
    long long steps = 0;   // Arnoldi steps taken (max over the batch per step), cumulative
    // restart length (at most the constructed one), cycles and tolerance of later solves
    void set_limits(int m, int restarts, double tol) {
        m_ = std::max(1, std::min(m, m_max_));
        restarts_ = restarts;
        tol_ = tol;
    }

   private:
    // orthogonalises w against V_0..V_j (H column j); returns the partial sums of |w|^2
    RedBuf ortho(int j, cplx<R>* w, uint64_t run, int nr);
    hipStream_t s_;
    size_t n_;
    int m_, restarts_, cap_, ortho_, m_max_;
    double tol_;
    bool flex_;
    DevMem V_, Z_, r_, w_, H_, nrm_, eta_, nit_, part_, ctl_;
    PinMem Hh_, nrmh_, etah_, nith_;
    RedBuf rb_;
};

// ---- hierarchy -------------------------------------------------------------------------
struct LevelGeom {
    int Nx = 0, Nt = 0, dof = 0;
    int ax = 1, at = 1, nv = 0;   // aggregation to the next level
    int bx = 1, bt = 1;           // SAP blocks
    size_t n() const { return (size_t)Nx * Nt * dof; }
    int V() const { return Nx * Nt; }
};

std::vector<LevelGeom> make_geometry(int Nx, int Nt, const dd_params& p);

// Level-0 test vectors drawn as AlgebraicMG::setUpPhase does (src/amg.cpp:6-21, 33-39):
// one std::mt19937(rank) per virtual rank of an Rx x Rt grid, loop cc -> local x ->
// local t -> dof, each entry taking two uniform(-1, 1) draws.  The reference writes
// `d(g) + I*d(g)`, whose operands g++ evaluates right to left: the first draw is the
// imaginary part.
class RankStreams {
   public:
    RankStreams(int Rx, int Rt);
    void fill(std::complex<double>* out, int first, int count, int Nx, int Nt, int dof);   // out[cc][V][dof]

   private:
    int Rx_, Rt_;
    std::vector<std::mt19937> g_;
};

template <typename R>
class Multigrid {
   public:
    Multigrid(hipStream_t s, int Nx, int Nt, const dd_params& p, const cplx<R>* U);
    // setup phase from raw level-0 test vectors [nv0][V0][2] (device, precision R)
    void setup(const cplx<R>* tv0, RankStreams& rng);
    // bootstrap passes after setup: per level l (finest first) every test vector w (as
    // it was before its orthonormalisation) is replaced by the current method's
    // approximate solution of D x = w - D w (one cycle from level l, zero initial guess),
    // the prolongator is rebuilt from the new vectors and the coarse operators of levels
    // l+1 .. from the prolongators.  Needs a setup with p.bootstrap > 0 (which keeps
    // those vectors).
    void bootstrap(int passes);
    // the per-level steps of the setup phase, for callers that build the hierarchy step
    // by step (Level::orthonormalize / makeCoarseLinks, src/level.cpp:193-284, 431-542):
    // zero every level's test vectors and coarse operators; install level l's test
    // vectors from tv [nv][V][dof]; orthonormalise them per aggregate and spin; build
    // level l+1's operator as the Galerkin product of level l's
    void clear_storage();
    void set_test_vectors(int l, const cplx<R>* tv);
    void orthonormalize_level(int l);
    void galerkin_level(int l);
    // level-0 link matrices of the Wilson-Dirac operator in the coarse-link layout
    // [V][5][2][2] (A = 0), as Level::makeDirac builds them (src/level.cpp:286-331)
    void dirac_links_level0(cplx<R>* out);
    void kcycle(int l, const cplx<R>* eta, cplx<R>* psi, uint64_t mask, bool zero_guess = false);
    // zero_guess (both cycles): the initial guess is zero and psi need not be set; psi is
    // zeroed only where it is read before being written (the coarsest solve, the
    // pre-smoothing), and without pre-smoothing the residual before the restriction is
    // eta itself and the prolongation writes psi = P p1
    void vcycle(int l, const cplx<R>* eta, cplx<R>* psi, uint64_t mask, bool zero_guess = false);
    void coarsest(const cplx<R>* eta, cplx<R>* psi, uint64_t mask);
    // psi = V-cycle(eta) from level 0 with zero initial guess, replayed from a HIP graph
    // captured on the first call for these buffers and this mask.  false (nothing done)
    // when the cycle cannot be captured (no persistent coarsest solver, kernel timers on,
    // or the runtime refused the capture once); the caller then runs it eagerly.
    bool vcycle_graph(const cplx<R>* eta, cplx<R>* psi, uint64_t mask);
    // the same for the K-cycle: its FGMRES solves at levels 1 .. L-2 run device-controlled
    // (Krylov::solve_dc) inside the capture
    bool kcycle_graph(const cplx<R>* eta, cplx<R>* psi, uint64_t mask);
    // reports (throws) a timed-out persistent coarsest solve among the graph replays
    void check_graph_failures();
    // reports (throws) a timed-out wait inside a fused SAP iteration since the last check
    void check_sap_failures();
    int graph_count() const { return (int)graphs_.size(); }
    // roots of the coarsest GMRES's polynomial preconditioner in their order of
    // application (empty without p.coarse_poly or before the coarsest operator exists)
    const std::vector<std::complex<double>>& poly_roots() const { return poly_theta_; }
    // out = p(D_c) in on the coarsest level for the systems of mask (in != out)
    void poly_apply(const cplx<R>* in, cplx<R>* out, uint64_t mask);
    ~Multigrid();
    void apply(int l, const cplx<R>* in, cplx<R>* out, const cplx<R>* b, uint64_t mask);
    void restrict_(int l, const cplx<R>* in, cplx<R>* out, uint64_t mask);
    void prolong(int l, const cplx<R>* vc, const cplx<R>* y, cplx<R>* out, uint64_t mask);
    void sap(int l, const cplx<R>* v, cplx<R>* x, int nu, double tol, uint64_t mask);
    const LevelGeom& geom(int l) const { return g_[l]; }
    int levels() const { return (int)g_.size(); }
    const cplx<R>* links(int l) const { return links_[l].template as<cplx<R>>(); }
    const cplx<R>* W(int l) const { return W_[l].template as<cplx<R>>(); }
    // binary16 copy of level l's links for the coarse SAP, or nullptr
    const _Float16* hlinks(int l) const { return hlinks_[l].template as<_Float16>(); }
    int cap(int l) const { return cap_[l]; }
    // times every compiled coarse-SAP variant on the built hierarchy (a batch of
    // max_rhs pseudo-random systems, one colour) and keeps the fastest per level
    void autotune();
    // uses coarse-SAP kernel variant v on every coarse level (BadShape if unavailable)
    void set_sap_variant(int v);
    void tuning(int l, int* variant, double* ms) const {
        *variant = (l >= 0 && l < (int)sapc_var_.size()) ? sapc_var_[l] : -1;
        *ms = (l >= 0 && l < (int)sapc_ms_.size()) ? sapc_ms_[l] : 0.0;
    }
    std::vector<long long> coarse_iters;   // coarsest GMRES iterations per call (stats)
    // Arnoldi steps of the coarsest solver, its per-phase clocks (persistent kernel only,
    // else nullptr) and the number of coarsest solves
    void coarsest_stats(const long long** steps, const double** clk, long long* calls) const {
        *steps = use_dev_ ? &coarse_dev_->steps : (coarse_ ? &coarse_->steps : nullptr);
        *clk = use_dev_ ? coarse_dev_->phase_clk : nullptr;
        *calls = (long long)coarse_iters.size();
    }

   private:
    hipStream_t s_;
    dd_params p_;
    std::vector<LevelGeom> g_;
    std::vector<int> cap_;
    const cplx<R>* U_;
    std::vector<DevMem> links_, W_, hlinks_;
    std::vector<DevMem> tvraw_;   // smoothed test vectors [nv][n] before orthonormalisation (bootstrap)
    // per-level work: residual, restricted rhs / coarse solution of the next level,
    // prolongated correction, SAP residual and gates
    std::vector<DevMem> res_, eta_, psi_, sapr_, gate_, nrm_;
    std::vector<RedBuf> rb_;
    std::vector<DevMem> part_;
    std::vector<std::unique_ptr<Krylov<R>>> kry_;   // K-cycle FGMRES of level l (index l)
    std::unique_ptr<Krylov<R>> coarse_;
    std::unique_ptr<CoarsestGmres<R>> coarse_dev_;   // persistent kernel (ortho_coarse 2, 3)
    bool use_dev_ = false;                           // coarsest solves run on coarse_dev_
    double coarse_ms_[2] = {0, 0};                   // setup timing: launch-per-step, persistent
    std::vector<int> iters_scratch_;
    std::vector<int> sapc_var_;
    std::vector<double> sapc_ms_;
    bool capturing_ = false, graphs_ok_ = true;
    bool dc_ = false;   // K-cycle FGMRES solves device-controlled (set while a K-cycle is captured)
    bool capture_(const cplx<R>* eta, cplx<R>* psi, uint64_t mask, bool kcyc);
    // fused SAP iterations (sap_fused_iteration) per level: block epochs, counters, norm
    // partials, launch epoch, and whether the level's shape allows them
    std::vector<DevMem> fuse_done_, fuse_ctr_, fuse_part_;
    DevMem fuse_fail_;
    std::vector<unsigned> fuse_epoch_;
    std::vector<char> fuse_ok_;
    int fuse_slices_ = 256;
    std::map<std::tuple<const void*, const void*, uint64_t, bool>, hipGraphExec_t> graphs_;
    // GMRES polynomial preconditioner of the coarsest solve (p.coarse_poly = degree):
    // roots, per-root lincomb coefficients [root][3 forms][64 systems][3], work vectors
    std::vector<std::complex<double>> poly_theta_;
    DevMem poly_coef_, poly_r_, poly_t_;
    // roots from p.coarse_poly Arnoldi steps on the coarsest operator as it is now
    void poly_build();
};

// ---- top-level solver --------------------------------------------------------------------
class Solver {
   public:
    Solver(int Nx, int Nt, int device);
    ~Solver();
    void set_gauge(const cplx<double>* U_dev);
    void setup(const dd_params& p);
    // allocates the hierarchy of p without running the setup phase (test vectors and
    // coarse operators zero); the caller fills it with the Multigrid per-level steps
    void setup_storage(const dd_params& p);
    // b, x: device pointers, nrhs systems of V*2 entries back to back
    // use_x0: x holds the initial guess on entry (otherwise x0 = 0)
    int solve(int nrhs, const cplx<double>* b, cplx<double>* x, int* iters, dd_solve_info* info,
              bool use_x0 = false);
    void apply_D(const cplx<double>* in, cplx<double>* out, int nrhs, double m0, bool dagger, int precision);
    // Comparison solvers of the reference on the level-0 FP64 operator (methods.hip), each
    // batched over nrhs independent systems in lock-step; device pointers, iters[r] per
    // system in the reference's count convention, DD_ERR_NOT_CONVERGED if any system
    // did not reach its tolerance.
    //   solve_cg:        conjugate_gradient on D^dagger D x = phi from x0 = phi
    //                    (src/conjugate_gradient.cpp:4-84); relres of the normal equations
    //   solve_bicgstab:  bi_cgstab on D x = b from x0 (src/conjugate_gradient.cpp:87-176)
    //   solve_gmres:     FGMRES with the identity (FGMRES_fine_level, include/fgmres.h:153-234)
    //                    or, with sap_bx > 0, one level-0 SAP iteration over sap_bx x sap_bt
    //                    site blocks from a zero guess as the flexible preconditioner
    //                    (FGMRES_SAP, include/sap.h:236-333)
    int solve_cg(int nrhs, const cplx<double>* phi, cplx<double>* x, double m0, int max_iter, double tol, int* iters,
                 dd_solve_info* info);
    int solve_bicgstab(int nrhs, const cplx<double>* b, const cplx<double>* x0, cplx<double>* x, double m0,
                       int max_iter, double tol, int* iters, dd_solve_info* info);
    int solve_gmres(int nrhs, const cplx<double>* b, const cplx<double>* x0, cplx<double>* x, double m0, int m,
                    int restarts, double tol, int sap_bx, int sap_bt, int* iters, dd_solve_info* info);
    //   sap_fine:        SAP_fine_level::SAP(v, x, nu, tol) as a stand-alone solver on
    //                    sap_bx x sap_bt site blocks (src/sap.cpp:81-165), x in/out;
    //                    iters[r] is the reference's return value, relres[r] the last
    //                    ||v - D x|| / ||v||
    int sap_fine(int nrhs, const cplx<double>* v, cplx<double>* x, int nu, double tol, int sap_bx, int sap_bt,
                 double m0, int* iters, double* relres);
    int Nx, Nt, device;
    hipStream_t s = nullptr;
    DevMem U64, U32, theta;
    dd_params prm{};
    bool have_gauge = false, have_setup = false;
    double setup_seconds = 0;
    std::unique_ptr<Multigrid<double>> mg64;
    std::unique_ptr<Multigrid<float>> mg32;

   private:
    std::unique_ptr<Krylov<double>> outer_;
    DevMem x32_, y32_, x64_, y64_;
    int outer_cap_ = 0;
};

}  // namespace dd

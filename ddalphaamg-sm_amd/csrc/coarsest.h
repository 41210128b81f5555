// Coarsest-level GMRES as one persistent cooperative kernel (coarsest.hip).
//
// Solves D_c psi = eta on the coarsest level for a batch of systems with the GMRES of
// GMRES_level_l (include/level.h:70-140; FGMRES::fgmres src/fgmres.cpp:5-127 with the
// identity preconditioner): x0 = eta (src/amg.cpp:69,124), restart length m, `restarts`
// cycles, break on |g_{j+1}| < tol ||eta||, maxIt kept across restarts, true residual
// after each cycle, count k*m + maxIt.  The Arnoldi basis is orthogonalised by classical
// Gram-Schmidt with delayed re-orthogonalisation (DCGS2: one global reduction and two
// passes over the basis per step); the Hessenberg matrix, Givens rotations and the
// triangular solve run on the device, so one launch performs the whole solve.
#pragma once
#include <complex>
#include <vector>

#include "devmem.h"
#include "kernels.h"

namespace dd {

template <typename R>
class CoarsestGmres {
   public:
    // half_basis (FP32 engine only): the Krylov basis is stored in IEEE binary16
    CoarsestGmres(hipStream_t s, int Nx, int Nt, int dof, int m, int restarts, double tol, int cap,
                  bool half_basis = false);
    // false when the lattice is too large for one co-resident grid (the caller then
    // uses the launch-per-step Krylov engine)
    bool supported() const { return ok_; }
    // GMRES polynomial right preconditioner p(D) = (1 - prod_i (1 - D / theta_i)) / D of
    // degree theta.size() (at most 32; empty: none): later solves run the Arnoldi process
    // on D p(D) and return x0 + p(D) Q y.  The roots are applied in the given order.
    void set_poly(const std::vector<std::complex<double>>& theta);
    // psi = D_c^-1 eta for the systems in mask (x0 = eta); iters[r] per system
    void solve(const cplx<R>* L, double m0, const cplx<R>* eta, cplx<R>* psi, uint64_t mask, int* iters);
    // deferred: solve() only enqueues device work (no host read-back, no stream
    // synchronisation; iters and the statistics are not updated), so that it can be
    // captured in a graph.  A timed-out wait is then reported by check_failure().
    bool deferred = false;
    // throws HipFailure when any launch since the last call timed out (synchronises)
    void check_failure();
    long long steps = 0;   // Arnoldi steps taken (cumulative)
    // restart length (at most the constructed one), cycles and tolerance of later solves
    void set_limits(int m, int restarts, double tol) {
        m_ = std::max(1, std::min(m, m_max_));
        restarts_ = restarts;
        tol_ = tol;
    }
    // wall-clock ticks of work-group 0 per phase (cumulative): reduction, publication of
    // the partial sums, scalar phase (after the totals), wait for the coefficient flags,
    // update, publication of the update, cycle end, neighbour wait and stencil, totals
    // (gather with its waits) of the scalar phase
    double phase_clk[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};

   private:
    template <typename S>
    void run(const cplx<R>* L, double m0, const cplx<R>* eta, cplx<R>* psi, uint64_t mask, int nsys, int* iters);
    hipStream_t s_;
    int Nx_, Nt_, dof_, m_, restarts_, cap_, G_ = 0, spw_ = 0, m_max_ = 0;
    double tol_;
    bool half_ = false, ok_ = false;
    DevMem Q_, u_, w_, part_, Hrot_, etals_, coef_, nrm_, ctrl_, bar_, wgclk_;
    DevMem sticky_;   // 1 once a launch's fail flag was set (kept across launches)
    DevMem pc_, z_;   // polynomial preconditioner: 1/theta_i, two stage vectors per system
    int npoly_ = 0;
    long long solves_ = 0;
    PinMem ctrlh_;
};

}  // namespace dd

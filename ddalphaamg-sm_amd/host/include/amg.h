// Algebraic multigrid and the AMG-preconditioned FGMRES of the reference
// (include/amg.h:6-270), backed by the GPU hierarchy of libddamg.so.
#ifndef AMG_H
#define AMG_H
#include "level.h"

struct dd_ctx;

class AlgebraicMG {
   public:
    // K-cycle FGMRES of level l (include/amg.h:16-94): the reference's host FGMRES over
    // the level operator, preconditioned by the K-cycle from level l.  k_cycle() itself
    // runs the whole cycle on the device; this class is the hook-level form of it.
    class FGMRES_k_cycle : public FGMRES {
       public:
        FGMRES_k_cycle(const int& m, const int& restarts, const double& tol, AlgebraicMG* parent, int l);
        ~FGMRES_k_cycle() {}

       private:
        AlgebraicMG* parent;
        int l;
        void func(const spinor& in, spinor& out) override { parent->levels[l]->D_operator(in, out); }
        void preconditioner(const spinor& in, spinor& out) override;
        c_double dot(c_double* A, c_double* B) override;
        void axpy(const spinor& X, const spinor& Y, const c_double& lambda, spinor& out) override;
        void scal(const c_double& lambda, const spinor& X, spinor& Y) override;
    };

    AlgebraicMG(const spinor& U, const double& m0, const int& nu1, const int& nu2);
    ~AlgebraicMG();
    void setUpPhase(const int& Nit);                                     // src/amg.cpp:4-62
    void testSetUp();                                                    // src/amg.cpp:249-298
    void testSAP();                                                      // src/amg.cpp:300-349
    void v_cycle(const int& l, const spinor& eta_l, spinor& psi_l);     // src/amg.cpp:64-116
    void k_cycle(const int& l, const spinor& eta_l, spinor& psi_l);     // src/amg.cpp:119-174
    void applyMultilevel(const int& it, const spinor& rhs, spinor& out, const double tol, const bool print_message);
    dd_ctx* context() const { return ctx; }
    // outer FGMRES knobs the solve uses (set by FGMRES_AMG)
    int fgmres_m = 20, fgmres_restarts = 50, cycle = 1;
    double fgmres_tol = 1e-10;

   private:
    dd_ctx* ctx;
    const spinor U;
    double m0;
    const int nu1, nu2;
    std::vector<Level*> levels;
    std::vector<FGMRES_k_cycle*> fgmres_k_cycle_l;
    bool setup_done = false;
};

// FGMRES with the AMG preconditioner (include/amg.h:147-270): the constructor runs the
// setup phase; fgmres() solves on the GPU (dd_solve_x0).  The hooks (D_phi and one cycle
// from level 0 with a zero initial guess, as FGMRES_AMG::preconditioner) are the
// hook-level form of the same solve.
class FGMRES_AMG : public FGMRES {
   public:
    FGMRES_AMG(const spinor& U, const int& m, const int& restarts, const double& tol, const int nu1, const int nu2,
               const double& m0, int cycle);
    ~FGMRES_AMG() {}
    int fgmres(const spinor& phi, const spinor& x0, spinor& x, const bool& print_message = false) override;
    AlgebraicMG amg;
    double last_relres = 0, last_seconds = 0;

   private:
    const spinor& U;
    const double m0;
    const int cyc;
    void func(const spinor& in, spinor& out) override { D_phi(U, in, out, m0); }
    void preconditioner(const spinor& in, spinor& out) override;
    c_double dot(c_double* A, c_double* B) override { return ::dot(A, B); }
    void axpy(const spinor& X, const spinor& Y, const c_double& lambda, spinor& out) override {
        ::axpy(X, Y, lambda, out);
    }
    void scal(const c_double& lambda, const spinor& X, spinor& Y) override { ::scal(lambda, X, Y); }
};

class FGMRES_AMG_k_cycle : public FGMRES_AMG {
   public:
    FGMRES_AMG_k_cycle(const spinor& U, const int& m, const int& restarts, const double& tol, const int nu1,
                       const int nu2, const double& m0)
        : FGMRES_AMG(U, m, restarts, tol, nu1, nu2, m0, 1) {}
};

class FGMRES_AMG_v_cycle : public FGMRES_AMG {
   public:
    FGMRES_AMG_v_cycle(const spinor& U, const int& m, const int& restarts, const double& tol, const int nu1,
                       const int nu2, const double& m0)
        : FGMRES_AMG(U, m, restarts, tol, nu1, nu2, m0, 0) {}
};

#endif

// Algebraic multigrid and the AMG-preconditioned FGMRES of the reference
// (include/amg.h:6-270), backed by the GPU hierarchy of libddamg.so.
#ifndef AMG_H
#define AMG_H
#include "level.h"

struct dd_ctx;

class AlgebraicMG {
   public:
    AlgebraicMG(const spinor& U, const double& m0, const int& nu1, const int& nu2);
    ~AlgebraicMG();
    void setUpPhase(const int& Nit);                                        // src/amg.cpp:4-62
    void testSetUp();                                                       // src/amg.cpp:249-298
    void testSAP();                                                         // src/amg.cpp:300-349
    void v_cycle(const int& l, const spinor& eta_l, spinor& psi_l);        // src/amg.cpp:64-116
    void k_cycle(const int& l, const spinor& eta_l, spinor& psi_l);        // src/amg.cpp:119-174
    void applyMultilevel(const int& it, const spinor& rhs, spinor& out, const double tol, const bool print_message);
    dd_ctx* context() const { return ctx; }
    // outer FGMRES knobs the solve uses (set by FGMRES_AMG)
    int fgmres_m = 20, fgmres_restarts = 50, cycle = 1;
    double fgmres_tol = 1e-10;

   private:
    dd_ctx* ctx;
    double m0;
    const int nu1, nu2;
    std::vector<Level*> levels;
    bool setup_done = false;
};

class FGMRES_AMG : public FGMRES {
   public:
    FGMRES_AMG(const spinor& U, const int& m, const int& restarts, const double& tol, const int nu1, const int nu2,
               const double& m0, int cycle);
    ~FGMRES_AMG() {}
    int fgmres(const spinor& phi, const spinor& x0, spinor& x, const bool& print_message = false) override;
    AlgebraicMG amg;
    double last_relres = 0, last_seconds = 0;
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

// The reference's comparison driver class (include/methods.h:8-62), declared here so that
// its implementation, src/methods.cpp, compiles unchanged against this host API: every
// method it calls (bi_cgstab, FGMRES_fine_level, conjugate_gradient, SAP_fine_level,
// FGMRES_SAP, AlgebraicMG, FGMRES_AMG_{k,v}_cycle, D_phi) runs on the GPU.
#ifndef METHODS_H
#define METHODS_H

#include "amg.h"
#include "conjugate_gradient.h"

class Methods {
   public:
    Methods(const spinor& U, const spinor& rhs, const spinor& x0, const double m0, const double tol)
        : U(U), rhs(rhs), x0(x0), m0(m0), tol(tol) {
        xBiCG = spinor(mpi::maxSizeH);
        xCG = spinor(mpi::maxSizeH);
        xGMRES = spinor(mpi::maxSizeH);
        xSAP = spinor(mpi::maxSizeH);
        xFGMRES_AMG_kcycle = spinor(mpi::maxSizeH);
        xFGMRES_AMG_vcycle = spinor(mpi::maxSizeH);
        xFGMRES_SAP = spinor(mpi::maxSizeH);
        xVcycle = spinor(mpi::maxSizeH);
        xKcycle = spinor(mpi::maxSizeH);
    }
    ~Methods() {}

    void BiCG(const int max_it, const bool print);
    void GMRES(const int len, const int restarts, const bool print);
    void CG(const bool print);
    void SAP(const int iterations, const int xblocks, const int tblocks, const bool print);
    void FGMRES_sap(const int len, const int restarts, const bool print);
    void FGMRES_amg_kcycle(const int nu1, const int nu2, const bool print);
    void FGMRES_amg_vcycle(const int nu1, const int nu2, const bool print);
    void Vcycle(const int iterations, const bool print);
    void Kcycle(const int iterations, const bool print);
    void check_solution(const spinor& x_sol);

    spinor xBiCG, xCG, xGMRES, xSAP, xFGMRES_AMG_kcycle, xFGMRES_AMG_vcycle, xFGMRES_SAP, xVcycle, xKcycle;

   private:
    const spinor U;
    const spinor rhs;
    const spinor x0;
    const double m0;
    const double tol;
    double start, end;
};

#endif

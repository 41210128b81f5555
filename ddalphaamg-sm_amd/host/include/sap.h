// FGMRES preconditioned by one level-0 SAP iteration (include/sap.h:236-333), run on the
// GPU (dd_solve_gmres with SAP blocks of SAPV::sap_block_x x SAPV::sap_block_t blocks per
// rank).
#ifndef SAP_H
#define SAP_H
#include "fgmres.h"

class FGMRES_SAP : public FGMRES {
   public:
    FGMRES_SAP(const int Ntot, const int dofs, const int maxSizeH, const int x_ini, const int t_ini, const int x_fin,
               const int t_fin, const int m, const int restarts, const double tol, const spinor& U, const double& m0)
        : FGMRES(Ntot, dofs, maxSizeH, x_ini, t_ini, x_fin, t_fin, m, restarts, tol), U(U), m0(m0) {}
    ~FGMRES_SAP() {}
    int fgmres(const spinor& phi, const spinor& x0, spinor& x, const bool& print_message = false) override;

   private:
    const spinor& U;
    const double m0;
};

#endif

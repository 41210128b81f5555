// FGMRES interface of the reference (include/fgmres.h:23-148).  The solver engine runs
// on the GPU for the operators the library provides (the AMG-preconditioned Wilson-Dirac
// operator, FGMRES_AMG_{k,v}_cycle in amg.h; plain GMRES below; FGMRES_SAP in sap.h);
// the constructors keep the reference's argument lists.
#ifndef FGMRES_H
#define FGMRES_H
#include <vector>

#include "dirac_operator.h"

typedef std::vector<c_double> c_vector;
typedef std::vector<c_vector> c_matrix;

class FGMRES {
   public:
    FGMRES(const int Ntot, const int dofs, const int maxSizeH, const int x_ini, const int t_ini, const int x_fin,
           const int t_fin, const int m, const int restarts, const double tol)
        : Ntot(Ntot), dofs(dofs), maxSizeH(maxSizeH), x_ini(x_ini), x_fin(x_fin), t_ini(t_ini), t_fin(t_fin), m(m),
          restarts(restarts), tol(tol) {}
    virtual ~FGMRES() {}
    // returns the iteration count k*m + maxIt (src/fgmres.cpp:89) or restarts*m
    virtual int fgmres(const spinor& phi, const spinor& x0, spinor& x, const bool& print_message = false) = 0;

   protected:
    const int Ntot, dofs, maxSizeH;
    const int x_ini, x_fin, t_ini, t_fin;
    const int m, restarts;
    const double tol;
};

// GMRES without a preconditioner on D_phi (include/fgmres.h:153-234), run on the GPU
// (dd_solve_gmres).
class FGMRES_fine_level : public FGMRES {
   public:
    FGMRES_fine_level(const int Ntot, const int dofs, const int maxSizeH, const int x_ini, const int t_ini,
                      const int x_fin, const int t_fin, const int m, const int restarts, const double tol,
                      const spinor& U, const double& m0)
        : FGMRES(Ntot, dofs, maxSizeH, x_ini, t_ini, x_fin, t_fin, m, restarts, tol), U(U), m0(m0) {}
    ~FGMRES_fine_level() {}
    int fgmres(const spinor& phi, const spinor& x0, spinor& x, const bool& print_message = false) override;

   private:
    const spinor& U;
    const double m0;
};

#endif

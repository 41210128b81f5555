// FGMRES interface of the reference (include/fgmres.h:23-148).
//
// FGMRES::fgmres is the reference's flexible GMRES on the host over five hooks that a
// subclass defines -- func (the operator), preconditioner, dot, axpy, scal -- with the
// reference's stopping test (|g_{j+1}| < tol ||phi|| inside a cycle, the true residual
// after it) and iteration count k*m + maxIt.  The hooks of the library's own subclasses
// run on the GPU; those subclasses (FGMRES_fine_level here, FGMRES_SAP in sap.h,
// FGMRES_AMG_{k,v}_cycle in amg.h) also override fgmres() itself with the batched solve
// on the device.  A caller's own subclass, as in the reference, implements the hooks and
// inherits the host algorithm.
#ifndef FGMRES_H
#define FGMRES_H
#include <utility>
#include <vector>

#include "dirac_operator.h"

typedef std::vector<c_double> c_vector;
typedef std::vector<c_vector> c_matrix;

class FGMRES {
   public:
    // Ntot: sites, dofs: entries per site, maxSizeH: array size (halo-padded),
    // [x_ini, x_fin] x [t_ini, t_fin]: the halo-padded coordinates the vector operations
    // loop over, m: restart length, restarts, tol
    FGMRES(const int Ntot, const int dofs, const int maxSizeH, const int x_ini, const int t_ini, const int x_fin,
           const int t_fin, const int m, const int restarts, const double tol);
    virtual ~FGMRES() {}
    // returns k*m + maxIt (src/fgmres.cpp:89), or restarts*m without convergence
    virtual int fgmres(const spinor& phi, const spinor& x0, spinor& x, const bool& print_message = false);

   protected:
    const int Ntot, dofs, maxSizeH;
    const int x_ini, x_fin, t_ini, t_fin;
    const int m, restarts;
    const double tol;
    int Nx, Nt;

   private:
    // the operator and the (flexible) preconditioner
    virtual void func(const spinor& in, spinor& out) = 0;
    virtual void preconditioner(const spinor& in, spinor& out) = 0;
    // sum A conj(B), out = X + lambda Y, Y = lambda X over the solver's sites
    virtual c_double dot(c_double* A, c_double* B) = 0;
    virtual void axpy(const spinor& X, const spinor& Y, const c_double& lambda, spinor& out) = 0;
    virtual void scal(const c_double& lambda, const spinor& X, spinor& Y) = 0;

    // Givens rotation of Hessenberg column j (src/fgmres.cpp:99-115) and the back
    // substitution (src/fgmres.cpp:118-127)
    void rotation(const int& j);
    void solve_upper_triangular(const c_matrix& A, const c_vector& b, const int& n, c_vector& out);
    void setZeros();

    std::vector<spinor> VmT, ZmT;
    c_matrix Hm;
    c_vector gm, sn, cn, eta;
    spinor r, w, Dx;
};

// GMRES without a preconditioner on D_phi (include/fgmres.h:153-234); fgmres() runs on
// the GPU (dd_solve_gmres), the hooks are D_phi on the GPU and the interior vector
// operations of utils.h.
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
    void func(const spinor& in, spinor& out) override { D_phi(U, in, out, m0); }
    void preconditioner(const spinor& in, spinor& out) override { out = in; }
    c_double dot(c_double* A, c_double* B) override { return ::dot(A, B); }
    void axpy(const spinor& X, const spinor& Y, const c_double& lambda, spinor& out) override {
        ::axpy(X, Y, lambda, out);
    }
    void scal(const c_double& lambda, const spinor& X, spinor& Y) override { ::scal(lambda, X, Y); }
};

#endif

// Schwarz alternating procedure of the reference (include/sap.h:11-333).
//
// SAP_C::SAP is the reference's red-black multiplicative Schwarz on the host over three
// hooks that a subclass defines -- funcGlobal (the operator on the whole lattice),
// funcLocal (the operator restricted to block current_block, zero outside) and dot --
// with each block solved by GMRES(sap_gmres_restart_length) x sap_gmres_restarts to
// sap_gmres_tolerance from the block right-hand side (SAP_C::GMRES_D_B).  SAP_fine_level
// overrides SAP() with the device solve (dd_sap_fine); FGMRES_SAP runs on the device
// (dd_solve_gmres with the SAP preconditioner).
#ifndef SAP_H
#define SAP_H
#include "fgmres.h"

class SAP_C {
   public:
    // block GMRES of one Schwarz block; vectors are the block with a one-site halo of
    // zeros, (x_elements+2)(t_elements+2) sites of dofs entries (include/sap.h:14-83)
    class GMRES_D_B : public FGMRES {
       public:
        GMRES_D_B(const int Ntot, const int dofs, const int maxSizeH, const int x_ini, const int t_ini,
                  const int x_fin, const int t_fin, const int m, const int restarts, const double tol, SAP_C* parent)
            : FGMRES(Ntot, dofs, maxSizeH, x_ini, t_ini, x_fin, t_fin, m, restarts, tol), parent(parent),
              xfin(x_fin), tfin(t_fin), nd(dofs) {}
        ~GMRES_D_B() {}

       private:
        SAP_C* parent;
        int xfin, tfin, nd;
        void func(const spinor& in, spinor& out) override { parent->funcLocal(in, out); }
        void preconditioner(const spinor& in, spinor& out) override { out = in; }
        c_double dot(c_double* A, c_double* B) override;
        void axpy(const spinor& X, const spinor& Y, const c_double& lambda, spinor& out) override;
        void scal(const c_double& lambda, const spinor& X, spinor& Y) override;
    };

    // Nx x Nt sites split into block_x x block_t Schwarz blocks; spins*colors entries per
    // site (include/sap.h:87-126)
    SAP_C(const int Nx, const int Nt, const int block_x, const int block_t, const int spins, const int colors);
    virtual ~SAP_C() {}

    // x (initial guess in, result out) after at most nu iterations on D x = v; stops when
    // ||v - D x|| < tol ||v|| and returns that iteration's index, else nu
    // (src/sap.cpp:81-165)
    virtual int SAP(const spinor& v, spinor& x, const int& nu, const double& tol, const bool print);

    GMRES_D_B gmres_DB;
    int nu = 0;
    int Nt, Nx;
    int dofs;
    double tol = 0;
    int Block_x, Block_t;
    int x_elements, t_elements;
    int NBlocks, coloring_blocks;
    int colors, spins;
    int Ntot_no_halo, Ntot_w_halo, Nvars_no_halo, Nvars_w_halo, Ntot_original_w_halo;
    int current_block = 0;
    // Blocks[block][site of the halo-padded block] = halo-padded lattice site
    std::vector<std::vector<int>> Blocks;
    std::vector<int> RedBlocks, BlackBlocks;

   private:
    void SchwarzBlocks();
    // x = I_B D_B^-1 I_B^T v on block `block` (src/sap.cpp:48-77)
    void I_D_B_1_It(const spinor& v, spinor& x, const int& block);
    virtual void funcGlobal(spinor& in, spinor& out) = 0;
    virtual void funcLocal(const spinor& in, spinor& out) = 0;
    virtual c_double dot(const spinor& X, const spinor& Y) = 0;
};

// SAP on the level-0 Wilson-Dirac operator (include/sap.h:192-234).  SAP() runs on the
// GPU (dd_sap_fine); the hooks are D_phi, the block operator D_B on the host and the
// interior dot.
class SAP_fine_level : public SAP_C {
   public:
    SAP_fine_level(const int Nx, const int Nt, const int block_x, const int block_t, const int spins,
                   const int colors)
        : SAP_C(Nx, Nt, block_x, block_t, spins, colors) {}
    void set_params(const spinor& conf, const double& bare_mass) {
        U = &conf;
        m0 = bare_mass;
    }
    int SAP(const spinor& v, spinor& x, const int& nu, const double& tol, const bool print) override;

   private:
    const spinor* U = nullptr;
    double m0 = 0;
    // Wilson-Dirac operator of block `block`, neighbours outside the block taken as zero
    // (src/sap.cpp:170-225)
    void D_B(const spinor& U, const spinor& v, spinor& x, const double& m0, const int& block);
    void funcGlobal(spinor& in, spinor& out) override { D_phi(*U, in, out, m0); }
    void funcLocal(const spinor& in, spinor& out) override { D_B(*U, in, out, m0, current_block); }
    c_double dot(const spinor& X, const spinor& Y) override { return ::dot(X.val, Y.val); }
};

// FGMRES preconditioned by one SAP iteration from a zero guess over SAPV::sap_block_x x
// sap_block_t blocks per (virtual) rank (include/sap.h:236-333): the one process holds
// every rank's blocks, so the lattice is split into sap_block_x*ranks_x x
// sap_block_t*ranks_t blocks, each as large as on the reference's rank grid.  fgmres() runs on the GPU
// (dd_solve_gmres); the hooks are D_phi and SAP_fine_level.
class FGMRES_SAP : public FGMRES {
   public:
    FGMRES_SAP(const int Ntot, const int dofs, const int maxSizeH, const int x_ini, const int t_ini, const int x_fin,
               const int t_fin, const int m, const int restarts, const double tol, const spinor& U, const double& m0)
        : FGMRES(Ntot, dofs, maxSizeH, x_ini, t_ini, x_fin, t_fin, m, restarts, tol), U(U), m0(m0),
          sap_preconditioner(x_fin - x_ini + 1, t_fin - t_ini + 1, SAPV::sap_block_x * mpi::ranks_x,
                             SAPV::sap_block_t * mpi::ranks_t, 2, 1) {
        sap_preconditioner.set_params(U, m0);
    }
    ~FGMRES_SAP() {}
    int fgmres(const spinor& phi, const spinor& x0, spinor& x, const bool& print_message = false) override;

   private:
    const spinor& U;
    const double m0;
    SAP_fine_level sap_preconditioner;
    void func(const spinor& in, spinor& out) override { D_phi(U, in, out, m0); }
    void preconditioner(const spinor& in, spinor& out) override {
        out.clearBuffer();
        sap_preconditioner.SAP(in, out, 1, SAPV::sap_tolerance, false);
    }
    c_double dot(c_double* A, c_double* B) override { return ::dot(A, B); }
    void axpy(const spinor& X, const spinor& Y, const c_double& lambda, spinor& out) override {
        ::axpy(X, Y, lambda, out);
    }
    void scal(const c_double& lambda, const spinor& X, spinor& Y) override { ::scal(lambda, X, Y); }
};

#endif

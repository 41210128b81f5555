// One level of the hierarchy (include/level.h:8-471), backed by the hierarchy that the
// library keeps on the GPU.  Spinors are halo-padded host arrays of the level,
// (Nx+2)(Nt+2)*DOF entries, DOF = 2*colors.
//
// Two ways to get a Level:
//  - Level(l, U), the reference's constructor: all Levels built this way for one gauge
//    field share a hierarchy whose storage follows LevelV (the parameters file) and
//    whose content the caller builds step by step, as AlgebraicMG::setUpPhase does
//    (src/amg.cpp:4-62): fill tvec, orthonormalize(), makeCoarseLinks(next).  Here tvec
//    on the host is the master copy: the operations that read the test vectors upload
//    it first.  Level 0's operator is the Wilson-Dirac operator of U (makeDirac()
//    fetches its link matrices into G1/G2/G3).
//  - Level(l, ctx): level l of a hierarchy that dd_setup built (AlgebraicMG), whose test
//    vectors live on the device (fetchTestVectors() copies them into tvec).
#ifndef LEVEL_H
#define LEVEL_H
#include "sap.h"

struct dd_ctx;

class Level {
   public:
    // SAP smoother of the level (include/level.h:14-60): SAP() on the device (dd_sap);
    // the hooks are the level operator, the block-restricted level operator and the
    // level dot
    class SAP_level_l : public SAP_C {
       public:
        explicit SAP_level_l(Level* parent);
        int SAP(const spinor& v, spinor& x, const int& nu, const double& tol, const bool print) override;

       private:
        Level* parent;
        void funcGlobal(spinor& in, spinor& out) override { parent->D_operator(in, out); }
        void funcLocal(const spinor& in, spinor& out) override;
        c_double dot(const spinor& X, const spinor& Y) override;
    };
    SAP_level_l* sap_l;

    // GMRES of the level, used on the coarsest one (include/level.h:70-140): the
    // reference's host FGMRES over the level operator (on the device), no
    // preconditioner, GMRES(gmres_restart_length_coarse_level) x
    // gmres_restarts_coarse_level to gmres_tol_coarse_level
    class GMRES_level_l : public FGMRES {
       public:
        explicit GMRES_level_l(Level* parent);
        ~GMRES_level_l() {}

       private:
        Level* parent;
        void func(const spinor& in, spinor& out) override { parent->D_operator(in, out); }
        void preconditioner(const spinor& in, spinor& out) override { out = in; }
        c_double dot(c_double* A, c_double* B) override;
        void axpy(const spinor& X, const spinor& Y, const c_double& lambda, spinor& out) override;
        void scal(const c_double& lambda, const spinor& X, spinor& Y) override;
    };
    GMRES_level_l* gmres_l = nullptr;

    Level(const int& level, const spinor& U);
    Level(const int& level, dd_ctx* ctx);
    ~Level();

    const int level;
    const spinor U;   // gauge configuration (empty for a view of a built hierarchy)
    const int Nx, Nt, Ntot, colors, DOF, Ntest;
    // aggregates: blocks per (virtual) rank and sites per aggregate (include/level.h:290-315)
    const int xblocks_per_rank, tblocks_per_rank, blocks_per_rank;
    const int x_elements, t_elements, sites_per_block, NBlocks;

    std::vector<spinor> tvec;   // [Ntest] test vectors of the level, halo-padded
    // link matrices of the level's operator in the reference's layout
    // (include/level.h:349-365): G1[n][alf][bet][c][b], G2/G3[n][alf][bet][c][b][mu]
    // over halo-padded sites n; filled by makeDirac() (level 0) and by the previous
    // level's makeCoarseLinks(*this)
    spinor G1, G2, G3;

    void D_operator(const spinor& v, spinor& out);   // src/level.cpp:394-426
    void P_vc(const spinor& vc, spinor& out);        // src/level.cpp:92-141
    void Pdagg_v(const spinor& v, spinor& out);      // src/level.cpp:145-189
    void orthonormalize();                           // src/level.cpp:193-284
    void checkOrthogonality();                       // src/level.cpp:544-604
    void makeDirac();                                // src/level.cpp:286-331
    void makeCoarseLinks(Level& next_level);         // src/level.cpp:431-542
    void fetchTestVectors();                         // device test vectors -> tvec
    dd_ctx* context() const { return ctx; }

   private:
    dd_ctx* ctx;
    bool owns_tvec;   // Level(l, U): tvec is the master copy
    void upload_tvec();
    void fetch_links(spinor& g1, spinor& g2, spinor& g3);
};

#endif

// One level of the hierarchy (include/level.h:8-471): the level's operator, transfer
// operators and smoother, backed by the hierarchy built on the GPU.  Spinors are
// halo-padded host arrays of the level, (Nx+2)(Nt+2)*DOF entries.
#ifndef LEVEL_H
#define LEVEL_H
#include "fgmres.h"

struct dd_ctx;

class Level {
   public:
    // SAP smoother of the level (SAP_C::SAP, src/sap.cpp:81-165)
    class SAP_level_l {
       public:
        explicit SAP_level_l(Level* parent) : parent(parent) {}
        int SAP(const spinor& v, spinor& x, const int& nu, const double& tol, const bool print);

       private:
        Level* parent;
    };
    SAP_level_l* sap_l;

    Level(const int& level, dd_ctx* ctx);
    ~Level();

    const int level;
    const int Nx, Nt, Ntot, colors, DOF, Ntest;

    void D_operator(const spinor& v, spinor& out);   // src/level.cpp:394-426
    void P_vc(const spinor& vc, spinor& out);        // src/level.cpp:92-141
    void Pdagg_v(const spinor& v, spinor& out);      // src/level.cpp:145-189
    void checkOrthogonality();                       // src/level.cpp:544-604
    // test vectors of this level after setup, tvec[cc] halo-padded (include/level.h:241)
    std::vector<spinor> tvec;
    void fetchTestVectors();

   private:
    dd_ctx* ctx;
};

#endif

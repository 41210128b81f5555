// Test-infrastructure driver around the reference DD-alphaAMG code (CPU, MPI).
//
// Only the oracle recipe (oracle/Makefile) compiles this file, against the
// reference sources where they lie; the product never links it.  It calls the
// reference's own functions and classes and writes their outputs per rank, so
// that tests/golden/make_golden.py can turn them into fixtures and bench.py can
// time the reference solve on the host cores.
//
// Reference entry points exercised (paths relative to the reference root):
//   D_phi / D_dagger_phi            src/dirac_operator.cpp:24-85
//   Level::D_operator, P_vc, Pdagg_v, orthonormalize, makeCoarseLinks
//                                   src/level.cpp:92-542
//   SAP_C::SAP                      src/sap.cpp:81-165
//   AlgebraicMG::setUpPhase         src/amg.cpp:4-62
//   FGMRES_AMG_{k,v}_cycle          include/amg.h:240-270, src/fgmres.cpp:5-97
//   conjugate_gradient, bi_cgstab   src/conjugate_gradient.cpp:4-176 (as Methods::CG/BiCG
//                                   call them, src/methods.cpp:3-59)
//   FGMRES_fine_level, FGMRES_SAP   include/fgmres.h:153-234, include/sap.h:236-333
//
// Access to the reference classes' private members (the level list of
// AlgebraicMG) is needed to dump the hierarchy; the access specifiers are
// lifted for this translation unit only (standard and MPI headers come first so
// that only the reference's own classes see the lifted specifiers).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <map>
#include <random>
#include <sstream>
#include <string>
#include <utility>
#include <vector>
#include "mpi.h"
#define private public
#define protected public
#include "amg.h"
#include "conjugate_gradient.h"
#include "params.h"
#include "io.h"
#undef private
#undef protected

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>

static std::map<std::string, std::string> g_opts;

static std::string opt(const char* key, const char* dflt) {
    auto it = g_opts.find(key);
    return it == g_opts.end() ? std::string(dflt) : it->second;
}

// Writes the interior of a halo-padded level array: header {cx, ct, nx, nt, per_site}
// then nx*nt*per_site complex128 values, x outer, t inner.
static void dump(const std::string& dir, const std::string& name, const c_double* val,
                 int nx, int nt, int per_site) {
    char path[1024];
    snprintf(path, sizeof(path), "%s/%s.r%d.bin", dir.c_str(), name.c_str(), mpi::rank2d);
    FILE* f = fopen(path, "wb");
    if (!f) { fprintf(stderr, "cannot open %s\n", path); MPI_Abort(MPI_COMM_WORLD, 1); }
    int hdr[5] = {mpi::coords[0], mpi::coords[1], nx, nt, per_site};
    fwrite(hdr, sizeof(int), 5, f);
    for (int x = 1; x <= nx; x++)
        for (int t = 1; t <= nt; t++)
            fwrite(&val[(size_t)(x * (nt + 2) + t) * per_site], sizeof(c_double), per_site, f);
    fclose(f);
}

static void write_scalar_file(const std::string& dir, const std::string& text) {
    if (mpi::rank2d != 0) return;
    std::string p = dir + "/result.txt";
    FILE* f = fopen(p.c_str(), "a");
    fputs(text.c_str(), f);
    fclose(f);
}

static double global_norm(const spinor& v, int nx, int nt, int dof) {
    double loc = 0;
    for (int x = 1; x <= nx; x++)
        for (int t = 1; t <= nt; t++)
            for (int d = 0; d < dof; d++) loc += std::norm(v.val[(x * (nt + 2) + t) * dof + d]);
    double g = 0;
    MPI_Allreduce(&loc, &g, 1, MPI_DOUBLE, MPI_SUM, mpi::cart_comm);
    return std::sqrt(g);
}

int main(int argc, char** argv) {
    MPI_Init(&argc, &argv);
    MPI_Comm_size(MPI_COMM_WORLD, &mpi::size);
    MPI_Comm_rank(MPI_COMM_WORLD, &mpi::rank);
    for (int i = 1; i < argc; i++) {
        const char* eq = strchr(argv[i], '=');
        if (eq) g_opts[std::string(argv[i], eq - argv[i])] = std::string(eq + 1);
    }
    const std::string mode = opt("mode", "dirac");
    const std::string out = opt("out", ".");
    mpi::ranks_x = atoi(opt("ranks_x", "2").c_str());
    mpi::ranks_t = atoi(opt("ranks_t", "2").c_str());
    LevelV::levels = atoi(opt("levels", "2").c_str());
    const double m0 = atof(opt("m0", "-0.1").c_str());
    mass::m0 = m0;
    // Solver knobs are mutable globals in the reference (src/variables.cpp:90-134).
    AMGV::nu1 = atoi(opt("nu1", "0").c_str());
    AMGV::nu2 = atoi(opt("nu2", "2").c_str());
    FGMRESV::fgmres_restart_length = atoi(opt("fgmres_m", "20").c_str());
    FGMRESV::fgmres_restarts = atoi(opt("fgmres_restarts", "50").c_str());
    const double tol = atof(opt("tol", "1e-10").c_str());

    LevelV::maxLevel = LevelV::levels - 1;
    allocate_lattice_arrays();
    initializeMPI();
    readParameters(opt("params", "parameters"));
    boundaries();

    spinor U(mpi::maxSizeH);
    spinor rhs(mpi::maxSizeH);
    spinor x0(mpi::maxSizeH);
    read_binary(opt("conf", "conf.bin"), U);
    read_binary(opt("rhs", "rhs.bin"), rhs);
    const int wx = mpi::width_x, wt = mpi::width_t;

    if (mode == "dirac") {
        spinor o1(mpi::maxSizeH), o2(mpi::maxSizeH), o3(mpi::maxSizeH);
        D_phi(U, rhs, o1, m0);
        D_dagger_phi(U, rhs, o2, m0);
        Level lev0(0, U);
        lev0.D_operator(rhs, o3);
        dump(out, "Dphi", o1.val, wx, wt, 2);
        dump(out, "Ddagphi", o2.val, wx, wt, 2);
        dump(out, "Dop0", o3.val, wx, wt, 2);
    } else if (mode == "setup") {
        // Hierarchy: setup phase exactly as the outer solver's constructor runs it.
        AlgebraicMG amg(U, m0, AMGV::nu1, AMGV::nu2);
        amg.setUpPhase(AMGV::Nit);
        const int L = LevelV::levels;
        for (int l = 0; l < L; l++) {
            Level* lv = amg.levels[l];
            if (l < L - 1)
                for (int cc = 0; cc < lv->Ntest; cc++)
                    dump(out, "tvec_l" + std::to_string(l) + "_" + std::to_string(cc), lv->tvec[cc].val,
                         lv->Nx, lv->Nt, lv->DOF);
            if (l >= 1) {
                dump(out, "G1_l" + std::to_string(l), lv->G1.val, lv->Nx, lv->Nt, lv->DOF * lv->DOF);
                dump(out, "G2_l" + std::to_string(l), lv->G2.val, lv->Nx, lv->Nt, 2 * lv->DOF * lv->DOF);
                dump(out, "G3_l" + std::to_string(l), lv->G3.val, lv->Nx, lv->Nt, 2 * lv->DOF * lv->DOF);
            }
        }
        // Transfer operators and the level-1 stencil on the restricted right-hand side.
        Level* l0 = amg.levels[0];
        Level* l1 = amg.levels[1];
        const int n1 = (l1->Nx + 2) * (l1->Nt + 2) * l1->DOF;
        spinor vc(n1), Dvc(n1), Pvc(mpi::maxSizeH);
        l0->Pdagg_v(rhs, vc);
        dump(out, "Pdagg_rhs", vc.val, l1->Nx, l1->Nt, l1->DOF);
        l1->D_operator(vc, Dvc);
        dump(out, "Dop1_Pdagg_rhs", Dvc.val, l1->Nx, l1->Nt, l1->DOF);
        l0->P_vc(vc, Pvc);
        dump(out, "P_Pdagg_rhs", Pvc.val, wx, wt, 2);
        // One SAP iteration on each non-coarsest level, zero initial guess.
        spinor s0(mpi::maxSizeH);
        l0->sap_l->SAP(rhs, s0, atoi(opt("sap_nu", "1").c_str()), 1e-10, false);
        dump(out, "SAP0_rhs", s0.val, wx, wt, 2);
        if (L > 2) {
            spinor s1(n1);
            l1->sap_l->SAP(vc, s1, atoi(opt("sap_nu", "1").c_str()), 1e-10, false);
            dump(out, "SAP1_Pdagg_rhs", s1.val, l1->Nx, l1->Nt, l1->DOF);
        }
    } else if (mode == "solve") {
        const std::string cyc = opt("cycle", "K");
        spinor x(mpi::maxSizeH);
        int its = 0;
        double t_setup = 0, t_total = 0;
        MPI_Barrier(MPI_COMM_WORLD);
        double t0 = MPI_Wtime();
        if (cyc == "K") {
            FGMRES_AMG_k_cycle f(U, FGMRESV::fgmres_restart_length, FGMRESV::fgmres_restarts, tol,
                                 AMGV::nu1, AMGV::nu2, m0);
            MPI_Barrier(MPI_COMM_WORLD);
            t_setup = MPI_Wtime() - t0;
            its = f.fgmres(rhs, x0, x, false);
        } else {
            FGMRES_AMG_v_cycle f(U, FGMRESV::fgmres_restart_length, FGMRESV::fgmres_restarts, tol,
                                 AMGV::nu1, AMGV::nu2, m0);
            MPI_Barrier(MPI_COMM_WORLD);
            t_setup = MPI_Wtime() - t0;
            its = f.fgmres(rhs, x0, x, false);
        }
        MPI_Barrier(MPI_COMM_WORLD);
        t_total = MPI_Wtime() - t0;
        spinor Dx(mpi::maxSizeH), r(mpi::maxSizeH);
        D_phi(U, x, Dx, m0);
        for (int i = 0; i < mpi::maxSizeH; i++) r.val[i] = rhs.val[i] - Dx.val[i];
        double rel = global_norm(r, wx, wt, 2) / global_norm(rhs, wx, wt, 2);
        if (opt("dump_x", "1") == "1") dump(out, "x", x.val, wx, wt, 2);
        char buf[512];
        snprintf(buf, sizeof(buf), "iterations %d\nrelres %.17g\nsetup_s %.6f\ntotal_s %.6f\nsolve_s %.6f\nranks %d\n",
                 its, rel, t_setup, t_total, t_total - t_setup, mpi::size);
        write_scalar_file(out, buf);
    } else if (mode == "methods") {
        // The comparison solvers of src/methods.cpp on the same system, each from x0 = 0
        // (CG from D^dagger b, as conjugate_gradient starts).
        CG::max_iter = BiCG::max_iter = atoi(opt("max_iter", "10000").c_str());
        CG::tol = BiCG::tol = tol;
        const int gm = atoi(opt("gmres_m", "20").c_str()), gr = atoi(opt("gmres_restarts", "100").c_str());
        spinor xc(mpi::maxSizeH), xb(mpi::maxSizeH), xg(mpi::maxSizeH), xs(mpi::maxSizeH), Ddag(mpi::maxSizeH);
        D_dagger_phi(U, rhs, Ddag, m0);
        const int it_cg = conjugate_gradient(U, Ddag, xc, m0, false);
        const int it_bi = bi_cgstab(U, rhs, x0, xb, m0, false);
        FGMRES_fine_level g(wx * wt, LV::dof, mpi::maxSizeH, 1, 1, wx, wt, gm, gr, tol, U, m0);
        const int it_gm = g.fgmres(rhs, x0, xg, false);
        FGMRES_SAP fs(wx * wt, LV::dof, mpi::maxSizeH, 1, 1, wx, wt, gm, gr, tol, U, m0);
        const int it_fs = fs.fgmres(rhs, x0, xs, false);
        auto rel = [&](const spinor& x) {
            spinor Dx(mpi::maxSizeH), r(mpi::maxSizeH);
            D_phi(U, x, Dx, m0);
            for (int i = 0; i < mpi::maxSizeH; i++) r.val[i] = rhs.val[i] - Dx.val[i];
            return global_norm(r, wx, wt, 2) / global_norm(rhs, wx, wt, 2);
        };
        dump(out, "x_cg", xc.val, wx, wt, 2);
        dump(out, "x_bicgstab", xb.val, wx, wt, 2);
        dump(out, "x_gmres", xg.val, wx, wt, 2);
        dump(out, "x_fgmres_sap", xs.val, wx, wt, 2);
        char buf[512];
        snprintf(buf, sizeof(buf),
                 "cg_iterations %d\nbicgstab_iterations %d\ngmres_iterations %d\nfgmres_sap_iterations %d\n"
                 "cg_relres %.17g\nbicgstab_relres %.17g\ngmres_relres %.17g\nfgmres_sap_relres %.17g\n",
                 it_cg, it_bi, it_gm, it_fs, rel(xc), rel(xb), rel(xg), rel(xs));
        write_scalar_file(out, buf);
    } else {
        if (mpi::rank == 0) fprintf(stderr, "unknown mode %s\n", mode.c_str());
        MPI_Abort(MPI_COMM_WORLD, 2);
    }
    free_lattice_arrays();
    MPI_Finalize();
    return 0;
}

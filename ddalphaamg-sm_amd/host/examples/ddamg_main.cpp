// Example driver with the reference's flow (src/main.cpp:10-105, src/methods.cpp:3-191):
// reads the run parameters on stdin, the gauge field / right-hand side in the reference's
// binary record format and the per-level parameters file; runs the comparison solvers
// (CG on the normal equations as the reference's main does, then BiCGStab and
// FGMRES + SAP), then FGMRES + V- and K-cycle, all on the GPU, and checks each solution
// with D_phi.
#include <chrono>
#include <cstdio>
#include <string>

#include "amg.h"
#include "conjugate_gradient.h"
#include "io.h"
#include "params.h"
#include "sap.h"

static void check_solution(const spinor& U, const spinor& rhs, const spinor& x_sol, double m0) {
    spinor x(mpi::maxSizeH);
    D_phi(U, x_sol, x, m0);
    for (int nx = 1; nx <= mpi::width_x; nx++)
        for (int nt = 1; nt <= mpi::width_t; nt++)
            for (int mu = 0; mu < 2; mu++)
                if (std::abs(x.val[idx(nx, nt, mu)] - rhs.val[idx(nx, nt, mu)]) > 1e-6)
                    std::cout << "Solution differs at nx" << nx << " nt " << nt << " mu " << mu << std::endl;
    std::cout << "Solution verified" << std::endl;
}

int main() {
    std::string confFile, rhsFile, pFile;
    double m0;
    std::cout << "Nx " << LV::Nx << " Nt " << LV::Nt << std::endl;
    std::cout << "ranks_x: ";
    std::cin >> mpi::ranks_x;
    std::cout << "ranks_t: ";
    std::cin >> mpi::ranks_t;
    std::cout << "Levels: ";
    std::cin >> LevelV::levels;
    std::cout << "m0: ";
    std::cin >> m0;
    std::cout << "Configuration file path: ";
    std::cin >> confFile;
    std::cout << "RHS file path: ";
    std::cin >> rhsFile;
    std::cout << "Method parameters file path: ";
    std::cin >> pFile;
    mass::m0 = m0;
    LevelV::maxLevel = LevelV::levels - 1;
    allocate_lattice_arrays();
    initializeMPI();
    readParameters(pFile);
    spinor U(mpi::maxSizeH), rhs(mpi::maxSizeH), x0(mpi::maxSizeH);
    read_binary(confFile, U);
    read_binary(rhsFile, rhs);
    const double tol = 1e-10;
    {   // Methods::CG, Methods::BiCG, Methods::FGMRES_sap (src/methods.cpp:3-104)
        std::cout << "--------Inverting the normal equations with CG----------" << std::endl;
        spinor Ddag(mpi::maxSizeH), xcg(mpi::maxSizeH);
        D_dagger_phi(U, rhs, Ddag, m0);
        auto t0 = std::chrono::steady_clock::now();
        const int itc = conjugate_gradient(U, Ddag, xcg, m0, true);
        double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::printf("time elapsed CG: %.4fs (%d iterations).\n", el, itc);
        check_solution(U, rhs, xcg, m0);
        std::cout << "--------------Bi-CGstab inversion--------------" << std::endl;
        spinor xb(mpi::maxSizeH);
        t0 = std::chrono::steady_clock::now();
        const int itb = bi_cgstab(U, rhs, x0, xb, m0, true);
        el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::printf("time elapsed Bi-CGstab: %.4fs (%d iterations).\n", el, itb);
        check_solution(U, rhs, xb, m0);
        std::cout << "--------------Flexible GMRES with SAP preconditioning--------------" << std::endl;
        spinor xs(mpi::maxSizeH);
        FGMRES_SAP fs(LV::Ntot, LV::dof, mpi::maxSizeH, 1, 1, mpi::width_x, mpi::width_t,
                      FGMRESV::fgmres_restart_length, FGMRESV::fgmres_restarts, tol, U, m0);
        t0 = std::chrono::steady_clock::now();
        const int its = fs.fgmres(rhs, x0, xs, true);
        el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::printf("time elapsed FGMRES_SAP: %.4fs (%d iterations).\n", el, its);
        check_solution(U, rhs, xs, m0);
    }
    for (int cycle : {0, 1}) {
        spinor x(mpi::maxSizeH);
        auto t0 = std::chrono::steady_clock::now();
        int its;
        if (cycle == 0) {
            std::cout << "--------------FGMRES with AMG V-cycle--------------" << std::endl;
            FGMRES_AMG_v_cycle f(U, FGMRESV::fgmres_restart_length, FGMRESV::fgmres_restarts, tol, AMGV::nu1,
                                 AMGV::nu2, m0);
            its = f.fgmres(rhs, x0, x, true);
        } else {
            std::cout << "--------------FGMRES with AMG K-cycle--------------" << std::endl;
            FGMRES_AMG_k_cycle f(U, FGMRESV::fgmres_restart_length, FGMRESV::fgmres_restarts, tol, AMGV::nu1,
                                 AMGV::nu2, m0);
            its = f.fgmres(rhs, x0, x, true);
        }
        const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::printf("time elapsed FGMRES AMG %c-cycle: %.4fs (%d iterations).\n", cycle ? 'K' : 'V', el, its);
        check_solution(U, rhs, x, m0);
    }
    free_lattice_arrays();
    return 0;
}

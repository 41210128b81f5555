// Example driver with the reference's flow (src/main.cpp:10-105, src/methods.cpp:133-191):
// reads the run parameters on stdin, the gauge field / right-hand side in the reference's
// binary record format and the per-level parameters file, solves with FGMRES + V- and
// K-cycle on the GPU and checks the solution with D_phi.
#include <chrono>
#include <cstdio>
#include <string>

#include "amg.h"
#include "io.h"
#include "params.h"

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

// Check helpers of the reference's include/tests.h that need no access to the internals
// of its MPI decomposition.  check_sol compares D x with the right-hand side site by
// site (include/tests.h:12-31).
#ifndef TESTS_H
#define TESTS_H
#include <random>

#include "amg.h"
#include "dirac_operator.h"
#include "io.h"

inline void check_sol(const spinor& U, const spinor& rhs, const spinor& inverse, const double& m0) {
    spinor x(mpi::maxSizeH);
    D_phi(U, inverse, x, m0);
    bool ok = true;
    for (int nx = 1; nx <= mpi::width_x; nx++)
        for (int nt = 1; nt <= mpi::width_t; nt++)
            for (int mu = 0; mu < 2; mu++) {
                const int index = idx(nx, nt, mu);
                if (std::abs(x.val[index] - rhs.val[index]) > 1e-8) {
                    ok = false;
                    std::cout << "Solution differs at nx" << nx << " nt " << nt << " mu " << mu << std::endl;
                }
            }
    if (ok && mpi::rank2d == 0) std::cout << "All good with the solution!" << std::endl;
}

#endif

// Helpers of the reference's include/utils.h on halo-padded level-0 host spinors: the
// site index, the interior dot/axpy/scal, and the FLOP-count reduction and printout.
// One process holds the lattice, so the reductions need no MPI.
#ifndef UTILS_H
#define UTILS_H
#include <iomanip>
#include <sstream>
#include <string>

#include "variables.h"

// halo-padded index of interior (x, t), x in [1, width_x], t in [1, width_t], spin mu
// (include/utils.h:46-52)
inline int idx(int x, int t, int mu) { return ((x * (mpi::width_t + 2) + t) * LV::dof + mu); }

// sum over the interior of A conj(B) (include/utils.h:68-84)
inline c_double dot(c_double* A, c_double* B) {
    c_double z = 0;
    for (int x = 1; x <= mpi::width_x; x++)
        for (int t = 1; t <= mpi::width_t; t++)
            for (int mu = 0; mu < LV::dof; mu++) {
                const int i = idx(x, t, mu);
                z += A[i] * std::conj(B[i]);
            }
    localFLOPS += (ca + cm) * (long long)mpi::width_x * mpi::width_t * LV::dof;
    return z;
}

// out = X + lambda Y over the interior (include/utils.h:90-103)
template <typename T>
inline void axpy(const spinor& X, const spinor& Y, const T& lambda, spinor& out) {
    for (int x = 1; x <= mpi::width_x; x++)
        for (int t = 1; t <= mpi::width_t; t++)
            for (int mu = 0; mu < LV::dof; mu++) {
                const int i = idx(x, t, mu);
                out.val[i] = X.val[i] + lambda * Y.val[i];
            }
    localFLOPS += (ca + cm) * (long long)mpi::width_x * mpi::width_t * LV::dof;
}

// Y = lambda X over the interior (include/utils.h:108-120)
template <typename T>
inline void scal(const T& lambda, const spinor& X, spinor& Y) {
    for (int x = 1; x <= mpi::width_x; x++)
        for (int t = 1; t <= mpi::width_t; t++)
            for (int mu = 0; mu < LV::dof; mu++) {
                const int i = idx(x, t, mu);
                Y.val[i] = lambda * X.val[i];
            }
    localFLOPS += cm * (long long)mpi::width_x * mpi::width_t * LV::dof;
}

// number with four decimals and the decimal point removed (file-name tags)
inline std::string format(const double& number) {
    std::ostringstream oss;
    oss << std::fixed << std::setprecision(4) << number;
    std::string str = oss.str();
    const size_t dot_at = str.find('.');
    if (dot_at != std::string::npos) str.erase(dot_at, 1);
    return str;
}

void printParameters();

// FLOPS = sum of localFLOPS over the processes: here the one process (include/utils.h:141-144)
inline void mpi_reduceFLOPS() { FLOPS = localFLOPS; }

inline void printFLOPS(const long long int& x) {
    if (mpi::rank2d == 0) std::cout << "GFLOPS = " << (long double)x / 1e9 << std::endl;
}

#endif

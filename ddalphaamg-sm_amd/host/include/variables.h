// Host API of the MI355X DD-alphaAMG library: global knobs and the spinor type with the
// names, types and defaults of the reference's include/variables.h:1-253 and
// src/variables.cpp:90-134, so that code written against the reference compiles.
// One process drives one GPU holding the whole lattice; mpi::ranks_x/ranks_t remain
// as the virtual rank grid whose per-rank random streams draw the test vectors and
// whose per-rank SAP block counts the parameters file gives (include/params.h:7-57).
#ifndef VARIABLES_H_INCLUDED
#define VARIABLES_H_INCLUDED
#include <complex>
#include <cstdlib>
#include <iomanip>
#include <iostream>
#include <string>
#include <vector>

#include "config.h"
#include "ddamg_mpi.h"

typedef std::complex<double> c_double;
extern double pi;
extern c_double I_number;

// Floating-point operation counts per complex128 operation (include/variables.h:16-25);
// the host-side algorithms (FGMRES and SAP_C over caller-defined hooks) add them to
// localFLOPS as the reference does.  The device paths add nothing.
constexpr long long int ca = 2;    // complex addition
constexpr long long int cm = 6;    // complex multiplication
constexpr long long int cd = 11;   // complex division
constexpr long long int da = 1;    // double addition
constexpr long long int dm = 1;    // double multiplication
constexpr long long int dd = 1;    // double division
constexpr long long int dcm = 2;   // double * complex
constexpr long long int dcd = 7;   // double / complex
constexpr long long int dsq = 1;   // sqrt(double)

extern long long int FLOPS;
extern long long int localFLOPS;

// Iteration and FLOP counters of the comparison runs (include/variables.h:30-50).
namespace iter_counters {
extern long long int kCycleIt, vCycleIt, CGIt, BiCGIt, GMRESIt, SAPIt, FGMRES_SAPIt;
}
namespace flop_counters {
extern long long int kCycleFlops, vCycleFlops, kCycleSetUpFlops, vCycleSetUpFlops, CGFlops, BiCGFLops, GMRESFlops,
    SAPFlops, FGMRES_SAPFlops;
}

namespace mass {
extern double m0;
}

namespace mpi {
extern int rank, size, maxSize, maxSizeH, sitesH, ranks_x, ranks_t, width_x, width_t, rank2d;
extern MPI_Comm cart_comm;   // MPI_COMM_WORLD (one process holds the lattice)
}

namespace LV {
constexpr int Nx = NS;
constexpr int Nt = NT;
constexpr int Ntot = Nx * Nt;
constexpr int dof = 2;
}  // namespace LV

namespace SAPV {
constexpr int sap_block_x = 4;
constexpr int sap_block_t = 4;
extern int sap_gmres_restart_length;
extern int sap_gmres_restarts;
extern double sap_gmres_tolerance;
extern double sap_tolerance;
}  // namespace SAPV

namespace AMGV {
extern int SAP_test_vectors_iterations;
extern int gmres_restarts_coarse_level;
extern int gmres_restart_length_coarse_level;
extern double gmres_tol_coarse_level;
extern int nu1, nu2, Nit;
extern int fgmres_k_cycle_restart_length;
extern int fgmres_k_cycle_restarts;
extern double fgmres_k_cycle_tol;
extern int cycle;
}  // namespace AMGV

namespace FGMRESV {
extern double fgmres_tolerance;
extern int fgmres_restart_length;
extern int fgmres_restarts;
}  // namespace FGMRESV

// CG on the normal equations and BiCGStab (include/variables.h:116-126)
namespace CG {
extern int max_iter;
extern double tol;
}  // namespace CG
namespace BiCG {
extern int max_iter;
extern double tol;
}  // namespace BiCG

namespace LevelV {
extern int levels;
extern int maxLevel;
extern int* BlocksX;   // aggregate blocks (global count) per level
extern int* BlocksT;
extern int* Ntest;
extern int* SAP_Block_x;   // SAP blocks per (virtual) rank per level
extern int* SAP_Block_t;
extern int* NxSites;
extern int* NtSites;
extern int* DOF;
extern int* Colors;
}  // namespace LevelV

// Additions of this library (no reference counterpart): device, precision, batch.
namespace DDV {
extern int device;      // HIP device
extern int precision;   // 0: FP64 throughout, 1: FP32 hierarchy inside FP64 FGMRES
extern int max_rhs;     // largest batch a solver object will be asked for
}  // namespace DDV

// Flattened halo-padded spinor (include/variables.h:201-238).
struct spinor {
    c_double* val;
    int size;
    spinor(int N = LV::Ntot) : size(N) { val = new c_double[N](); }
    spinor(const spinor& other) : size(other.size) {
        val = new c_double[size];
        std::copy(other.val, other.val + size, val);
    }
    spinor& operator=(const spinor& other) {
        if (this != &other) {
            if (size != other.size) {
                delete[] val;
                size = other.size;
                val = new c_double[size];
            }
            std::copy(other.val, other.val + size, val);
        }
        return *this;
    }
    ~spinor() { delete[] val; }
    inline void clearBuffer() {
        for (int n = 0; n < size; n++) val[n] = 0;
    }
};

void allocate_lattice_arrays();
void free_lattice_arrays();
// Process set-up (src/main.cpp:57-62): widths, halo sizes.  No MPI.
void initializeMPI();

#endif

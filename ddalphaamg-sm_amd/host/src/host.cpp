// Host API of the MI355X DD-alphaAMG library: the reference's classes and free
// functions (variables.h, utils.h, dirac_operator.h, fgmres.h, sap.h, level.h, amg.h,
// params.h, io.h) implemented over the C ABI of libddamg.so (include/ddamg.h).
// Halo-padded host spinors are packed to the boundary format (interior sites, x outer,
// t inner, dof) and back.  The generic FGMRES and SAP_C algorithms over caller-defined
// hooks run on the host, as in the reference; every operator the library provides runs
// on the GPU.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <functional>
#include <random>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>

#include "../../../include/ddamg.h"
#include "amg.h"
#include "conjugate_gradient.h"
#include "io.h"
#include "params.h"
#include "sap.h"

double pi = 3.14159265359;
c_double I_number(0, 1);
long long int FLOPS = 0;
long long int localFLOPS = 0;
namespace iter_counters {
long long int kCycleIt = 0, vCycleIt = 0, CGIt = 0, BiCGIt = 0, GMRESIt = 0, SAPIt = 0, FGMRES_SAPIt = 0;
}
namespace flop_counters {
long long int kCycleFlops = 0, vCycleFlops = 0, kCycleSetUpFlops = 0, vCycleSetUpFlops = 0, CGFlops = 0, BiCGFLops = 0,
              GMRESFlops = 0, SAPFlops = 0, FGMRES_SAPFlops = 0;
}
namespace mass {
double m0 = -0.1;
}
namespace mpi {
int rank = 0, size = 1, maxSize = LV::Ntot, maxSizeH = LV::dof * (LV::Nx + 2) * (LV::Nt + 2),
    sitesH = (LV::Nx + 2) * (LV::Nt + 2), ranks_x = 2, ranks_t = 2, width_x = LV::Nx, width_t = LV::Nt, rank2d = 0;
MPI_Comm cart_comm = MPI_COMM_WORLD;
}
namespace SAPV {
int sap_gmres_restart_length = 5;
int sap_gmres_restarts = 5;
double sap_gmres_tolerance = 1e-3;
double sap_tolerance = 1e-10;
}  // namespace SAPV
namespace AMGV {
int SAP_test_vectors_iterations = 4;
int gmres_restarts_coarse_level = 10;
int gmres_restart_length_coarse_level = 400;
double gmres_tol_coarse_level = 0.1;
int nu1 = 0, nu2 = 2, Nit = 1;
int fgmres_k_cycle_restart_length = 5;
int fgmres_k_cycle_restarts = 2;
double fgmres_k_cycle_tol = 0.1;
int cycle = 0;
}  // namespace AMGV
namespace FGMRESV {
double fgmres_tolerance = 1e-10;
int fgmres_restart_length = 20;
int fgmres_restarts = 50;
}  // namespace FGMRESV
namespace CG {
int max_iter = 10000;
double tol = 1e-10;
}  // namespace CG
namespace BiCG {
int max_iter = 10000;
double tol = 1e-10;
}  // namespace BiCG
namespace LevelV {
int levels = 2, maxLevel = 1;
int* BlocksX = nullptr;
int* BlocksT = nullptr;
int* Ntest = nullptr;
int* SAP_Block_x = nullptr;
int* SAP_Block_t = nullptr;
int* NxSites = nullptr;
int* NtSites = nullptr;
int* DOF = nullptr;
int* Colors = nullptr;
}  // namespace LevelV
namespace DDV {
int device = 0;
int precision = 0;
int max_rhs = 1;
}  // namespace DDV

namespace {
void check(int st, const char* what) {
    if (st != DD_OK && st != DD_ERR_NOT_CONVERGED) {
        std::fprintf(stderr, "libddamg: %s failed (%d): %s\n", what, st, dd_last_error());
        throw std::runtime_error(what);
    }
}

// interior of a halo-padded level array -> boundary format and back
std::vector<double> pack(const spinor& s, int Nx, int Nt, int dof) {
    std::vector<double> out((size_t)2 * Nx * Nt * dof);
    for (int x = 1; x <= Nx; x++)
        for (int t = 1; t <= Nt; t++)
            for (int d = 0; d < dof; d++) {
                const c_double v = s.val[(size_t)(x * (Nt + 2) + t) * dof + d];
                const size_t o = ((size_t)((x - 1) * Nt + (t - 1)) * dof + d) * 2;
                out[o] = v.real();
                out[o + 1] = v.imag();
            }
    return out;
}
void unpack(const std::vector<double>& in, spinor& s, int Nx, int Nt, int dof) {
    for (int x = 1; x <= Nx; x++)
        for (int t = 1; t <= Nt; t++)
            for (int d = 0; d < dof; d++) {
                const size_t o = ((size_t)((x - 1) * Nt + (t - 1)) * dof + d) * 2;
                s.val[(size_t)(x * (Nt + 2) + t) * dof + d] = c_double(in[o], in[o + 1]);
            }
}

dd_ctx* operator_ctx() {   // process-wide context for the free Wilson-Dirac functions
    static dd_ctx* c = nullptr;
    if (!c) check(dd_create(&c, LV::Nx, LV::Nt, DDV::device), "dd_create");
    return c;
}

void apply(const spinor& U, const spinor& phi, spinor& out, double m0, int dagger) {
    dd_ctx* c = operator_ctx();
    std::vector<double> u = pack(U, LV::Nx, LV::Nt, 2), p = pack(phi, LV::Nx, LV::Nt, 2), o(p.size());
    check(dd_set_gauge(c, u.data()), "dd_set_gauge");
    check(dd_apply_D(c, p.data(), o.data(), 1, m0, dagger, 0), "dd_apply_D");
    unpack(o, out, LV::Nx, LV::Nt, 2);
}

dd_params make_params(double m0, int nu1, int nu2, int m, int restarts, double tol, int cycle) {
    dd_params p;
    check(dd_default_params(&p, LevelV::levels), "dd_default_params");
    for (int l = 0; l < LevelV::levels - 1; l++) {
        p.agg_x[l] = LevelV::NxSites[l] / LevelV::BlocksX[l];
        p.agg_t[l] = LevelV::NtSites[l] / LevelV::BlocksT[l];
        p.sap_x[l] = (LevelV::NxSites[l] / mpi::ranks_x) / LevelV::SAP_Block_x[l];
        p.sap_t[l] = (LevelV::NtSites[l] / mpi::ranks_t) / LevelV::SAP_Block_t[l];
        p.nv[l] = LevelV::Ntest[l];
    }
    p.m0 = m0;
    p.nu1 = nu1;
    p.nu2 = nu2;
    p.setup_sap_iters = AMGV::SAP_test_vectors_iterations;
    p.sap_m = SAPV::sap_gmres_restart_length;
    p.sap_restarts = SAPV::sap_gmres_restarts;
    p.sap_tol = SAPV::sap_gmres_tolerance;
    p.coarse_m = AMGV::gmres_restart_length_coarse_level;
    p.coarse_restarts = AMGV::gmres_restarts_coarse_level;
    p.coarse_tol = AMGV::gmres_tol_coarse_level;
    p.kcycle_m = AMGV::fgmres_k_cycle_restart_length;
    p.kcycle_restarts = AMGV::fgmres_k_cycle_restarts;
    p.kcycle_tol = AMGV::fgmres_k_cycle_tol;
    p.fgmres_m = m;
    p.fgmres_restarts = restarts;
    p.tol = tol;
    p.cycle = cycle;
    p.precision = DDV::precision;
    p.rng_ranks_x = mpi::ranks_x;
    p.rng_ranks_t = mpi::ranks_t;
    p.max_rhs = DDV::max_rhs;
    return p;
}
}  // namespace

// ---- global set-up -------------------------------------------------------------------
void allocate_lattice_arrays() {
    const int L = 8;
    LevelV::BlocksX = new int[L]();
    LevelV::BlocksT = new int[L]();
    LevelV::Ntest = new int[L]();
    LevelV::SAP_Block_x = new int[L]();
    LevelV::SAP_Block_t = new int[L]();
    LevelV::NxSites = new int[L]();
    LevelV::NtSites = new int[L]();
    LevelV::DOF = new int[L]();
    LevelV::Colors = new int[L]();
}

void free_lattice_arrays() {
    for (int* p : {LevelV::BlocksX, LevelV::BlocksT, LevelV::Ntest, LevelV::SAP_Block_x, LevelV::SAP_Block_t,
                   LevelV::NxSites, LevelV::NtSites, LevelV::DOF, LevelV::Colors})
        delete[] p;
}

void initializeMPI() {
    mpi::width_x = LV::Nx;
    mpi::width_t = LV::Nt;
    mpi::maxSize = LV::Ntot;
    mpi::maxSizeH = LV::dof * (LV::Nx + 2) * (LV::Nt + 2);
    mpi::sitesH = (LV::Nx + 2) * (LV::Nt + 2);
}

void readParameters(const std::string& inputFile) {
    std::ifstream in(inputFile);
    if (!in) {
        std::cerr << "File " << inputFile << " not found" << std::endl;
        std::exit(1);
    }
    LevelV::maxLevel = LevelV::levels - 1;
    int level, bx, bt, nt, sx, st, count = 1;
    while (in >> level >> bx >> bt >> nt >> sx >> st) {
        LevelV::BlocksX[level] = bx;
        LevelV::BlocksT[level] = bt;
        LevelV::Ntest[level] = nt;
        LevelV::SAP_Block_x[level] = sx;
        LevelV::SAP_Block_t[level] = st;
        count++;
    }
    for (int l = 0; l < LevelV::levels; l++) {
        LevelV::NxSites[l] = (l == 0) ? LV::Nx : LevelV::BlocksX[l - 1];
        LevelV::NtSites[l] = (l == 0) ? LV::Nt : LevelV::BlocksT[l - 1];
        LevelV::Colors[l] = (l == 0) ? 1 : LevelV::Ntest[l - 1];
        LevelV::DOF[l] = 2 * LevelV::Colors[l];
    }
    if (count != LevelV::levels) {
        std::cout << "The number of levels in the inputs file does not match the number of levels assigned on the "
                     "terminal"
                  << std::endl;
        std::exit(1);
    }
}

// ---- I/O ----------------------------------------------------------------------------
void read_binary(const std::string& name, const spinor& U) {
    std::ifstream f(name, std::ios::binary);
    if (!f) {
        std::cerr << "File " << name << " not found " << std::endl;
        std::exit(1);
    }
    for (int x = 1; x <= LV::Nx; x++)
        for (int t = 1; t <= LV::Nt; t++)
            for (int mu = 0; mu < 2; mu++) {
                int xr, tr, mr;
                double re, im;
                f.read(reinterpret_cast<char*>(&xr), sizeof(int));
                f.read(reinterpret_cast<char*>(&tr), sizeof(int));
                f.read(reinterpret_cast<char*>(&mr), sizeof(int));
                f.read(reinterpret_cast<char*>(&re), sizeof(double));
                f.read(reinterpret_cast<char*>(&im), sizeof(double));
                U.val[2 * (x * (LV::Nt + 2) + t) + mu] = c_double(re, im);
            }
}

void write_binary(const std::string& name, const spinor& U) {
    std::ofstream f(name, std::ios::binary);
    for (int x = 1; x <= LV::Nx; x++)
        for (int t = 1; t <= LV::Nt; t++)
            for (int mu = 0; mu < 2; mu++) {
                int xr = x - 1, tr = t - 1, mr = mu;
                const c_double v = U.val[2 * (x * (LV::Nt + 2) + t) + mu];
                double re = v.real(), im = v.imag();
                f.write(reinterpret_cast<const char*>(&xr), sizeof(int));
                f.write(reinterpret_cast<const char*>(&tr), sizeof(int));
                f.write(reinterpret_cast<const char*>(&mr), sizeof(int));
                f.write(reinterpret_cast<const char*>(&re), sizeof(double));
                f.write(reinterpret_cast<const char*>(&im), sizeof(double));
            }
}

// ---- Wilson-Dirac ------------------------------------------------------------------------
void D_phi(const spinor& U, const spinor& phi, spinor& Dphi, const double& m0) { apply(U, phi, Dphi, m0, 0); }
void D_dagger_phi(const spinor& U, const spinor& phi, spinor& Dphi, const double& m0) { apply(U, phi, Dphi, m0, 1); }
void D_D_dagger_phi(const spinor& U, const spinor& phi, spinor& Dphi, const double& m0) {
    spinor buf(mpi::maxSizeH);
    D_phi(U, phi, buf, m0);
    D_dagger_phi(U, buf, Dphi, m0);
}

// ---- generic FGMRES over the hooks (src/fgmres.cpp:5-127) ---------------------------------
FGMRES::FGMRES(const int Ntot, const int dofs, const int maxSizeH, const int x_ini, const int t_ini, const int x_fin,
               const int t_fin, const int m, const int restarts, const double tol)
    : Ntot(Ntot), dofs(dofs), maxSizeH(maxSizeH), x_ini(x_ini), x_fin(x_fin), t_ini(t_ini), t_fin(t_fin), m(m),
      restarts(restarts), tol(tol), Nx(x_fin - x_ini + 1), Nt(t_fin - t_ini + 1) {}

void FGMRES::setZeros() {
    if ((int)VmT.size() != m + 1) {   // buffers on first use (the library's subclasses never need them)
        VmT.assign(m + 1, spinor(maxSizeH));
        ZmT.assign(m, spinor(maxSizeH));
        r = spinor(maxSizeH);
        w = spinor(maxSizeH);
        Dx = spinor(maxSizeH);
    }
    for (auto& v : VmT) v.clearBuffer();
    for (auto& v : ZmT) v.clearBuffer();
    r.clearBuffer();
    w.clearBuffer();
    Dx.clearBuffer();
    Hm.assign(m + 1, c_vector(m, 0.0));
    gm.assign(m + 1, 0.0);
    sn.assign(m, 0.0);
    cn.assign(m, 0.0);
    eta.assign(m, 0.0);
}

int FGMRES::fgmres(const spinor& phi, const spinor& x0, spinor& x, const bool& print_message) {
    setZeros();
    x = x0;
    func(x, Dx);
    axpy(phi, Dx, -1.0, r);
    const double norm_phi = std::sqrt(std::real(dot(phi.val, phi.val)));
    double err = std::sqrt(std::real(dot(r.val, r.val)));
    localFLOPS += 2 * dsq;
    // a lane of the interior: halo-padded site (nx, nt), entry mu
    auto at = [&](int nx, int nt, int mu) { return dofs * (nx * (Nt + 2) + nt) + mu; };
    int maxIt = m;   // kept across restarts, as the reference does
    for (int k = 0; k < restarts; k++) {
        const c_double beta = err;
        scal(1.0 / beta, r, VmT[0]);
        gm[0] = beta;
        localFLOPS += dd + dcm + dcd;
        for (int j = 0; j < m; j++) {
            preconditioner(VmT[j], ZmT[j]);
            func(ZmT[j], w);
            for (int i = 0; i <= j; i++) {   // modified Gram-Schmidt
                Hm[i][j] = dot(w.val, VmT[i].val);
                for (int nx = x_ini; nx <= x_fin; nx++)
                    for (int nt = t_ini; nt <= t_fin; nt++)
                        for (int mu = 0; mu < dofs; mu++) w.val[at(nx, nt, mu)] -= Hm[i][j] * VmT[i].val[at(nx, nt, mu)];
                localFLOPS += (ca + cm) * (long long)Nx * Nt * dofs;
            }
            Hm[j + 1][j] = std::sqrt(std::real(dot(w.val, w.val)));
            localFLOPS += dsq;
            if (std::real(Hm[j + 1][j]) > 0) {   // breakdown guard (src/fgmres.cpp:44-47)
                scal(1.0 / Hm[j + 1][j], w, VmT[j + 1]);
                localFLOPS += dcd;
            }
            rotation(j);
            gm[j + 1] = -sn[j] * gm[j];
            gm[j] = std::conj(cn[j]) * gm[j];
            localFLOPS += 2 * cm;
            if (std::abs(gm[j + 1]) < tol * norm_phi) {
                maxIt = j + 1;
                break;
            }
        }
        solve_upper_triangular(Hm, gm, maxIt, eta);
        for (int nx = x_ini; nx <= x_fin; nx++)
            for (int nt = t_ini; nt <= t_fin; nt++)
                for (int mu = 0; mu < dofs; mu++)
                    for (int j = 0; j < maxIt; j++) x.val[at(nx, nt, mu)] += eta[j] * ZmT[j].val[at(nx, nt, mu)];
        localFLOPS += (ca + cm) * (long long)Nx * Nt * dofs * maxIt;
        func(x, Dx);
        axpy(phi, Dx, -1.0, r);
        err = std::sqrt(std::real(dot(r.val, r.val)));
        localFLOPS += dsq;
        if (err < tol * norm_phi) {
            if (print_message && mpi::rank2d == 0) {
                std::cout << "FGMRES converged in " << k + 1 << " cycles Error " << err << std::endl;
                std::cout << "With " << k * m + maxIt << " iterations" << std::endl;
            }
            return k * m + maxIt;
        }
    }
    if (print_message && mpi::rank2d == 0)
        std::cout << "FGMRES did not converge in " << restarts << " restarts of lenght " << m << " Error " << err
                  << std::endl;
    return restarts * m;
}

void FGMRES::rotation(const int& j) {
    for (int i = 0; i < j; i++) {
        const c_double temp = std::conj(cn[i]) * Hm[i][j] + std::conj(sn[i]) * Hm[i + 1][j];
        Hm[i + 1][j] = -sn[i] * Hm[i][j] + cn[i] * Hm[i + 1][j];
        Hm[i][j] = temp;
    }
    const c_double den = std::sqrt(std::real(std::conj(Hm[j][j]) * Hm[j][j] + std::conj(Hm[j + 1][j]) * Hm[j + 1][j]));
    sn[j] = Hm[j + 1][j] / den;
    cn[j] = Hm[j][j] / den;
    Hm[j][j] = std::conj(cn[j]) * Hm[j][j] + std::conj(sn[j]) * Hm[j + 1][j];
    Hm[j + 1][j] = 0.0;
    localFLOPS += (ca + 2 * cm) * 2 * j + dsq + 2 * (ca + 2 * cm) + 2 * cd;
}

void FGMRES::solve_upper_triangular(const c_matrix& A, const c_vector& b, const int& n, c_vector& out) {
    for (int i = n - 1; i >= 0; i--) {
        out[i] = b[i];
        for (int j = i + 1; j < n; j++) out[i] -= A[i][j] * out[j];
        out[i] /= A[i][i];
    }
    localFLOPS += (ca + cm) * (long long)n * (n - 1) / 2 + cd * n;
}

// ---- generic SAP over the hooks (src/sap.cpp:3-165, include/sap.h:14-126) -----------------
c_double SAP_C::GMRES_D_B::dot(c_double* A, c_double* B) {
    c_double z = 0;
    for (int x = 1; x <= xfin; x++)
        for (int t = 1; t <= tfin; t++)
            for (int d = 0; d < nd; d++) {
                const int i = nd * (x * (tfin + 2) + t) + d;
                z += A[i] * std::conj(B[i]);
            }
    localFLOPS += (ca + cm) * (long long)xfin * tfin * nd;
    return z;
}

void SAP_C::GMRES_D_B::axpy(const spinor& X, const spinor& Y, const c_double& lambda, spinor& out) {
    for (int x = 1; x <= xfin; x++)
        for (int t = 1; t <= tfin; t++)
            for (int d = 0; d < nd; d++) {
                const int i = nd * (x * (tfin + 2) + t) + d;
                out.val[i] = X.val[i] + lambda * Y.val[i];
            }
    localFLOPS += (ca + cm) * (long long)xfin * tfin * nd;
}

void SAP_C::GMRES_D_B::scal(const c_double& lambda, const spinor& X, spinor& Y) {
    for (int x = 1; x <= xfin; x++)
        for (int t = 1; t <= tfin; t++)
            for (int d = 0; d < nd; d++) {
                const int i = nd * (x * (tfin + 2) + t) + d;
                Y.val[i] = lambda * X.val[i];
            }
    localFLOPS += cm * (long long)xfin * tfin * nd;
}

SAP_C::SAP_C(const int Nx_, const int Nt_, const int block_x, const int block_t, const int spins_, const int colors_)
    : gmres_DB(Nx_ / block_x * Nt_ / block_t, spins_ * colors_, (Nx_ / block_x + 2) * (Nt_ / block_t + 2) * spins_ * colors_,
               1, 1, Nx_ / block_x, Nt_ / block_t, SAPV::sap_gmres_restart_length, SAPV::sap_gmres_restarts,
               SAPV::sap_gmres_tolerance, this),
      Nt(Nt_), Nx(Nx_), Block_x(block_x), Block_t(block_t), colors(colors_), spins(spins_) {
    x_elements = Nx / block_x;
    t_elements = Nt / block_t;
    NBlocks = Block_x * Block_t;
    dofs = spins * colors;
    Ntot_original_w_halo = (Nx + 2) * (Nt + 2);
    Ntot_no_halo = x_elements * t_elements;
    Ntot_w_halo = (x_elements + 2) * (t_elements + 2);
    Nvars_no_halo = Ntot_no_halo * dofs;
    Nvars_w_halo = Ntot_w_halo * dofs;
    coloring_blocks = (NBlocks != 1) ? NBlocks / 2 : 1;
    Blocks.assign(NBlocks, std::vector<int>(Ntot_w_halo, 0));
    RedBlocks.assign(coloring_blocks, 0);
    BlackBlocks.assign(coloring_blocks, 0);
    SchwarzBlocks();
}

// block b = bx*Block_t + bt covers sites [bx*x_elements, (bx+1)*x_elements) x [...);
// red when bx + bt is even (the reference's rule at src/sap.cpp:28-38 for both parities
// of Block_t), each colour listed in ascending block order
void SAP_C::SchwarzBlocks() {
    int nred = 0, nblack = 0;
    for (int bx = 0; bx < Block_x; bx++)
        for (int bt = 0; bt < Block_t; bt++) {
            const int block = bx * Block_t + bt;
            for (int mx = 1; mx <= x_elements; mx++)
                for (int mt = 1; mt <= t_elements; mt++)
                    Blocks[block][mx * (t_elements + 2) + mt] =
                        (bx * x_elements + mx) * (Nt + 2) + (bt * t_elements + mt);
            if ((bx + bt) % 2 == 0) {
                if (nred < coloring_blocks) RedBlocks[nred++] = block;
            } else if (nblack < coloring_blocks) {
                BlackBlocks[nblack++] = block;
            }
        }
}

void SAP_C::I_D_B_1_It(const spinor& v, spinor& x, const int& block) {
    spinor vb(Nvars_w_halo), xb(Nvars_w_halo);
    for (int mx = 1; mx <= x_elements; mx++)
        for (int mt = 1; mt <= t_elements; mt++) {
            const int m = mx * (t_elements + 2) + mt;
            for (int d = 0; d < dofs; d++) vb.val[dofs * m + d] = v.val[dofs * Blocks[block][m] + d];
        }
    current_block = block;
    gmres_DB.fgmres(vb, vb, xb, false);   // initial guess = the block right-hand side
    for (int mx = 1; mx <= x_elements; mx++)
        for (int mt = 1; mt <= t_elements; mt++) {
            const int m = mx * (t_elements + 2) + mt;
            for (int d = 0; d < dofs; d++) x.val[dofs * Blocks[block][m] + d] = xb.val[dofs * m + d];
        }
}

int SAP_C::SAP(const spinor& v, spinor& x, const int& nu_, const double& tol_, const bool print) {
    nu = nu_;
    tol = tol_;
    const double v_norm = std::sqrt(std::real(dot(v, v)));
    spinor DB_1_r(dofs * Ntot_original_w_halo), r(dofs * Ntot_original_w_halo), Dphi(dofs * Ntot_original_w_halo);
    auto residual = [&] {   // r = v - D x
        funcGlobal(x, Dphi);
        for (int nx = 1; nx <= Nx; nx++)
            for (int nt = 1; nt <= Nt; nt++)
                for (int d = 0; d < dofs; d++) {
                    const int i = dofs * (nx * (Nt + 2) + nt) + d;
                    r.val[i] = v.val[i] - Dphi.val[i];
                }
        localFLOPS += ca * (long long)Nx * Nt * dofs;
    };
    auto sweep = [&](const std::vector<int>& colour) {   // additive within a colour
        for (int block : colour) {
            I_D_B_1_It(r, DB_1_r, block);
            for (int mx = 1; mx <= x_elements; mx++)
                for (int mt = 1; mt <= t_elements; mt++) {
                    const int m = mx * (t_elements + 2) + mt;
                    for (int d = 0; d < dofs; d++) x.val[dofs * Blocks[block][m] + d] += DB_1_r.val[dofs * Blocks[block][m] + d];
                }
        }
    };
    residual();
    double err = 0;
    for (int i = 0; i < nu; i++) {
        sweep(RedBlocks);
        residual();
        sweep(BlackBlocks);
        residual();
        err = std::sqrt(std::real(dot(r, r)));
        if (err < tol * v_norm) {
            if (print && mpi::rank2d == 0) std::cout << "SAP converged in " << i << " iterations, error: " << err << std::endl;
            return i;
        }
    }
    if (print && mpi::rank2d == 0) std::cout << "SAP didn't converge in " << nu << " iterations, error: " << err << std::endl;
    return nu;
}

// Wilson-Dirac operator of one block, neighbours outside it zero (the block never wraps
// around the lattice, so the antiperiodic sign never enters)
void SAP_fine_level::D_B(const spinor& Ug, const spinor& v, spinor& xo, const double& mass, const int& block) {
    const int W = t_elements + 2;
    const c_double I(0, 1);
    for (int mx = 1; mx <= x_elements; mx++)
        for (int mt = 1; mt <= t_elements; mt++) {
            const int m = mx * W + mt, n = Blocks[block][m];
            const int fp0 = m + 1, fp1 = m + W, bm0 = m - 1, bm1 = m - W;   // t+1, x+1, t-1, x-1 in the block
            const c_double u0 = Ug.val[2 * n], u1 = Ug.val[2 * n + 1];
            // U_mu(n - mu) of the backward neighbours (zero contribution when outside)
            const c_double ub0 = (mt > 1) ? std::conj(Ug.val[2 * (n - 1)]) : 0.0;
            const c_double ub1 = (mx > 1) ? std::conj(Ug.val[2 * (n - (Nt + 2)) + 1]) : 0.0;
            const c_double* a = v.val;
            xo.val[2 * m] = (mass + 2) * a[2 * m] -
                            0.5 * (u0 * (a[2 * fp0] - a[2 * fp0 + 1]) + u1 * (a[2 * fp1] + I * a[2 * fp1 + 1]) +
                                   ub0 * (a[2 * bm0] + a[2 * bm0 + 1]) + ub1 * (a[2 * bm1] - I * a[2 * bm1 + 1]));
            xo.val[2 * m + 1] = (mass + 2) * a[2 * m + 1] -
                                0.5 * (u0 * (-a[2 * fp0] + a[2 * fp0 + 1]) + u1 * (-I * a[2 * fp1] + a[2 * fp1 + 1]) +
                                       ub0 * (a[2 * bm0] + a[2 * bm0 + 1]) + ub1 * (I * a[2 * bm1] + a[2 * bm1 + 1]));
        }
    localFLOPS += 2 * 81 * (long long)x_elements * t_elements;
}

// ---- Level ------------------------------------------------------------------------------
namespace {
int level_shape(dd_ctx* c, int l, int which) {
    int nx = 0, nt = 0, dof = 0;
    check(dd_level_shape(c, l, &nx, &nt, &dof), "dd_level_shape");
    return which == 0 ? nx : which == 1 ? nt : dof;
}

// The hierarchy that Level(l, U) objects share: storage for the LevelV levels on the
// gauge field U, built step by step by the Levels' own methods.
struct SharedLevels {
    dd_ctx* ctx = nullptr;
    std::vector<double> u;   // the field it was made for
    int users = 0;
};
SharedLevels& shared_levels() {
    static SharedLevels s;
    return s;
}

dd_ctx* acquire_shared(const spinor& U) {
    SharedLevels& sh = shared_levels();
    std::vector<double> u = pack(U, LV::Nx, LV::Nt, 2);
    if (sh.ctx && u != sh.u) {
        if (sh.users > 0) {
            std::cerr << "Level(l, U): the Levels alive were built on another gauge field" << std::endl;
            std::exit(1);
        }
        dd_destroy(sh.ctx);
        sh.ctx = nullptr;
    }
    if (!sh.ctx) {
        check(dd_create(&sh.ctx, LV::Nx, LV::Nt, DDV::device), "dd_create");
        check(dd_set_gauge(sh.ctx, u.data()), "dd_set_gauge");
        dd_params p = make_params(mass::m0, AMGV::nu1, AMGV::nu2, FGMRESV::fgmres_restart_length,
                                  FGMRESV::fgmres_restarts, FGMRESV::fgmres_tolerance, AMGV::cycle);
        check(dd_setup_storage(sh.ctx, &p), "dd_setup_storage");
        sh.u = std::move(u);
    }
    sh.users++;
    return sh.ctx;
}

void release_shared(dd_ctx* c) {
    SharedLevels& sh = shared_levels();
    if (c == sh.ctx && --sh.users == 0) {
        dd_destroy(sh.ctx);
        sh.ctx = nullptr;
        sh.u.clear();
    }
}

int lv_nx(int l) { return LevelV::NxSites[l]; }
int lv_nt(int l) { return LevelV::NtSites[l]; }
int lv_ntest(int l) { return l < LevelV::maxLevel ? LevelV::Ntest[l] : 1; }
int lv_bx(int l) { return l < LevelV::maxLevel ? LevelV::BlocksX[l] : 1; }
int lv_bt(int l) { return l < LevelV::maxLevel ? LevelV::BlocksT[l] : 1; }
}  // namespace

Level::Level(const int& l, const spinor& Ug)
    : level(l), U(Ug), Nx(lv_nx(l)), Nt(lv_nt(l)), Ntot(Nx * Nt), colors(LevelV::Colors[l]), DOF(2 * colors),
      Ntest(lv_ntest(l)), xblocks_per_rank(lv_bx(l)), tblocks_per_rank(lv_bt(l)),
      blocks_per_rank(xblocks_per_rank * tblocks_per_rank), x_elements(Nx / xblocks_per_rank),
      t_elements(Nt / tblocks_per_rank), sites_per_block(x_elements * t_elements), NBlocks(blocks_per_rank),
      G1(0), G2(0), G3(0), ctx(acquire_shared(Ug)), owns_tvec(true) {
    tvec.assign(Ntest, spinor((Nx + 2) * (Nt + 2) * DOF));
    if (level == 0) makeDirac();
    if (level == LevelV::maxLevel) gmres_l = new GMRES_level_l(this);
    sap_l = new SAP_level_l(this);
}

Level::Level(const int& l, dd_ctx* c)
    : level(l), U(0), Nx(level_shape(c, l, 0)), Nt(level_shape(c, l, 1)), Ntot(Nx * Nt),
      colors(level_shape(c, l, 2) / 2), DOF(level_shape(c, l, 2)), Ntest(lv_ntest(l)),
      xblocks_per_rank(lv_bx(l)), tblocks_per_rank(lv_bt(l)), blocks_per_rank(xblocks_per_rank * tblocks_per_rank),
      x_elements(Nx / xblocks_per_rank), t_elements(Nt / tblocks_per_rank), sites_per_block(x_elements * t_elements),
      NBlocks(blocks_per_rank), G1(0), G2(0), G3(0), ctx(c), owns_tvec(false) {
    if (level == LevelV::maxLevel) gmres_l = new GMRES_level_l(this);
    sap_l = new SAP_level_l(this);
}

Level::~Level() {
    delete sap_l;
    delete gmres_l;
    if (owns_tvec) release_shared(ctx);
}

void Level::upload_tvec() {
    if (!owns_tvec || level >= LevelV::maxLevel) return;
    const size_t one = (size_t)2 * Nx * Nt * DOF;
    std::vector<double> t(one * Ntest);
    for (int cc = 0; cc < Ntest; cc++) {
        std::vector<double> p = pack(tvec[cc], Nx, Nt, DOF);
        std::copy(p.begin(), p.end(), t.begin() + cc * one);
    }
    check(dd_set_test_vectors(ctx, level, t.data(), DDV::precision), "dd_set_test_vectors");
}

void Level::D_operator(const spinor& v, spinor& out) {
    std::vector<double> a = pack(v, Nx, Nt, DOF), b(a.size());
    check(dd_level_apply(ctx, level, a.data(), b.data(), 1, DDV::precision), "dd_level_apply");
    unpack(b, out, Nx, Nt, DOF);
}

void Level::Pdagg_v(const spinor& v, spinor& out) {
    upload_tvec();
    const int nx = level_shape(ctx, level + 1, 0), nt = level_shape(ctx, level + 1, 1), d = level_shape(ctx, level + 1, 2);
    std::vector<double> a = pack(v, Nx, Nt, DOF), b((size_t)2 * nx * nt * d);
    check(dd_restrict(ctx, level, a.data(), b.data(), 1, DDV::precision), "dd_restrict");
    for (int i = 0; i < out.size; i++) out.val[i] = 0;
    unpack(b, out, nx, nt, d);
}

void Level::P_vc(const spinor& vc, spinor& out) {
    upload_tvec();
    const int nx = level_shape(ctx, level + 1, 0), nt = level_shape(ctx, level + 1, 1), d = level_shape(ctx, level + 1, 2);
    std::vector<double> a = pack(vc, nx, nt, d), b((size_t)2 * Nx * Nt * DOF);
    check(dd_prolong(ctx, level, a.data(), b.data(), 1, DDV::precision), "dd_prolong");
    for (int i = 0; i < out.size; i++) out.val[i] = 0;
    unpack(b, out, Nx, Nt, DOF);
}

void Level::fetchTestVectors() {
    std::vector<double> t((size_t)2 * Ntest * Nx * Nt * DOF);
    check(dd_get_test_vectors(ctx, level, t.data(), DDV::precision), "dd_get_test_vectors");
    tvec.assign(Ntest, spinor((Nx + 2) * (Nt + 2) * DOF));
    for (int cc = 0; cc < Ntest; cc++) {
        std::vector<double> one(t.begin() + (size_t)cc * 2 * Nx * Nt * DOF, t.begin() + (size_t)(cc + 1) * 2 * Nx * Nt * DOF);
        unpack(one, tvec[cc], Nx, Nt, DOF);
    }
}

void Level::orthonormalize() {
    upload_tvec();
    check(dd_orthonormalize(ctx, level, DDV::precision), "dd_orthonormalize");
    fetchTestVectors();
}

// links of this level's operator from the device into the reference's layout
void Level::fetch_links(spinor& g1, spinor& g2, spinor& g3) {
    const int d = DOF, c = colors, H = (Nx + 2) * (Nt + 2);
    std::vector<double> L((size_t)2 * Nx * Nt * 5 * d * d);
    check(dd_get_links(ctx, level, L.data(), DDV::precision), "dd_get_links");
    g1 = spinor(H * 4 * c * c);
    g2 = spinor(H * 8 * c * c);
    g3 = spinor(H * 8 * c * c);
    for (int x = 1; x <= Nx; x++)
        for (int t = 1; t <= Nt; t++) {
            const int n = x * (Nt + 2) + t;
            const size_t site = (size_t)(x - 1) * Nt + (t - 1);
            for (int alf = 0; alf < 2; alf++)
                for (int bet = 0; bet < 2; bet++)
                    for (int cc = 0; cc < c; cc++)
                        for (int b = 0; b < c; b++) {
                            const int row = cc * 2 + alf, col = b * 2 + bet;
                            auto get = [&](int k) {
                                const size_t o = (((site * 5 + k) * d + row) * d + col) * 2;
                                return c_double(L[o], L[o + 1]);
                            };
                            const int i1 = (((n * 2 + alf) * 2 + bet) * c + cc) * c + b;
                            g1.val[i1] = get(0);
                            for (int mu = 0; mu < 2; mu++) {
                                g2.val[i1 * 2 + mu] = get(1 + mu);
                                g3.val[i1 * 2 + mu] = get(3 + mu);
                            }
                        }
        }
}

void Level::makeDirac() {
    if (level == 0) fetch_links(G1, G2, G3);
}

void Level::makeCoarseLinks(Level& next) {
    upload_tvec();
    check(dd_make_coarse_links(ctx, level, DDV::precision), "dd_make_coarse_links");
    next.fetch_links(next.G1, next.G2, next.G3);
}

void Level::checkOrthogonality() {
    if (owns_tvec) upload_tvec();
    fetchTestVectors();
    bool ok = true;
    for (int X = 0; X < xblocks_per_rank; X++)
        for (int T = 0; T < tblocks_per_rank; T++)
            for (int s = 0; s < 2; s++)
                for (int a = 0; a < Ntest; a++)
                    for (int b = 0; b < Ntest; b++) {
                        c_double d = 0;
                        for (int x = X * x_elements + 1; x <= (X + 1) * x_elements; x++)
                            for (int t = T * t_elements + 1; t <= (T + 1) * t_elements; t++)
                                for (int c = 0; c < colors; c++) {
                                    const int i = (x * (Nt + 2) + t) * DOF + c * 2 + s;
                                    d += std::conj(tvec[a].val[i]) * tvec[b].val[i];
                                }
                        if (std::abs(d - (a == b ? 1.0 : 0.0)) > 1e-8) ok = false;
                    }
    std::cout << "Test vectors on level " << level << (ok ? " are orthonormalized" : " are NOT orthonormal")
              << std::endl;
}

// SAP blocks per (virtual) rank from the parameters file; the one process holds all ranks'
// blocks, so each block keeps the sites it has in the reference on that rank grid
Level::SAP_level_l::SAP_level_l(Level* p)
    : SAP_C(p->Nx, p->Nt,
            p->level < LevelV::maxLevel ? std::max(1, LevelV::SAP_Block_x[p->level] * mpi::ranks_x) : 1,
            p->level < LevelV::maxLevel ? std::max(1, LevelV::SAP_Block_t[p->level] * mpi::ranks_t) : 1, 2, p->colors),
      parent(p) {}

int Level::SAP_level_l::SAP(const spinor& v, spinor& x, const int& nu_, const double& tol_, const bool print) {
    Level* p = parent;
    std::vector<double> a = pack(v, p->Nx, p->Nt, p->DOF), b = pack(x, p->Nx, p->Nt, p->DOF);
    check(dd_sap(p->ctx, p->level, a.data(), b.data(), nu_, tol_, 1, DDV::precision), "dd_sap");
    unpack(b, x, p->Nx, p->Nt, p->DOF);
    if (print && mpi::rank2d == 0) std::cout << "SAP ran " << nu_ << " iterations on level " << p->level << std::endl;
    return nu_;
}

// the level operator restricted to block current_block: I_B^T D I_B (the stencil's
// neighbours outside the block are zero, as in D_local, src/level.cpp:607-653)
void Level::SAP_level_l::funcLocal(const spinor& in, spinor& out) {
    Level* p = parent;
    spinor full((p->Nx + 2) * (p->Nt + 2) * p->DOF), Dfull(full.size);
    const std::vector<int>& blk = Blocks[current_block];
    for (int mx = 1; mx <= x_elements; mx++)
        for (int mt = 1; mt <= t_elements; mt++) {
            const int m = mx * (t_elements + 2) + mt;
            for (int d = 0; d < dofs; d++) full.val[dofs * blk[m] + d] = in.val[dofs * m + d];
        }
    p->D_operator(full, Dfull);
    out.clearBuffer();
    for (int mx = 1; mx <= x_elements; mx++)
        for (int mt = 1; mt <= t_elements; mt++) {
            const int m = mx * (t_elements + 2) + mt;
            for (int d = 0; d < dofs; d++) out.val[dofs * m + d] = Dfull.val[dofs * blk[m] + d];
        }
}

c_double Level::SAP_level_l::dot(const spinor& X, const spinor& Y) {
    Level* p = parent;
    c_double z = 0;
    for (int x = 1; x <= p->Nx; x++)
        for (int t = 1; t <= p->Nt; t++)
            for (int mu = 0; mu < p->DOF; mu++) {
                const int i = (x * (p->Nt + 2) + t) * p->DOF + mu;
                z += X.val[i] * std::conj(Y.val[i]);
            }
    return z;
}

Level::GMRES_level_l::GMRES_level_l(Level* p)
    : FGMRES(p->Ntot, p->DOF, (p->Nx + 2) * (p->Nt + 2) * p->DOF, 1, 1, p->Nx, p->Nt,
             AMGV::gmres_restart_length_coarse_level, AMGV::gmres_restarts_coarse_level, AMGV::gmres_tol_coarse_level),
      parent(p) {}

c_double Level::GMRES_level_l::dot(c_double* A, c_double* B) {
    c_double z = 0;
    for (int x = 1; x <= Nx; x++)
        for (int t = 1; t <= Nt; t++)
            for (int mu = 0; mu < dofs; mu++) {
                const int i = (x * (Nt + 2) + t) * dofs + mu;
                z += A[i] * std::conj(B[i]);
            }
    localFLOPS += (ca + cm) * (long long)Nx * Nt * dofs;
    return z;
}

void Level::GMRES_level_l::axpy(const spinor& X, const spinor& Y, const c_double& lambda, spinor& out) {
    for (int x = 1; x <= Nx; x++)
        for (int t = 1; t <= Nt; t++)
            for (int mu = 0; mu < dofs; mu++) {
                const int i = (x * (Nt + 2) + t) * dofs + mu;
                out.val[i] = X.val[i] + lambda * Y.val[i];
            }
    localFLOPS += (ca + cm) * (long long)Nx * Nt * dofs;
}

void Level::GMRES_level_l::scal(const c_double& lambda, const spinor& X, spinor& Y) {
    for (int x = 1; x <= Nx; x++)
        for (int t = 1; t <= Nt; t++)
            for (int mu = 0; mu < dofs; mu++) {
                const int i = (x * (Nt + 2) + t) * dofs + mu;
                Y.val[i] = lambda * X.val[i];
            }
    localFLOPS += cm * (long long)Nx * Nt * dofs;
}

// ---- AlgebraicMG ------------------------------------------------------------------------
AlgebraicMG::FGMRES_k_cycle::FGMRES_k_cycle(const int& m_, const int& restarts_, const double& tol_, AlgebraicMG* p,
                                            int l_)
    : FGMRES(p->levels[l_]->Ntot, p->levels[l_]->DOF, (p->levels[l_]->Nx + 2) * (p->levels[l_]->Nt + 2) * p->levels[l_]->DOF,
             1, 1, p->levels[l_]->Nx, p->levels[l_]->Nt, m_, restarts_, tol_),
      parent(p), l(l_) {}

void AlgebraicMG::FGMRES_k_cycle::preconditioner(const spinor& in, spinor& out) { parent->k_cycle(l, in, out); }

c_double AlgebraicMG::FGMRES_k_cycle::dot(c_double* A, c_double* B) {
    c_double z = 0;
    for (int x = 1; x <= Nx; x++)
        for (int t = 1; t <= Nt; t++)
            for (int mu = 0; mu < dofs; mu++) {
                const int i = (x * (Nt + 2) + t) * dofs + mu;
                z += A[i] * std::conj(B[i]);
            }
    localFLOPS += (ca + cm) * (long long)Nx * Nt * dofs;
    return z;
}

void AlgebraicMG::FGMRES_k_cycle::axpy(const spinor& X, const spinor& Y, const c_double& lambda, spinor& out) {
    for (int x = 1; x <= Nx; x++)
        for (int t = 1; t <= Nt; t++)
            for (int mu = 0; mu < dofs; mu++) {
                const int i = (x * (Nt + 2) + t) * dofs + mu;
                out.val[i] = X.val[i] + lambda * Y.val[i];
            }
    localFLOPS += (ca + cm) * (long long)Nx * Nt * dofs;
}

void AlgebraicMG::FGMRES_k_cycle::scal(const c_double& lambda, const spinor& X, spinor& Y) {
    for (int x = 1; x <= Nx; x++)
        for (int t = 1; t <= Nt; t++)
            for (int mu = 0; mu < dofs; mu++) {
                const int i = (x * (Nt + 2) + t) * dofs + mu;
                Y.val[i] = lambda * X.val[i];
            }
    localFLOPS += cm * (long long)Nx * Nt * dofs;
}

AlgebraicMG::AlgebraicMG(const spinor& Ug, const double& m0_, const int& nu1_, const int& nu2_)
    : ctx(nullptr), U(Ug), m0(m0_), nu1(nu1_), nu2(nu2_) {
    check(dd_create(&ctx, LV::Nx, LV::Nt, DDV::device), "dd_create");
    std::vector<double> u = pack(U, LV::Nx, LV::Nt, 2);
    check(dd_set_gauge(ctx, u.data()), "dd_set_gauge");
}

AlgebraicMG::~AlgebraicMG() {
    for (auto* f : fgmres_k_cycle_l) delete f;
    for (Level* l : levels) delete l;
    dd_destroy(ctx);
}

// Nit counts the setup passes (AMGV::Nit, default 1): the first is the reference's setup,
// each further one a bootstrap pass (dd_params::bootstrap).  The reference declares the
// knob but does not use it (src/amg.cpp:4).
void AlgebraicMG::setUpPhase(const int& Nit) {
    dd_params p = make_params(m0, nu1, nu2, fgmres_m, fgmres_restarts, fgmres_tol, cycle);
    p.bootstrap = Nit > 1 ? Nit - 1 : 0;
    check(dd_setup(ctx, &p), "dd_setup");
    for (auto* f : fgmres_k_cycle_l) delete f;
    fgmres_k_cycle_l.clear();
    for (Level* l : levels) delete l;
    levels.clear();
    for (int l = 0; l < LevelV::levels; l++) levels.push_back(new Level(l, ctx));
    for (int l = 0; l < LevelV::levels; l++)
        fgmres_k_cycle_l.push_back(new FGMRES_k_cycle(AMGV::fgmres_k_cycle_restart_length,
                                                      AMGV::fgmres_k_cycle_restarts, AMGV::fgmres_k_cycle_tol, this, l));
    setup_done = true;
    if (mpi::rank2d == 0) std::cout << "Set-up phase finished" << std::endl;
}

void AlgebraicMG::k_cycle(const int& l, const spinor& eta_l, spinor& psi_l) {
    Level* lv = levels[l];
    std::vector<double> a = pack(eta_l, lv->Nx, lv->Nt, lv->DOF), b = pack(psi_l, lv->Nx, lv->Nt, lv->DOF);
    check(dd_cycle(ctx, l, 1, a.data(), b.data(), 1, DDV::precision), "dd_cycle");
    unpack(b, psi_l, lv->Nx, lv->Nt, lv->DOF);
}

void AlgebraicMG::v_cycle(const int& l, const spinor& eta_l, spinor& psi_l) {
    Level* lv = levels[l];
    std::vector<double> a = pack(eta_l, lv->Nx, lv->Nt, lv->DOF), b = pack(psi_l, lv->Nx, lv->Nt, lv->DOF);
    check(dd_cycle(ctx, l, 0, a.data(), b.data(), 1, DDV::precision), "dd_cycle");
    unpack(b, psi_l, lv->Nx, lv->Nt, lv->DOF);
}

void AlgebraicMG::applyMultilevel(const int& it, const spinor& rhs, spinor& out, const double tol,
                                  const bool print_message) {
    spinor Dx(mpi::maxSizeH);
    double nb = 0;
    for (int x = 1; x <= LV::Nx; x++)
        for (int t = 1; t <= LV::Nt; t++)
            for (int d = 0; d < 2; d++) nb += std::norm(rhs.val[idx(x, t, d)]);
    nb = std::sqrt(nb);
    double err = 0;
    for (int i = 0; i < it; i++) {
        if (AMGV::cycle == 1) k_cycle(0, rhs, out);
        else v_cycle(0, rhs, out);
        levels[0]->D_operator(out, Dx);
        err = 0;
        for (int x = 1; x <= LV::Nx; x++)
            for (int t = 1; t <= LV::Nt; t++)
                for (int d = 0; d < 2; d++) err += std::norm(rhs.val[idx(x, t, d)] - Dx.val[idx(x, t, d)]);
        err = std::sqrt(err);
        if (err < tol * nb) {
            if (print_message) std::cout << "cycle converged in " << i + 1 << " cycles Error " << err << std::endl;
            return;
        }
    }
    if (print_message) std::cout << "cycle did not converge in " << it << " cycles Error " << err << std::endl;
}

void AlgebraicMG::testSetUp() {
    std::mt19937 g(1);
    std::uniform_real_distribution<double> d(0, 2 * M_PI);
    for (int l = 0; l < LevelV::maxLevel; l++) {
        Level* f = levels[l];
        Level* c = levels[l + 1];
        spinor vc((c->Nx + 2) * (c->Nt + 2) * c->DOF), o1(vc.size), o2(vc.size), v((f->Nx + 2) * (f->Nt + 2) * f->DOF),
            tmp(v.size);
        for (int x = 1; x <= c->Nx; x++)
            for (int t = 1; t <= c->Nt; t++)
                for (int k = 0; k < c->DOF; k++) vc.val[(x * (c->Nt + 2) + t) * c->DOF + k] = std::polar(1.0, d(g));
        c->D_operator(vc, o1);
        f->P_vc(vc, v);
        f->D_operator(v, tmp);
        f->Pdagg_v(tmp, o2);
        double diff = 0;
        for (int i = 0; i < vc.size; i++) diff = std::max(diff, std::abs(o1.val[i] - o2.val[i]));
        std::cout << (diff < 1e-8 ? "Dc coincides with P^+ D P at level " : "Dc differs from P^+ D P at level ")
                  << l + 1 << " (max diff " << diff << ")" << std::endl;
    }
}

void AlgebraicMG::testSAP() {
    std::mt19937 g(2);
    std::uniform_real_distribution<double> d(0, 2 * M_PI);
    for (int l = 0; l < LevelV::maxLevel; l++) {
        Level* lv = levels[l];
        spinor rhs((lv->Nx + 2) * (lv->Nt + 2) * lv->DOF), x(rhs.size), Dx(rhs.size);
        for (int xx = 1; xx <= lv->Nx; xx++)
            for (int t = 1; t <= lv->Nt; t++)
                for (int k = 0; k < lv->DOF; k++) rhs.val[(xx * (lv->Nt + 2) + t) * lv->DOF + k] = std::polar(1.0, d(g));
        lv->sap_l->SAP(rhs, x, 500, 1e-12, false);
        lv->D_operator(x, Dx);
        double diff = 0;
        for (int i = 0; i < rhs.size; i++) diff = std::max(diff, std::abs(Dx.val[i] - rhs.val[i]));
        std::cout << "D_operator (SAP_l rhs) = rhs at level " << l << (diff < 1e-8 ? ": solution is correct" : ": FAILED")
                  << " (max diff " << diff << ")" << std::endl;
    }
}

// ---- FGMRES_AMG -------------------------------------------------------------------------
FGMRES_AMG::FGMRES_AMG(const spinor& Ug, const int& m, const int& restarts, const double& tol, const int nu1,
                       const int nu2, const double& mass, int cycle)
    : FGMRES(LV::Ntot, LV::dof, mpi::maxSizeH, 1, 1, mpi::width_x, mpi::width_t, m, restarts, tol),
      amg(Ug, mass, nu1, nu2), U(Ug), m0(mass), cyc(cycle) {
    amg.fgmres_m = m;
    amg.fgmres_restarts = restarts;
    amg.fgmres_tol = tol;
    amg.cycle = cycle;
    auto t0 = std::chrono::steady_clock::now();
    amg.setUpPhase(AMGV::Nit);
    const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (mpi::rank2d == 0)
        std::cout << "[rank " << mpi::rank2d << "] Elapsed time for Set-up phase = " << el << " seconds" << std::endl;
}

// one cycle from level 0 with a zero initial guess (FGMRES_AMG::preconditioner,
// include/amg.h:186-196)
void FGMRES_AMG::preconditioner(const spinor& in, spinor& out) {
    out.clearBuffer();
    if (cyc == 1) amg.k_cycle(0, in, out);
    else amg.v_cycle(0, in, out);
}

int FGMRES_AMG::fgmres(const spinor& phi, const spinor& x0, spinor& x, const bool& print_message) {
    std::vector<double> b = pack(phi, LV::Nx, LV::Nt, 2), xv = pack(x0, LV::Nx, LV::Nt, 2);
    int its = 0;
    dd_solve_info info{};
    const int st = dd_solve_x0(amg.context(), 1, b.data(), xv.data(), &its, &info);
    check(st, "dd_solve_x0");
    x = x0;
    unpack(xv, x, LV::Nx, LV::Nt, 2);
    last_relres = info.relres;
    last_seconds = info.seconds;
    if (print_message && mpi::rank2d == 0) {
        if (st == DD_OK) {
            std::cout << "FGMRES converged in " << (its + m - 1) / m << " cycles Error " << info.relres << std::endl;
            std::cout << "With " << its << " iterations" << std::endl;
        } else {
            std::cout << "FGMRES did not converge in " << restarts << " restarts of lenght " << m << " Error "
                      << info.relres << std::endl;
        }
    }
    return its;
}

// ---- comparison solvers (src/conjugate_gradient.cpp, include/fgmres.h:153-234,
// include/sap.h:236-333) -------------------------------------------------------------------
namespace {
// ||v|| over the interior sites (the reference's sqrt(real(dot(v, v))))
double interior_norm(const spinor& v) {
    double s2 = 0;
    for (int x = 1; x <= mpi::width_x; x++)
        for (int t = 1; t <= mpi::width_t; t++)
            for (int mu = 0; mu < LV::dof; mu++) s2 += std::norm(v.val[idx(x, t, mu)]);
    return std::sqrt(s2);
}

// solver on the process-wide context with U installed; returns (status, iterations)
std::pair<int, int> run_method(const spinor& U, const char* name, const std::function<int(dd_ctx*, int*, dd_solve_info*)>& f,
                               dd_solve_info& info) {
    dd_ctx* c = operator_ctx();
    std::vector<double> u = pack(U, LV::Nx, LV::Nt, 2);
    check(dd_set_gauge(c, u.data()), "dd_set_gauge");
    int its = 0;
    const int st = f(c, &its, &info);
    if (st != DD_OK && st != DD_ERR_NOT_CONVERGED) check(st, name);
    return {st, its};
}
}  // namespace

int conjugate_gradient(const spinor& U, const spinor& phi, spinor& x, const double& m0, const bool print) {
    std::vector<double> p = pack(phi, LV::Nx, LV::Nt, 2), xv(p.size());
    dd_solve_info info{};
    auto r = run_method(U, "dd_solve_cg", [&](dd_ctx* c, int* its, dd_solve_info* in) {
        return dd_solve_cg(c, 1, p.data(), xv.data(), m0, CG::max_iter, CG::tol, its, in);
    }, info);
    unpack(xv, x, LV::Nx, LV::Nt, 2);
    if (mpi::rank2d == 0 && (print || r.first != DD_OK)) {
        if (r.first == DD_OK)
            std::cout << "CG for D^+D converged in " << r.second << " iterations Error "
                      << info.relres * interior_norm(phi) << std::endl;
        else
            std::cout << "CG for D^+D did not converge in " << CG::max_iter << " iterations" << std::endl;
    }
    return r.first == DD_OK ? r.second : 0;
}

int bi_cgstab(const spinor& U, const spinor& phi, const spinor& x0, spinor& x, const double& m0,
              const bool& print_message) {
    std::vector<double> p = pack(phi, LV::Nx, LV::Nt, 2), x0v = pack(x0, LV::Nx, LV::Nt, 2), xv(p.size());
    dd_solve_info info{};
    auto r = run_method(U, "dd_solve_bicgstab", [&](dd_ctx* c, int* its, dd_solve_info* in) {
        return dd_solve_bicgstab(c, 1, p.data(), x0v.data(), xv.data(), m0, BiCG::max_iter, BiCG::tol, its, in);
    }, info);
    x = x0;
    unpack(xv, x, LV::Nx, LV::Nt, 2);
    if (print_message && mpi::rank2d == 0) {
        if (r.first == DD_OK)
            std::cout << "Bi-CG-stab for D converged in " << r.second << " iterations Error "
                      << info.relres * interior_norm(phi) << std::endl;
        else
            std::cout << "Bi-CG-stab for D did not converge in " << BiCG::max_iter << " iterations" << std::endl;
    }
    return r.second;
}

namespace {
int gmres_common(const spinor& U, double m0, const spinor& phi, const spinor& x0, spinor& x, int m, int restarts,
                 double tol, int bx, int bt, bool print, const char* label) {
    std::vector<double> p = pack(phi, LV::Nx, LV::Nt, 2), x0v = pack(x0, LV::Nx, LV::Nt, 2), xv(p.size());
    dd_solve_info info{};
    auto r = run_method(U, "dd_solve_gmres", [&](dd_ctx* c, int* its, dd_solve_info* in) {
        return dd_solve_gmres(c, 1, p.data(), x0v.data(), xv.data(), m0, m, restarts, tol, bx, bt, its, in);
    }, info);
    x = x0;
    unpack(xv, x, LV::Nx, LV::Nt, 2);
    if (print && mpi::rank2d == 0) {
        if (r.first == DD_OK)
            std::cout << label << " converged in " << (r.second + m - 1) / m << " cycles Error " << info.relres
                      << std::endl << "With " << r.second << " iterations" << std::endl;
        else
            std::cout << label << " did not converge in " << restarts << " restarts of lenght " << m << " Error "
                      << info.relres << std::endl;
    }
    return r.second;
}
}  // namespace

int FGMRES_fine_level::fgmres(const spinor& phi, const spinor& x0, spinor& x, const bool& print_message) {
    return gmres_common(U, m0, phi, x0, x, m, restarts, tol, 0, 0, print_message, "FGMRES");
}

int FGMRES_SAP::fgmres(const spinor& phi, const spinor& x0, spinor& x, const bool& print_message) {
    // SAPV::sap_block_x x sap_block_t blocks per (virtual) rank, as SAP_fine_level is built
    const int bx = (LV::Nx / mpi::ranks_x) / SAPV::sap_block_x, bt = (LV::Nt / mpi::ranks_t) / SAPV::sap_block_t;
    return gmres_common(U, m0, phi, x0, x, m, restarts, tol, bx, bt, print_message, "FGMRES");
}

// ---- SAP_fine_level on the GPU (src/sap.cpp:81-165; Methods::SAP, src/methods.cpp:61-78) --
int SAP_fine_level::SAP(const spinor& v, spinor& x, const int& nu_, const double& tol_, const bool print) {
    if (!U) throw std::runtime_error("SAP_fine_level: set_params before SAP");
    if (Nx != LV::Nx || Nt != LV::Nt || dofs != LV::dof)
        throw std::runtime_error("SAP_fine_level on the GPU covers the whole level-0 lattice");
    nu = nu_;
    tol = tol_;
    dd_ctx* c = operator_ctx();
    std::vector<double> u = pack(*U, LV::Nx, LV::Nt, 2), a = pack(v, LV::Nx, LV::Nt, 2), b = pack(x, LV::Nx, LV::Nt, 2);
    check(dd_set_gauge(c, u.data()), "dd_set_gauge");
    int its = 0;
    double rr = 0;
    check(dd_sap_fine(c, 1, a.data(), b.data(), nu_, tol_, x_elements, t_elements, m0, &its, &rr), "dd_sap_fine");
    unpack(b, x, LV::Nx, LV::Nt, 2);
    if (print && mpi::rank2d == 0) {
        if (its < nu_) std::cout << "SAP converged in " << its << " iterations, error: " << rr << " (relative)" << std::endl;
        else std::cout << "SAP didn't converge in " << nu_ << " iterations, error: " << rr << " (relative)" << std::endl;
    }
    return its;
}

// ---- parameters summary and run metadata (src/utils.cpp:55-95, src/io.cpp:58-140) ---------
void printParameters() {
    if (mpi::rank2d != 0) return;
    std::cout << "******************* Parameters summary *******************" << std::endl;
    std::cout << "| Nx = " << LV::Nx << " Nt = " << LV::Nt << std::endl;
    std::cout << "| Lattice dimension = " << LV::Ntot << std::endl;
    std::cout << "| Bare mass m0 = " << mass::m0 << std::endl;
    std::cout << "| Ranks_x = " << mpi::ranks_x << "  Ranks_t = " << mpi::ranks_t
              << " (virtual rank grid: test-vector streams and SAP blocks per rank)" << std::endl;
    for (int l = 0; l < LevelV::levels - 1; l++)
        std::cout << "| Level " << l << " Block X " << LevelV::BlocksX[l] << " Block T " << LevelV::BlocksT[l]
                  << " Ntest " << LevelV::Ntest[l] << " Schwarz Block X " << LevelV::SAP_Block_x[l]
                  << " Schwarz Block T " << LevelV::SAP_Block_t[l] << std::endl;
    for (int l = 0; l < LevelV::levels; l++)
        std::cout << "| Level " << l << " Nxsites " << LevelV::NxSites[l] << " NtSites " << LevelV::NtSites[l]
                  << " DOF " << LevelV::DOF[l] << " Colors " << LevelV::Colors[l] << std::endl;
    std::cout << "**********************************************************" << std::endl;
}

void broadcast_file_name(std::string&) {}

void writeMetadata(int cycle, double tol) {
    if (mpi::rank2d != 0) return;
    std::ostringstream name;
    name << "metadata_" << LV::Nx << "x" << LV::Nt << "_Levels" << LevelV::levels << (cycle == 0 ? "_Vcycle" : "_Kcycle")
         << ".dat";
    std::ofstream f(name.str());
    if (!f) {
        std::cerr << "Error opening " << name.str() << " for writing." << std::endl;
        return;
    }
    const char* host = std::getenv("HOSTNAME");
    const std::time_t now = std::chrono::system_clock::to_time_t(std::chrono::system_clock::now());
    char date[64];
    std::strftime(date, sizeof(date), "%Y-%m-%d %H:%M:%S", std::localtime(&now));
    f << "Host: " << (host ? host : "") << "\n";
    f << "Date: " << date << "\n";
    f << "x ranks: " << mpi::ranks_x << "  t_ranks: " << mpi::ranks_t << "\n";
    f << "m0: " << std::setprecision(7) << mass::m0 << "\n";
    f << "solvers tolerance: " << std::setprecision(4) << tol << "\n";
    f << "#level, blocks_x (global), blocks_t (global), Ntest, SAP_blocks_x (local), SAP_blocks_t (local)\n";
    for (int l = 0; l < LevelV::levels - 1; l++)
        f << l << std::setw(5) << LevelV::BlocksX[l] << std::setw(5) << LevelV::BlocksT[l] << std::setw(5)
          << LevelV::Ntest[l] << std::setw(5) << LevelV::SAP_Block_x[l] << std::setw(5) << LevelV::SAP_Block_t[l] << "\n";
    f << "#SAP_test_vectors_iterations\n" << AMGV::SAP_test_vectors_iterations << "\n";
    f << "#cycle (0 = V, 1 = K)\n" << cycle << "\n";
    f << "#Pre-smoothing nu1, Post-smoothing nu2\n" << AMGV::nu1 << std::setw(5) << AMGV::nu2 << "\n";
    f << "#sap gmres restart length, sap gmres restarts, sap gmres tolerance\n"
      << SAPV::sap_gmres_restart_length << std::setw(5) << SAPV::sap_gmres_restarts << std::setw(15) << std::scientific
      << std::setprecision(7) << SAPV::sap_gmres_tolerance << "\n";
    f << "#GMRES restart length coarse level, GMRES restarts coarse level, GMRES tolerance coarse level\n"
      << AMGV::gmres_restart_length_coarse_level << std::setw(5) << AMGV::gmres_restarts_coarse_level << std::setw(15)
      << AMGV::gmres_tol_coarse_level << "\n";
    f << "#FGMRES restart length, FGMRES restarts, FGMRES tolerance\n"
      << FGMRESV::fgmres_restart_length << std::setw(5) << FGMRESV::fgmres_restarts << std::setw(15)
      << FGMRESV::fgmres_tolerance << "\n";
    f << "#CG iterations,   FLOPS\n" << iter_counters::CGIt << std::setw(30) << flop_counters::CGFlops << "\n";
    f << "#BiCGStab iterations,   FLOPS\n" << iter_counters::BiCGIt << std::setw(30) << flop_counters::BiCGFLops << "\n";
    f << "#GMRES iterations,   FLOPS\n" << iter_counters::GMRESIt << std::setw(30) << flop_counters::GMRESFlops << "\n";
    f << "#SAP iterations,   FLOPS\n" << iter_counters::SAPIt << std::setw(30) << flop_counters::SAPFlops << "\n";
    f << "#FGMRES_SAP iterations,   FLOPS\n"
      << iter_counters::FGMRES_SAPIt << std::setw(30) << flop_counters::FGMRES_SAPFlops << "\n";
    if (cycle == 0)
        f << "#V cycle fgmres iterations,  Setup phase FLOPS, Total FLOPS\n"
          << iter_counters::vCycleIt << std::setw(30) << flop_counters::vCycleSetUpFlops << std::setw(30)
          << flop_counters::vCycleFlops << "\n";
    else
        f << "#K cycle fgmres iterations,  Setup phase FLOPS, Total FLOPS\n"
          << iter_counters::kCycleIt << std::setw(30) << flop_counters::kCycleSetUpFlops << std::setw(30)
          << flop_counters::kCycleFlops << "\n";
}

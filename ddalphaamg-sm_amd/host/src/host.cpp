// Host API of the MI355X DD-alphaAMG library: the reference's classes and free
// functions (variables.h, dirac_operator.h, level.h, amg.h, params.h, io.h) implemented
// over the C ABI of libddamg.so (include/ddamg.h).  Halo-padded host spinors are
// packed to the boundary format (interior sites, x outer, t inner, dof) and back.
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
namespace mass {
double m0 = -0.1;
}
namespace mpi {
int rank = 0, size = 1, maxSize = LV::Ntot, maxSizeH = LV::dof * (LV::Nx + 2) * (LV::Nt + 2),
    sitesH = (LV::Nx + 2) * (LV::Nt + 2), ranks_x = 2, ranks_t = 2, width_x = LV::Nx, width_t = LV::Nt, rank2d = 0;
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

// ---- Level ------------------------------------------------------------------------------
static int level_shape(dd_ctx* c, int l, int which) {
    int nx = 0, nt = 0, dof = 0;
    check(dd_level_shape(c, l, &nx, &nt, &dof), "dd_level_shape");
    return which == 0 ? nx : which == 1 ? nt : dof;
}

Level::Level(const int& l, dd_ctx* c)
    : level(l), Nx(level_shape(c, l, 0)), Nt(level_shape(c, l, 1)), Ntot(Nx * Nt), colors(level_shape(c, l, 2) / 2),
      DOF(level_shape(c, l, 2)), Ntest(l < LevelV::maxLevel ? LevelV::Ntest[l] : 1), ctx(c) {
    sap_l = new SAP_level_l(this);
}

Level::~Level() { delete sap_l; }

void Level::D_operator(const spinor& v, spinor& out) {
    std::vector<double> a = pack(v, Nx, Nt, DOF), b(a.size());
    check(dd_level_apply(ctx, level, a.data(), b.data(), 1, DDV::precision), "dd_level_apply");
    unpack(b, out, Nx, Nt, DOF);
}

void Level::Pdagg_v(const spinor& v, spinor& out) {
    const int nx = level_shape(ctx, level + 1, 0), nt = level_shape(ctx, level + 1, 1), d = level_shape(ctx, level + 1, 2);
    std::vector<double> a = pack(v, Nx, Nt, DOF), b((size_t)2 * nx * nt * d);
    check(dd_restrict(ctx, level, a.data(), b.data(), 1, DDV::precision), "dd_restrict");
    for (int i = 0; i < out.size; i++) out.val[i] = 0;
    unpack(b, out, nx, nt, d);
}

void Level::P_vc(const spinor& vc, spinor& out) {
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

void Level::checkOrthogonality() {
    fetchTestVectors();
    const int bx = Nx / level_shape(ctx, level + 1, 0), bt = Nt / level_shape(ctx, level + 1, 1);
    bool ok = true;
    for (int X = 0; X < Nx / bx; X++)
        for (int T = 0; T < Nt / bt; T++)
            for (int s = 0; s < 2; s++)
                for (int a = 0; a < Ntest; a++)
                    for (int b = 0; b < Ntest; b++) {
                        c_double d = 0;
                        for (int x = X * bx + 1; x <= (X + 1) * bx; x++)
                            for (int t = T * bt + 1; t <= (T + 1) * bt; t++)
                                for (int c = 0; c < colors; c++) {
                                    const int i = (x * (Nt + 2) + t) * DOF + c * 2 + s;
                                    d += std::conj(tvec[a].val[i]) * tvec[b].val[i];
                                }
                        if (std::abs(d - (a == b ? 1.0 : 0.0)) > 1e-8) ok = false;
                    }
    std::cout << "Test vectors on level " << level << (ok ? " are orthonormalized" : " are NOT orthonormal")
              << std::endl;
}

int Level::SAP_level_l::SAP(const spinor& v, spinor& x, const int& nu, const double& tol, const bool print) {
    Level* p = parent;
    std::vector<double> a = pack(v, p->Nx, p->Nt, p->DOF), b = pack(x, p->Nx, p->Nt, p->DOF);
    check(dd_sap(p->ctx, p->level, a.data(), b.data(), nu, tol, 1, DDV::precision), "dd_sap");
    unpack(b, x, p->Nx, p->Nt, p->DOF);
    if (print) std::cout << "SAP ran " << nu << " iterations on level " << p->level << std::endl;
    return nu;
}

// ---- AlgebraicMG ------------------------------------------------------------------------
AlgebraicMG::AlgebraicMG(const spinor& U, const double& m0_, const int& nu1_, const int& nu2_)
    : ctx(nullptr), m0(m0_), nu1(nu1_), nu2(nu2_) {
    check(dd_create(&ctx, LV::Nx, LV::Nt, DDV::device), "dd_create");
    std::vector<double> u = pack(U, LV::Nx, LV::Nt, 2);
    check(dd_set_gauge(ctx, u.data()), "dd_set_gauge");
}

AlgebraicMG::~AlgebraicMG() {
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
    for (Level* l : levels) delete l;
    levels.clear();
    for (int l = 0; l < LevelV::levels; l++) levels.push_back(new Level(l, ctx));
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
FGMRES_AMG::FGMRES_AMG(const spinor& U, const int& m, const int& restarts, const double& tol, const int nu1,
                       const int nu2, const double& m0, int cycle)
    : FGMRES(LV::Ntot, LV::dof, mpi::maxSizeH, 1, 1, mpi::width_x, mpi::width_t, m, restarts, tol),
      amg(U, m0, nu1, nu2) {
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

// Example: the setup of a two-level hierarchy built step by step with the reference's
// Level objects (include/level.h:147-233, 334-467), as AlgebraicMG::setUpPhase does it
// (src/amg.cpp:42-58) but with test vectors that the caller supplies:
//   Level l0(0, U), l1(1, U); fill l0.tvec; l0.orthonormalize(); l0.makeCoarseLinks(l1);
// then the level operator and the transfer operators on a right-hand side.
//
// stdin: ranks_x ranks_t m0 conf_file rhs_file params_file tvec_prefix out_prefix.
// The test vectors are <tvec_prefix><cc>.bin in the record format of read_binary.
// Writes (raw complex128, interior sites, x outer): <out>G1.bin, <out>G2.bin, <out>G3.bin
// (level-1 links in the reference's layout), <out>Pdagg.bin (P^dagger rhs),
// <out>D1.bin (D_1 P^dagger rhs), <out>P.bin (P P^dagger rhs), <out>tvec.bin (the
// orthonormalised level-0 test vectors).
#include <cstdio>
#include <fstream>
#include <string>

#include "amg.h"
#include "io.h"
#include "params.h"

static void write_interior(const std::string& name, const spinor& s, int Nx, int Nt, int per_site) {
    std::ofstream f(name, std::ios::binary);
    for (int x = 1; x <= Nx; x++)
        for (int t = 1; t <= Nt; t++)
            for (int k = 0; k < per_site; k++) {
                const c_double v = s.val[(size_t)(x * (Nt + 2) + t) * per_site + k];
                const double re_im[2] = {v.real(), v.imag()};
                f.write(reinterpret_cast<const char*>(re_im), sizeof(re_im));
            }
}

int main() {
    std::string conf, rhsf, pfile, tvp, out;
    double m0;
    std::cin >> mpi::ranks_x >> mpi::ranks_t >> m0 >> conf >> rhsf >> pfile >> tvp >> out;
    LevelV::levels = 2;
    LevelV::maxLevel = 1;
    mass::m0 = m0;
    allocate_lattice_arrays();
    initializeMPI();
    readParameters(pfile);
    spinor U(mpi::maxSizeH), rhs(mpi::maxSizeH);
    read_binary(conf, U);
    read_binary(rhsf, rhs);
    {
        Level l0(0, U), l1(1, U);
        for (int cc = 0; cc < l0.Ntest; cc++) read_binary(tvp + std::to_string(cc) + ".bin", l0.tvec[cc]);
        l0.orthonormalize();
        l0.checkOrthogonality();
        l0.makeCoarseLinks(l1);
        const int c = l1.colors;
        write_interior(out + "G1.bin", l1.G1, l1.Nx, l1.Nt, 4 * c * c);
        write_interior(out + "G2.bin", l1.G2, l1.Nx, l1.Nt, 8 * c * c);
        write_interior(out + "G3.bin", l1.G3, l1.Nx, l1.Nt, 8 * c * c);
        spinor vc((l1.Nx + 2) * (l1.Nt + 2) * l1.DOF), d1(vc.size), pv(mpi::maxSizeH);
        l0.Pdagg_v(rhs, vc);
        l1.D_operator(vc, d1);
        l0.P_vc(vc, pv);
        write_interior(out + "Pdagg.bin", vc, l1.Nx, l1.Nt, l1.DOF);
        write_interior(out + "D1.bin", d1, l1.Nx, l1.Nt, l1.DOF);
        write_interior(out + "P.bin", pv, l0.Nx, l0.Nt, l0.DOF);
        std::ofstream f(out + "tvec.bin", std::ios::binary);
        for (int cc = 0; cc < l0.Ntest; cc++) {
            write_interior(out + "tv.tmp", l0.tvec[cc], l0.Nx, l0.Nt, l0.DOF);
            std::ifstream g(out + "tv.tmp", std::ios::binary);
            f << g.rdbuf();
        }
        std::remove((out + "tv.tmp").c_str());
    }
    free_lattice_arrays();
    std::cout << "level steps done" << std::endl;
    return 0;
}

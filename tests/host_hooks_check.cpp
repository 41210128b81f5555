// Test program for the host API's generic algorithms over caller-defined hooks: the
// FGMRES base class (include/fgmres.h) and the SAP_C base class (include/sap.h), driven by
// subclasses whose hooks are a host Wilson-Dirac operator written here.  No GPU call is
// made.  tests/test_host_dropin.py compares the output with the oracle's fgmres and sap.
//
// stdin (binary): int Lx, int Lt, double m0, int nbx, int nbt, int nu, then U and b as
// complex128 [Lx][Lt][2] (interior order).  stdout (binary): int fgmres_iterations,
// x_fgmres, int sap_iterations, x_sap (complex128 [Lx][Lt][2]).
#include <cstdio>
#include <vector>

#include "sap.h"

namespace {
int Lx, Lt;
double mass0;
std::vector<c_double> Ug;   // halo-padded [(Lx+2)(Lt+2)][2]

int site(int x, int t) { return x * (Lt + 2) + t; }

// D_phi on the interior, periodic in x and antiperiodic in t (src/dirac_operator.cpp:24-54)
void wilson(const spinor& in, spinor& out) {
    const c_double I(0, 1);
    for (int x = 1; x <= Lx; x++)
        for (int t = 1; t <= Lt; t++) {
            const int n = site(x, t);
            const int tp = t == Lt ? 1 : t + 1, tm = t == 1 ? Lt : t - 1;
            const int xp = x == Lx ? 1 : x + 1, xm = x == 1 ? Lx : x - 1;
            const double sp = t == Lt ? -1.0 : 1.0, sm = t == 1 ? -1.0 : 1.0;
            const int f0 = site(x, tp), f1 = site(xp, t), b0 = site(x, tm), b1 = site(xm, t);
            const c_double u0 = sp * Ug[2 * n], u1 = Ug[2 * n + 1];
            const c_double ub0 = sm * std::conj(Ug[2 * b0]), ub1 = std::conj(Ug[2 * b1 + 1]);
            const c_double* a = in.val;
            out.val[2 * n] = (mass0 + 2) * a[2 * n] -
                             0.5 * (u0 * (a[2 * f0] - a[2 * f0 + 1]) + u1 * (a[2 * f1] + I * a[2 * f1 + 1]) +
                                    ub0 * (a[2 * b0] + a[2 * b0 + 1]) + ub1 * (a[2 * b1] - I * a[2 * b1 + 1]));
            out.val[2 * n + 1] = (mass0 + 2) * a[2 * n + 1] -
                                 0.5 * (u0 * (-a[2 * f0] + a[2 * f0 + 1]) + u1 * (-I * a[2 * f1] + a[2 * f1 + 1]) +
                                        ub0 * (a[2 * b0] + a[2 * b0 + 1]) + ub1 * (I * a[2 * b1] + a[2 * b1 + 1]));
        }
}

c_double interior_dot(const c_double* A, const c_double* B) {
    c_double z = 0;
    for (int x = 1; x <= Lx; x++)
        for (int t = 1; t <= Lt; t++)
            for (int d = 0; d < 2; d++) z += A[2 * site(x, t) + d] * std::conj(B[2 * site(x, t) + d]);
    return z;
}

class HostGmres : public FGMRES {
   public:
    HostGmres(int m, int restarts, double tol)
        : FGMRES(Lx * Lt, 2, 2 * (Lx + 2) * (Lt + 2), 1, 1, Lx, Lt, m, restarts, tol) {}

   private:
    void func(const spinor& in, spinor& out) override { wilson(in, out); }
    void preconditioner(const spinor& in, spinor& out) override { out = in; }
    c_double dot(c_double* A, c_double* B) override { return interior_dot(A, B); }
    void axpy(const spinor& X, const spinor& Y, const c_double& lambda, spinor& out) override {
        for (int x = 1; x <= Lx; x++)
            for (int t = 1; t <= Lt; t++)
                for (int d = 0; d < 2; d++) {
                    const int i = 2 * site(x, t) + d;
                    out.val[i] = X.val[i] + lambda * Y.val[i];
                }
    }
    void scal(const c_double& lambda, const spinor& X, spinor& Y) override {
        for (int x = 1; x <= Lx; x++)
            for (int t = 1; t <= Lt; t++)
                for (int d = 0; d < 2; d++) Y.val[2 * site(x, t) + d] = lambda * X.val[2 * site(x, t) + d];
    }
};

// SAP over the same operator; the block operator embeds the block vector in the lattice,
// applies the operator and keeps the block's rows (zero outside the block)
class HostSap : public SAP_C {
   public:
    HostSap(int nbx, int nbt) : SAP_C(Lx, Lt, nbx, nbt, 2, 1) {}

   private:
    void funcGlobal(spinor& in, spinor& out) override { wilson(in, out); }
    void funcLocal(const spinor& in, spinor& out) override {
        spinor full(2 * (Lx + 2) * (Lt + 2)), Dfull(full.size);
        const std::vector<int>& blk = Blocks[current_block];
        for (int mx = 1; mx <= x_elements; mx++)
            for (int mt = 1; mt <= t_elements; mt++) {
                const int m = mx * (t_elements + 2) + mt;
                for (int d = 0; d < 2; d++) full.val[2 * blk[m] + d] = in.val[2 * m + d];
            }
        wilson(full, Dfull);
        out.clearBuffer();
        for (int mx = 1; mx <= x_elements; mx++)
            for (int mt = 1; mt <= t_elements; mt++) {
                const int m = mx * (t_elements + 2) + mt;
                for (int d = 0; d < 2; d++) out.val[2 * m + d] = Dfull.val[2 * blk[m] + d];
            }
    }
    c_double dot(const spinor& X, const spinor& Y) override { return interior_dot(X.val, Y.val); }
};

void read_field(spinor& s) {
    for (int x = 1; x <= Lx; x++)
        for (int t = 1; t <= Lt; t++)
            for (int d = 0; d < 2; d++) {
                double re_im[2];
                if (std::fread(re_im, sizeof(double), 2, stdin) != 2) std::exit(2);
                s.val[2 * site(x, t) + d] = c_double(re_im[0], re_im[1]);
            }
}

void write_field(const spinor& s) {
    for (int x = 1; x <= Lx; x++)
        for (int t = 1; t <= Lt; t++)
            for (int d = 0; d < 2; d++) {
                const double re_im[2] = {s.val[2 * site(x, t) + d].real(), s.val[2 * site(x, t) + d].imag()};
                std::fwrite(re_im, sizeof(double), 2, stdout);
            }
}
}  // namespace

int main() {
    int hdr[2], blk[3];
    if (std::fread(hdr, sizeof(int), 2, stdin) != 2 || std::fread(&mass0, sizeof(double), 1, stdin) != 1 ||
        std::fread(blk, sizeof(int), 3, stdin) != 3)
        return 2;
    Lx = hdr[0];
    Lt = hdr[1];
    const int H = 2 * (Lx + 2) * (Lt + 2);
    spinor U(H), b(H), x0(H), x(H), xs(H);
    read_field(U);
    read_field(b);
    Ug.assign(U.val, U.val + H);
    HostGmres gm(10, 20, 1e-10);
    const int its = gm.fgmres(b, x0, x, false);
    HostSap sap(blk[0], blk[1]);
    const int sits = sap.SAP(b, xs, blk[2], 1e-12, false);
    std::fwrite(&its, sizeof(int), 1, stdout);
    write_field(x);
    std::fwrite(&sits, sizeof(int), 1, stdout);
    write_field(xs);
    return 0;
}

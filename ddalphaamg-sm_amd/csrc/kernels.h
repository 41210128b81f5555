// Launchers of the gfx950 kernels (kernels.hip).  All launches go to the given stream
// and check the launch status; shapes are validated by the callers (solver.hip).
#pragma once
#include <vector>

#include "common.h"

namespace dd {

// ---- live kernel timing (HIP events on the launching stream) -----------------------
enum KernelId {
    K_WILSON = 0, K_COARSE, K_SAP0, K_SAPC, K_MULTIDOT, K_MULTIAXPY, K_RESTRICT, K_PROLONG, K_COARSEST,
    K_DOT, K_NUM
};
void timer_enable(int id, bool on);
void timer_reset();
void timer_begin(hipStream_t s, int id);
// `bytes`: algorithmic HBM bytes of the launch (each operand counted once)
void timer_end(hipStream_t s, int id, double bytes);
// adds algorithmic bytes to kernel `id` after the fact (kernels that count their own work)
void timer_add_bytes(int id, double bytes);
// destroys the timers' event pools (re-created on demand)
void timer_release();
// true when any kernel timer is on (their event records stay out of graph captures)
bool timer_any();
// total elapsed ms, number of timed launches and algorithmic bytes of kernel `id`
// since the last reset (synchronises the events)
double timer_read(int id, long long* count, double* bytes);
// device word that the kernels of family `id` add their algorithmic floating-point
// operations to while its timer is on (nullptr while it is off: the kernels then count
// nothing), and its total since the last reset (synchronises the device)
unsigned long long* timer_flops_dev(int id);
double timer_read_flops(int id);

// ---- device-side system masks ---------------------------------------------------------
// While a DevMaskScope is alive on the calling host thread, every launcher below that takes
// a system mask also hands its kernel the device word m, and the kernel ANDs its mask with
// *m when it starts: the systems a launch serves can then be decided on the device (the
// run masks of a device-controlled Krylov solve inside a captured graph).  Scopes nest;
// without one (nullptr) launches use their host masks only.
const uint64_t* dev_mask();
struct DevMaskScope {
    const uint64_t* prev;
    explicit DevMaskScope(const uint64_t* m);
    ~DevMaskScope();
    DevMaskScope(const DevMaskScope&) = delete;
    DevMaskScope& operator=(const DevMaskScope&) = delete;
};

// Maximum block-GMRES restart length compiled into the SAP block kernels.
constexpr int kSapM = 5;

struct RedBuf {            // scratch for two-pass reductions
    void* partial = nullptr;   // [nrhs][nblk] complex<double>
    int nblk = 0;
};
// most partial sums a reduction keeps per system (blocks of dot_partial)
constexpr int kMaxRedBlocks = 256;
// bytes of RedBuf::partial that dot_partial needs for nrhs systems
inline size_t dot_scratch_bytes(int nrhs) { return sizeof(cplx<double>) * (size_t)kMaxRedBlocks * nrhs; }

// ---- level-0 Wilson-Dirac ---------------------------------------------------------
// out = D in  (b == nullptr)   or   out = b - D in.   dagger selects D^dagger.
template <typename R>
void wilson(hipStream_t s, const cplx<R>* U, const cplx<R>* in, cplx<R>* out, const cplx<R>* b,
            int Nx, int Nt, double m0, bool dagger, int nrhs, uint64_t mask, size_t ld);

// Level-0 link matrices in the generic coarse-link layout [V][5][2][2] (col-major).
template <typename R>
void dirac_links(hipStream_t s, const cplx<R>* U, cplx<R>* L, int Nx, int Nt);

// ---- generic level stencil ----------------------------------------------------------
template <typename R>
void coarse_apply(hipStream_t s, const cplx<R>* L, const cplx<R>* in, cplx<R>* out, const cplx<R>* b,
                  int Nx, int Nt, int dof, double m0, int nrhs, uint64_t mask, size_t ld);

// ---- SAP block solves (one colour, x += D_B^-1 r on the blocks of that colour) -------
template <typename R>
// form_residual: r is the right-hand side v, and each block forms its right-hand side
// v - D x at its own sites from the current x (the full-lattice operator); valid only
// where no two blocks of one colour are neighbours (even block counts in both directions)
void sap_blocks0(hipStream_t s, const cplx<R>* U, const cplx<R>* r, cplx<R>* x, int Nx, int Nt, int bx, int bt,
                 int colour, double m0, int m, int restarts, double tol, int nrhs, uint64_t mask, size_t ld,
                 const int* gate, bool form_residual = false);
// variant: kernel variant (systems per work-group, load batching), -1 = default;
// sapc_variants<R>() lists the compiled ones (autotuned in solver.hip).  Lh: binary16
// copy of L (to_half), read by variant 33 (the matrix-core kernel on binary16 links);
// nullptr where there is none.
template <typename R>
void sap_blocksc(hipStream_t s, const cplx<R>* L, const cplx<R>* r, cplx<R>* x, int Nx, int Nt, int dof, int bx,
                 int bt, int colour, double m0, int m, int restarts, double tol, int nrhs, uint64_t mask,
                 size_t ld, const int* gate, int variant = -1, const _Float16* Lh = nullptr);
// State of one fused coarse-SAP iteration (sap_fused_iteration): per (block, tile) the
// launch epoch at which the block's red solve stored its iterate, two counters that are
// zero when a launch starts (black solves done, norm slices done), a sticky failure word
// (a bounded wait timed out), the norm partial sums, the per-system gate of the
// iterations and ||v||^2 per system (stride 2).
struct SapFuse {
    unsigned long long* flops = nullptr;   // matrix-core flops are added here (nullptr: not counted)
    unsigned* done = nullptr;           // [nbx*nbt][ntile]
    unsigned* ctr = nullptr;            // [2]
    unsigned* fail = nullptr;
    float* part = nullptr;              // [slices][64]
    int* gate = nullptr;                // [64]
    const cplx<double>* vnorm = nullptr;
    unsigned epoch = 0;
    double tol = 0;
};
// One SAP iteration on the matrix cores in one launch (FP32, coarse-SAP variants 32/33):
// red blocks, black blocks, ||v - D x|| and the gate update, with each block's
// right-hand side v - D x formed from the iterate (src/sap.cpp:103-156).  Returns false,
// launching nothing, where it does not apply (FP64, another variant, odd numbers of
// blocks in a direction, blocks above 16 sites or dof above 24).  nslice: work-groups of
// the norm; 0 launches no norm stage and leaves the gate unchanged.
template <typename R>
bool sap_fused_iteration(hipStream_t s, const cplx<R>* L, const _Float16* Lh, const cplx<R>* v, cplx<R>* x, int Nx,
                         int Nt, int dof, int bx, int bt, double m0, int m, int restarts, double tol, int nrhs,
                         uint64_t mask, size_t ld, int variant, const SapFuse& fz, int nslice);
template <typename R>
std::vector<int> sapc_variants();
template <typename R>
bool sapc_supported(int variant, int dof, int S);

// ---- transfers ----------------------------------------------------------------------
// out(coarse) = P^dagger v (fine);  W: [V][nv][dof]
template <typename R>
void restrict_(hipStream_t s, const cplx<R>* W, const cplx<R>* v, cplx<R>* out, int Nx, int Nt, int dof,
               int ax, int at, int nv, int nrhs, uint64_t mask, size_t ld_f, size_t ld_c);
// out(fine) = y + P vc;  y may alias out, y == nullptr means 0
template <typename R>
void prolong(hipStream_t s, const cplx<R>* W, const cplx<R>* vc, const cplx<R>* y, cplx<R>* out, int Nx, int Nt,
             int dof, int ax, int at, int nv, int nrhs, uint64_t mask, size_t ld_f, size_t ld_c);
// block-local modified Gram-Schmidt of W [V][nv][dof] per (aggregate, spin)
template <typename R>
void orthonormalize(hipStream_t s, cplx<R>* W, int Nx, int Nt, int dof, int ax, int at, int nv);
// Galerkin links of the next level from links Lf of this level and test vectors W
template <typename R>
void galerkin(hipStream_t s, const cplx<R>* Lf, const cplx<R>* W, cplx<R>* Lc, int Nx, int Nt, int dof, int ax,
              int at, int nv, bool has_A);
// [nv][V][dof] <-> [V][nv][dof]
template <typename R>
void transpose_tv(hipStream_t s, const cplx<R>* in, cplx<R>* out, int V, int nv, int dof, bool to_site_major);

// ---- BLAS-1 -----------------------------------------------------------------------------
// partial sums of a . conj(b) per system into rb.partial; returns nothing
template <typename R>
void dot_partial(hipStream_t s, const cplx<R>* a, const cplx<R>* b, size_t n, size_t ld, int nrhs, uint64_t mask,
                 RedBuf& rb);
// out[r*ostride] = sum of partials (complex, precision R)
template <typename R>
void dot_final(hipStream_t s, const RedBuf& rb, cplx<double>* out, int ostride, int nrhs, uint64_t mask);
// h = sum of in's partials -> H[r*hstride]; w -= h * v; partials of <w, nxt> (nxt null:
// of |w|^2) -> out, the bits of the update pass followed by dot_partial(w, nxt)
template <typename R>
void mgs_update_dot(hipStream_t s, cplx<R>* w, const cplx<R>* v, const cplx<R>* nxt, size_t n, size_t ld, int nrhs,
                    uint64_t mask, const RedBuf& in, RedBuf& out, cplx<double>* H, int hstride);
// h = sqrt(sum partials.re) -> H[r*hstride]; if h > 0: out = (1/h) w
template <typename R>
void norm_scale(hipStream_t s, cplx<R>* out, const cplx<R>* w, size_t n, size_t ld, int nrhs, uint64_t mask,
                const RedBuf& rb, cplx<double>* H, int hstride);
// out = scale[r] * w, scale from device (1/beta computed in-kernel from beta[r*bstride].x)
template <typename R>
void scale_inv(hipStream_t s, cplx<R>* out, const cplx<R>* w, size_t n, size_t ld, int nrhs, uint64_t mask,
               const cplx<double>* beta, int bstride);
// out = x + alpha*y, alpha host scalar (real)
template <typename R>
void axpy(hipStream_t s, cplx<R>* out, const cplx<R>* x, const cplx<R>* y, double alpha, size_t n, size_t ld,
          int nrhs, uint64_t mask);
// out = c[3r] x + c[3r+1] y + c[3r+2] z per system r (c on the device; null operands skipped;
// out may alias an operand)
template <typename R>
void lincomb(hipStream_t s, cplx<R>* out, const cplx<R>* x, const cplx<R>* y, const cplx<R>* z,
             const cplx<double>* c, size_t n, size_t ld, int nrhs, uint64_t mask);
// x += sum_{j < nit[r]} eta[r*m + j] * Z_j   (Z_j = Z + j*zstride)
template <typename R>
void update_x(hipStream_t s, cplx<R>* x, const cplx<R>* Z, size_t zstride, const cplx<double>* eta,
              const int* nit, int m, size_t n, size_t ld, int nrhs, uint64_t mask);
template <typename R>
void copy(hipStream_t s, cplx<R>* out, const cplx<R>* in, size_t n, size_t ld_out, size_t ld_in, int nrhs,
          uint64_t mask);
template <typename R>
void zero(hipStream_t s, cplx<R>* out, size_t n, size_t ld, int nrhs, uint64_t mask);
template <typename To, typename From>
void convert(hipStream_t s, cplx<To>* out, const cplx<From>* in, size_t n, size_t ld_out, size_t ld_in, int nrhs,
             uint64_t mask);
// multiply the spin-1 components by -1 (sigma_3 on dof = colour*2 + spin)
// out[i] = in[i] rounded to binary16 (n reals)
void to_half(hipStream_t s, const float* in, _Float16* out, size_t n);
// out[i] = in[i] (exact)
void from_half(hipStream_t s, const _Float16* in, double* out, size_t n);
template <typename R>
void sigma3(hipStream_t s, cplx<R>* out, const cplx<R>* in, size_t n, size_t ld, int nrhs, uint64_t mask);
// gate[r] = 1 for the 64 systems, or (inside a DevMaskScope) bit r of the device mask
void gate_init(hipStream_t s, int* gate);
// gate[r] = 0 where sqrt(err2[r*stride].x) < tol * sqrt(v2[r*stride].x)
void gate_update(hipStream_t s, int* gate, const cplx<double>* err2, const cplx<double>* v2, int stride, double tol,
                 int nrhs, uint64_t mask);

// ---- gauge field ------------------------------------------------------------------------
void metropolis_u1(hipStream_t s, double* theta, int Nx, int Nt, double beta, int sweeps, uint64_t seed,
                   double step);
template <typename R>
void theta_to_u(hipStream_t s, const double* theta, cplx<R>* U, size_t n);

// ---- derivative of the stencil --------------------------------------------------------
// grad[x][mu] += sign * Re( y(x)^H dD/dtheta_mu(x) x )  summed over systems, see kernels.hip
void dtheta_contract(hipStream_t s, const cplx<double>* U, const cplx<double>* y, const cplx<double>* x,
                     double* grad, int Nx, int Nt, int nrhs, size_t ld, double scale);

}  // namespace dd

namespace dd {
// Classical Gram-Schmidt pass against V_0..V_{nvec-1} (V_i = V + i*vstride):
// partials of <V_i, w> per block, then h_i = sum, w -= sum_i h_i V_i and
// H[r*hstride + i] (=|+=) h_i.
template <typename R>
void multidot(hipStream_t s, const cplx<R>* w, const cplx<R>* V, size_t vstride, int nvec, size_t n, size_t ld,
              int nrhs, uint64_t mask, RedBuf& rb);
template <typename R>
void multi_axpy(hipStream_t s, cplx<R>* w, const cplx<R>* V, size_t vstride, int nvec, size_t n, size_t ld, int nrhs,
                uint64_t mask, const RedBuf& rb, cplx<double>* H, int hstride, bool accumulate);
}  // namespace dd

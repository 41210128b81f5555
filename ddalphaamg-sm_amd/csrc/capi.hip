// extern "C" boundary of libddamg.so (declarations and reference citations in
// include/ddamg.h).  Converts between the boundary format (complex128, reference
// interior order) and the device layouts, and maps exceptions to status codes.
#include <cstring>
#include <mutex>
#include <string>
#include <thread>

#include "poly.h"
#include "solver.h"

// Device scratch of dd_dot_dev_noalloc: the partial sums of up to 64 systems x 256
// blocks (256 KiB per GPU, held until the process ends).  Host code never uses this
// name as an address: it takes the per-device address from hipGetSymbolAddress.
static constexpr int kNoAllocSystems = 64;
static constexpr size_t kNoAllocScratchBytes = sizeof(dd::cplx<double>) * dd::kMaxRedBlocks * kNoAllocSystems;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

struct dd_ctx {
    std::unique_ptr<dd::Solver> s;
    void* caller = nullptr;        // stream whose queued work the _dev calls wait for
    hipEvent_t ordered = nullptr;  // recorded on `caller`, waited for by the context's stream
    ~dd_ctx() {
        if (ordered) (void)hipEventDestroy(ordered);
    }
};

namespace {
std::string g_err;

void set_err(const std::string& e) { g_err = e; }

template <typename F>
int guard(F&& f) {
    try {
        return f();
    } catch (const dd::HipFailure& e) {
        set_err(std::string("HIP: ") + hipGetErrorString(e.err));
        return DD_ERR_HIP;
    } catch (const dd::BadShape& e) {
        set_err(e.what);
        return DD_ERR_BAD_SHAPE;
    } catch (const std::exception& e) {
        set_err(e.what());
        return DD_ERR_HIP;
    }
}

using dd::cplx;
using dd::DevMem;

size_t vsize(dd_ctx* c) { return (size_t)c->s->Nx * c->s->Nt * 2; }

// host <-> device staging of complex128 data
struct Staged {
    DevMem d;
    Staged(const double* h, size_t ncplx, hipStream_t s) : d(sizeof(cplx<double>) * ncplx) {
        if (h) DD_HIP(hipMemcpyAsync(d.as<void>(), h, sizeof(cplx<double>) * ncplx, hipMemcpyHostToDevice, s));
    }
    cplx<double>* p() { return d.as<cplx<double>>(); }
};

void download(double* h, const cplx<double>* d, size_t ncplx, hipStream_t s) {
    DD_HIP(hipMemcpyAsync(h, d, sizeof(cplx<double>) * ncplx, hipMemcpyDeviceToHost, s));
    DD_HIP(hipStreamSynchronize(s));
}

int check(dd_ctx* c) {
    if (!c || !c->s) {
        set_err("null context");
        return DD_ERR_ARG;
    }
    return DD_OK;
}

// dd_params from a caller compiled against another layout of the struct
int check_params(const dd_params* p) {
    if (!p) {
        set_err("null dd_params");
        return DD_ERR_ARG;
    }
    if (p->size != (int)sizeof(dd_params)) {
        set_err("dd_params of " + std::to_string(p->size) + " bytes; this library's is " +
                std::to_string(sizeof(dd_params)) + " (fill it with dd_default_params of the same header)");
        return DD_ERR_ARG;
    }
    return DD_OK;
}

// the context's stream waits for the work queued so far on the caller's stream
void order_after_caller(dd_ctx* c) {
    if (!c->ordered) DD_HIP(hipEventCreateWithFlags(&c->ordered, hipEventDisableTiming));
    DD_HIP(hipEventRecord(c->ordered, (hipStream_t)c->caller));
    DD_HIP(hipStreamWaitEvent(c->s->s, c->ordered, 0));
}

// address of `scratch` on the current device, looked up once per device
void* scratch_on_current_device() {
    static std::mutex mu;
    static void* ptr[256] = {};
    int dev = 0;
    DD_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 256) throw dd::BadShape("device ordinal out of range");
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[dev]) {
        void* p = nullptr;
        DD_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        ptr[dev] = p;
    }
    return ptr[dev];
}

template <typename R>
dd::Multigrid<R>* mg(dd_ctx* c);
template <>
dd::Multigrid<double>* mg<double>(dd_ctx* c) {
    if (!c->s->mg64) throw dd::BadShape("no FP64 hierarchy (setup with precision 0 first)");
    return c->s->mg64.get();
}
template <>
dd::Multigrid<float>* mg<float>(dd_ctx* c) {
    if (!c->s->mg32) throw dd::BadShape("no FP32 hierarchy (setup with precision 1 first)");
    return c->s->mg32.get();
}

// runs f(in_R, out_R) on level-shaped data converted to precision R
template <typename R, typename F>
void with_level_io(dd_ctx* c, const double* in, double* out, size_t n_in, size_t n_out, int nrhs, F&& f,
                   const double* inout_init = nullptr) {
    hipStream_t s = c->s->s;
    Staged a(in, n_in * nrhs, s);
    DevMem ar(sizeof(cplx<R>) * n_in * nrhs), br(sizeof(cplx<R>) * n_out * nrhs);
    dd::convert<R, double>(s, ar.as<cplx<R>>(), a.p(), n_in * nrhs, 0, 0, 1, 1);
    if (inout_init) {
        Staged b0(inout_init, n_out * nrhs, s);
        dd::convert<R, double>(s, br.as<cplx<R>>(), b0.p(), n_out * nrhs, 0, 0, 1, 1);
        DD_HIP(hipStreamSynchronize(s));
    } else {
        dd::zero<R>(s, br.as<cplx<R>>(), n_out * nrhs, 0, 1, 1);
    }
    f(ar.as<cplx<R>>(), br.as<cplx<R>>());
    DevMem o(sizeof(cplx<double>) * n_out * nrhs);
    dd::convert<double, R>(s, o.as<cplx<double>>(), br.as<cplx<R>>(), n_out * nrhs, 0, 0, 1, 1);
    download(out, o.as<cplx<double>>(), n_out * nrhs, s);
}
}  // namespace

extern "C" {

const char* dd_last_error(void) { return g_err.c_str(); }

int dd_params_size(void) { return (int)sizeof(dd_params); }

int dd_default_params_sized(dd_params* p, int size, int levels) {
    if (!p || levels < 2 || levels > DD_MAX_LEVELS) {
        set_err("dd_default_params: null pointer or levels outside [2, 8]");
        return DD_ERR_ARG;
    }
    if (size != (int)sizeof(dd_params)) {
        set_err("dd_default_params: caller's dd_params has " + std::to_string(size) + " bytes, this library's " +
                std::to_string(sizeof(dd_params)) + " (rebuild against the library's ddamg.h)");
        return DD_ERR_ARG;
    }
    std::memset(p, 0, sizeof(*p));
    p->size = (int)sizeof(dd_params);
    p->levels = levels;
    for (int l = 0; l < DD_MAX_LEVELS; l++) {
        p->agg_x[l] = p->agg_t[l] = 4;
        p->sap_x[l] = p->sap_t[l] = 4;
        p->nv[l] = 8;
    }
    p->m0 = -0.1;
    p->nu1 = 0;
    p->nu2 = 2;
    p->setup_sap_iters = 4;
    p->sap_m = 5;
    p->sap_restarts = 5;
    p->sap_tol = 1e-3;
    p->coarse_m = 400;
    p->coarse_restarts = 10;
    p->coarse_tol = 0.1;
    p->kcycle_m = 5;
    p->kcycle_restarts = 2;
    p->kcycle_tol = 0.1;
    p->fgmres_m = 20;
    p->fgmres_restarts = 50;
    p->tol = 1e-10;
    p->cycle = 1;
    p->precision = 0;
    p->rng_ranks_x = 2;
    p->rng_ranks_t = 2;
    p->ortho_outer = 0;
    p->ortho_inner = 0;
    p->ortho_coarse = 3;
    p->max_rhs = 1;
    p->autotune = 1;
    p->sap_variant = -1;
    p->coarse_basis = 1;
    p->coarse_sap_links = 0;
    p->bootstrap = 0;
    p->graph = 0;
    p->sap_fused = 1;
    p->coarse_poly = 0;
    return DD_OK;
}

int dd_finalize(void) {
    return guard([&] {
        dd::timer_release();
        return DD_OK;
    });
}

int dd_set_stream(dd_ctx* c, void* stream) {
    if (check(c)) return DD_ERR_ARG;
    c->caller = stream;
    return DD_OK;
}

int dd_create(dd_ctx** ctx, int Nx, int Nt, int device) {
    if (!ctx) return DD_ERR_ARG;
    return guard([&] {
        *ctx = nullptr;
        auto* c = new dd_ctx;
        try {
            c->s.reset(new dd::Solver(Nx, Nt, device));
        } catch (...) {
            delete c;
            throw;
        }
        *ctx = c;
        return DD_OK;
    });
}

int dd_destroy(dd_ctx* ctx) {
    if (!ctx) return DD_OK;
    return guard([&] {
        delete ctx;
        return DD_OK;
    });
}

int dd_set_gauge(dd_ctx* c, const double* U) {
    if (check(c) || !U) return DD_ERR_ARG;
    return guard([&] {
        Staged u(U, vsize(c), c->s->s);
        c->s->set_gauge(u.p());
        return DD_OK;
    });
}

int dd_set_gauge_theta(dd_ctx* c, const double* theta) {
    if (check(c) || !theta) return DD_ERR_ARG;
    return guard([&] {
        hipStream_t s = c->s->s;
        const size_t n = vsize(c);
        DD_HIP(hipMemcpyAsync(c->s->theta.as<void>(), theta, sizeof(double) * n, hipMemcpyHostToDevice, s));
        DevMem u(sizeof(cplx<double>) * n);
        dd::theta_to_u<double>(s, c->s->theta.as<double>(), u.as<cplx<double>>(), n);
        c->s->set_gauge(u.as<cplx<double>>());
        return DD_OK;
    });
}

int dd_generate_gauge(dd_ctx* c, double beta, int sweeps, unsigned long long seed, double step, double* theta_out) {
    if (check(c)) return DD_ERR_ARG;
    return guard([&] {
        hipStream_t s = c->s->s;
        const size_t n = vsize(c);
        dd::metropolis_u1(s, c->s->theta.as<double>(), c->s->Nx, c->s->Nt, beta, sweeps, seed, step);
        DevMem u(sizeof(cplx<double>) * n);
        dd::theta_to_u<double>(s, c->s->theta.as<double>(), u.as<cplx<double>>(), n);
        c->s->set_gauge(u.as<cplx<double>>());
        if (theta_out) {
            DD_HIP(hipMemcpyAsync(theta_out, c->s->theta.as<void>(), sizeof(double) * n, hipMemcpyDeviceToHost, s));
            DD_HIP(hipStreamSynchronize(s));
        }
        return DD_OK;
    });
}

int dd_get_gauge(dd_ctx* c, double* U) {
    if (check(c) || !U) return DD_ERR_ARG;
    return guard([&] {
        download(U, c->s->U64.as<cplx<double>>(), vsize(c), c->s->s);
        return DD_OK;
    });
}

int dd_apply_D_dev(dd_ctx* c, const double* in, double* out, int nrhs, double m0, int dagger, int precision) {
    if (check(c) || !in || !out || nrhs < 1 || nrhs > 64) return DD_ERR_ARG;
    return guard([&] {
        if (!c->s->have_gauge) throw dd::BadShape("no gauge field");
        order_after_caller(c);
        c->s->apply_D((const cplx<double>*)in, (cplx<double>*)out, nrhs, m0, dagger != 0, precision);
        DD_HIP(hipStreamSynchronize(c->s->s));
        return DD_OK;
    });
}

int dd_apply_D(dd_ctx* c, const double* in, double* out, int nrhs, double m0, int dagger, int precision) {
    if (check(c) || !in || !out || nrhs < 1 || nrhs > 64) return DD_ERR_ARG;
    return guard([&] {
        if (!c->s->have_gauge) throw dd::BadShape("no gauge field");
        hipStream_t s = c->s->s;
        const size_t n = vsize(c) * nrhs;
        Staged a(in, n, s), b(nullptr, n, s);
        c->s->apply_D(a.p(), b.p(), nrhs, m0, dagger != 0, precision);
        download(out, b.p(), n, s);
        return DD_OK;
    });
}

int dd_setup(dd_ctx* c, const dd_params* p) {
    if (check_params(p) || check(c)) return DD_ERR_ARG;
    return guard([&] {
        c->s->setup(*p);
        return DD_OK;
    });
}

int dd_setup_storage(dd_ctx* c, const dd_params* p) {
    if (check_params(p) || check(c)) return DD_ERR_ARG;
    return guard([&] {
        c->s->setup_storage(*p);
        return DD_OK;
    });
}

int dd_set_test_vectors(dd_ctx* c, int level, const double* tv, int precision) {
    if (check(c) || !tv) return DD_ERR_ARG;
    return guard([&] {
        auto run = [&](auto tag) {
            using R = decltype(tag);
            auto* m = mg<R>(c);
            if (level < 0 || level >= m->levels() - 1) throw dd::BadShape("test vectors exist below the coarsest level");
            const dd::LevelGeom& G = m->geom(level);
            hipStream_t s = c->s->s;
            const size_t tot = (size_t)G.V() * G.nv * G.dof;
            Staged a(tv, tot, s);
            DevMem t(sizeof(cplx<R>) * tot);
            dd::convert<R, double>(s, t.as<cplx<R>>(), a.p(), tot, 0, 0, 1, 1);
            m->set_test_vectors(level, t.as<cplx<R>>());
        };
        if (precision == 0) run(double());
        else run(float());
        return DD_OK;
    });
}

int dd_orthonormalize(dd_ctx* c, int level, int precision) {
    if (check(c)) return DD_ERR_ARG;
    return guard([&] {
        if (precision == 0) mg<double>(c)->orthonormalize_level(level);
        else mg<float>(c)->orthonormalize_level(level);
        return DD_OK;
    });
}

int dd_make_coarse_links(dd_ctx* c, int level, int precision) {
    if (check(c)) return DD_ERR_ARG;
    return guard([&] {
        if (precision == 0) mg<double>(c)->galerkin_level(level);
        else mg<float>(c)->galerkin_level(level);
        return DD_OK;
    });
}

double dd_setup_seconds(dd_ctx* c) { return (c && c->s) ? c->s->setup_seconds : -1.0; }

int dd_solve_dev(dd_ctx* c, int nrhs, const double* b, double* x, int* iters, dd_solve_info* info) {
    if (check(c) || !b || !x) return DD_ERR_ARG;
    return guard([&] {
        order_after_caller(c);
        return c->s->solve(nrhs, (const cplx<double>*)b, (cplx<double>*)x, iters, info);
    });
}

int dd_solve_many_dev(int n, dd_ctx* const* ctx, const int* nrhs, const double* const* b, double* const* x,
                      dd_solve_info* info) {
    if (n < 1 || !ctx || !nrhs || !b || !x) return DD_ERR_ARG;
    for (int i = 0; i < n; i++)
        if (check(ctx[i]) || !b[i] || !x[i] || nrhs[i] < 1) return DD_ERR_ARG;
    std::vector<int> st(n, DD_OK);
    std::vector<std::string> err(n);
    std::vector<std::thread> th;
    for (int i = 0; i < n; i++)
        th.emplace_back([&, i] {
            try {
                DD_HIP(hipSetDevice(ctx[i]->s->device));
                order_after_caller(ctx[i]);
                dd_solve_info inf{};
                st[i] = ctx[i]->s->solve(nrhs[i], (const cplx<double>*)b[i], (cplx<double>*)x[i], nullptr, &inf);
                if (info) info[i] = inf;
            } catch (const dd::BadShape& e) {
                st[i] = DD_ERR_BAD_SHAPE;
                err[i] = e.what;
            } catch (const dd::HipFailure& e) {
                st[i] = DD_ERR_HIP;
                err[i] = hipGetErrorString(e.err);
            } catch (...) {
                st[i] = DD_ERR_STATE;
                err[i] = "unexpected error";
            }
        });
    for (auto& t : th) t.join();
    int rc = DD_OK;
    for (int i = 0; i < n; i++)
        if (st[i] != DD_OK && (rc == DD_OK || rc == DD_ERR_NOT_CONVERGED)) {
            rc = st[i];
            if (!err[i].empty()) set_err("context " + std::to_string(i) + ": " + err[i]);
        }
    return rc;
}

int dd_solve(dd_ctx* c, int nrhs, const double* b, double* x, int* iters, dd_solve_info* info) {
    if (check(c) || !b || !x || nrhs < 1) return DD_ERR_ARG;
    return guard([&] {
        hipStream_t s = c->s->s;
        const size_t n = vsize(c) * nrhs;
        Staged bb(b, n, s), xx(nullptr, n, s);
        const int st = c->s->solve(nrhs, bb.p(), xx.p(), iters, info);
        download(x, xx.p(), n, s);
        return st;
    });
}

int dd_solve_cg(dd_ctx* c, int nrhs, const double* phi, double* x, double m0, int max_iter, double tol, int* iters,
                dd_solve_info* info) {
    if (check(c) || !phi || !x || nrhs < 1 || nrhs > 64 || max_iter < 1 || !(tol > 0)) return DD_ERR_ARG;
    return guard([&] {
        hipStream_t s = c->s->s;
        const size_t n = vsize(c) * nrhs;
        Staged bb(phi, n, s), xx(nullptr, n, s);
        const int st = c->s->solve_cg(nrhs, bb.p(), xx.p(), m0, max_iter, tol, iters, info);
        download(x, xx.p(), n, s);
        return st;
    });
}

int dd_solve_bicgstab(dd_ctx* c, int nrhs, const double* b, const double* x0, double* x, double m0, int max_iter,
                      double tol, int* iters, dd_solve_info* info) {
    if (check(c) || !b || !x || nrhs < 1 || nrhs > 64 || max_iter < 1 || !(tol > 0)) return DD_ERR_ARG;
    return guard([&] {
        hipStream_t s = c->s->s;
        const size_t n = vsize(c) * nrhs;
        Staged bb(b, n, s), xx(x0, n, s);
        if (!x0) DD_HIP(hipMemsetAsync(xx.p(), 0, sizeof(double) * 2 * n, s));
        const int st = c->s->solve_bicgstab(nrhs, bb.p(), xx.p(), xx.p(), m0, max_iter, tol, iters, info);
        download(x, xx.p(), n, s);
        return st;
    });
}

int dd_solve_gmres(dd_ctx* c, int nrhs, const double* b, const double* x0, double* x, double m0, int m,
                   int restarts, double tol, int sap_bx, int sap_bt, int* iters, dd_solve_info* info) {
    if (check(c) || !b || !x || nrhs < 1 || nrhs > 64 || m < 1 || restarts < 1 || !(tol > 0) || sap_bx < 0 ||
        (sap_bx > 0 && sap_bt < 1))
        return DD_ERR_ARG;
    return guard([&] {
        hipStream_t s = c->s->s;
        const size_t n = vsize(c) * nrhs;
        Staged bb(b, n, s), xx(x0, n, s);
        if (!x0) DD_HIP(hipMemsetAsync(xx.p(), 0, sizeof(double) * 2 * n, s));
        const int st = c->s->solve_gmres(nrhs, bb.p(), xx.p(), xx.p(), m0, m, restarts, tol, sap_bx, sap_bt, iters, info);
        download(x, xx.p(), n, s);
        return st;
    });
}

int dd_sap_fine(dd_ctx* c, int nrhs, const double* v, double* x, int nu, double tol, int sap_bx, int sap_bt,
                double m0, int* iters, double* relres) {
    if (check(c) || !v || !x || !iters || nrhs < 1 || nrhs > 64 || nu < 0 || !(tol >= 0)) return DD_ERR_ARG;
    return guard([&] {
        hipStream_t s = c->s->s;
        const size_t n = vsize(c) * nrhs;
        Staged vv(v, n, s), xx(x, n, s);
        const int st = c->s->sap_fine(nrhs, vv.p(), xx.p(), nu, tol, sap_bx, sap_bt, m0, iters, relres);
        download(x, xx.p(), n, s);
        return st;
    });
}

int dd_solve_x0(dd_ctx* c, int nrhs, const double* b, double* x, int* iters, dd_solve_info* info) {
    if (check(c) || !b || !x || nrhs < 1) return DD_ERR_ARG;
    return guard([&] {
        hipStream_t s = c->s->s;
        const size_t n = vsize(c) * nrhs;
        Staged bb(b, n, s), xx(x, n, s);
        const int st = c->s->solve(nrhs, bb.p(), xx.p(), iters, info, true);
        download(x, xx.p(), n, s);
        return st;
    });
}

int dd_solve_adjoint_dev(dd_ctx* c, int nrhs, const double* eta, double* y, int* iters, dd_solve_info* info) {
    if (check(c) || !eta || !y || nrhs < 1 || nrhs > 64) return DD_ERR_ARG;
    return guard([&] {
        hipStream_t s = c->s->s;
        const size_t n = vsize(c);
        order_after_caller(c);
        DevMem t(sizeof(cplx<double>) * n * nrhs);
        dd::sigma3<double>(s, t.as<cplx<double>>(), (const cplx<double>*)eta, n, n, nrhs, dd::full_mask(nrhs));
        const int st = c->s->solve(nrhs, t.as<cplx<double>>(), (cplx<double>*)y, iters, info);
        dd::sigma3<double>(s, (cplx<double>*)y, (const cplx<double>*)y, n, n, nrhs, dd::full_mask(nrhs));
        DD_HIP(hipStreamSynchronize(s));
        return st;
    });
}

int dd_solve_adjoint(dd_ctx* c, int nrhs, const double* eta, double* y, int* iters, dd_solve_info* info) {
    if (check(c) || !eta || !y || nrhs < 1) return DD_ERR_ARG;
    return guard([&] {
        hipStream_t s = c->s->s;
        const size_t n = vsize(c) * nrhs;
        Staged e(eta, n, s), yy(nullptr, n, s);
        const int st = dd_solve_adjoint_dev(c, nrhs, (const double*)e.p(), (double*)yy.p(), iters, info);
        if (st != DD_OK && st != DD_ERR_NOT_CONVERGED) return st;
        download(y, yy.p(), n, s);
        return st;
    });
}

int dd_apply_sigma3(dd_ctx* c, const double* in, double* out, int nrhs) {
    if (check(c) || !in || !out || nrhs < 1 || nrhs > 64) return DD_ERR_ARG;
    return guard([&] {
        hipStream_t s = c->s->s;
        const size_t n = vsize(c);
        Staged a(in, n * nrhs, s), b(nullptr, n * nrhs, s);
        dd::sigma3<double>(s, b.p(), a.p(), n, n, nrhs, dd::full_mask(nrhs));
        download(out, b.p(), n * nrhs, s);
        return DD_OK;
    });
}

int dd_dot_dev(dd_ctx* c, long long n, int nrhs, const double* a, const double* b, double* out) {
    if (check(c) || !a || !b || !out || n < 1 || nrhs < 1 || nrhs > 64) return DD_ERR_ARG;
    return guard([&] {
        hipStream_t s = c->s->s;
        order_after_caller(c);
        DevMem part(dd::dot_scratch_bytes(nrhs));
        dd::RedBuf rb;
        rb.partial = part.as<void>();
        dd::dot_partial<double>(s, (const cplx<double>*)a, (const cplx<double>*)b, (size_t)n, (size_t)n, nrhs,
                                dd::full_mask(nrhs), rb);
        dd::dot_final<double>(s, rb, (cplx<double>*)out, 1, nrhs, dd::full_mask(nrhs));
        DD_HIP(hipStreamSynchronize(s));
        return DD_OK;
    });
}

int dd_dot_dev_noalloc(dd_ctx* c, long long n, int nrhs, const double* a, const double* b, double* out) {
    if (check(c) || !a || !b || !out || n < 1) return DD_ERR_ARG;
    if (nrhs < 1 || nrhs > kNoAllocSystems) {
        set_err("dd_dot_dev_noalloc: nrhs " + std::to_string(nrhs) + " outside [1, " +
                std::to_string(kNoAllocSystems) + "] (the fixed device scratch holds " +
                std::to_string(kNoAllocSystems) + " systems)");
        return DD_ERR_ARG;
    }
    return guard([&] {
        DD_HIP(hipSetDevice(c->s->device));
        dd::RedBuf rb;
        rb.partial = scratch_on_current_device();
        // every partial slot the final pass reads is written by the first pass: nothing
        // in the scratch needs clearing between calls
        hipStream_t s = (hipStream_t)c->caller;
        dd::dot_partial<double>(s, (const cplx<double>*)a, (const cplx<double>*)b, (size_t)n, (size_t)n, nrhs,
                                dd::full_mask(nrhs), rb);
        dd::dot_final<double>(s, rb, (cplx<double>*)out, 1, nrhs, dd::full_mask(nrhs));
        return DD_OK;
    });
}

int dd_dot(dd_ctx* c, long long n, int nrhs, const double* a, const double* b, double* out) {
    if (check(c) || !a || !b || !out || n < 1 || nrhs < 1 || nrhs > 64) return DD_ERR_ARG;
    return guard([&] {
        hipStream_t s = c->s->s;
        Staged aa(a, (size_t)n * nrhs, s), bb(b, (size_t)n * nrhs, s), oo(nullptr, nrhs, s);
        const int st = dd_dot_dev(c, n, nrhs, (const double*)aa.p(), (const double*)bb.p(), (double*)oo.p());
        if (st != DD_OK) return st;
        download(out, oo.p(), nrhs, s);
        return DD_OK;
    });
}

int dd_dD_dtheta(dd_ctx* c, int nrhs, const double* y, const double* x, double* grad, double scale) {
    if (check(c) || !y || !x || !grad || nrhs < 1) return DD_ERR_ARG;
    return guard([&] {
        hipStream_t s = c->s->s;
        const size_t n = vsize(c);
        Staged yy(y, n * nrhs, s), xx(x, n * nrhs, s);
        DevMem g(sizeof(double) * n);
        DD_HIP(hipMemcpyAsync(g.as<void>(), grad, sizeof(double) * n, hipMemcpyHostToDevice, s));
        // the kernel's bracket omits the -1/2 of the hopping term
        dd::dtheta_contract(s, c->s->U64.as<cplx<double>>(), yy.p(), xx.p(), g.as<double>(), c->s->Nx, c->s->Nt,
                            nrhs, n, -0.5 * scale);
        DD_HIP(hipMemcpyAsync(grad, g.as<void>(), sizeof(double) * n, hipMemcpyDeviceToHost, s));
        DD_HIP(hipStreamSynchronize(s));
        return DD_OK;
    });
}

int dd_get_tuning(dd_ctx* c, int level, int* variant, double* ms) {
    if (check(c)) return DD_ERR_ARG;
    return guard([&] {
        if (!c->s->have_setup) throw dd::BadShape("no hierarchy");
        int v = -1;
        double t = 0;
        if (c->s->mg64) c->s->mg64->tuning(level, &v, &t);
        else c->s->mg32->tuning(level, &v, &t);
        if (variant) *variant = v;
        if (ms) *ms = t;
        return DD_OK;
    });
}

int dd_coarsest_stats(dd_ctx* c, double* out) {
    if (check(c) || !out) return DD_ERR_ARG;
    return guard([&] {
        if (!c->s->have_setup) throw dd::BadShape("no hierarchy");
        for (int k = 0; k < 12; k++) out[k] = 0;
        int dev = 0, khz = 0;
        DD_HIP(hipGetDevice(&dev));
        DD_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev));
        auto fill = [&](auto* m) {
            const long long* st = nullptr;
            const double* clk = nullptr;
            long long calls = 0;
            m->coarsest_stats(&st, &clk, &calls);
            out[0] = (double)calls;
            out[1] = st ? (double)*st : 0.0;
            for (int k = 0; k < 9 && clk; k++) out[2 + k] = khz > 0 ? clk[k] / khz : 0.0;
            out[11] = clk ? 1.0 : 0.0;
        };
        if (c->s->mg64) fill(c->s->mg64.get());
        else fill(c->s->mg32.get());
        return DD_OK;
    });
}

int dd_graph_count(dd_ctx* c, int* count) {
    if (check(c) || !count) return DD_ERR_ARG;
    return guard([&] {
        if (!c->s->have_setup) throw dd::BadShape("no hierarchy");
        *count = c->s->mg64 ? c->s->mg64->graph_count() : c->s->mg32->graph_count();
        return DD_OK;
    });
}

int dd_level_shape(dd_ctx* c, int level, int* Nx, int* Nt, int* dof) {
    if (check(c)) return DD_ERR_ARG;
    return guard([&] {
        if (!c->s->have_setup) throw dd::BadShape("no hierarchy");
        std::vector<dd::LevelGeom> g = dd::make_geometry(c->s->Nx, c->s->Nt, c->s->prm);
        if (level < 0 || level >= (int)g.size()) return (int)DD_ERR_ARG;
        if (Nx) *Nx = g[level].Nx;
        if (Nt) *Nt = g[level].Nt;
        if (dof) *dof = g[level].dof;
        return (int)DD_OK;
    });
}

int dd_get_coarse_poly(dd_ctx* c, int* degree, double* roots) {
    if (check(c) || !degree) return DD_ERR_ARG;
    return guard([&] {
        if (!c->s->have_setup) throw dd::BadShape("no hierarchy");
        const std::vector<std::complex<double>>& th =
            c->s->mg64 ? c->s->mg64->poly_roots() : c->s->mg32->poly_roots();
        *degree = (int)th.size();
        if (roots)
            for (size_t i = 0; i < th.size(); i++) {
                roots[2 * i] = th[i].real();
                roots[2 * i + 1] = th[i].imag();
            }
        return (int)DD_OK;
    });
}

int dd_poly_roots(const double* H, int d, double* roots) {
    if (!H || !roots || d < 1 || d > 64) return DD_ERR_ARG;
    return guard([&] {
        std::vector<std::complex<double>> h((size_t)(d + 1) * d);
        for (size_t i = 0; i < h.size(); i++) h[i] = {H[2 * i], H[2 * i + 1]};
        std::vector<std::complex<double>> th;
        try {
            th = dd::gmres_poly_roots(h, d);
        } catch (const std::runtime_error& e) {   // singular H_d, or the QR iteration failed
            set_err(e.what());
            return (int)DD_ERR_ARG;
        }
        for (int i = 0; i < d; i++) {
            roots[2 * i] = th[i].real();
            roots[2 * i + 1] = th[i].imag();
        }
        return (int)DD_OK;
    });
}

int dd_level_apply(dd_ctx* c, int level, const double* in, double* out, int nrhs, int precision) {
    if (check(c) || !in || !out || nrhs < 1 || nrhs > 64) return DD_ERR_ARG;
    return guard([&] {
        auto run = [&](auto tag) {
            using R = decltype(tag);
            auto* m = mg<R>(c);
            if (level < 0 || level >= m->levels()) throw dd::BadShape("bad level");
            if (nrhs > m->cap(level)) throw dd::BadShape("nrhs above the level's batch capacity");
            const size_t n = m->geom(level).n();
            with_level_io<R>(c, in, out, n, n, nrhs, [&](const cplx<R>* a, cplx<R>* b) {
                m->apply(level, a, b, nullptr, dd::full_mask(nrhs));
            });
        };
        if (precision == 0) run(double());
        else run(float());
        return DD_OK;
    });
}

int dd_cycle(dd_ctx* c, int level, int cycle, const double* eta, double* psi, int nrhs, int precision) {
    if (check(c) || !eta || !psi || nrhs < 1 || nrhs > 64) return DD_ERR_ARG;
    return guard([&] {
        auto run = [&](auto tag) {
            using R = decltype(tag);
            auto* m = mg<R>(c);
            if (level < 0 || level >= m->levels()) throw dd::BadShape("bad level");
            if (nrhs > std::max(1, c->s->prm.max_rhs)) throw dd::BadShape("nrhs above max_rhs");
            const size_t n = m->geom(level).n();
            with_level_io<R>(
                c, eta, psi, n, n, nrhs,
                [&](const cplx<R>* a, cplx<R>* b) {
                    if (cycle == 1) m->kcycle(level, a, b, dd::full_mask(nrhs));
                    else m->vcycle(level, a, b, dd::full_mask(nrhs));
                },
                psi);
        };
        if (precision == 0) run(double());
        else run(float());
        return DD_OK;
    });
}

int dd_restrict(dd_ctx* c, int level, const double* in, double* out, int nrhs, int precision) {
    if (check(c) || !in || !out || nrhs < 1 || nrhs > 64) return DD_ERR_ARG;
    return guard([&] {
        auto run = [&](auto tag) {
            using R = decltype(tag);
            auto* m = mg<R>(c);
            if (level < 0 || level >= m->levels() - 1) throw dd::BadShape("bad level");
            with_level_io<R>(c, in, out, m->geom(level).n(), m->geom(level + 1).n(), nrhs,
                             [&](const cplx<R>* a, cplx<R>* b) { m->restrict_(level, a, b, dd::full_mask(nrhs)); });
        };
        if (precision == 0) run(double());
        else run(float());
        return DD_OK;
    });
}

int dd_prolong(dd_ctx* c, int level, const double* in, double* out, int nrhs, int precision) {
    if (check(c) || !in || !out || nrhs < 1 || nrhs > 64) return DD_ERR_ARG;
    return guard([&] {
        auto run = [&](auto tag) {
            using R = decltype(tag);
            auto* m = mg<R>(c);
            if (level < 0 || level >= m->levels() - 1) throw dd::BadShape("bad level");
            with_level_io<R>(c, in, out, m->geom(level + 1).n(), m->geom(level).n(), nrhs,
                             [&](const cplx<R>* a, cplx<R>* b) {
                                 m->prolong(level, a, nullptr, b, dd::full_mask(nrhs));
                             });
        };
        if (precision == 0) run(double());
        else run(float());
        return DD_OK;
    });
}

int dd_sap(dd_ctx* c, int level, const double* v, double* x, int nu, double tol, int nrhs, int precision) {
    if (check(c) || !v || !x || nrhs < 1 || nrhs > 64) return DD_ERR_ARG;
    return guard([&] {
        auto run = [&](auto tag) {
            using R = decltype(tag);
            auto* m = mg<R>(c);
            if (level < 0 || level >= m->levels() - 1) throw dd::BadShape("no SAP on the coarsest level");
            if (nrhs > m->cap(level)) throw dd::BadShape("nrhs above the level's batch capacity");
            const size_t n = m->geom(level).n();
            with_level_io<R>(
                c, v, x, n, n, nrhs,
                [&](const cplx<R>* a, cplx<R>* b) { m->sap(level, a, b, nu, tol, dd::full_mask(nrhs)); }, x);
            m->check_sap_failures();
        };
        if (precision == 0) run(double());
        else run(float());
        return DD_OK;
    });
}

int dd_get_test_vectors(dd_ctx* c, int level, double* out, int precision) {
    if (check(c) || !out) return DD_ERR_ARG;
    return guard([&] {
        auto run = [&](auto tag) {
            using R = decltype(tag);
            auto* m = mg<R>(c);
            if (level < 0 || level >= m->levels() - 1) throw dd::BadShape("bad level");
            const dd::LevelGeom& G = m->geom(level);
            hipStream_t s = c->s->s;
            const size_t tot = (size_t)G.V() * G.nv * G.dof;
            DevMem t(sizeof(cplx<R>) * tot), o(sizeof(cplx<double>) * tot);
            dd::transpose_tv<R>(s, m->W(level), t.as<cplx<R>>(), G.V(), G.nv, G.dof, false);
            dd::convert<double, R>(s, o.as<cplx<double>>(), t.as<cplx<R>>(), tot, 0, 0, 1, 1);
            download(out, o.as<cplx<double>>(), tot, s);
        };
        if (precision == 0) run(double());
        else run(float());
        return DD_OK;
    });
}

int dd_get_links(dd_ctx* c, int level, double* out, int precision) {
    if (check(c) || !out) return DD_ERR_ARG;
    return guard([&] {
        auto run = [&](auto tag) {
            using R = decltype(tag);
            auto* m = mg<R>(c);
            if (level < 0 || level >= m->levels()) throw dd::BadShape("bad level");
            const dd::LevelGeom& G = m->geom(level);
            hipStream_t s = c->s->s;
            const size_t tot = (size_t)G.V() * 5 * G.dof * G.dof;
            DevMem o(sizeof(cplx<double>) * tot);
            if (level == 0) {   // the Wilson-Dirac link matrices (no A)
                DevMem l0(sizeof(cplx<R>) * tot);
                m->dirac_links_level0(l0.as<cplx<R>>());
                dd::convert<double, R>(s, o.as<cplx<double>>(), l0.as<cplx<R>>(), tot, 0, 0, 1, 1);
            } else if (precision == 2) {
                if constexpr (sizeof(R) == 4) {
                    if (!m->hlinks(level)) throw dd::BadShape("no binary16 link copy on this level");
                    dd::from_half(s, m->hlinks(level), o.as<double>(), 2 * tot);
                }
            } else {
                dd::convert<double, R>(s, o.as<cplx<double>>(), m->links(level), tot, 0, 0, 1, 1);
            }
            std::vector<double> h(2 * tot);
            download(h.data(), o.as<cplx<double>>(), tot, s);
            const int d = G.dof;
            for (size_t blk = 0; blk < (size_t)G.V() * 5; blk++)
                for (int row = 0; row < d; row++)
                    for (int col = 0; col < d; col++) {
                        const size_t src = (blk * d + col) * d + row, dst = (blk * d + row) * d + col;
                        out[2 * dst] = h[2 * src];
                        out[2 * dst + 1] = h[2 * src + 1];
                    }
        };
        if (precision == 0) run(double());
        else run(float());
        return DD_OK;
    });
}

int dd_timer_enable(int kernel, int on) {
    if (kernel < 0 || kernel >= dd::K_NUM) return DD_ERR_ARG;
    dd::timer_enable(kernel, on != 0);
    return DD_OK;
}

int dd_timer_reset(void) {
    dd::timer_reset();
    return DD_OK;
}

int dd_timer_flops(int kernel, double* flops) {
    if (kernel < 0 || kernel >= dd::K_NUM || !flops) return DD_ERR_ARG;
    return guard([&] {
        *flops = dd::timer_read_flops(kernel);
        return DD_OK;
    });
}

int dd_timer_read(int kernel, double* total_ms, long long* launches, double* bytes) {
    if (kernel < 0 || kernel >= dd::K_NUM) return DD_ERR_ARG;
    return guard([&] {
        long long c = 0;
        double by = 0;
        const double t = dd::timer_read(kernel, &c, &by);
        if (total_ms) *total_ms = t;
        if (launches) *launches = c;
        if (bytes) *bytes = by;
        return DD_OK;
    });
}

}  // extern "C"

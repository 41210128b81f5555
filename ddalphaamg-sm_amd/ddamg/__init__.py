"""Python binding (ctypes) over the C ABI of libddamg.so (include/ddamg.h).

The binding is a thin shell: every call goes to the HIP library; there is no CPU
fallback.  If the library is missing the import of this module raises.

Vectors use the reference's interior site order without halos: numpy complex128
arrays of shape (Nx, Nt, dof) or (nrhs, Nx, Nt, dof) with dof = colour*2 + spin.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import weakref

import numpy as np

try:
    # torch ships its own HIP runtime under the same soname (libamdhip64.so.7); whichever
    # is loaded first serves the whole process.  Loading torch's first keeps torch working
    # next to this library (the reverse order leaves torch without devices).
    import torch  # noqa: F401
except ImportError:
    pass

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get('DDAMG_LIB', os.path.join(os.path.dirname(_HERE), 'lib', 'libddamg.so'))
HEADER = os.path.join(os.path.dirname(os.path.dirname(_HERE)), 'include', 'ddamg.h')

if not os.path.exists(LIB_PATH):
    raise ImportError(f'libddamg.so not built ({LIB_PATH}); run python ddalphaamg-sm_amd/build.py')
lib = C.CDLL(LIB_PATH)

DD_OK, DD_ERR_NOT_CONVERGED, DD_ERR_BAD_SHAPE, DD_ERR_HIP, DD_ERR_STATE, DD_ERR_ARG = 0, 1, 2, 3, 4, 5
MAX_LEVELS = 8


class Params(C.Structure):
    """Mirror of dd_params (include/ddamg.h)."""
    _fields_ = [('size', C.c_int), ('levels', C.c_int),
                ('agg_x', C.c_int * MAX_LEVELS), ('agg_t', C.c_int * MAX_LEVELS),
                ('sap_x', C.c_int * MAX_LEVELS), ('sap_t', C.c_int * MAX_LEVELS),
                ('nv', C.c_int * MAX_LEVELS),
                ('m0', C.c_double), ('nu1', C.c_int), ('nu2', C.c_int), ('setup_sap_iters', C.c_int),
                ('sap_m', C.c_int), ('sap_restarts', C.c_int), ('sap_tol', C.c_double),
                ('coarse_m', C.c_int), ('coarse_restarts', C.c_int), ('coarse_tol', C.c_double),
                ('kcycle_m', C.c_int), ('kcycle_restarts', C.c_int), ('kcycle_tol', C.c_double),
                ('fgmres_m', C.c_int), ('fgmres_restarts', C.c_int), ('tol', C.c_double),
                ('cycle', C.c_int), ('precision', C.c_int),
                ('rng_ranks_x', C.c_int), ('rng_ranks_t', C.c_int),
                ('ortho_outer', C.c_int), ('ortho_inner', C.c_int), ('ortho_coarse', C.c_int),
                ('max_rhs', C.c_int), ('autotune', C.c_int),
                ('sap_variant', C.c_int), ('coarse_basis', C.c_int), ('coarse_sap_links', C.c_int),
                ('bootstrap', C.c_int), ('graph', C.c_int), ('sap_fused', C.c_int),
                ('coarse_poly', C.c_int)]


class SolveInfo(C.Structure):
    _fields_ = [('iterations', C.c_int), ('relres', C.c_double), ('seconds', C.c_double)]


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_ctx = C.c_void_p

_SIGS = {
    'dd_default_params_sized': (C.c_int, [C.POINTER(Params), C.c_int, C.c_int]),
    'dd_params_size': (C.c_int, []),
    'dd_finalize': (C.c_int, []),
    'dd_set_stream': (C.c_int, [_ctx, C.c_void_p]),
    'dd_apply_sigma3': (C.c_int, [_ctx, _dp, _dp, C.c_int]),
    'dd_dot': (C.c_int, [_ctx, C.c_longlong, C.c_int, _dp, _dp, _dp]),
    'dd_dot_dev': (C.c_int, [_ctx, C.c_longlong, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    'dd_setup_storage': (C.c_int, [_ctx, C.POINTER(Params)]),
    'dd_set_test_vectors': (C.c_int, [_ctx, C.c_int, _dp, C.c_int]),
    'dd_orthonormalize': (C.c_int, [_ctx, C.c_int, C.c_int]),
    'dd_make_coarse_links': (C.c_int, [_ctx, C.c_int, C.c_int]),
    'dd_sap_fine': (C.c_int, [_ctx, C.c_int, _dp, _dp, C.c_int, C.c_double, C.c_int, C.c_int, C.c_double, _ip, _dp]),
    'dd_dot_dev_noalloc': (C.c_int, [_ctx, C.c_longlong, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    'dd_create': (C.c_int, [C.POINTER(_ctx), C.c_int, C.c_int, C.c_int]),
    'dd_destroy': (C.c_int, [_ctx]),
    'dd_set_gauge': (C.c_int, [_ctx, _dp]),
    'dd_set_gauge_theta': (C.c_int, [_ctx, _dp]),
    'dd_generate_gauge': (C.c_int, [_ctx, C.c_double, C.c_int, C.c_ulonglong, C.c_double, _dp]),
    'dd_get_gauge': (C.c_int, [_ctx, _dp]),
    'dd_apply_D': (C.c_int, [_ctx, _dp, _dp, C.c_int, C.c_double, C.c_int, C.c_int]),
    'dd_apply_D_dev': (C.c_int, [_ctx, C.c_void_p, C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_int]),
    'dd_setup': (C.c_int, [_ctx, C.POINTER(Params)]),
    'dd_setup_seconds': (C.c_double, [_ctx]),
    'dd_solve': (C.c_int, [_ctx, C.c_int, _dp, _dp, _ip, C.POINTER(SolveInfo)]),
    'dd_solve_dev': (C.c_int, [_ctx, C.c_int, C.c_void_p, C.c_void_p, _ip, C.POINTER(SolveInfo)]),
    'dd_solve_adjoint': (C.c_int, [_ctx, C.c_int, _dp, _dp, _ip, C.POINTER(SolveInfo)]),
    'dd_solve_x0': (C.c_int, [_ctx, C.c_int, _dp, _dp, _ip, C.POINTER(SolveInfo)]),
    'dd_solve_adjoint_dev': (C.c_int, [_ctx, C.c_int, C.c_void_p, C.c_void_p, _ip, C.POINTER(SolveInfo)]),
    'dd_dD_dtheta': (C.c_int, [_ctx, C.c_int, _dp, _dp, _dp, C.c_double]),
    'dd_level_apply': (C.c_int, [_ctx, C.c_int, _dp, _dp, C.c_int, C.c_int]),
    'dd_cycle': (C.c_int, [_ctx, C.c_int, C.c_int, _dp, _dp, C.c_int, C.c_int]),
    'dd_restrict': (C.c_int, [_ctx, C.c_int, _dp, _dp, C.c_int, C.c_int]),
    'dd_prolong': (C.c_int, [_ctx, C.c_int, _dp, _dp, C.c_int, C.c_int]),
    'dd_sap': (C.c_int, [_ctx, C.c_int, _dp, _dp, C.c_int, C.c_double, C.c_int, C.c_int]),
    'dd_get_test_vectors': (C.c_int, [_ctx, C.c_int, _dp, C.c_int]),
    'dd_get_links': (C.c_int, [_ctx, C.c_int, _dp, C.c_int]),
    'dd_level_shape': (C.c_int, [_ctx, C.c_int, _ip, _ip, _ip]),
    'dd_get_coarse_poly': (C.c_int, [_ctx, _ip, _dp]),
    'dd_poly_roots': (C.c_int, [_dp, C.c_int, _dp]),
    'dd_get_tuning': (C.c_int, [_ctx, C.c_int, _ip, _dp]),
    'dd_coarsest_stats': (C.c_int, [_ctx, _dp]),
    'dd_graph_count': (C.c_int, [_ctx, C.POINTER(C.c_int)]),
    'dd_solve_cg': (C.c_int, [_ctx, C.c_int, _dp, _dp, C.c_double, C.c_int, C.c_double, _ip, C.POINTER(SolveInfo)]),
    'dd_solve_bicgstab': (C.c_int, [_ctx, C.c_int, _dp, _dp, _dp, C.c_double, C.c_int, C.c_double, _ip,
                                    C.POINTER(SolveInfo)]),
    'dd_solve_gmres': (C.c_int, [_ctx, C.c_int, _dp, _dp, _dp, C.c_double, C.c_int, C.c_int, C.c_double, C.c_int,
                                 C.c_int, _ip, C.POINTER(SolveInfo)]),
    'dd_solve_many_dev': (C.c_int, [C.c_int, C.POINTER(_ctx), _ip, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                    C.POINTER(SolveInfo)]),
    'dd_last_error': (C.c_char_p, []),
    'dd_timer_enable': (C.c_int, [C.c_int, C.c_int]),
    'dd_timer_reset': (C.c_int, []),
    'dd_timer_read': (C.c_int, [C.c_int, _dp, C.POINTER(C.c_longlong), _dp]),
    'dd_timer_flops': (C.c_int, [C.c_int, _dp]),
}
K_WILSON, K_COARSE, K_SAP0, K_SAPC, K_MULTIDOT, K_MULTIAXPY, K_RESTRICT, K_PROLONG, K_COARSEST, K_DOT = range(10)


def timer_enable(kernel: int, on: bool = True):
    _chk(lib.dd_timer_enable(kernel, int(on)))


def timer_reset():
    _chk(lib.dd_timer_reset())


def timer_read(kernel: int):
    """(total ms, launches, algorithmic bytes) of the timed launches of `kernel` since
    the last reset."""
    t, n, b = C.c_double(), C.c_longlong(), C.c_double()
    _chk(lib.dd_timer_read(kernel, C.byref(t), C.byref(n), C.byref(b)))
    return t.value, n.value, b.value


def timer_flops(kernel: int) -> float:
    """Algorithmic floating-point operations the timed launches of `kernel` counted on
    the device since the last reset (the coarse SAP's matrix-core products; 0 for the
    other families)."""
    f = C.c_double()
    _chk(lib.dd_timer_flops(kernel, C.byref(f)))
    return f.value
for _name, (_res, _args) in _SIGS.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args

if C.sizeof(Params) != lib.dd_params_size():
    raise ImportError(f'Params mirror has {C.sizeof(Params)} bytes, the library\'s dd_params '
                      f'{lib.dd_params_size()}: the binding and include/ddamg.h disagree')
_LIVE = weakref.WeakSet()   # open contexts


@atexit.register
def _finalize():
    """Destroys the open contexts and the library's process-wide HIP objects while the
    HIP runtime is still up (interpreter exit runs this before the runtime's own
    teardown)."""
    for c in list(_LIVE):
        c.close()
    lib.dd_finalize()


class DDError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f'{msg} (status {code})')
        self.code = code


def _chk(code, allow=(DD_OK,)):
    if code not in allow:
        raise DDError(code, lib.dd_last_error().decode())
    return code


def _cptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


def _c128(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.complex128)


def params(levels: int = 2, **kw) -> Params:
    """dd_params with the reference defaults; per-level lists (agg, sap_block, nv) and
    scalar knobs as keywords: agg=[(4,4),...], sap_block=[(4,4),...], nv=[8,...]."""
    p = Params()
    _chk(lib.dd_default_params_sized(C.byref(p), C.sizeof(Params), levels))
    for l, (ax, at) in enumerate(kw.pop('agg', [])):
        p.agg_x[l], p.agg_t[l] = ax, at
    for l, (bx, bt) in enumerate(kw.pop('sap_block', [])):
        p.sap_x[l], p.sap_t[l] = bx, bt
    for l, nv in enumerate(kw.pop('nv', [])):
        p.nv[l] = nv
    if 'rng_ranks' in kw:
        p.rng_ranks_x, p.rng_ranks_t = kw.pop('rng_ranks')
    for k, v in kw.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


class Context:
    """One lattice (Nx x Nt) on one GPU: gauge field, hierarchy and solves."""

    def __init__(self, Nx: int, Nt: int, device: int = 0):
        self.Nx, self.Nt = Nx, Nt
        self._h = _ctx()
        _chk(lib.dd_create(C.byref(self._h), Nx, Nt, device))
        self.prm: Params | None = None
        _LIVE.add(self)

    def close(self):
        if self._h:
            lib.dd_destroy(self._h)
            self._h = _ctx()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- gauge ------------------------------------------------------------------------
    def set_gauge(self, U: np.ndarray):
        U = _c128(U)
        assert U.shape == (self.Nx, self.Nt, 2)
        _chk(lib.dd_set_gauge(self._h, _cptr(U)))

    def set_gauge_theta(self, theta: np.ndarray):
        th = np.ascontiguousarray(theta, dtype=np.float64)
        assert th.shape == (self.Nx, self.Nt, 2)
        _chk(lib.dd_set_gauge_theta(self._h, _cptr(th)))

    def generate_gauge(self, beta: float, sweeps: int, seed: int, step: float = 0.5) -> np.ndarray:
        th = np.zeros((self.Nx, self.Nt, 2))
        _chk(lib.dd_generate_gauge(self._h, beta, sweeps, seed, step, _cptr(th)))
        return th

    def gauge(self) -> np.ndarray:
        U = np.zeros((self.Nx, self.Nt, 2), complex)
        _chk(lib.dd_get_gauge(self._h, _cptr(U)))
        return U

    # -- operator -----------------------------------------------------------------------
    def apply_D(self, v: np.ndarray, m0: float, dagger: bool = False, precision: int = 0) -> np.ndarray:
        v = _c128(v)
        batch = v.ndim == 4
        vb = v if batch else v[None]
        out = np.zeros_like(vb)
        _chk(lib.dd_apply_D(self._h, _cptr(vb), _cptr(out), vb.shape[0], m0, int(dagger), precision))
        return out if batch else out[0]

    # -- hierarchy and solves -----------------------------------------------------------------
    def setup(self, prm: Params) -> float:
        _chk(lib.dd_setup(self._h, C.byref(prm)))
        self.prm = prm
        return lib.dd_setup_seconds(self._h)

    def solve(self, b: np.ndarray, adjoint: bool = False, check: bool = True):
        """Returns (x, iterations per system, info)."""
        b = _c128(b)
        batch = b.ndim == 4
        bb = b if batch else b[None]
        nr = bb.shape[0]
        x = np.zeros_like(bb)
        its = (C.c_int * nr)()
        info = SolveInfo()
        f = lib.dd_solve_adjoint if adjoint else lib.dd_solve
        code = f(self._h, nr, _cptr(bb), _cptr(x), its, C.byref(info))
        _chk(code, allow=(DD_OK,) if check else (DD_OK, DD_ERR_NOT_CONVERGED))
        its = np.array(list(its))
        return (x if batch else x[0]), its, info

    def solve_method(self, method: str, b: np.ndarray, m0: float, x0: np.ndarray | None = None,
                     tol: float = 1e-10, max_iter: int = 10000, m: int = 20, restarts: int = 50,
                     sap_block: tuple[int, int] = (4, 4), check: bool = True):
        """The reference's comparison solvers (src/methods.cpp) on the FP64 Wilson-Dirac
        operator, for D x = b: method 'cg' (CG on D^dagger D x = D^dagger b from
        D^dagger b, as Methods::CG; x0 ignored), 'bicgstab', 'gmres' (no preconditioner) or
        'fgmres_sap' (one SAP iteration over sap_block-site blocks per step).
        b: [nrhs] + lattice shape or lattice shape.  Returns (x, iterations per system,
        info); info.relres is the residual of the system the method solves."""
        b = _c128(b)
        batch = b.ndim == 4
        bb = b if batch else b[None]
        nr = bb.shape[0]
        x = np.zeros_like(bb)
        x0p = None
        if x0 is not None:
            x0c = _c128(x0)
            x0c = x0c if x0c.ndim == 4 else x0c[None]
            x0p = _cptr(x0c)
        its = (C.c_int * nr)()
        info = SolveInfo()
        if method == 'cg':
            phi = self.apply_D(bb, m0, dagger=True)
            code = lib.dd_solve_cg(self._h, nr, _cptr(phi), _cptr(x), m0, max_iter, tol, its, C.byref(info))
        elif method == 'bicgstab':
            code = lib.dd_solve_bicgstab(self._h, nr, _cptr(bb), x0p, _cptr(x), m0, max_iter, tol, its,
                                         C.byref(info))
        elif method in ('gmres', 'fgmres_sap'):
            sb = sap_block if method == 'fgmres_sap' else (0, 0)
            code = lib.dd_solve_gmres(self._h, nr, _cptr(bb), x0p, _cptr(x), m0, m, restarts, tol, sb[0], sb[1],
                                      its, C.byref(info))
        else:
            raise ValueError(f'unknown method {method!r}')
        _chk(code, allow=(DD_OK,) if check else (DD_OK, DD_ERR_NOT_CONVERGED))
        return (x if batch else x[0]), np.array(list(its)), info

    def _order_after_torch(self):
        """The next _dev call waits for the work torch has queued on its current stream."""
        s = None
        try:
            import torch
            if torch.cuda.is_initialized():
                s = torch.cuda.current_stream().cuda_stream or None
        except Exception:  # noqa: BLE001  (no torch: the null stream)
            pass
        _chk(lib.dd_set_stream(self._h, s))

    def solve_dev(self, b_ptr: int, x_ptr: int, nrhs: int, adjoint: bool = False, check: bool = True):
        """Solve with device pointers (complex128 interleaved, nrhs systems back to back),
        e.g. torch tensors' data_ptr().  The solve first waits for the work queued on
        torch's current stream.  Returns (iterations per system, info)."""
        self._order_after_torch()
        its = (C.c_int * nrhs)()
        info = SolveInfo()
        f = lib.dd_solve_adjoint_dev if adjoint else lib.dd_solve_dev
        code = f(self._h, nrhs, C.c_void_p(b_ptr), C.c_void_p(x_ptr), its, C.byref(info))
        _chk(code, allow=(DD_OK,) if check else (DD_OK, DD_ERR_NOT_CONVERGED))
        return np.array(list(its)), info

    def apply_D_dev(self, in_ptr: int, out_ptr: int, nrhs: int, m0: float, dagger: bool = False,
                    precision: int = 0):
        self._order_after_torch()
        _chk(lib.dd_apply_D_dev(self._h, C.c_void_p(in_ptr), C.c_void_p(out_ptr), nrhs, m0, int(dagger),
                                precision))

    def sigma3(self, v: np.ndarray) -> np.ndarray:
        """sigma3 v (spin-1 components negated) of level-0 vectors."""
        v = _c128(v)
        vb = v if v.ndim == 4 else v[None]
        out = np.zeros_like(vb)
        _chk(lib.dd_apply_sigma3(self._h, _cptr(vb), _cptr(out), vb.shape[0]))
        return out if v.ndim == 4 else out[0]

    def dot(self, a: np.ndarray, b: np.ndarray) -> np.ndarray:
        """sum a conj(b) per system: a, b of shape (nrhs, ...) (or one vector)."""
        a, b = _c128(a), _c128(b)
        single = a.ndim < 4
        ab = a.reshape(1 if single else a.shape[0], -1)
        bb = b.reshape(ab.shape)
        out = np.zeros(ab.shape[0], complex)
        _chk(lib.dd_dot(self._h, ab.shape[1], ab.shape[0], _cptr(np.ascontiguousarray(ab)),
                        _cptr(np.ascontiguousarray(bb)), _cptr(out)))
        return out[0] if single else out

    def dot_dev(self, a_ptr: int, b_ptr: int, out_ptr: int, n: int, nrhs: int):
        """Device dot products: out[r] = sum a_r conj(b_r), n complex entries per system."""
        self._order_after_torch()
        _chk(lib.dd_dot_dev(self._h, n, nrhs, C.c_void_p(a_ptr), C.c_void_p(b_ptr), C.c_void_p(out_ptr)))

    def dD_dtheta(self, y: np.ndarray, x: np.ndarray, scale: float = -1.0) -> np.ndarray:
        y, x = _c128(y), _c128(x)
        yb = y if y.ndim == 4 else y[None]
        xb = x if x.ndim == 4 else x[None]
        g = np.zeros((self.Nx, self.Nt, 2))
        _chk(lib.dd_dD_dtheta(self._h, yb.shape[0], _cptr(yb), _cptr(xb), _cptr(g), scale))
        return g

    def tuning(self, level: int):
        """(variant, ms) the autotuner chose for the coarse SAP of `level`."""
        v, t = C.c_int(), C.c_double()
        _chk(lib.dd_get_tuning(self._h, level, C.byref(v), C.byref(t)))
        return v.value, t.value

    def coarsest_stats(self) -> dict:
        """Coarsest-solver counters since setup (solves, Arnoldi steps, phase ms)."""
        out = (C.c_double * 12)()
        _chk(lib.dd_coarsest_stats(self._h, out))
        keys = ['solves', 'steps', 'ms_reduce', 'ms_publish_partials', 'ms_scalar', 'ms_wait_coefs', 'ms_update',
                'ms_publish_update', 'ms_cycle_end', 'ms_wait_neighbours_stencil', 'ms_gather', 'persistent']
        return dict(zip(keys, list(out)))

    def graph_count(self) -> int:
        """V-cycle graphs captured by the hierarchy (params graph = 1)."""
        k = C.c_int()
        _chk(lib.dd_graph_count(self._h, C.byref(k)))
        return k.value

    def level_shape(self, level: int):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        _chk(lib.dd_level_shape(self._h, level, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def _level_op(self, fn, level, v, out_level, precision, *extra):
        v = _c128(v)
        batch = v.ndim == 4
        vb = v if batch else v[None]
        out = np.zeros((vb.shape[0],) + self.level_shape(out_level), complex)
        _chk(fn(self._h, level, _cptr(vb), _cptr(out), *extra, vb.shape[0], precision))
        return out if batch else out[0]

    def level_apply(self, level, v, precision=0):
        return self._level_op(lib.dd_level_apply, level, v, level, precision)

    def restrict(self, level, v, precision=0):
        return self._level_op(lib.dd_restrict, level, v, level + 1, precision)

    def prolong(self, level, v, precision=0):
        """P v for a level+1 vector v; the result lives on `level`."""
        v = _c128(v)
        batch = v.ndim == 4
        vb = v if batch else v[None]
        out = np.zeros((vb.shape[0],) + self.level_shape(level), complex)
        _chk(lib.dd_prolong(self._h, level, _cptr(vb), _cptr(out), vb.shape[0], precision))
        return out if batch else out[0]

    def sap(self, level, v, x0, nu, tol=1e-10, precision=0):
        v, x = _c128(v), _c128(x0).copy()
        batch = v.ndim == 4
        vb = v if batch else v[None]
        xb = x if batch else x[None]
        _chk(lib.dd_sap(self._h, level, _cptr(vb), _cptr(xb), nu, tol, vb.shape[0], precision))
        return xb if batch else xb[0]

    def cycle(self, level, eta, psi0, kcycle=True, precision=0):
        """One K- or V-cycle from `level` with initial guess psi0 (returns psi)."""
        eta, psi = _c128(eta), _c128(psi0).copy()
        batch = eta.ndim == 4
        eb = eta if batch else eta[None]
        pb = psi if batch else psi[None]
        _chk(lib.dd_cycle(self._h, level, int(kcycle), _cptr(eb), _cptr(pb), eb.shape[0], precision))
        return pb if batch else pb[0]

    def test_vectors(self, level, precision=0):
        Nx, Nt, d = self.level_shape(level)
        nv = self.prm.nv[level]
        out = np.zeros((nv, Nx, Nt, d), complex)
        _chk(lib.dd_get_test_vectors(self._h, level, _cptr(out), precision))
        return out

    def setup_storage(self, prm: Params):
        """Allocates the hierarchy of prm without running the setup phase."""
        _chk(lib.dd_setup_storage(self._h, C.byref(prm)))
        self.prm = prm

    def set_test_vectors(self, level, tv, precision=0):
        tv = _c128(tv)
        _chk(lib.dd_set_test_vectors(self._h, level, _cptr(tv), precision))

    def orthonormalize(self, level, precision=0):
        _chk(lib.dd_orthonormalize(self._h, level, precision))

    def make_coarse_links(self, level, precision=0):
        _chk(lib.dd_make_coarse_links(self._h, level, precision))

    def sap_fine(self, v, x0, nu, tol, block, m0):
        """Stand-alone red-black SAP (dd_sap_fine) on level-0 vectors; returns (x, iters,
        relres) per system."""
        v, x = _c128(v), _c128(x0).copy()
        vb = v if v.ndim == 4 else v[None]
        xb = x if x.ndim == 4 else x[None]
        its = (C.c_int * vb.shape[0])()
        rr = np.zeros(vb.shape[0])
        _chk(lib.dd_sap_fine(self._h, vb.shape[0], _cptr(vb), _cptr(xb), nu, tol, block[0], block[1], m0, its,
                             _cptr(rr)), allow=(DD_OK, DD_ERR_NOT_CONVERGED))
        return (xb if v.ndim == 4 else xb[0]), np.array(list(its)), rr

    def coarse_poly_roots(self):
        """Roots of the coarsest GMRES's polynomial preconditioner in their order of
        application (empty without dd_params.coarse_poly)."""
        d = C.c_int()
        out = np.zeros(64, complex)
        _chk(lib.dd_get_coarse_poly(self._h, C.byref(d), _cptr(out)))
        return out[:d.value].copy()

    def links(self, level, precision=0):
        """(A, B[2], C[2]) of level >= 1, each (Nx, Nt, dof, dof); precision 0 FP64, 1 FP32,
        2 the binary16 copy of the FP32 links that the matrix-core coarse SAP reads."""
        Nx, Nt, d = self.level_shape(level)
        out = np.zeros((Nx, Nt, 5, d, d), complex)
        _chk(lib.dd_get_links(self._h, level, _cptr(out), precision))
        return out[:, :, 0], np.stack([out[:, :, 1], out[:, :, 2]]), np.stack([out[:, :, 3], out[:, :, 4]])


def poly_roots(H):
    """Harmonic Ritz values of the Arnoldi relation H ((d+1) x d) in Leja order, as the
    library orders the coarsest polynomial preconditioner's roots (dd_poly_roots; host
    code, no device needed)."""
    H = np.ascontiguousarray(H, dtype=complex)
    d = H.shape[1]
    if H.shape != (d + 1, d):
        raise ValueError('H must be (d+1) x d')
    out = np.zeros(d, complex)
    _chk(lib.dd_poly_roots(_cptr(H), d, _cptr(out)))
    return out


def solve_many_dev(contexts, nrhs, b_ptrs, x_ptrs):
    """Ragged batch (dd_solve_many_dev): context i solves nrhs[i] systems with device
    pointers b_ptrs[i], x_ptrs[i]; all contexts at the same time.  Returns the SolveInfo
    list."""
    n = len(contexts)
    for c in contexts:
        c._order_after_torch()
    ctxs = (C.c_void_p * n)(*[c._h.value for c in contexts])
    nr = (C.c_int * n)(*nrhs)
    bp = (C.c_void_p * n)(*b_ptrs)
    xp = (C.c_void_p * n)(*x_ptrs)
    info = (SolveInfo * n)()
    _chk(lib.dd_solve_many_dev(n, ctxs, nr, bp, xp, info), allow=(DD_OK, DD_ERR_NOT_CONVERGED))
    return list(info)


def header_symbols(path: str = HEADER) -> list[str]:
    """Function names declared in include/ddamg.h."""
    import re
    with open(path) as f:
        txt = f.read()
    return sorted(set(re.findall(r'^\s*(?:int|double|const char\*)\s+(dd_\w+)\s*\(', txt, re.M)))

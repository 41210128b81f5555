"""CPU checks of the C ABI: the library loads, exports every function include/ddamg.h
declares, and the host-only parameter defaults are the reference's knobs
(src/variables.cpp:90-134)."""
import ctypes as C

import pytest

ddamg = pytest.importorskip('ddamg')


def test_exports_every_declared_symbol():
    names = ddamg.header_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(ddamg.lib, n), n


def test_default_params_are_reference_knobs():
    p = ddamg.params(3)
    assert (p.levels, p.nu1, p.nu2, p.setup_sap_iters) == (3, 0, 2, 4)
    assert (p.sap_m, p.sap_restarts, p.sap_tol) == (5, 5, 1e-3)
    assert (p.coarse_m, p.coarse_restarts, p.coarse_tol) == (400, 10, 0.1)
    assert (p.kcycle_m, p.kcycle_restarts, p.kcycle_tol) == (5, 2, 0.1)
    assert (p.fgmres_m, p.fgmres_restarts, p.tol) == (20, 50, 1e-10)


def test_params_keywords():
    p = ddamg.params(3, agg=[(4, 4), (2, 2)], sap_block=[(4, 4), (2, 2)], nv=[12, 12], m0=-0.18,
                     rng_ranks=(2, 4), fgmres_m=30)
    assert (p.agg_x[1], p.sap_t[1], p.nv[0], p.rng_ranks_t, p.fgmres_m) == (2, 2, 12, 4, 30)
    with pytest.raises(KeyError):
        ddamg.params(2, nonsense=1)


def test_bad_arguments_are_status_codes():
    sz = C.sizeof(ddamg.Params)
    assert ddamg.lib.dd_default_params_sized(None, sz, 2) == ddamg.DD_ERR_ARG
    p = ddamg.Params()
    assert ddamg.lib.dd_default_params_sized(C.byref(p), sz, 1) == ddamg.DD_ERR_ARG
    assert ddamg.lib.dd_create(None, 8, 8, 0) == ddamg.DD_ERR_ARG
    assert ddamg.lib.dd_solve(None, 1, None, None, None, None) == ddamg.DD_ERR_ARG


def test_params_mirror_matches_library():
    assert C.sizeof(ddamg.Params) == ddamg.lib.dd_params_size()
    assert ddamg.params(2).size == C.sizeof(ddamg.Params)


def test_params_of_another_layout_are_rejected():
    """A caller compiled against a dd_params of another size gets DD_ERR_ARG and its
    struct is left untouched (no write past its end); dd_setup refuses a struct whose
    size field is not the library's."""
    sz = C.sizeof(ddamg.Params)
    for wrong in (sz - 8, sz - 4, sz + 8):
        buf = (C.c_ubyte * (sz + 64))(*([0xA5] * (sz + 64)))
        assert ddamg.lib.dd_default_params_sized(C.cast(buf, C.POINTER(ddamg.Params)), wrong, 3) == ddamg.DD_ERR_ARG
        assert bytes(buf) == bytes([0xA5] * (sz + 64))
        assert 'dd_params' in ddamg.lib.dd_last_error().decode()
    p = ddamg.params(3)
    p.size = sz - 4
    assert ddamg.lib.dd_setup(None, C.byref(p)) == ddamg.DD_ERR_ARG
    assert 'bytes' in ddamg.lib.dd_last_error().decode()

"""dd_dot_dev_noalloc (the entry point that allocates nothing) against dd_dot_dev, through
the C ABI: two calls in a row, the second on a smaller input, bit for bit equal; and the
size limit of its fixed device scratch is a clean error."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
ddamg = pytest.importorskip('ddamg')
torch = pytest.importorskip('torch')


def _dot_pair(ctx, a, b, n, nrhs):
    ref = torch.zeros(2 * nrhs, dtype=torch.float64, device='cuda')
    got = torch.full((2 * nrhs,), float('nan'), dtype=torch.float64, device='cuda')
    ctx.dot_dev(a.data_ptr(), b.data_ptr(), ref.data_ptr(), n, nrhs)
    ddamg._chk(ddamg.lib.dd_set_stream(ctx._h, None))   # the null stream, torch's current one
    ddamg._chk(ddamg.lib.dd_dot_dev_noalloc(ctx._h, n, nrhs, C.c_void_p(a.data_ptr()), C.c_void_p(b.data_ptr()),
                                            C.c_void_p(got.data_ptr())))
    torch.cuda.synchronize()
    return ref.cpu().numpy(), got.cpu().numpy()


def test_noalloc_dot_equals_dot_dev():
    rng = np.random.default_rng(77)
    N = 128
    with ddamg.Context(N, N) as ctx:
        for nrhs, n in ((8, N * N * 2), (3, 1000)):     # second call: smaller input
            a = rng.standard_normal((nrhs, n)) + 1j * rng.standard_normal((nrhs, n))
            b = rng.standard_normal((nrhs, n)) + 1j * rng.standard_normal((nrhs, n))
            ad = torch.from_numpy(a.view(np.float64).copy()).cuda()
            bd = torch.from_numpy(b.view(np.float64).copy()).cuda()
            ref, got = _dot_pair(ctx, ad, bd, n, nrhs)
            assert np.array_equal(ref, got), (nrhs, n)
            want = np.sum(a * np.conj(b), axis=1)
            assert np.abs(got.view(np.complex128) - want).max() < 1e-12 * np.abs(want).max()
        x = torch.zeros(2 * 65, dtype=torch.float64, device='cuda')
        st = ddamg.lib.dd_dot_dev_noalloc(ctx._h, 10, 65, C.c_void_p(x.data_ptr()), C.c_void_p(x.data_ptr()),
                                          C.c_void_p(x.data_ptr()))
        assert st == ddamg.DD_ERR_ARG
        assert 'nrhs 65' in ddamg.lib.dd_last_error().decode()

"""GPU checks of the small C-ABI entry points (dd_dot, dd_dot_dev, dd_apply_sigma3) and of
the stream ordering of the _dev calls after work that torch queued and did not wait for."""
import numpy as np
import pytest

from oracle import ddamg_oracle as orc

pytestmark = pytest.mark.gpu
ddamg = pytest.importorskip('ddamg')


def test_dot_and_sigma3():
    rng = np.random.default_rng(5)
    N = 48
    a = rng.standard_normal((3, N, N, 2)) + 1j * rng.standard_normal((3, N, N, 2))
    b = rng.standard_normal((3, N, N, 2)) + 1j * rng.standard_normal((3, N, N, 2))
    with ddamg.Context(N, N) as ctx:
        ctx.set_gauge(np.ones((N, N, 2), complex))
        got = ctx.dot(a, b)
        want = np.array([np.sum(a[r] * np.conj(b[r])) for r in range(3)])   # include/utils.h:68-84
        assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()
        s3 = ctx.sigma3(a)
        for r in range(3):
            assert np.array_equal(s3[r], orc.sigma3(a[r]))


def test_dev_calls_wait_for_torch_stream():
    """Inputs that torch writes on its stream right before a _dev call (no host
    synchronisation in between) are complete when the library reads them."""
    torch = pytest.importorskip('torch')
    rng = np.random.default_rng(6)
    N, m0 = 256, -0.18
    U = np.exp(1j * rng.uniform(-np.pi, np.pi, (N, N, 2)))
    v = rng.standard_normal((4, N, N, 2)) + 1j * rng.standard_normal((4, N, N, 2))
    with ddamg.Context(N, N) as ctx:
        ctx.set_gauge(U)
        want = ctx.apply_D(v, m0)
        src = torch.from_numpy(v.view(np.float64).copy()).cuda()
        torch.cuda.synchronize()
        for rep in range(3):
            x = torch.zeros_like(src)
            out = torch.full_like(src, float('nan'))
            for _ in range(20):            # a queue of torch work ending in the input
                x.mul_(0.5).add_(src, alpha=0.5)
            x.copy_(src)
            ctx.apply_D_dev(x.data_ptr(), out.data_ptr(), 4, m0)
            got = out.cpu().numpy().view(np.complex128).reshape(v.shape)
            assert np.abs(got - want).max() < 1e-13 * np.abs(want).max()
            d = torch.zeros(8, dtype=torch.float64, device='cuda')
            ctx.dot_dev(x.data_ptr(), x.data_ptr(), d.data_ptr(), N * N * 2, 4)
            dn = d.cpu().numpy().view(np.complex128)
            assert np.allclose(dn.real, np.sum(np.abs(v.reshape(4, -1)) ** 2, axis=1), rtol=1e-12)

"""CPU checks of the host computation behind the coarsest GMRES's polynomial
preconditioner (dd_poly_roots, csrc/poly.h; no device needed).

For Arnoldi relations of non-Hermitian complex test matrices (a shifted random matrix,
and a 1-D advection-diffusion stencil with a complex shift, degrees 1 to 32) the roots
equal the harmonic Ritz values found from their definition (the generalized
eigenproblem over the basis and its image, tests/poly_ref.py) to 1e-9, and come in Leja
order.  The polynomial they define, applied in the product form, minimises the residual
of d GMRES steps from the start vector: ||v - D p(D) v|| equals the GMRES residual from
the Hessenberg least-squares problem.  A singular H_d is refused with DD_ERR_ARG.
"""
import numpy as np
import pytest

from poly_ref import arnoldi, assert_leja_order, harmonic_ritz_from_images, match_distance

ddamg = pytest.importorskip('ddamg')


def matrices():
    rng = np.random.default_rng(31)
    n = 80
    M1 = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n) + 2.5 * np.eye(n)
    main = (2.0 + 0.3j) * np.ones(n)
    M2 = np.diag(main) + np.diag(-1.3 * np.ones(n - 1), 1) + np.diag(-0.7 * np.ones(n - 1), -1)
    return [('random', M1), ('advection', M2)]


@pytest.mark.parametrize('d', [1, 2, 5, 12, 32])
@pytest.mark.parametrize('which', [0, 1])
def test_roots_are_harmonic_ritz_values_in_leja_order(d, which):
    name, M = matrices()[which]
    rng = np.random.default_rng(7 + d)
    v0 = rng.standard_normal(M.shape[0]) + 1j * rng.standard_normal(M.shape[0])
    V, W, H = arnoldi(lambda x: M @ x, v0, d)
    th = ddamg.poly_roots(H)
    ref = harmonic_ritz_from_images(V, W)
    dist = max(match_distance(th, ref), match_distance(ref, th))
    print(f'{name} d={d}: roots vs definition {dist:.2e}')
    assert dist < 1e-9
    assert_leja_order(th)

    # the product form of p gives the GMRES residual of d steps from v0
    def p_apply(v):
        y, r = np.zeros_like(v), v.copy()
        for i, t in enumerate(th):
            y = y + r / t
            if i + 1 < len(th):
                r = r - (M @ r) / t
        return y
    res_poly = np.linalg.norm(v0 - M @ p_apply(v0))
    beta = np.linalg.norm(v0)
    e1 = np.zeros(d + 1, complex)
    e1[0] = beta
    y, *_ = np.linalg.lstsq(H, e1, rcond=None)
    res_gmres = np.linalg.norm(e1 - H @ y)
    print(f'  ||v - D p(D) v|| = {res_poly:.6e}, GMRES({d}) residual {res_gmres:.6e}')
    assert abs(res_poly - res_gmres) <= 1e-8 * beta


def test_singular_hessenberg_is_refused():
    H = np.zeros((4, 3), complex)
    H[1, 0] = H[2, 1] = H[3, 2] = 1.0   # H_d singular (zero first row)
    with pytest.raises(ddamg.DDError):
        ddamg.poly_roots(H)

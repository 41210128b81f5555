"""numpy side of the polynomial-preconditioner checks (shared by the CPU and GPU tests).

The library's roots are the harmonic Ritz values of d Arnoldi steps, found from the
Hessenberg matrix as eigenvalues of H_d + |h_{d+1,d}|^2 H_d^{-H} e_d e_d^T and put in
Leja order (csrc/poly.h).  The checks here find them from their definition instead:
theta is a harmonic Ritz value of D on span(V) when some u = V g has
(D u - theta u) orthogonal to D V, i.e. (D V)^H (D V) g = theta (D V)^H V g, a
generalized eigenproblem over the basis V and its image D V; and the Leja order is
checked as the property it names, not recomputed.
"""
import numpy as np


def fill_hash(n, seed):
    """The library's deterministic start vector (k_fill_hash): splitmix64 of seed +
    golden-ratio increments, the top 24 bits giving re and the low 24 bits im in [-1, 1)."""
    with np.errstate(over='ignore'):
        i = np.arange(n, dtype=np.uint64)
        z = np.uint64(seed) + np.uint64(0x9E3779B97F4A7C15) * (i + np.uint64(1))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z ^= z >> np.uint64(31)
    re = (z >> np.uint64(40)).astype(np.float64) / 8388608.0 - 1.0
    im = (z & np.uint64(0xFFFFFF)).astype(np.float64) / 8388608.0 - 1.0
    return re + 1j * im


def arnoldi(apply, v0, d, round32=False):
    """d Arnoldi steps (classical Gram-Schmidt, two passes): basis V (n x d), the images
    D V (n x d) and H ((d+1) x d).  round32: the operator sees each basis vector rounded
    to complex64, as an FP32 hierarchy applies it."""
    n = v0.size
    V = np.zeros((n, d + 1), complex)
    W = np.zeros((n, d), complex)
    H = np.zeros((d + 1, d), complex)
    V[:, 0] = v0 / np.linalg.norm(v0)
    for j in range(d):
        x = V[:, j].astype(np.complex64).astype(complex) if round32 else V[:, j]
        w = apply(x)
        W[:, j] = w
        for _ in range(2):
            h = V[:, :j + 1].conj().T @ w
            w = w - V[:, :j + 1] @ h
            H[:j + 1, j] += h
        H[j + 1, j] = np.linalg.norm(w)
        V[:, j + 1] = w / H[j + 1, j]
    return V[:, :d], W, H


def harmonic_ritz_from_images(V, W):
    """Eigenvalues theta of (W^H W) g = theta (W^H V) g with W = D V."""
    return np.linalg.eigvals(np.linalg.solve(W.conj().T @ V, W.conj().T @ W))


def assert_leja_order(th, rtol=1e-9):
    """th[0] has the largest modulus; every later th[k] has, among th[k:], the largest sum
    of log-distances to th[:k] (up to rtol of the spread)."""
    th = np.asarray(th)
    scale = np.max(np.abs(th))
    assert abs(th[0]) >= np.max(np.abs(th)) * (1 - rtol)
    for k in range(1, len(th)):
        score = [np.sum(np.log(np.abs(th[m] - th[:k]))) for m in range(k, len(th))]
        assert score[0] >= max(score) - rtol * max(1.0, abs(max(score))) - 1e-12 * scale, (k, score)


def match_distance(a, b):
    """max over a of the distance to the nearest element of b, relative to max |b|."""
    a, b = np.asarray(a), np.asarray(b)
    return max(np.min(np.abs(b - x)) for x in a) / np.max(np.abs(b))

"""Independent Orr-Sommerfeld eigen-solver (Chebyshev collocation) used to pin
the full time step: a small-amplitude Tollmien-Schlichting wave in plane
Poiseuille flow must grow at the rate of the leading OS eigenvalue.

Scaling: U = 1 - y'^2 on y' in [-1, 1] (the DNS's u = y(2-y) with y' = y-1),
Re = 1/nu.  Orszag (1971): Re=1e4, alpha=1 -> c = 0.23752649 + 0.00373967i.
"""
import numpy as np


def cheb(N):
    """Chebyshev differentiation matrix and points x_j = cos(pi j/N)."""
    x = np.cos(np.pi * np.arange(N + 1) / N)
    c = np.hstack([2.0, np.ones(N - 1), 2.0]) * (-1.0) ** np.arange(N + 1)
    X = np.tile(x, (N + 1, 1)).T
    dX = X - X.T
    D = np.outer(c, 1.0 / c) / (dX + np.eye(N + 1))
    D = D - np.diag(D.sum(axis=1))
    return D, x


def cheb4c(N):
    """Fourth derivative with clamped BCs p(+-1) = p'(+-1) = 0 on the interior
    points, via p = (1-x^2) q (Trefethen, Spectral Methods in MATLAB, ch. 14)."""
    D, x = cheb(N)
    xi = x[1:-1]
    S = np.diag(np.hstack([0.0, 1.0 / (1.0 - xi ** 2), 0.0]))
    D2, D3, D4 = D @ D, D @ D @ D, D @ D @ D @ D
    L = np.diag(1.0 - x ** 2) @ D4 - 8.0 * np.diag(x) @ D3 - 12.0 * D2
    return (L @ S)[1:-1, 1:-1], xi


def leading_mode(alpha, Re, N=120, beta=0.0):
    """Returns (c, P, dP): phase speed c of the least-stable OS mode of the
    (alpha, beta) wave and Chebyshev interpolants of its v(y') and dv/dy',
    normalised to max |v| = 1.  With beta != 0 the operator uses
    k^2 = alpha^2 + beta^2 and advection by alpha U."""
    D, x = cheb(N)
    D2 = (D @ D)[1:-1, 1:-1]
    D4, xi = cheb4c(N)
    I = np.eye(N - 1)
    U = np.diag(1.0 - xi ** 2)
    k2 = alpha ** 2 + beta ** 2
    L2 = D2 - k2 * I
    L4 = D4 - 2 * k2 * D2 + k2 ** 2 * I
    A = L4 / Re - 1j * alpha * U @ L2 + 1j * alpha * (-2.0) * I
    B = L2
    lam, vec = np.linalg.eig(np.linalg.solve(B, A))
    k = np.argmax(lam.real)
    c = 1j * lam[k] / alpha
    v = np.hstack([0.0, vec[:, k], 0.0])
    v = v / v[np.argmax(np.abs(v))]
    # Chebyshev coefficients of the interpolant for evaluation anywhere
    coef = np.polynomial.chebyshev.chebfit(x, v.real, N) + 1j * np.polynomial.chebyshev.chebfit(x, v.imag, N)
    P = np.polynomial.chebyshev.Chebyshev(coef)
    return c, P, P.deriv()

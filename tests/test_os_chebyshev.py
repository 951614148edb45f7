"""The independent Orr-Sommerfeld solver that pins the whole time step
(tests/os_chebyshev.py) against the published eigenvalue: Orszag (1971),
plane Poiseuille flow at Re = 10^4, alpha = 1, least-stable mode
c = 0.23752649 + 0.00373967i (8 digits).  A wrong sign or factor in the
operator (the U'' term, the k^2 powers, the i alpha Re scaling) moves c far
beyond that."""
import numpy as np
import pytest

from os_chebyshev import leading_mode

ORSZAG = 0.23752649 + 0.00373967j


@pytest.mark.parametrize("N", [80, 120])
def test_orszag_eigenvalue(N):
    c, P, dP = leading_mode(1.0, 1e4, N)
    assert abs(c.real - ORSZAG.real) < 1e-8 and abs(c.imag - ORSZAG.imag) < 1e-8, c
    # the eigenfunction is normalised and clamped at the walls
    y = np.array([-1.0, 1.0])
    assert np.max(np.abs(P(y))) < 1e-10 and np.max(np.abs(dP(y))) < 1e-8


def test_oblique_mode_is_squire_equivalent():
    """Squire's transformation: the (alpha, beta) wave at Re has the phase
    speed of the 2-D wave k = sqrt(alpha^2 + beta^2) at Re alpha / k."""
    alpha, beta, Re = 1.0, 0.5, 8000.0
    k = np.hypot(alpha, beta)
    c3, _, _ = leading_mode(alpha, Re, 100, beta=beta)
    c2, _, _ = leading_mode(k, Re * alpha / k, 100)
    assert abs(c3 - c2) < 1e-9, (c3, c2)

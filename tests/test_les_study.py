"""CPU checks of the closure study's filter and subgrid-stress targets
(channel_amd/les_study.py) on small grids: the top-hat transfer function,
the LES rows and modes of G(u), and tau = G(u_i u_j) - G(u_i) G(u_j)
against a direct evaluation with the complex-FFT transforms of the closure
checker (oracle/les_oracle.py)."""
import numpy as np

from channel_amd import les_study as ls
from channel_amd.dns import DnsConfig
from channel_amd.initial import perturbed_laminar
from oracle import les_oracle as lo


def grids():
    cf = DnsConfig(nx=15, ny=16, nz=15, alfa0=0.5, beta0=1.0, ni=1.0 / 12431.0, a=1.5)
    cl = DnsConfig(nx=7, ny=8, nz=7, alfa0=0.5, beta0=1.0, ni=1.0 / 12431.0, a=1.5)
    return cf, cl


def test_tophat():
    m = np.arange(-5, 6)
    H = ls.tophat(m, 48, 4)
    t = np.pi * m * 4 / 48
    np.testing.assert_allclose(H[m != 0], np.sin(t[m != 0]) / t[m != 0], rtol=1e-15)
    assert H[m == 0][0] == 1.0
    np.testing.assert_array_equal(H, H[::-1])


def test_ratios_and_rows():
    cf, cl = grids()
    rx, rz, ys = ls.ratios(cf, cl)
    assert (rx, rz, ys) == (2, 2, 2)
    from channel_amd.dns import y_grid
    yf, yl = y_grid(cf), y_grid(cl)
    np.testing.assert_allclose(yl[1:cl.ny + 2], yf[1 + ys * np.arange(cl.ny + 1)], rtol=0, atol=1e-15)


def test_targets_match_direct_evaluation():
    cf, cl = grids()
    V = perturbed_laminar(cf, seed=3, amp=0.2, kmax=6)
    tau = ls.targets(cf, cl, V, workers=1)
    rx, rz, ys = ls.ratios(cf, cl)
    Mxf, Mzf, Mxl, Mzl = 2 * cf.nxd, cf.nzd, 2 * cl.nxd, cl.nzd
    Hx = ls.tophat(np.arange(cl.nx + 1), Mxf, rx)
    Hz = ls.tophat(np.arange(-cl.nz, cl.nz + 1), Mzf, rz)
    H = Hz[:, None] * Hx[None, :]
    for j in (0, 3, cl.ny):
        row = 1 + ys * j
        u = np.stack([lo.physical(cf, V[row, :, :, c], Mxf, Mzf) for c in range(3)])  # (3, Mx, Mz)
        Ul = V[row, cf.nz - cl.nz:cf.nz + cl.nz + 1, :cl.nx + 1, :] * H[:, :, None]
        gu = np.stack([lo.physical(cl, Ul[..., c], Mxl, Mzl) for c in range(3)])
        for q, (a, b) in enumerate(ls.PAIRS):
            P = lo.spectral(cf, u[a] * u[b])  # fine modes
            Pl = P[cf.nz - cl.nz:cf.nz + cl.nz + 1, :cl.nx + 1] * H
            want = lo.physical(cl, Pl, Mxl, Mzl) - gu[a] * gu[b]
            np.testing.assert_allclose(tau[q, j], want, rtol=0, atol=1e-13 * max(1.0, np.abs(want).max()))
    # the filtered field carries the same LES modes
    from channel_amd.dns import y_grid
    W = ls.filtered(cf, cl, V, y_grid(cl))
    np.testing.assert_allclose(W[1 + 2, :, :, 0], V[1 + ys * 2, cf.nz - cl.nz:cf.nz + cl.nz + 1, :cl.nx + 1, 0] * H)

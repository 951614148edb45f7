"""Pins of the closure checker oracle/les_oracle.py by closed forms: the
y differences are exact on quadratics, the x and z derivatives spectral
(exact on single Fourier modes), y+ of the laminar profile u = y(2-y) is
(1 - |y - 1|) sqrt(2 nu)/nu, and the neighbour blocks are the point's
gradients of the rows iy-2..iy+2 clamped to the walls."""
import numpy as np
import pytest

from channel_amd import lib
from channel_amd.dns import DnsConfig
from oracle import les_oracle as lo


def cfg_small():
    return DnsConfig(nx=7, ny=16, nz=7, alfa0=0.5, beta0=1.0, ni=1.0 / 12431.0, a=1.5)


def _setup(cfg):
    S = lib.setup_host(cfg)
    return S["y"], S["bnd"]


def test_laminar_features():
    cfg = cfg_small()
    y, bnd = _setup(cfg)
    V = np.zeros(cfg.field_shape, dtype=complex, order="F")
    V[:, cfg.nz, 0, 0] = y * (2 - y)
    X = lo.features(cfg, V, y, bnd[1], bnd[6])
    Mx, Mz = 2 * cfg.nxd, cfg.nzd
    X = X.reshape(cfg.ny + 1, Mx, Mz, 46)
    yr = y[1:cfg.ny + 2]
    np.testing.assert_allclose(X[..., 1], np.broadcast_to((2 - 2 * yr)[:, None, None], X.shape[:3]), atol=1e-13)
    others = [j for j in range(9) if j != 1]
    assert np.max(np.abs(X[..., others])) < 1e-13
    yplus = np.minimum(yr, 2 - yr) * np.sqrt(2 * cfg.ni) / cfg.ni
    np.testing.assert_allclose(X[:, 0, 0, 45], yplus, rtol=1e-12)
    # neighbour blocks: rows iy-2, iy-1, iy+1, iy+2, clamped
    rows = np.arange(cfg.ny + 1)
    for k, off in enumerate((-2, -1, 1, 2)):
        np.testing.assert_array_equal(X[:, 0, 0, 9 * (k + 1) + 1], X[np.clip(rows + off, 0, cfg.ny), 0, 0, 1])


@pytest.mark.parametrize("ix,iz", [(1, 0), (2, 3), (0, 2)])
def test_single_mode_derivatives(ix, iz):
    cfg = cfg_small()
    y, bnd = _setup(cfg)
    yr = y[1:cfg.ny + 2]
    f = yr * (2 - yr)
    V = np.zeros(cfg.field_shape, dtype=complex, order="F")
    # w = 2 Re[0.5 f e^{i(kx x + kz z)}] = f cos(kx x + kz z) in component w
    V[1:cfg.ny + 2, cfg.nz + iz, ix, 2] = 0.5 * f
    if ix == 0:
        V[1:cfg.ny + 2, cfg.nz - iz, 0, 2] = 0.5 * f
    X = lo.features(cfg, V, y, bnd[1], bnd[6])
    Mx, Mz = 2 * cfg.nxd, cfg.nzd
    X = X.reshape(cfg.ny + 1, Mx, Mz, 46)
    kx, kz = cfg.alfa0 * ix, cfg.beta0 * iz
    x = np.arange(Mx) * (2 * np.pi / cfg.alfa0) / Mx
    z = np.arange(Mz) * (2 * np.pi / cfg.beta0) / Mz
    ph = kx * x[:, None] + kz * z[None, :]
    np.testing.assert_allclose(X[..., 6], -kx * f[:, None, None] * np.sin(ph)[None], atol=1e-12)
    np.testing.assert_allclose(X[..., 8], -kz * f[:, None, None] * np.sin(ph)[None], atol=1e-12)
    np.testing.assert_allclose(X[..., 7], (2 - 2 * yr)[:, None, None] * np.cos(ph)[None], atol=1e-12)


def test_divergence_closed_forms():
    cfg = cfg_small()
    y, _ = _setup(cfg)
    yr = y[1:cfg.ny + 2]
    Mx, Mz = 2 * cfg.nxd, cfg.nzd
    x = np.arange(Mx) * (2 * np.pi / cfg.alfa0) / Mx
    z = np.arange(Mz) * (2 * np.pi / cfg.beta0) / Mz
    kx, kz = cfg.alfa0 * 2, cfg.beta0 * 1
    tau = np.zeros((6, cfg.ny + 1, Mx, Mz))
    tau[3] = (yr ** 2)[:, None, None]                           # tau_xy = y^2
    tau[0] = np.cos(kx * x)[None, :, None] * np.ones((cfg.ny + 1, 1, 1))  # tau_xx
    tau[5] = np.sin(kz * z)[None, None, :] * yr[:, None, None]   # tau_yz
    F = lo.divergence(cfg, tau, y)
    nz = cfg.nz
    # F_u = -(d_x tau_xx + d_y tau_xy): mean -2y, mode (2,0) from -d_x cos = kx sin
    np.testing.assert_allclose(F[1:-1, nz, 0, 0].real, -2 * yr, atol=1e-12)
    # kx sin(kx x) = 2 Re[(-i kx / 2) e^{i kx x}] -> coefficient -i kx/2 at (ix=2, iz=0)
    np.testing.assert_allclose(F[1:-1, nz, 2, 0], -0.5j * kx * np.ones(cfg.ny + 1), atol=1e-12)
    # F_v = -d_z tau_yz: -kz y cos(kz z) -> coefficient -kz y / 2 at (0, +-1)
    np.testing.assert_allclose(F[1:-1, nz + 1, 0, 1], -0.5 * kz * yr, atol=1e-12)
    # F_w = -d_y tau_yz = -sin(kz z) -> coefficient i/2 at (0, +1)
    np.testing.assert_allclose(F[1:-1, nz + 1, 0, 2], 0.5j * np.ones(cfg.ny + 1), atol=1e-12)
    assert np.all(F[0] == 0) and np.all(F[-1] == 0)

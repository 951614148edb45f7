"""CPU float64 checker for the device closure kernels of
include/channel_closure.h (chan_les_features, chan_les_closure_force).

Test infrastructure only: the tests compare the device kernels with it; the
product never calls it.  The closure has no counterpart in the reference.

It restates, in numpy, what the device computes on the dealiased physical
grid Mx x Mz = 2nxd x nzd of rows iy = 0..ny:
- gradients du_c/dx_d: x and z derivatives spectral (i kx, i kz), y
  derivatives the second-order differences of numpy.gradient (edge_order 2)
  on the stretched rows, applied to the spectral coefficients;
- features per sample (iy, x, z): the nine gradients at rows iy, iy-2, iy-1,
  iy+1, iy+2 (rows clamped to 0..ny), then y+ = min(y-ymin, ymax-y) u_tau/nu
  with u_tau = sqrt(nu (|dU/dy|_bottom + |dU/dy|_top) / 2) from the
  wall-gradient stencils d140, d14n of the (0,0) mode;
- the body force F_i = -(d_x tau_xi + d_y tau_yi + d_z tau_zi) of the six
  stresses (xx, yy, zz, xy, xz, yz) given on the same grid, spectral in x and
  z, the same y differences.
Fields use the reference layout V(iy, iz, ix, c), shape (ny+3, 2nz+1,
nx+1, 3).
"""
import numpy as np


def physical(cfg, S, Mx, Mz):
    """Spectral (..., 2nz+1, nx+1) -> real physical (..., Mx, Mz), with the
    Hermitian extension in x (the sum over kx >= 0 with weight 2 off kx = 0)."""
    nz, nx = cfg.nz, cfg.nx
    F = np.zeros(S.shape[:-2] + (Mx, Mz), dtype=np.complex128)
    kz = np.arange(-nz, nz + 1) % Mz
    for ix in range(nx + 1):
        F[..., ix, kz] = S[..., :, ix]
        if ix > 0:
            F[..., (-ix) % Mx, (-np.arange(-nz, nz + 1)) % Mz] = np.conj(S[..., :, ix])
    return np.fft.ifft2(F, axes=(-2, -1)).real * (Mx * Mz)


def spectral(cfg, f):
    """Real physical (..., Mx, Mz) -> spectral (..., 2nz+1, nx+1)."""
    Mx, Mz = f.shape[-2], f.shape[-1]
    F = np.fft.fft2(f, axes=(-2, -1)) / (Mx * Mz)
    iz = np.arange(-cfg.nz, cfg.nz + 1) % Mz
    return np.swapaxes(F[..., :cfg.nx + 1, :][..., iz], -1, -2)


def ddy(y, F):
    """d/dy along axis 0 (rows 0..ny) by numpy.gradient, second order."""
    return np.gradient(F, y, axis=0, edge_order=2)


def features(cfg, V, y, d140, d14n):
    """X (n, 46) float64 for samples in (iy, x, z) order; y = y(-1..ny+1)."""
    Mx, Mz = 2 * cfg.nxd, cfg.nzd
    ny = cfg.ny
    yr = y[1:ny + 2]
    ix = np.arange(cfg.nx + 1)
    iz = np.arange(-cfg.nz, cfg.nz + 1)
    kx = (1j * cfg.alfa0 * ix)[None, None, :]
    kz = (1j * cfg.beta0 * iz)[None, :, None]
    G = np.empty((ny + 1, Mx, Mz, 9))
    for c in range(3):
        S = V[1:ny + 2, :, :, c]
        G[..., 3 * c + 0] = physical(cfg, S * kx, Mx, Mz)
        G[..., 3 * c + 1] = physical(cfg, ddy(yr, S), Mx, Mz)
        G[..., 3 * c + 2] = physical(cfg, S * kz, Mx, Mz)
    X = np.empty((ny + 1, Mx, Mz, 46))
    rows = np.arange(ny + 1)
    for k, off in enumerate((0, -2, -1, 1, 2)):
        X[..., 9 * k:9 * k + 9] = G[np.clip(rows + off, 0, ny)]
    u = V[:, cfg.nz, 0, 0].real
    db = np.dot(d140, u[0:5])
    dt = np.dot(d14n, u[ny - 2:ny + 3])
    utau = np.sqrt(cfg.ni * 0.5 * (abs(db) + abs(dt)))
    X[..., 45] = (np.minimum(yr - cfg.ymin, cfg.ymax - yr) * utau / cfg.ni)[:, None, None]
    return X.reshape(-1, 46)


def divergence(cfg, tau, y):
    """F in the reference layout (rows -1..ny+1; the ghost rows zero) from
    tau (6, ny+1, Mx, Mz) = (xx, yy, zz, xy, xz, yz)."""
    ny = cfg.ny
    yr = y[1:ny + 2]
    T = spectral(cfg, tau)  # (6, ny+1, 2nz+1, nx+1)
    ikx = (1j * cfg.alfa0 * np.arange(cfg.nx + 1))[None, None, :]
    ikz = (1j * cfg.beta0 * np.arange(-cfg.nz, cfg.nz + 1))[None, :, None]
    xx, yy, zz, xy, xz, yz = T
    F = np.zeros((ny + 3, 2 * cfg.nz + 1, cfg.nx + 1, 3), dtype=np.complex128)
    F[1:ny + 2, :, :, 0] = -(ikx[0] * xx + ddy(yr, xy) + ikz[0] * xz)
    F[1:ny + 2, :, :, 1] = -(ikx[0] * xy + ddy(yr, yy) + ikz[0] * yz)
    F[1:ny + 2, :, :, 2] = -(ikx[0] * xz + ddy(yr, yz) + ikz[0] * zz)
    return F

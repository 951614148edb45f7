"""Seeded perturbed-laminar initial field (synthetic input of the benchmark
and the parity tests).

The reference generates ``u = 1.5*y*(2-y)`` with zero perturbation when
Dati.cart.out is missing (dnsdata.f90:704-718).  The workloads here need a
reproducible perturbed start, so this generator writes a field that both the
HIP path and the oracle read:

* mean flow u = y(2-y) (laminar Poiseuille, centreline 1), w = 0;
* for 0 <= ix <= 4, -4 <= iz <= 4, (ix,iz) != (0,0) and iz > 0 when ix = 0:
  v = A r1 e^{i phi1} (1-(y-1)^2)^2 and eta = A r2 e^{i phi2} (1-(y-1)^2),
  with r, phi uniform from splitmix64(seed); u and w follow from continuity
  the way linsolve reconstructs them (linsolve_blocking.inc:101-103):
  u = (i kx v' - i kz eta)/k^2, w = (i kz v' + i kx eta)/k^2;
* the kx = 0 column is made Hermitian in z: V(-iz,0) = conj(V(iz,0)).
"""
import math

import numpy as np

from .dns import y_grid

_MASK = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.s = seed & _MASK

    def next(self):
        self.s = (self.s + 0x9E3779B97F4A7C15) & _MASK
        z = self.s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next() >> 11) * (1.0 / 9007199254740992.0)


def perturbed_laminar(cfg, seed=20261202, amp=1e-2, kmax=4, mean=True):
    """Returns V (ny+3, 2nz+1, nx+1, 3) complex128, Fortran order."""
    ny, nz, nx = cfg.ny, cfg.nz, cfg.nx
    V = np.zeros(cfg.field_shape, dtype=np.complex128, order="F")
    y = y_grid(cfg)
    if mean:
        V[:, nz, 0, 0] = y * (2.0 - y)
    rng = SplitMix64(seed)
    s = 1.0 - (y - 1.0) ** 2
    fv, dfv = s * s, -4.0 * (y - 1.0) * s
    for ix in range(0, min(kmax, nx) + 1):
        for iz in range(-min(kmax, nz), min(kmax, nz) + 1):
            if ix == 0 and iz <= 0:
                continue
            r1, p1, r2, p2 = (rng.uniform() for _ in range(4))
            cv = amp * r1 * complex(math.cos(2 * math.pi * p1), math.sin(2 * math.pi * p1))
            ce = amp * r2 * complex(math.cos(2 * math.pi * p2), math.sin(2 * math.pi * p2))
            kx, kz = cfg.alfa0 * ix, cfg.beta0 * iz
            k2 = kx * kx + kz * kz
            v, dv, eta = cv * fv, cv * dfv, ce * s
            u = (1j * kx * dv - 1j * kz * eta) / k2
            w = (1j * kz * dv + 1j * kx * eta) / k2
            V[:, nz + iz, ix, 0] = u
            V[:, nz + iz, ix, 1] = v
            V[:, nz + iz, ix, 2] = w
            if ix == 0:
                V[:, nz - iz, 0, :] = np.conj(V[:, nz + iz, 0, :])
    return V


def single_mode(cfg, ix, iz, v_profile, dv_profile, eta_profile=None):
    """Field with one (ix,iz) perturbation mode (and its Hermitian partner
    when ix = 0) given by its v, dv/dy and eta profiles on y(-1..ny+1)."""
    V = np.zeros(cfg.field_shape, dtype=np.complex128, order="F")
    nz = cfg.nz
    kx, kz = cfg.alfa0 * ix, cfg.beta0 * iz
    k2 = kx * kx + kz * kz
    eta = np.zeros_like(v_profile) if eta_profile is None else eta_profile
    V[:, nz + iz, ix, 0] = (1j * kx * dv_profile - 1j * kz * eta) / k2
    V[:, nz + iz, ix, 1] = v_profile
    V[:, nz + iz, ix, 2] = (1j * kz * dv_profile + 1j * kx * eta) / k2
    if ix == 0:
        V[:, nz - iz, 0, :] = np.conj(V[:, nz + iz, 0, :])
    return V

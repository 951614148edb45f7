"""Workload definitions shared by the oracle and GPU tests."""
import numpy as np

from channel_amd.dns import DnsConfig, y_grid
from channel_amd.initial import perturbed_laminar, single_mode


def config1(**kw):
    """Plain FP64 channel, small: nx=31, nz=31, ny=64, a=1.5, alfa0=1,
    beta0=2, Re=3000, constant flow rate, deltat=0.004."""
    c = DnsConfig(nx=31, ny=64, nz=31, alfa0=1.0, beta0=2.0, ni=1.0 / 3000.0, a=1.5,
                  meanflowx=4.0 / 3.0, deltat=0.004, cflmax=0.0)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def config2(**kw):
    """nx=95, nz=95, ny=128, alfa0=1, beta0=2, Re=3000, constant flow rate."""
    c = DnsConfig(nx=95, ny=128, nz=95, alfa0=1.0, beta0=2.0, ni=1.0 / 3000.0, a=1.5,
                  meanflowx=4.0 / 3.0, deltat=0.004, cflmax=0.0)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def config3(**kw):
    """Production 384-class grid: nx=191, nz=189, ny=384, alfa0=0.5, beta0=1,
    Re=12431, CPI type 1 with gamma=0.161436, cflmax=1."""
    c = DnsConfig(nx=191, ny=384, nz=189, alfa0=0.5, beta0=1.0, ni=1.0 / 12431.0, a=1.5,
                  cpi=True, cpi_type=1, gamma=0.161436, deltat=0.0, cflmax=1.0)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def tiny(**kw):
    """A fast grid for CPU tests of the step's properties."""
    c = DnsConfig(nx=7, ny=32, nz=7, alfa0=1.0, beta0=2.0, ni=1.0 / 3000.0, a=1.5,
                  meanflowx=4.0 / 3.0, deltat=0.004, cflmax=0.0)
    for k, v in kw.items():
        setattr(c, k, v)
    return c


def seeded_field(cfg, seed=20261202):
    return perturbed_laminar(cfg, seed=seed)


def laminar_field(cfg):
    V = np.zeros(cfg.field_shape, dtype=np.complex128, order="F")
    y = y_grid(cfg)
    V[:, cfg.nz, 0, 0] = y * (2.0 - y)
    return V


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def os_case(alpha=1.0, beta=0.0, Re=1.0e4, ny=128, eps=1e-6, N=120):
    """Laminar Poiseuille (kept by dp/dx = 2 nu) plus eps times the leading
    Orr-Sommerfeld mode of the (alpha, beta) wave in mode (ix=1, iz=1 or 0)."""
    from os_chebyshev import leading_mode
    iz = 1 if beta else 0
    cfg = DnsConfig(nx=3, ny=ny, nz=2, alfa0=alpha, beta0=beta if beta else 1.0, ni=1.0 / Re, a=1.5,
                    meanpx=2.0 / Re, deltat=0.01, cflmax=0.0)
    c, P, dP = leading_mode(alpha, Re, N, beta=beta)
    y = y_grid(cfg)
    V = laminar_field(cfg) + eps * single_mode(cfg, 1, iz, P(y - 1.0), dP(y - 1.0))
    return cfg, V, c, P, iz

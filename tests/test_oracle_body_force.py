"""Pins the oracle's body force (dnsdata.f90:616-659, body_forces/*):

* a force profile c y(2-y) in x and z keeps the quartic laminar profile it
  balances (the mean-mode term rD0(F,1,3));
* steady forced Stokes flow of an oblique and a 2-D mode equals the solution
  of the forced v-eta equations by an independent Chebyshev solver (the D2v
  terms -k2 D0 F_v - i kx D1 F_u - i kz D1 F_w and the eta terms
  i kz D0 F_u - i kx D0 F_w, signs and factors);
* am_f1 on a Tollmien-Schlichting wave lowers its growth rate by amp and
  leaves its phase speed;
* the wall rows: zero on the walls, D4 of the first and last interior rows
  zero through the ghost rows;
* the module masks (literal restatements in the oracle) against the tables
  of channel_amd.body_force, built independently.
"""
import os

import numpy as np
import pytest

from cases import os_case, seeded_field, tiny
from channel_amd import body_force as bf
from channel_amd.dns import DnsConfig, y_grid
from oracle import oracle as orc
from os_chebyshev import cheb, cheb4c

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _quartic(y):
    # U'' = -q for q = y(2-y), U(0) = U(2) = 0
    return 2.0 * y / 3.0 - y ** 3 / 3.0 + y ** 4 / 12.0


def test_force_profile_keeps_its_laminar_flow():
    cfg = tiny(meanflowx=0.0, meanpx=0.0)
    nu = cfg.ni
    cx, cz = 12.0 * nu / 5.0, 6.0 * nu / 5.0  # centreline U = 1, W = 1/2
    y = y_grid(cfg)
    q = y * (2.0 - y)
    nz = cfg.nz
    V0 = np.zeros(cfg.field_shape, dtype=np.complex128, order="F")
    V0[:, nz, 0, 0] = cx / nu * _quartic(y)
    V0[:, nz, 0, 2] = cz / nu * _quartic(y)
    F = np.zeros_like(V0)
    F[:, nz, 0, 0] = cx * q
    F[:, nz, 0, 2] = cz * q
    # The quartic is exact for the interior stencils but not for the solver's
    # ghost-row closure (its fourth derivative is not zero), so the rows next
    # to the walls settle by ~1e-9; the central half of the channel keeps the
    # profile to rounding, where a missing or sign-flipped force term moves it
    # by the force times the elapsed time (~3e-5 here).
    mid = slice(cfg.ny // 4 + 1, 3 * cfg.ny // 4 + 2)
    o = orc.Oracle(cfg)
    o.set_field(V0)
    o.set_body_force(F)
    o.steps(10)
    V = o.get_field()
    assert np.max(np.abs(V[mid] - V0[mid])) < 1e-13
    assert np.max(np.abs(V[1:-1] - V0[1:-1])) < 1e-8
    for scale in (0.0, -1.0):
        o2 = orc.Oracle(cfg)
        o2.set_field(V0)
        o2.set_body_force(scale * F)
        o2.steps(10)
        assert np.max(np.abs(o2.get_field()[mid] - V0[mid])) > 1e-5


def _stokes_chebyshev(alpha, beta, nu, f, N=64):
    """Steady Stokes response of the mode (alpha, beta) to F = q(y') f with
    q = 1 - y'^2 on y' in [-1, 1]:
      nu (D^2-k^2)^2 v = k^2 F_v + i alpha D F_u + i beta D F_w,  v = Dv = 0,
      nu (D^2-k^2) eta = -(i beta F_u - i alpha F_w),              eta = 0,
    then u = (i alpha Dv - i beta eta)/k^2, w = (i beta Dv + i alpha eta)/k^2.
    Returns Chebyshev interpolants of u, v, w."""
    D, x = cheb(N)
    D2 = D @ D
    D4, xi = cheb4c(N)
    I = np.eye(N - 1)
    k2 = alpha ** 2 + beta ** 2
    q, dq = 1.0 - xi ** 2, -2.0 * xi
    rhs_v = k2 * f[1] * q + 1j * alpha * f[0] * dq + 1j * beta * f[2] * dq
    L4 = D4 - 2 * k2 * D2[1:-1, 1:-1] + k2 ** 2 * I
    v = np.hstack([0.0, np.linalg.solve(nu * L4, rhs_v), 0.0])
    rhs_e = -(1j * beta * f[0] * q - 1j * alpha * f[2] * q)
    eta = np.hstack([0.0, np.linalg.solve(nu * (D2[1:-1, 1:-1] - k2 * I), rhs_e), 0.0])
    dv = D @ v
    u = (1j * alpha * dv - 1j * beta * eta) / k2
    w = (1j * beta * dv + 1j * alpha * eta) / k2

    def interp(g):
        c = np.polynomial.chebyshev.chebfit(x, g.real, N) + 1j * np.polynomial.chebyshev.chebfit(x, g.imag, N)
        return np.polynomial.chebyshev.Chebyshev(c)

    return interp(u), interp(v), interp(w)


@pytest.mark.parametrize("iz", [1, 0])
def test_forced_stokes_steady_state(iz):
    cfg = DnsConfig(nx=3, ny=48, nz=2, alfa0=1.0, beta0=2.0, ni=0.25, a=1.5, deltat=0.02, cflmax=0.0)
    y = y_grid(cfg)
    eps = 1e-6
    f = eps * np.array([0.7 - 0.2j, -0.4 + 0.5j, 0.3 + 0.9j])
    nz = cfg.nz
    F = np.zeros(cfg.field_shape, dtype=np.complex128, order="F")
    for c in range(3):
        F[:, nz + iz, 1, c] = f[c] * y * (2.0 - y)
    o = orc.Oracle(cfg)
    o.set_field(np.zeros_like(F))
    o.set_body_force(F)
    o.steps(1500)
    V1 = o.get_field()
    o.steps(10)
    V = o.get_field()
    assert np.max(np.abs(V - V1)) < 1e-9 * np.max(np.abs(V))  # steady
    P = _stokes_chebyshev(cfg.alfa0, cfg.beta0 * iz, cfg.ni, f)
    yy = y[1:-1] - 1.0  # rows 0..ny
    for c in range(3):
        ref = P[c](yy)
        got = V[1:-1, nz + iz, 1, c]
        err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
        print("component", c, "max rel err", err)
        assert err < 1e-5
    # no other mode is forced (nonlinear products are O(eps^2))
    other = V.copy()
    other[:, nz + iz, 1, :] = 0
    assert np.max(np.abs(other)) < 1e-5 * np.max(np.abs(V))


def _amplitude(V, P, cfg, iz):
    y = y_grid(cfg)
    phi = P(y[1:-1] - 1.0)
    v = V[1:-1, cfg.nz + iz, 1, 1]
    return np.vdot(phi, v) / np.vdot(phi, phi)


def test_am_f1_damps_a_ts_wave_by_amp():
    cfg, V0, c, P, iz = os_case(beta=0.0, ny=128)
    amp = 0.02
    rates = []
    for module in (False, True):
        o = orc.Oracle(cfg)
        o.set_field(V0)
        if module:
            # iz_f = NINT(2 PI 1000 / (500 beta0)) exceeds nz; for iz = 0 the
            # mask lambda_z+ = 1e10 > 2.3 y+^2 holds on every row
            o.set_body_force_module("am_f1", amp, 500.0)
        o.steps(20)
        a0 = _amplitude(o.get_field(), P, cfg, iz)
        o.steps(200)
        a1 = _amplitude(o.get_field(), P, cfg, iz)
        T = 200 * cfg.deltat
        rates.append((np.log(abs(a1 / a0)) / T, -np.angle(a1 / a0) / T))
    (s0, w0), (s1, w1) = rates
    print("growth", s0, s1, "difference", s1 - s0, "phase", w0, w1)
    assert abs((s1 - s0) + amp) < 1e-3 * amp
    assert abs(w1 - w0) < 1e-6


def test_wall_rows():
    cfg = tiny()
    o = orc.Oracle(cfg)
    rng = np.random.default_rng(3)
    F = rng.standard_normal(cfg.field_shape) + 1j * rng.standard_normal(cfg.field_shape)
    o.set_body_force(np.asfortranarray(F))
    o.body_force_walls()
    G = o.get_body_force()
    ny = cfg.ny
    assert np.all(G[1] == 0) and np.all(G[ny + 1] == 0)  # rows 0 and ny
    np.testing.assert_array_equal(G[2:ny + 1], F[2:ny + 1])  # rows 1..ny-1 untouched
    d4 = o.der()[:, 3, :]
    top = np.einsum("j,jzxc->zxc", d4[0], G[0:5])  # row 1: rows -1..3
    bot = np.einsum("j,jzxc->zxc", d4[ny - 2], G[ny - 2:ny + 3])  # row ny-1
    scale = np.max(np.abs(d4[0])) * np.max(np.abs(F))
    assert np.max(np.abs(top)) < 1e-13 * scale and np.max(np.abs(bot)) < 1e-13 * scale


MODULES = [  # name, oracle parameters, table, whether the mask covers every row and mode
    ("am_f1", (50.0, 500.0), lambda cfg: bf.am_f1(cfg, 50.0, 500.0), False),
    ("am_butterfly", (50.0, 500.0), lambda cfg: bf.am_butterfly(cfg, 50.0, 500.0), False),
    ("coriolis", (0.02, 5.0, 0.3), lambda cfg: bf.coriolis(cfg, 0.02, 5.0, 0.3), False),
    ("coriolis", (0.02, 9999999.0, 1.0), lambda cfg: bf.coriolis(cfg, 0.02), True),  # coriolis.in defaults
]


@pytest.mark.parametrize("name,par,table,full", MODULES, ids=["am_f1", "am_butterfly", "coriolis", "coriolis_full"])
def test_module_tables_match_oracle_modules(name, par, table, full):
    cfg = tiny(nz=7, beta0=2.0)
    o = orc.Oracle(cfg)
    V = seeded_field(cfg)
    rng = np.random.default_rng(1)
    V = V + 1e-3 * (rng.standard_normal(V.shape) + 1j * rng.standard_normal(V.shape))
    o.set_field(V)
    o.set_body_force_module(name, *par)
    o.apply_body_force()
    F = o.get_body_force()
    src, gain = table(cfg)
    nz = cfg.nz
    G = np.zeros_like(F)
    col = np.arange(2 * nz + 1)
    for d in range(3):
        if src[d] < 0:
            continue
        for ix in range(cfg.nx + 1):
            g = gain[d][:, col].copy()  # (ny+3, 2nz+1)
            if ix == 0:
                g[:, nz] = gain[d][:, 2 * nz + 1]
            G[:, :, ix, d] = g * V[:, :, ix, src[d]]
    np.testing.assert_array_equal(F, G)
    on = gain[[d for d in range(3) if src[d] >= 0]] != 0
    assert on.any() and on.all() == full


def test_iz_f_and_read_coriolis_in():
    cfg = tiny(beta0=2.0)
    assert bf.iz_f(cfg, 500.0) == 6  # NINT(6.28)
    assert bf.iz_f(tiny(beta0=1.0), 500.0) == 13  # NINT(12.57)
    p = bf.read_coriolis_in(os.path.join(GOLDEN, "coriolis.in"))
    assert p == dict(omega2=0.02, kz_cutoff=9999999.0, y_threshold_bot=1.0)

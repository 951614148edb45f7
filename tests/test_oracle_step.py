"""Pins the oracle's whole step: dealiased products against the exact
truncated convolution, the laminar Poiseuille fixed point, Orr-Sommerfeld
growth of 2-D and oblique Tollmien-Schlichting waves (independent Chebyshev
solver), Runtimedata columns, and thread-count independence."""
import numpy as np
import pytest
from scipy.signal import convolve2d

from cases import laminar_field, os_case, rel_l2, seeded_field, tiny
from channel_amd.dns import y_grid
from oracle import oracle as orc

RK1 = (120.0 / 32.0, 2.0, 0.0)
RK2 = (120.0 / 8.0, 50.0 / 8.0, 34.0 / 8.0)
RK3 = (120.0 / 20.0, 90.0 / 20.0, 50.0 / 20.0)


def _random_plane_field(cfg, rng):
    V = np.zeros(cfg.field_shape, dtype=np.complex128, order="F")
    V[...] = rng.standard_normal(V.shape) + 1j * rng.standard_normal(V.shape)
    nz = cfg.nz
    V[:, nz, 0, :] = V[:, nz, 0, :].real
    for iz in range(1, nz + 1):
        V[:, nz - iz, 0, :] = np.conj(V[:, nz + iz, 0, :])
    return V


def _full_spectrum(plane, nx, nz):
    """(2nx+1, 2nz+1) array over kx=-nx..nx from the Hermitian half plane[iz, ix]."""
    F = np.zeros((2 * nx + 1, 2 * nz + 1), dtype=np.complex128)
    for ix in range(nx + 1):
        F[nx + ix, :] = plane[:, ix]
        if ix > 0:
            F[nx - ix, :] = np.conj(plane[::-1, ix])
    return F


@pytest.mark.parametrize("nx,nz", [(7, 7), (11, 5), (15, 9)])
def test_products_are_exact_truncated_convolutions(nx, nz):
    cfg = tiny(nx=nx, nz=nz, ny=16)
    o = orc.Oracle(cfg)
    assert 2 * o.nxd > 3 * nx and o.nzd > 3 * nz  # no aliasing into the kept modes
    V = _random_plane_field(cfg, np.random.default_rng(nx * 100 + nz))
    o.set_field(V)
    iy = 5
    P = o.convolutions(iy)  # [k][ix][iz]
    full = [_full_spectrum(V[iy + 1, :, :, c], nx, nz) for c in range(3)]
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (0, 2)]  # uu vv ww uv vw uw
    for k, (a, b) in enumerate(pairs):
        C = convolve2d(full[a], full[b])  # kx in -2nx..2nx, kz in -2nz..2nz
        ref = C[2 * nx: 3 * nx + 1, nz: 3 * nz + 1]  # kx=0..nx, kz=-nz..nz
        np.testing.assert_allclose(P[k], ref.T.T, atol=1e-11 * np.max(np.abs(ref)))


def test_laminar_poiseuille_is_a_fixed_point():
    # dp/dx = 2 nu balances nu U'' for U = y(2-y); D2 and D0 are exact on
    # quadratics and the RK3 coefficients satisfy ODE2 - ODE3 = 2
    cfg = tiny(meanflowx=0.0, meanpx=2.0 / 3000.0)
    o = orc.Oracle(cfg)
    V0 = laminar_field(cfg)
    o.set_field(V0)
    for _ in range(10):
        o.step()
    V = o.get_field()
    assert np.max(np.abs(V - V0)) < 1e-13


def test_constant_flow_rate_keeps_poiseuille():
    cfg = tiny()  # meanflowx = 4/3, no pressure gradient
    o = orc.Oracle(cfg)
    V0 = laminar_field(cfg)
    o.set_field(V0)
    rows = o.steps(5)
    V = o.get_field()
    assert np.max(np.abs(V - V0)) < 1e-12
    s = o.scalars()
    # the correction supplies the laminar pressure gradient: corrpx = 2 nu * (ODE2-ODE3)...
    assert abs(rows[-1, 5] - 4.0 / 3.0) < 1e-12  # fr_x after correction
    assert abs(rows[-1, 1] - 2.0) < 1e-10 and abs(rows[-1, 2] - 2.0) < 1e-10  # dudy bottom, -dudy top
    assert abs(s["corrpx"] - 4.0 / 3000.0) < 1e-12 * 4


def _mode_amplitude(V, P, cfg, iz):
    y = y_grid(cfg)
    phi = P(y[1:-1] - 1.0)  # walls and interior
    v = V[1:-1, cfg.nz + iz, 1, 1]
    return np.vdot(phi, v) / np.vdot(phi, phi)


@pytest.mark.parametrize("beta", [0.0, 1.0])
def test_orr_sommerfeld_growth(beta):
    cfg, V, c, P, iz = os_case(beta=beta, ny=128)
    o = orc.Oracle(cfg)
    o.set_field(V)
    n_warm, n = 20, 200
    o.steps(n_warm)
    a0 = _mode_amplitude(o.get_field(), P, cfg, iz)
    o.steps(n)
    a1 = _mode_amplitude(o.get_field(), P, cfg, iz)
    T = n * cfg.deltat
    sigma = np.log(np.abs(a1 / a0)) / T  # alpha c_i
    omega = -np.angle(a1 / a0) / T  # alpha c_r (mod 2pi/T)
    assert abs(sigma - cfg.alfa0 * c.imag) < 0.01 * abs(cfg.alfa0 * c.imag) + 2e-6
    assert abs(omega - cfg.alfa0 * c.real) < 1e-4


def test_hermitian_symmetry_preserved():
    cfg = tiny()
    o = orc.Oracle(cfg)
    o.set_field(seeded_field(cfg))
    o.steps(2)
    V = o.get_field()
    nz = cfg.nz
    for iz in range(1, nz + 1):
        np.testing.assert_allclose(V[:, nz - iz, 0, :], np.conj(V[:, nz + iz, 0, :]), atol=1e-15)


def test_thread_count_independent():
    cfg = tiny()
    V0 = seeded_field(cfg)
    out = []
    for t in (1, 4):
        orc.set_threads(t)
        o = orc.Oracle(cfg)
        o.set_field(V0)
        o.steps(2)
        out.append(o.get_field())
    orc.set_threads(1)
    np.testing.assert_array_equal(out[0], out[1])


def test_substeps_compose_into_step():
    cfg = tiny()
    V0 = seeded_field(cfg)
    a = orc.Oracle(cfg)
    a.set_field(V0)
    row = a.step()
    b = orc.Oracle(cfg)
    b.set_field(V0)
    dt = cfg.deltat
    for ode, cfl in ((RK1, False), (RK2, False), (RK3, True)):
        b.buildrhs(ode, cfl)
        b.linsolve(ode[0] / dt)
    np.testing.assert_array_equal(a.get_field(), b.get_field())
    assert row[0] == pytest.approx(dt, rel=1e-15)


def test_initial_cfl_and_cpi():
    cfg = tiny(cpi=True, cpi_type=1, gamma=0.161436, meanflowx=0.0, cflmax=1.0, deltat=0.0)
    o = orc.Oracle(cfg)
    V = laminar_field(cfg)
    o.set_field(V)
    o.initial()
    s = o.scalars()
    # laminar field: CFL = max u/dx = 1/dx (at the centreline, physical point
    # values are exact for a single mode), fr = 4/3, meanpx = 1.5/gamma*fr*nu
    pi32 = float(np.float32(np.pi))
    dx = pi32 / (cfg.alfa0 * o.nxd)
    y, _ = o.grid()
    umax = np.max((y * (2 - y))[2:-2])
    assert s["cfl"] == pytest.approx(umax / dx, rel=1e-14)
    assert s["fr1"] == pytest.approx(4.0 / 3.0, rel=1e-14)
    assert s["meanpx"] == pytest.approx(1.5 / 0.161436 * (4.0 / 3.0) * cfg.ni, rel=1e-14)
    row = o.outstats()
    assert row[10] == pytest.approx(1.0 / (umax / dx), rel=1e-14)  # deltat = cflmax / cfl
    assert row[9] == pytest.approx(1.0, rel=1e-14)


# ---------------------------------------------------------------- Couette --
# Wall velocities u0, uN (dns.in line 8) enter as the (0,0) wall values of the
# eta system (channel.f90:122-124, linsolve_blocking.inc:65-76); with no
# pressure gradient the linear profile between them is steady.
@pytest.mark.parametrize("u0,uN", [(-1.0, 1.0), (0.3, 1.7)])
def test_laminar_couette_is_a_fixed_point(u0, uN):
    cfg = tiny(meanflowx=0.0, meanpx=0.0, u0=u0, uN=uN)
    o = orc.Oracle(cfg)
    y, _ = o.grid()
    V0 = np.zeros(cfg.field_shape, dtype=np.complex128, order="F")
    V0[:, cfg.nz, 0, 0] = u0 + (uN - u0) * y / 2.0
    o.set_field(V0)
    rows = o.steps(10)
    V = o.get_field()
    assert np.max(np.abs(V - V0)) < 1e-13
    # Runtimedata: du/dy at the bottom wall and -du/dy at the top wall
    slope = (uN - u0) / 2.0
    assert rows[-1, 1] == pytest.approx(slope, rel=1e-12)
    assert rows[-1, 2] == pytest.approx(-slope, rel=1e-12)


def test_couette_walls_hold_with_perturbations():
    # perturbed Couette: the (0,0) u column takes the wall values exactly and
    # every other mode keeps no-slip
    cfg = tiny(meanflowx=0.0, meanpx=0.0, u0=-1.0, uN=1.0)
    o = orc.Oracle(cfg)
    y, _ = o.grid()
    V0 = seeded_field(cfg)
    V0[:, cfg.nz, 0, 0] = y - 1.0
    o.set_field(V0)
    o.steps(5)
    V = o.get_field()
    nz, ny = cfg.nz, cfg.ny
    assert V[1, nz, 0, 0].real == pytest.approx(-1.0, abs=1e-13)
    assert V[ny + 1, nz, 0, 0].real == pytest.approx(1.0, abs=1e-13)
    W = V.copy()
    W[:, nz, 0, 0] = 0.0
    assert np.max(np.abs(W[[1, ny + 1]])) < 1e-13
    assert np.max(np.abs(V[:, :, :, 1] - V0[:, :, :, 1])) > 1e-6  # the perturbation evolves


def test_couette_plus_poiseuille_is_a_fixed_point():
    # superposition: y(2-y) kept by dp/dx = 2 nu, plus the wall-driven line
    cfg = tiny(meanflowx=0.0, meanpx=2.0 / 3000.0, u0=-0.5, uN=0.5)
    o = orc.Oracle(cfg)
    y, _ = o.grid()
    V0 = np.zeros(cfg.field_shape, dtype=np.complex128, order="F")
    V0[:, cfg.nz, 0, 0] = y * (2.0 - y) + 0.5 * (y - 1.0)
    o.set_field(V0)
    o.steps(10)
    assert np.max(np.abs(o.get_field() - V0)) < 1e-13

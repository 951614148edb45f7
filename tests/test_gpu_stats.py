"""Statistics profiles (chan_stats) against physical-space plane averages.

The reference forms its profiles as spectral sums over the Hermitian half
plane (postpro/history/var_history.f90:36-87).  Parseval turns them into
plane averages of the physical fields, which numpy's inverse FFT of the full
spectrum gives independently: the plane mean of u, the mean of (u - m)^2 and
of (u - m_u)(v - m_v).  A wrong kx weight (1 vs 2) or a missed (0,0)
subtraction fails these checks."""
import numpy as np
import pytest

from cases import config1, seeded_field, tiny

pytestmark = pytest.mark.gpu


def physical_rows(cfg, V):
    """Physical values of all three components on a (2nx+1, 2nz+1) grid for
    every row iy = 0..ny: array (3, ny+1, Mx, Mz)."""
    nx, nz, ny = cfg.nx, cfg.nz, cfg.ny
    Mx, Mz = 2 * nx + 1, 2 * nz + 1
    out = np.zeros((3, ny + 1, Mx, Mz))
    iz = np.arange(-nz, nz + 1)
    ix = np.arange(nx + 1)
    IZ, IX = np.meshgrid(iz, ix, indexing="ij")  # shapes (2nz+1, nx+1) like a plane
    pos = IX > 0
    for c in range(3):
        for iy in range(ny + 1):
            plane = V[iy + 1, :, :, c]  # [iz+nz, ix]
            F = np.zeros((Mx, Mz), dtype=np.complex128)
            F[IX % Mx, IZ % Mz] = plane
            F[(-IX[pos]) % Mx, (-IZ[pos]) % Mz] = np.conj(plane[pos])
            u = np.fft.ifft2(F) * (Mx * Mz)
            assert np.max(np.abs(u.imag)) < 1e-12 * max(1.0, np.max(np.abs(u.real)))
            out[c, iy] = u.real
    return out


def check(cfg, V, st, avg=None):
    phys = physical_rows(cfg, V)
    mean = phys.mean(axis=(2, 3))
    np.testing.assert_allclose(st["mean"], mean, rtol=0, atol=1e-13)
    m = mean if avg is None else avg
    fl = phys - m[:, :, None, None]
    var = (fl ** 2).mean(axis=(2, 3))
    uv = (fl[0] * fl[1]).mean(axis=(1, 2))
    scale = max(1e-300, np.max(np.abs(var)))
    np.testing.assert_allclose(st["var"], var, rtol=1e-11, atol=1e-13 * scale)
    np.testing.assert_allclose(st["uv"], uv, rtol=1e-10, atol=1e-13 * scale)


@pytest.mark.parametrize("mk", [tiny, config1])
def test_stats_match_physical_averages(mk):
    from channel_amd.solver import Channel
    cfg = mk()
    print("[case] stats", cfg.nx, cfg.ny, cfg.nz, flush=True)
    g = Channel(cfg)
    g.set_field(seeded_field(cfg))
    g.steps(2)
    V = g.get_field()
    check(cfg, V, g.stats())
    # the reference subtracts a time average: any given mean profile
    rng = np.random.default_rng(7)
    avg = np.stack([V[1:cfg.ny + 2, cfg.nz, 0, c].real for c in range(3)]) + 1e-3 * rng.standard_normal((3, cfg.ny + 1))
    check(cfg, V, g.stats(avg), avg)

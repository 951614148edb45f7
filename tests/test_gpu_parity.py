"""Parity of the HIP path (through the C-ABI) with the CPU oracle on identical
seeded inputs.  FP64 bar: relative L2 field error <= 1e-12 after 1 and 10
RK3 steps; intermediate quantities (rhs, scalars, Runtimedata columns) are
held to the same order.  FFTs against numpy at <= 1e-15 relative L2."""

import numpy as np
import pytest

from cases import config1, config2, laminar_field, os_case, rel_l2, seeded_field, tiny
from channel_amd import lib
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

RK1 = (120.0 / 32.0, 2.0, 0.0)
RK2 = (120.0 / 8.0, 50.0 / 8.0, 34.0 / 8.0)
RK3 = (120.0 / 20.0, 90.0 / 20.0, 50.0 / 20.0)
TOL = 1e-12



def say(*a):
    print("[case]", *a, flush=True)


def pair(cfg, V):
    from channel_amd.solver import Channel
    g = Channel(cfg)
    g.set_field(V)
    o = orc.Oracle(cfg)
    o.set_field(V)
    return g, o


@pytest.mark.parametrize("n", [12, 24, 48, 96, 192, 384, 768, 1536, 16, 32, 64, 128, 256, 512, 1024])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft_matches_numpy(n, sign):
    say("fft", n, sign)
    rng = np.random.default_rng(n + (sign > 0))
    x = rng.standard_normal((37, n)) + 1j * rng.standard_normal((37, n))
    y = lib.fft_selftest(x, sign)
    ref = np.fft.fft(x, axis=1) if sign < 0 else np.fft.ifft(x, axis=1) * n
    assert rel_l2(y, ref) < 1e-15


def test_field_roundtrip_is_bitwise():
    cfg = config1()
    say("roundtrip", cfg.nx, cfg.ny, cfg.nz)
    from channel_amd.solver import Channel
    g = Channel(cfg)
    V = seeded_field(cfg)
    rng = np.random.default_rng(0)
    V = V + rng.standard_normal(V.shape) * 1e-3
    g.set_field(V)
    np.testing.assert_array_equal(g.get_field(), V)


@pytest.mark.parametrize("mk", [tiny, config1])
def test_buildrhs_matches_oracle(mk):
    cfg = mk()
    say("buildrhs", mk.__name__)
    V = seeded_field(cfg)
    g, o = pair(cfg, V)
    for ode, cfl in ((RK1, False), (RK3, True)):
        g.buildrhs(ode, cfl)
        o.buildrhs(ode, cfl)
        R = g.get_rhs()
        Vo = o.get_field()
        Ro = Vo[2:cfg.ny + 1, :, :, 0:2]
        assert rel_l2(R[..., 0], Ro[..., 0]) < 1e-13
        assert rel_l2(R[..., 1], Ro[..., 1]) < 1e-13
        # the next substep starts from the same (unchanged) field on both sides
        o.set_field(V)
    so, sg = o.scalars(), g.scalars()
    assert sg["cfl"] == pytest.approx(so["cfl"], rel=1e-13)


@pytest.mark.parametrize("mk", [tiny, config1])
def test_substep_matches_oracle(mk):
    cfg = mk()
    say("substep", mk.__name__)
    g, o = pair(cfg, seeded_field(cfg))
    g.buildrhs(RK1, False)
    o.buildrhs(RK1, False)
    g.linsolve(RK1[0] / cfg.deltat)
    o.linsolve(RK1[0] / cfg.deltat)
    Vg, Vo = g.get_field(), o.get_field()
    assert rel_l2(Vg, Vo) < 1e-13
    for c in range(3):
        assert rel_l2(Vg[..., c], Vo[..., c]) < 1e-13
    sg, so = g.scalars(), o.scalars()
    # flow rates are O(1) (fr2 is roundoff noise around 0); corrpx is a
    # difference of nearly equal flow rates divided by fr3
    scale = abs(so["fr1"])
    for k in ("fr1", "fr2", "fr3"):
        assert abs(sg[k] - so[k]) <= 1e-13 * scale, k
    assert abs(sg["corrpx"] - so["corrpx"]) <= 1e-13 * scale / abs(so["fr3"])


def assert_rows_close(rg, ro):
    """Runtimedata columns to 1e-11 of max(|value|, 1): the flow-rate,
    gradient and CFL columns are O(1); the pressure-gradient columns carry
    corrections formed by cancellation between nearly equal flow rates and
    the spanwise columns are roundoff around zero, so both are held to an
    absolute 1e-11."""
    scale = np.maximum(np.abs(ro), 1.0)
    np.testing.assert_array_less(np.abs(rg - ro), 1e-11 * scale)


def _steps_parity(cfg, V, nsteps, tol=TOL):
    g, o = pair(cfg, V)
    g.set_scalars(deltat=cfg.deltat, time=cfg.time)
    o.set_scalars(deltat=cfg.deltat, time=cfg.time)
    out = []
    for k in range(nsteps):
        rg, ro = g.step(), o.step()
        assert_rows_close(rg, ro)
        if k + 1 in (1, 10, nsteps):
            e = rel_l2(g.get_field(), o.get_field())
            out.append((k + 1, e))
            assert e <= tol, (k + 1, e)
    return out


def test_config1_1_and_10_steps():
    cfg = config1()
    say("config1 10 steps")
    errs = _steps_parity(cfg, seeded_field(cfg), 10)
    print("rel L2 per step count:", errs)


def test_cpi_adaptive_dt_steps():
    # config-3 physics (CPI type 1, cflmax=1, Re=12431) on a small grid
    cfg = tiny(nx=15, ny=48, nz=15, alfa0=0.5, beta0=1.0, ni=1 / 12431.0, cpi=True, cpi_type=1, gamma=0.161436,
               meanflowx=0.0, cflmax=1.0, deltat=0.0)
    say("cpi adaptive 10 steps")
    V = seeded_field(cfg)
    g, o = pair(cfg, V)
    g.initial()
    o.initial()
    rg, ro = g.outstats(), o.outstats()
    assert_rows_close(rg, ro)
    for k in range(10):
        rg, ro = g.step(), o.step()
        assert_rows_close(rg, ro)
    assert rel_l2(g.get_field(), o.get_field()) <= TOL


def test_laminar_fixed_point_gpu():
    cfg = tiny(meanflowx=0.0, meanpx=2.0 / 3000.0)
    say("laminar")
    from channel_amd.solver import Channel
    g = Channel(cfg)
    V0 = laminar_field(cfg)
    g.set_field(V0)
    g.steps(10)
    assert np.max(np.abs(g.get_field() - V0)) < 1e-13


def test_orr_sommerfeld_gpu():
    cfg, V, c, P, iz = os_case(ny=96)
    cfg.nz = 4  # nzd = 12
    V2 = np.zeros(cfg.field_shape, dtype=np.complex128, order="F")
    V2[:, 4 - 2:4 + 3, :, :] = V
    say("orr-sommerfeld")
    g, o = pair(cfg, V2)
    g.steps(50)
    o.steps(50)
    assert rel_l2(g.get_field(), o.get_field()) < 1e-12


@pytest.mark.parametrize("nx,nz", [(255, 7), (511, 7), (383, 7), (7, 511)])
def test_wide_transform_paths(nx, nz):
    """Grids with long x or z transforms: 2nxd = 768 (the four-step x-pass
    transform of the 191x384x189 grid), 2nxd = 1536 (128-lane teams of two
    waves: the two-wave four-step transform, with all 512 modes at nx = 511
    and fewer than it holds at nx = 383) and nzd = 1536 (two teams per z
    block) -- the paths of the 511x512x511 grid.  Every mode carries a small
    random part, so the extension and truncation at the highest modes are
    exercised."""
    cfg = tiny(nx=nx, ny=16, nz=nz)
    say("wide transform paths", nx, nz)
    V = seeded_field(cfg)
    rng = np.random.default_rng(nx + nz)
    R = 1e-4 * (rng.standard_normal(V.shape) + 1j * rng.standard_normal(V.shape))
    R[:, cfg.nz, 0, :] = R[:, cfg.nz, 0, :].real
    for iz in range(1, cfg.nz + 1):
        R[:, cfg.nz - iz, 0, :] = np.conj(R[:, cfg.nz + iz, 0, :])
    V = np.asfortranarray(V + R)
    errs = _steps_parity(cfg, V, 2)
    print("rel L2:", errs)


def test_config2_one_step():
    cfg = config2()
    say("config2 1 step")
    errs = _steps_parity(cfg, seeded_field(cfg), 1)
    print("config2 rel L2:", errs)


def test_hermitian_symmetry_gpu():
    cfg = config1()
    say("hermitian")
    from channel_amd.solver import Channel
    g = Channel(cfg)
    g.set_field(seeded_field(cfg))
    g.steps(3)
    V = g.get_field()
    nz = cfg.nz
    err = max(np.max(np.abs(V[:, nz - iz, 0, :] - np.conj(V[:, nz + iz, 0, :]))) for iz in range(1, nz + 1))
    assert err < 1e-14 * np.max(np.abs(V[:, :, 0, :]))


@pytest.mark.parametrize("sign", [-1, 1])
def test_four_step_768_matches_numpy(sign):
    """The x pass's four-step 768-point transform (permlane transpose, two
    LDS exchanges, permuted layout between backward and forward) in natural
    order against numpy."""
    say("four-step 768", sign)
    rng = np.random.default_rng(768 + sign)
    x = rng.standard_normal((37, 768)) + 1j * rng.standard_normal((37, 768))
    y, _ = lib.fft768_fs(x, sign)
    ref = np.fft.fft(x, axis=1) if sign < 0 else np.fft.ifft(x, axis=1) * 768
    assert rel_l2(y, ref) < 1e-15


@pytest.mark.parametrize("n", [384, 768, 1536])
@pytest.mark.parametrize("variant", range(12))
def test_fft_build_variants_match_numpy(n, variant):
    say("fft variant", n, variant)
    rng = np.random.default_rng(1000 * n + variant)
    x = rng.standard_normal((19, n)) + 1j * rng.standard_normal((19, n))
    for sign in (-1, 1):
        y, _ = lib.fft_bench(x, sign, variant, reps=1)
        ref = np.fft.fft(x, axis=1) if sign < 0 else np.fft.ifft(x, axis=1) * n
        assert rel_l2(y, ref) < 1e-15

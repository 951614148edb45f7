"""HIP path: Couette flow through the wall velocities u0, uN (dns.in line 8,
channel.f90:122-124), the explicit nonlinear terms of buildrhs against the
Kim-Moin-Moser form (dnsdata.f90:633-661), and bitwise reruns of the step."""

import numpy as np
import pytest

from cases import config1, rel_l2, seeded_field, tiny
from channel_amd import lib
from oracle import oracle as orc
from test_oracle_nonlinear import _case, _expected, _field

pytestmark = pytest.mark.gpu

TOL = 1e-12


def say(*a):
    print("[case]", *a, flush=True)


def _couette(cfg, u0, uN, y):
    V = np.zeros(cfg.field_shape, dtype=np.complex128, order="F")
    V[:, cfg.nz, 0, 0] = u0 + (uN - u0) * y / 2.0
    return V


@pytest.mark.parametrize("precision", [lib.CHAN_FP64, lib.CHAN_SPLIT])
@pytest.mark.parametrize("u0,uN", [(-1.0, 1.0), (0.3, 1.7)])
def test_laminar_couette_fixed_point(precision, u0, uN):
    from channel_amd.dns import y_grid
    from channel_amd.solver import Channel
    cfg = tiny(meanflowx=0.0, meanpx=0.0, u0=u0, uN=uN)
    say("couette laminar", precision, u0, uN)
    g = Channel(cfg, precision)
    V0 = _couette(cfg, u0, uN, y_grid(cfg))
    g.set_field(V0)
    rows = g.steps(10)
    tol = 1e-13 if precision == lib.CHAN_FP64 else 1e-7
    assert np.max(np.abs(g.get_field() - V0)) < tol
    slope = (uN - u0) / 2.0
    assert rows[-1, 1] == pytest.approx(slope, rel=1e-11)
    assert rows[-1, 2] == pytest.approx(-slope, rel=1e-11)


def test_perturbed_couette_matches_oracle():
    from channel_amd.dns import y_grid
    from channel_amd.solver import Channel
    cfg = config1(meanflowx=0.0, meanpx=0.0, u0=-1.0, uN=1.0)
    say("couette perturbed 10 steps")
    V = seeded_field(cfg)
    V[:, cfg.nz, 0, 0] = y_grid(cfg) - 1.0
    g = Channel(cfg)
    g.set_field(V)
    o = orc.Oracle(cfg)
    o.set_field(V)
    for k in range(10):
        rg, ro = g.step(), o.step()
        np.testing.assert_array_less(np.abs(rg - ro), 1e-11 * np.maximum(np.abs(ro), 1.0))
        if k + 1 in (1, 10):
            Vg, Vo = g.get_field(), o.get_field()
            for c in range(3):
                assert rel_l2(Vg[..., c], Vo[..., c]) <= TOL, (k + 1, c)
    nz, ny = cfg.nz, cfg.ny
    assert Vg[1, nz, 0, 0].real == pytest.approx(-1.0, abs=1e-13)
    assert Vg[ny + 1, nz, 0, 0].real == pytest.approx(1.0, abs=1e-13)


@pytest.mark.parametrize("precision", [lib.CHAN_FP64, lib.CHAN_SPLIT])
def test_explicit_terms_match_kmm_form(precision):
    """buildrhs(ODE=(0,1,0)) - buildrhs(ODE=(0,0,0)) of the HIP path on
    polynomial low modes equals D0 of the KMM terms formed in physical space
    (the check that pins the oracle's assembly, applied to the GPU)."""
    from channel_amd.solver import Channel
    cfg, S = _case(7)
    say("kmm form", precision)
    g = Channel(cfg, precision)
    g.set_field(_field(cfg, S))
    out = []
    for ode in ((0.0, 1.0, 0.0), (0.0, 0.0, 0.0)):
        g.buildrhs(ode, False)
        out.append(g.get_rhs())
    R = out[0] - out[1]
    d0 = g.setup()[1][:, 0, :]
    eta_x, d2v_x = _expected(cfg, S, d0)
    tol = 1e-12 if precision == lib.CHAN_FP64 else 1e-5
    nz = cfg.nz
    mean, mean_x = R[:, nz, 0, 0].copy(), eta_x[:, nz, 0].copy()
    eta, eta_xx = R[..., 0].copy(), eta_x.copy()
    eta[:, nz, 0] = eta_xx[:, nz, 0] = 0.0
    for got, want in ((eta, eta_xx), (R[..., 1], d2v_x), (mean, mean_x)):
        assert np.max(np.abs(got - want)) / np.max(np.abs(want)) < tol


@pytest.mark.parametrize("precision", [lib.CHAN_FP64, lib.CHAN_SPLIT])
def test_rerun_is_bitwise_identical(precision):
    """Two runs of 10 steps from the same field give the same bits: every
    reduction has a fixed order (the CFL maximum is order-free)."""
    from channel_amd.solver import Channel
    cfg = config1()
    say("bitwise rerun", precision)
    V = seeded_field(cfg)
    res = []
    for _ in range(2):
        g = Channel(cfg, precision)
        g.set_field(V)
        rows = g.steps(10)
        res.append((g.get_field(), rows))
        g.close()
    np.testing.assert_array_equal(res[0][0], res[1][0])
    np.testing.assert_array_equal(res[0][1], res[1][1])

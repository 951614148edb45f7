"""Body force on the HIP path (chan_set_body_force, chan_set_body_force_feedback,
chan_get_body_force) against the oracle, which restates the reference's
force terms in buildrhs (dnsdata.f90:616-659) and its three modules
(body_forces/*) literally.  FP64 bar as for the plain step: relative L2 field
error <= 1e-12 after 1 and 10 steps; the split mode to 1e-6."""

import numpy as np
import pytest

from cases import config1, config2, rel_l2, seeded_field, tiny
from channel_amd import body_force as bf
from channel_amd import lib
from channel_amd.dns import y_grid
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

TOL = 1e-12


def say(*a):
    print("[case]", *a, flush=True)


def smooth_force(cfg, seed, amp=1e-2, kmax=6):
    """Random force q(y) c_{ix,iz} on the low modes, Hermitian in the kx = 0
    column, plus a mean-mode part in x and z."""
    rng = np.random.default_rng(seed)
    y = y_grid(cfg)
    q = y * (2.0 - y)
    nz = cfg.nz
    F = np.zeros(cfg.field_shape, dtype=np.complex128, order="F")
    for ix in range(min(kmax, cfg.nx) + 1):
        for iz in range(-min(kmax, nz), min(kmax, nz) + 1):
            c = amp * (rng.standard_normal(3) + 1j * rng.standard_normal(3))
            F[:, nz + iz, ix, :] = q[:, None] * c[None, :]
    F[:, nz, 0, :] = F[:, nz, 0, :].real
    for iz in range(1, nz + 1):
        F[:, nz - iz, 0, :] = np.conj(F[:, nz + iz, 0, :])
    return F


def channel(cfg, V, precision=lib.CHAN_FP64):
    from channel_amd.solver import Channel
    g = Channel(cfg, precision)
    g.set_field(V)
    return g


def oracle(cfg, V):
    o = orc.Oracle(cfg)
    o.set_field(V)
    return o


def run_pair(g, o, n, tol):
    errs = []
    for k in range(n):
        rg, ro = g.step(), o.step()
        np.testing.assert_array_less(np.abs(rg - ro), 1e3 * tol * np.maximum(np.abs(ro), 1.0))
        if k + 1 in (1, n):
            e = rel_l2(g.get_field(), o.get_field())
            errs.append((k + 1, e))
            assert e <= tol, (k + 1, e)
    return errs


def test_static_force_parity():
    cfg = config1()
    say("static force config1 10 steps")
    V = seeded_field(cfg)
    F = smooth_force(cfg, 5)
    g, o = channel(cfg, V), oracle(cfg, V)
    g.set_body_force(F)
    o.set_body_force(F)
    errs = run_pair(g, o, 10, TOL)
    print("rel L2:", errs)
    # the force matters: without it the field differs far beyond the bar
    h = channel(cfg, V)
    h.steps(10)
    assert rel_l2(h.get_field(), g.get_field()) > 1e-6


MODULES = {
    "am_f1": ((50.0, 500.0), lambda cfg: bf.am_f1(cfg, 50.0, 500.0)),
    "am_butterfly": ((50.0, 500.0), lambda cfg: bf.am_butterfly(cfg, 50.0, 500.0)),
    "coriolis": ((0.5, 9.0, 0.4), lambda cfg: bf.coriolis(cfg, 0.5, 9.0, 0.4)),
}


@pytest.mark.parametrize("name", list(MODULES))
def test_feedback_module_parity(name):
    cfg = config1()
    say("feedback", name, "config1 10 steps")
    par, table = MODULES[name]
    V = seeded_field(cfg)
    g, o = channel(cfg, V), oracle(cfg, V)
    g.set_body_force_feedback(*table(cfg))
    o.set_body_force_module(name, *par)
    errs = run_pair(g, o, 10, TOL)
    print("rel L2:", errs)
    h = channel(cfg, V)
    h.steps(10)
    assert rel_l2(h.get_field(), g.get_field()) > 1e-8


@pytest.mark.parametrize("name", ["static", "coriolis", "am_butterfly"])
def test_get_body_force_matches_oracle(name):
    cfg = tiny(nx=15, nz=15, ny=48)
    say("get_body_force", name)
    V = seeded_field(cfg)
    g, o = channel(cfg, V), oracle(cfg, V)
    assert np.all(g.get_body_force() == 0)
    if name == "static":
        rng = np.random.default_rng(2)
        F = np.asfortranarray(rng.standard_normal(cfg.field_shape) + 1j * rng.standard_normal(cfg.field_shape))
        g.set_body_force(F)
        o.set_body_force(F)
    else:
        par, table = MODULES[name]
        g.set_body_force_feedback(*table(cfg))
        o.set_body_force_module(name, *par)
        o.apply_body_force()
    o.body_force_walls()
    Fg, Fo = g.get_body_force(), o.get_body_force()
    assert rel_l2(Fg, Fo) < 1e-14
    ny = cfg.ny
    assert np.all(Fg[1] == 0) and np.all(Fg[ny + 1] == 0)


def test_force_profile_keeps_its_laminar_flow():
    """c y(2-y) in x and z balances the quartic laminar profile (central rows
    to rounding; the wall closure of the solve moves the near-wall rows)."""
    cfg = config1(meanflowx=0.0, meanpx=0.0)
    say("laminar quartic with force")
    nu = cfg.ni
    y = y_grid(cfg)
    q = y * (2.0 - y)
    quart = 2.0 * y / 3.0 - y ** 3 / 3.0 + y ** 4 / 12.0
    nz = cfg.nz
    V0 = np.zeros(cfg.field_shape, dtype=np.complex128, order="F")
    V0[:, nz, 0, 0] = 12.0 / 5.0 * quart
    V0[:, nz, 0, 2] = 6.0 / 5.0 * quart
    F = np.zeros_like(V0)
    F[:, nz, 0, 0] = 12.0 * nu / 5.0 * q
    F[:, nz, 0, 2] = 6.0 * nu / 5.0 * q
    g = channel(cfg, V0)
    g.set_body_force(F)
    g.steps(10)
    V = g.get_field()
    mid = slice(cfg.ny // 4 + 1, 3 * cfg.ny // 4 + 2)
    assert np.max(np.abs(V[mid] - V0[mid])) < 1e-13
    assert np.max(np.abs(V[1:-1] - V0[1:-1])) < 1e-8


def test_force_removal_and_bad_source():
    cfg = tiny(nx=15, nz=15, ny=48)
    say("force removal, invalid source")
    V = seeded_field(cfg)
    g, o = channel(cfg, V), oracle(cfg, V)
    g.set_body_force(smooth_force(cfg, 1))
    g.set_body_force_feedback(*bf.coriolis(cfg, 0.3))
    g.set_body_force(None)
    g.set_body_force_feedback(None, None)
    assert np.all(g.get_body_force() == 0)
    run_pair(g, o, 3, TOL)
    with pytest.raises(lib.ChanError) as e:
        g.set_body_force_feedback((0, 3, -1), np.zeros((3, cfg.ny + 3, 2 * cfg.nz + 2)))
    assert e.value.code == lib.CHAN_E_ARG


def test_applied_force_is_the_third_substeps():
    """chan_keep_body_force / chan_get_applied_body_force: after each step the
    kept force equals the oracle's F, which its last buildrhs set from the
    field of the third substep and then gave the wall rows (what
    Force.cart.N.out holds, dnsdata.f90:904-906); the force of the resident
    field is a different one."""
    cfg = config1()
    say("applied force: keep, coriolis feedback, static force")
    V = seeded_field(cfg)
    g, o = channel(cfg, V), oracle(cfg, V)
    for keep in (False, True):
        with pytest.raises(lib.ChanError) as e:
            g.get_applied_body_force()  # off, then on with no buildrhs yet
        assert e.value.code == lib.CHAN_E_STATE
        if keep:
            g.keep_body_force()
    g.step()
    o.step()
    assert np.all(g.get_applied_body_force() == 0)  # no force in that step
    par, table = MODULES["coriolis"]
    g.set_body_force_feedback(*table(cfg))
    o.set_body_force_module("coriolis", *par)
    for k in range(3):
        g.step()
        o.step()
        Fa, Fo = g.get_applied_body_force(), o.get_body_force()
        e = rel_l2(Fa, Fo)
        print("step", k + 1, "applied force rel L2", e)
        assert e <= TOL
    assert rel_l2(g.get_body_force(), Fa) > 1e-8
    assert rel_l2(g.get_field(), o.get_field()) <= TOL
    # a static force: F after the step is F0 with the wall rows
    g.set_body_force_feedback(None, None)
    o.set_body_force_module(None)
    F = smooth_force(cfg, 4)
    g.set_body_force(F)
    o.set_body_force(F)
    g.step()
    o.step()
    assert rel_l2(g.get_applied_body_force(), o.get_body_force()) < 1e-14
    g.keep_body_force(False)
    with pytest.raises(lib.ChanError) as e:
        g.get_applied_body_force()
    assert e.value.code == lib.CHAN_E_STATE


def test_split_mode_with_feedback():
    cfg = config1()
    say("split mode with coriolis feedback, 10 steps")
    V = seeded_field(cfg)
    g, o = channel(cfg, V, lib.CHAN_SPLIT), oracle(cfg, V)
    par, table = MODULES["coriolis"]
    g.set_body_force_feedback(*table(cfg))
    o.set_body_force_module("coriolis", *par)
    errs = run_pair(g, o, 10, 1e-6)
    print("split rel L2:", errs)


def test_static_and_feedback_at_95x128x95():
    """Both parts at once: the library adds F0 and gain * V; the oracle gets
    that sum as a static force before each of its substeps."""
    cfg = config2()
    say("static force + am_f1 feedback at 95x128x95, 3 steps")
    V = seeded_field(cfg)
    F = smooth_force(cfg, 9)
    src, gain = bf.am_f1(cfg, 20.0, 500.0)
    g, o = channel(cfg, V), oracle(cfg, V)
    g.set_body_force(F)
    g.set_body_force_feedback(src, gain)
    errs = []
    for k in range(3):
        rg = g.step()
        ro = _oracle_step_with(o, F, gain, src, cfg)
        np.testing.assert_array_less(np.abs(rg - ro), 1e-9 * np.maximum(np.abs(ro), 1.0))
        errs.append(rel_l2(g.get_field(), o.get_field()))
    print("rel L2:", errs)
    assert max(errs) <= TOL


RK = ((120.0 / 32.0, 2.0, 0.0), (120.0 / 8.0, 50.0 / 8.0, 34.0 / 8.0), (120.0 / 20.0, 90.0 / 20.0, 50.0 / 20.0))


def _oracle_step_with(o, F0, gain, src, cfg):
    """One oracle time step whose force, before every buildrhs, is
    F0 + gain * V of the current field (the library's additive form)."""
    nz, dt = cfg.nz, cfg.deltat
    for k, ode in enumerate(RK):
        V = o.get_field()
        F = F0.copy()
        for d in range(3):
            gg = np.repeat(gain[d][:, :2 * nz + 1, None], cfg.nx + 1, axis=2)
            gg[:, nz, 0] = gain[d][:, 2 * nz + 1]
            F[..., d] += gg * V[..., src[d]]
        o.set_body_force(F)
        s = o.scalars()
        o.set_scalars(time=s["time"] + 2.0 / ode[0] * dt)
        o.buildrhs(ode, k == 2)
        o.linsolve(ode[0] / dt)
    return o.outstats()

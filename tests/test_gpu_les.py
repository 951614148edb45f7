"""The closure inside a coarse channel step on the device
(include/channel_closure.h: chan_les_features, chan_les_closure_force)
against the float64 checker oracle/les_oracle.py.  The device forms the
features and the force in FP32 (the features are the MLP's BF16 inputs), so
features are held to 2e-6 of each feature's largest value and the force,
which also carries the MLP's BF16 rounding, to a relative L2 of 2e-3 against
the checker run on the device's own features."""
import numpy as np
import pytest

from channel_amd.dns import DnsConfig
from channel_amd.initial import perturbed_laminar
from oracle import les_oracle as lo
from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu


def les_cfg():
    """A coarse grid of the closure's physics: Re 12431, alfa0 0.5, beta0 1."""
    return DnsConfig(nx=15, ny=32, nz=15, alfa0=0.5, beta0=1.0, ni=1.0 / 12431.0, a=1.5,
                     cpi=True, cpi_type=1, gamma=0.161436, deltat=0.0, cflmax=1.0)


def _field(cfg):
    return perturbed_laminar(cfg, seed=7, amp=0.05, kmax=8)


def test_features_match_checker():
    from channel_amd.solver import Channel
    cfg = les_cfg()
    print("[case] les features", flush=True)
    V = _field(cfg)
    g = Channel(cfg)
    g.set_field(V)
    X = g.les_features().astype(np.float64)
    y, _, bnd = g.setup()
    g.close()
    Xo = lo.features(cfg, V, y, bnd[1], bnd[6])
    assert X.shape == Xo.shape == ((cfg.ny + 1) * 2 * cfg.nxd * cfg.nzd, 46)
    scale = np.max(np.abs(Xo), axis=0)
    err = np.max(np.abs(X - Xo), axis=0) / scale
    print("worst feature error / scale", float(err.max()))
    assert np.all(err < 2e-6)
    # the neighbour blocks are the point's gradients of the rows iy-2..iy+2
    assert np.any(Xo[:, 9:18] != Xo[:, 0:9])


def test_closure_force_matches_checker():
    from channel_amd.closure import Mlp
    from channel_amd.solver import Channel
    cfg = les_cfg()
    print("[case] les closure force", flush=True)
    V = _field(cfg)
    g = Channel(cfg)
    g.set_field(V)
    X = g.les_features().astype(np.float64)
    rng = np.random.default_rng(3)
    p = (mo.init_params(5) + 0.05 * rng.standard_normal(mo.NPARAM)).astype(np.float32).astype(np.float64)
    xm, xs = X.mean(0), X.std(0) + 1e-3
    tm = 1e-3 * rng.standard_normal(6)
    ts = 1e-3 * (1.0 + rng.random(6))
    norm = np.concatenate([xm, xs, tm, ts]).astype(np.float32)
    m = Mlp(16)
    m.set_params(p)
    g.les_closure_force(m, norm)
    F = g.get_body_force()
    y, _, _ = g.setup()
    g.close()
    m.close()
    n32 = norm.astype(np.float64)
    Y = mo.mirrored(p, (X - n32[:46]) / n32[46:92])[0]
    tau = Y * n32[98:104] + n32[92:98]
    Mx, Mz = 2 * cfg.nxd, cfg.nzd
    tau = np.moveaxis(tau.reshape(cfg.ny + 1, Mx, Mz, 6), -1, 0)
    Fo = lo.divergence(cfg, tau, y)
    inner = slice(2, cfg.ny + 1)  # rows 1..ny-1; the wall and ghost rows are the library's
    e = np.linalg.norm((F[inner] - Fo[inner]).ravel()) / np.linalg.norm(Fo[inner].ravel())
    print("force rel L2 vs checker", e)
    assert np.linalg.norm(Fo[inner].ravel()) > 0
    assert e < 2e-3


def test_study_pipeline_small():
    """The closure study end to end on small grids (fine 47x64x47, LES
    15x32x15, Re 3000 so that the small fine grid resolves its flow): DNS, filtered snapshots, targets, training on
    the device, LES runs with and without the device closure, profiles and
    costs.  It checks the pipeline, not the physics."""
    from channel_amd import les_study as ls
    print("[case] les study small", flush=True)
    cf = DnsConfig(nx=47, ny=64, nz=47, alfa0=0.5, beta0=1.0, ni=1.0 / 3000.0, a=1.5,
                   cpi=True, cpi_type=1, gamma=0.161436, deltat=0.0, cflmax=1.0)
    cl = DnsConfig(nx=15, ny=32, nz=15, alfa0=0.5, beta0=1.0, ni=1.0 / 3000.0, a=1.5,
                   cpi=True, cpi_type=1, gamma=0.161436, deltat=0.0, cflmax=1.0)
    res = ls.run(cf, cl, amp=3e-3, kmax=8, warm_time=2.0, ntrain=3, train_every=0.2, window=0.6,
                 sample_every=0.2, train_steps=300, batch=4096, log=lambda s: print(s, flush=True))
    assert res["train"]["loss_after"] < res["train"]["loss_before"]
    for model in ("none", "mlp"):
        assert res["runs"][model]["finite"], model
        assert res["runs"][model]["steps"] > 0
        assert set(res["errors"][model]) == {"mean_u", "rms_u", "rms_v", "rms_w", "uv"}
    assert res["dns"]["ms_per_time_unit"] > 0
    ref = np.array(res["profiles"]["dns_filtered"])
    assert ref.shape == (5, cl.ny + 1) and np.all(np.isfinite(ref))
